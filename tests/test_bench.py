"""Host logic of bench.py, on the CPU: the seeded served mix, the one-GPU guard,
and the CPU baseline's binomial legs, whose parsed prices must equal the
reference kernels' golden values (tests/golden/ref_kernels.csv)."""
import csv
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
import bench  # noqa: E402

REF = os.path.join(ROOT, "oracle", "_ref")
STEPS = {"eur": 1, "basket": 1, "asia": 252, "amer": 50}


def test_mix_is_seeded_and_spans_the_four_kinds():
    a, b = bench.make_mix(2026, 64), bench.make_mix(2026, 64)
    key = lambda m: [(q["kind"], q["args"], q["seed"], q["steps"]) for q in m]  # noqa: E731
    assert key(a) == key(b)
    assert key(bench.make_mix(2027, 64)) != key(a)
    assert {q["kind"] for q in a} == set(STEPS)
    for q in a:
        N = q["args"][5]
        assert N % 2 == 0 and 2**14 <= N <= 2**22, N  # antithetic pairs; paths 2^14..2^22
        assert q["steps"] == STEPS[q["kind"]]
    assert bench.path_steps(a) == sum(q["args"][5] * q["steps"] for q in a)


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=300)
    assert r.returncode == 2
    assert "one GPU" in r.stderr
    assert r.stdout.strip() == ""  # no JSON line


def _golden(prog, payoff, N):
    with open(os.path.join(ROOT, "tests", "golden", "ref_kernels.csv")) as f:
        for row in csv.reader(f):
            if row and row[0] == prog and row[1] == payoff and int(row[7]) == N:
                return float(row[-1])
    raise KeyError((prog, payoff, N))


@pytest.mark.skipif(not all(os.path.exists(os.path.join(REF, e))
                            for e in ("binom_embar_omp", "binom_vanilla_amer", "binom_vanilla_eur")),
                    reason="the reference programs in oracle/_ref are not built")
def test_cpu_binomial_legs_parse_the_reference_programs():
    threads = max(1, min(8, os.cpu_count() or 1))
    out = bench.cpu_binomial(REF, dict(os.environ, OMP_NUM_THREADS=str(threads)), threads)
    want = {"binom_embar_N100000": _golden("binom_embar", "call", 100000),
            "lattice_amer_put_N1e4": _golden("binom_vanilla_amer", "put", 10000),
            "lattice_eur_call_N1e4": _golden("binom_vanilla_eur", "call", 10000)}
    for name, v in want.items():
        leg = out[name]
        # the programs print 10 significant digits (include/common.h:231-248)
        assert abs(leg["price"] - v) <= 5e-10 * v, (name, leg["price"], v)
        assert leg["seconds"] > 0
    assert out["binom_embar_N100000"]["log_adds_per_s_chain"] == pytest.approx(0.75 * 1e10 / out[
        "binom_embar_N100000"]["seconds"])
    N = 10000
    assert out["lattice_amer_put_N1e4"]["node_updates_per_s"] == pytest.approx(
        0.5 * (N + 1) * (N + 2) / out["lattice_amer_put_N1e4"]["seconds"])
