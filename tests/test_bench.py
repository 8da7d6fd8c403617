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
                            for e in ("mc_eur_omp", "mc_asia_omp", "mc_amer_omp", "binom_embar_omp",
                                      "binom_vanilla_amer", "binom_vanilla_eur")),
                    reason="the reference programs in oracle/_ref are not built")
def test_cpu_reference_legs_at_the_configurations_sizes():
    """Each leg runs the reference program at its configuration's own size and
    parses its row; the deterministic programs print the golden values."""
    cpu = bench.cpu_reference(full=False, repeats=1)
    legs = cpu["legs"]
    assert cpu["threads"] >= 1 and cpu["nproc"] >= 1 and cpu["cpu_model"]
    want = {"binom_embar_N100000": _golden("binom_embar", "call", 100000),
            "lattice_amer_put_N1e4": _golden("binom_vanilla_amer", "put", 10000),
            "lattice_eur_call_N1e4": _golden("binom_vanilla_eur", "call", 10000)}
    for name, v in want.items():
        # the programs print 10 significant digits (include/common.h:231-248)
        assert abs(legs[name]["result"] - v) <= 5e-10 * v, (name, legs[name]["result"], v)
    assert "1048576 50" in legs["mc_amer_put_2^20x50"]["argv"]
    assert "4194304 252" in legs["mc_asia_2^22x252"]["argv"]
    for name, leg in legs.items():
        assert leg["seconds_best"] > 0 and leg["per_s"] == pytest.approx(leg["work"] / leg["seconds_best"])
    assert legs["binom_embar_N100000"]["work"] == 0.75 * 1e10
    # the headline figure: the mix's paths x steps over the time these rates give
    mix = bench.make_mix(2026, 64)
    line = bench.cpu_baseline_line(cpu, mix)
    sample = [q for q in mix if q["kind"] != "basket"]
    t = sum(q["args"][5] * q["steps"] / {"eur": legs["mc_eur_1e6"], "asia": legs["mc_asia_2^22x252"],
                                          "amer": legs["mc_amer_put_2^20x50"]}[q["kind"]]["per_s"] for q in sample)
    assert line["value"] == pytest.approx(bench.path_steps(sample) / t)
    assert line["kind"] == "reference" and line["cores"] == cpu["threads"]
    # beside the GPU figures of the same configuration and size
    configs = {"mc_eur_1e6": {"paths_steps_per_s_fp32k": 2e11, "paths_steps_per_s_fp64": 1e11},
               "binom_embar_N100000": {"log_adds_per_s_chain": 1e13}}
    bench.attach_cpu(configs, cpu)
    assert configs["mc_eur_1e6"]["gpu_over_cpu"]["gpu_fp32k"] == pytest.approx(2e11 / legs["mc_eur_1e6"]["per_s"])
    assert configs["binom_embar_N100000"]["gpu_over_cpu"]["gpu_fp64"] == pytest.approx(
        1e13 / legs["binom_embar_N100000"]["per_s"])
