"""The drop-in programs (parcompfin_amd/bin/*): reference argv, reference row.

CPU tests cover argument handling (it runs before any device is opened); the
GPU tests run each program in verification mode (--rng mt19937 --seed S, the
reference's own libstdc++ normal stream) and compare the printed row with the
row the reference program printed for the same seed (tests/golden/ref_rows.csv,
made with time() pinned): the Result, Abs_Error and Error columns must be the
same text; Method, Parallel and the two time columns differ by design.
"""
import csv
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "parcompfin_amd", "bin")
HERE = os.path.dirname(os.path.abspath(__file__))
COLUMNS = ["Method", "Payoff", "S0", "E", "r", "sigma", "T", "N", "M", "Parallel", "Nr_of_assets", "T_overall",
           "T_calculation", "Result", "Abs_Error", "Error"]


def run(prog, *args, env=None):
    return subprocess.run([os.path.join(BIN, prog)] + [str(a) for a in args], capture_output=True, text=True,
                          timeout=120, env=env)


PROGS = ["mc_eur", "binom_embar", "mc_asia", "mc_amer", "mc_eur_multi", "binom_vanilla_eur", "binom_vanilla_amer"]
# the names the runscripts' OpenMP sweeps call (runscript_mc_eur.sh:35, runscript_mc_asia.sh:43,
# runscript_mc_amer.sh:41, runscript_binom_embar.sh:37; src/mc_eur_multi_omp.cpp)
OMP_PROGS = ["mc_eur_omp", "binom_embar_omp", "mc_asia_omp", "mc_amer_omp", "mc_eur_multi_omp"]


@pytest.mark.parametrize("prog", PROGS + OMP_PROGS)
def test_programs_built(prog):
    assert os.access(os.path.join(BIN, prog), os.X_OK), "run __graft_entry__.build() first"


@pytest.mark.parametrize("prog,extra", [("mc_eur", []), ("mc_asia", [252]), ("mc_amer", [50]),
                                        ("mc_eur_multi", [8, 0.3]), ("binom_embar", []),
                                        ("binom_vanilla_amer", [])])
def test_unknown_payoff_is_rejected_like_the_reference(prog, extra):
    # src/mc_eur.cpp:39-42 throws std::invalid_argument("Unknown payoff function").
    p = run(prog, "straddle", 100, 100, 0.05, 0.2, 1, 1000, *extra)
    assert p.returncode == 1 and "Unknown payoff function" in p.stderr


def test_usage_and_bad_numbers():
    p = run("mc_eur", "call", 100, 100)
    assert p.returncode == 1 and "usage" in p.stderr
    p = run("mc_asia", "call", 100, 100, 0.05, "x", 1, 1000, 252)
    assert p.returncode == 1 and "sigma" in p.stderr
    p = run("mc_eur", "call", 100, 100, 0.05, 0.2, 1, 1000, "--precision", "fp16")
    assert p.returncode == 1


def parse_row(text):
    lines = [l for l in text.strip().splitlines() if l.strip()]
    assert len(lines) == 1, text
    f = lines[0].split(",")
    assert len(f) == 16, lines[0]
    return dict(zip(COLUMNS, f))


def ref_rows():
    return list(csv.DictReader(open(os.path.join(HERE, "golden", "ref_rows.csv"))))


@pytest.mark.gpu
def test_verification_rows_match_the_reference_programs():
    rows = ref_rows()
    assert len(rows) >= 20
    for r in rows:
        prog = r["prog"]
        args = [r["payoff"], r["S0"], r["E"], r["r"], r["sigma"], r["T"], r["N"]]
        if prog in ("mc_asia", "mc_amer"):
            args.append(r["M"])
        deterministic = prog.startswith("binom")  # no random stream: no flags
        flags = [] if deterministic else ["--rng", "mt19937", "--seed", r["seed"]]
        if prog == "mc_amer":
            flags.append("--ref-exact")  # the reference's payoff(S-E,E) exercise test
        print("[case]", prog, *args, *flags, flush=True)
        p = run(prog, *args, *flags)
        assert p.returncode == 0, p.stderr
        got = parse_row(p.stdout)
        want = dict(zip(COLUMNS, r["row"].split(",")))
        method = {"binom_embar": "HIP_fp64"}.get(prog, "HIP_vanilla" if deterministic else "HIP_verify")
        assert got["Method"] == method and got["Parallel"] == "1"
        for c in ("Payoff", "S0", "E", "r", "sigma", "T", "N", "M", "Nr_of_assets", "Result", "Abs_Error", "Error"):
            assert got[c] == want[c], (prog, c, got[c], want[c])


@pytest.mark.gpu
def test_binomial_programs_print_the_reference_values():
    fx = {(r["prog"], r["payoff"], int(r["N"]), r["E"]): float(r["result"])
          for r in csv.DictReader(open(os.path.join(HERE, "golden", "ref_kernels.csv")))}
    for prog, pf, N, E, r_, s in (("binom_embar", "call", 1000, "100", 0.05, 0.2),
                                  ("binom_vanilla_eur", "call", 2000, "100", 0.05, 0.2),
                                  ("binom_vanilla_amer", "put", 2000, "100", 0.06, 0.2)):
        p = run(prog, pf, 100, E, r_, s, 1, N)
        assert p.returncode == 0, p.stderr
        got = parse_row(p.stdout)
        assert got["Method"] in ("HIP_fp64", "HIP_vanilla")
        assert got["Result"] == "%.10g" % fx[(prog, pf, N, E)], (prog, got["Result"])


@pytest.mark.gpu
def test_comparison_column_from_the_environment():
    env = dict(os.environ, PCF_COMPARISON="10.45")
    p = run("binom_vanilla_eur", "call", 100, 100, 0.05, 0.2, 1, 1000, env=env)
    assert p.returncode == 0, p.stderr
    got = parse_row(p.stdout)
    res = float(got["Result"])  # printed with 10 significant digits
    assert abs(float(got["Error"]) - (res - 10.45)) < 1e-8
    assert abs(float(got["Abs_Error"]) - abs(res - 10.45)) < 1e-8


def test_comparison_header_errors(tmp_path):
    # --comparison-file is read before any device is opened
    p = run("mc_eur", "call", 100, 100, 0.05, 0.2, 1, 1000, "--comparison-file", str(tmp_path / "missing.h"))
    assert p.returncode == 1 and "cannot read comparison header" in p.stderr
    h = tmp_path / "comparison.h"
    h.write_text("#pragma once\n")
    p = run("mc_eur", "call", 100, 100, 0.05, 0.2, 1, 1000, "--comparison-file", str(h))
    assert p.returncode == 1 and "no 'comparison = value;'" in p.stderr
    h.write_text("#pragma once\ndouble comparison = ten;\n")
    p = run("mc_eur", "call", 100, 100, 0.05, 0.2, 1, 1000, "--comparison-file", str(h))
    assert p.returncode == 1 and "bad numeric argument" in p.stderr


@pytest.mark.gpu
def test_runscript_header_and_comparison_file(tmp_path):
    # What the runscripts write before rebuilding (runscript_mc_eur.sh:10-13): the
    # header line, and include/comparison.h with the comparison value.
    h = tmp_path / "comparison.h"
    h.write_text("#pragma once\ndouble comparison = 26.61316121;\n")
    for how in (["--comparison-file", str(h)], []):
        env = dict(os.environ, PCF_COMPARISON_FILE=str(h)) if not how else None
        p = run("binom_embar", "call", 100, 110, 0.02, 0.75, 1, 10000, "--header", *how, env=env)
        assert p.returncode == 0, p.stderr
        lines = p.stdout.strip().splitlines()
        assert len(lines) == 2 and lines[0] == ",".join(COLUMNS), lines
        got = parse_row(lines[1])
        res = float(got["Result"])
        assert got["Result"] == "26.61316121"  # results/results_binom_embar.csv:16
        # Error is the full-precision result minus the comparison (the row prints
        # Result with 10 digits, so res - 26.61316121 is only good to ~1e-9)
        assert abs(float(got["Error"]) - (res - 26.61316121)) < 1e-9
        assert float(got["Abs_Error"]) == abs(float(got["Error"]))


@pytest.mark.gpu
def test_amer_basis_flag():
    # --basis K: the K-term regression; with --stats the fallback count is printed.
    p = run("mc_amer", "put", 100, 100, 0.06, 0.2, 1, 1 << 16, 50, "--seed", 5, "--basis", 4, "--stats")
    assert p.returncode == 0, p.stderr
    got = parse_row(p.stdout)
    assert got["Method"] == "HIP" and abs(float(got["Result"]) - 5.7989) < 0.05
    q = run("mc_amer", "put", 100, 100, 0.06, 0.2, 1, 1 << 16, 50, "--seed", 5)
    assert parse_row(q.stdout)["Result"] != got["Result"]  # a different regression
    for bad in (["--basis", 9], ["--basis", 4, "--ref-exact"], ["--basis", 4, "--rng", "mt19937", "--seed", 1]):
        p = run("mc_amer", "put", 100, 100, 0.06, 0.2, 1, 4096, 50, *bad)
        assert p.returncode == 1 and "basis" in p.stderr, (bad, p.stderr)


@pytest.mark.gpu
def test_stats_line_states_the_fp64_error():
    p = run("mc_asia", "call", 100, 100, 0.05, 0.25, 1, 1 << 16, 252, "--seed", 3, "--stats")
    assert p.returncode == 0, p.stderr
    line = [l for l in p.stderr.splitlines() if l.startswith("stats:")][0]
    kv = dict(t.split("=", 1) for t in line.split()[1:])
    assert float(kv["fp64_err_se"]) < 0.05 and float(kv["fp64_err"]) >= 0.0
    q = run("mc_asia", "call", 100, 100, 0.05, 0.25, 1, 1 << 16, 252, "--seed", 3, "--precision", "fp64")
    d = abs(float(parse_row(p.stdout)["Result"]) - float(parse_row(q.stdout)["Result"]))
    assert abs(d - float(kv["fp64_err"])) < 1e-8  # the rows print 10 significant digits


@pytest.mark.gpu
def test_trailing_thread_count_is_accepted():
    # the _omp programs' argv (src/mc_eur_omp.cpp:41): a trailing thread count
    p = run("mc_eur", "call", 100, 100, 0.05, 0.2, 1, 100000, 16)
    assert p.returncode == 0, p.stderr
    got = parse_row(p.stdout)
    assert got["Method"] == "HIP" and abs(float(got["Result"]) - 10.45) < 0.2


@pytest.mark.gpu
@pytest.mark.parametrize("prog", OMP_PROGS)
def test_omp_names_print_the_same_row(prog):
    """Each <prog>_omp takes the reference OpenMP program's argv, trailing
    thread count included, and prints the row of <prog> for the same input."""
    base = prog[:-len("_omp")]
    pos = {"mc_eur": [100000], "binom_embar": [2000], "mc_asia": [4096, 50], "mc_amer": [4096, 20],
           "mc_eur_multi": [1 << 14, 8, 0.3]}[base]
    args = ["call", 100, 100, 0.05, 0.2, 1] + pos
    flags = [] if base == "binom_embar" else ["--seed", 7]
    print("[case]", prog, *args, flush=True)
    p = run(prog, *args, 16, *flags)
    q = run(base, *args, *flags)
    assert p.returncode == 0 and q.returncode == 0, (p.stderr, q.stderr)
    a, b = parse_row(p.stdout), parse_row(q.stdout)
    for c in COLUMNS:
        if c not in ("T_overall", "T_calculation"):
            assert a[c] == b[c], (prog, c, a[c], b[c])
