"""GPU parity: libpcf (through the C-ABI) against the oracle.

Bars:
  * verification mode (host normals, FP64 replay of the reference order) and
    the chain binomial: relative error <= 1e-12 against the oracle;
  * FP64 Philox production vs the oracle's FP64 Philox restatement on the
    same counters: relative error <= 1e-12;
  * FP32/Kahan vs FP64 on the same Philox paths: |difference| <= 1 standard
    error (observed differences are ~1e-3 SE or less), and both within 4 SE
    of the closed form where one exists.
"""
import csv
import ctypes as C
import math
import os

import numpy as np
import pytest

import _oracle as O

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
REL = 1e-12


def say(*a):
    print("[case]", *a, flush=True)


def rel(a, b):
    return abs(a - b) / max(abs(b), 1e-300)


# ---------------------------------------------------------------- generator
def test_device_normals_match_oracle_philox(engine):
    say("pcf_normals fp64/fp32 vs oracle, seed 99, streams 1..4")
    for stream in (1, 2, 3, 4):
        want = O.philox_normals(99, stream, 1000, 512, 37)
        got64 = engine.normals(99, stream, 1000, 512, 37, "fp64")
        got32 = engine.normals(99, stream, 1000, 512, 37, "fp32k")
        assert np.abs(got64 - want).max() < 1e-13
        # FP32 Box-Muller on the hardware log/sqrt/sin/cos units
        assert np.abs(got32 - want).max() < 2e-5


def test_device_normals_launch_shape_invariance(engine):
    say("pcf_normals path window invariance")
    a = engine.normals(5, 2, 0, 4096, 16, "fp32k")
    b = engine.normals(5, 2, 1000, 100, 16, "fp32k")
    assert np.array_equal(a[1000:1100], b)


# ---------------------------------------------------------------- mc_eur
@pytest.mark.parametrize("N,pf,seed", [(1000, "call", 42), (100000, "put", 1), (1000000, "call", 1234567)])
def test_eur_verification_vs_oracle(engine, N, pf, seed):
    say("mc_eur verify", N, pf, seed)
    z = O.normals_mt(seed, 1.0, N)  # sqrt(T) = 1
    p = 1.0 if pf == "call" else -1.0
    want = O.lib().orc_mc_eur_z(100, 100, 0.05, 0.2, 1.0, N, p, O.ptr(z))
    got = engine.mc_eur(100, 100, 0.05, 0.2, 1.0, N, pf, normals=z, precision="fp64")
    assert rel(got.price, want) < REL


def test_eur_verification_vs_reference_fixture(engine):
    """Same seed as a fixture row of the reference's own mc_eur."""
    rows = [r for r in csv.DictReader(open(os.path.join(HERE, "golden", "ref_kernels.csv")))
            if r["prog"] == "mc_eur"]
    for r in rows[:6]:
        N, seed = int(r["N"]), int(r["seed"])
        say("mc_eur fixture", N, seed, r["payoff"])
        T = float(r["T"])
        z = O.normals_mt(seed, math.sqrt(T), N)
        got = engine.mc_eur(float(r["S0"]), float(r["E"]), float(r["r"]), float(r["sigma"]), T, N,
                            r["payoff"], normals=z)
        assert rel(got.price, float(r["result"])) < REL


def test_eur_fp64_philox_vs_oracle(engine):
    say("mc_eur fp64 philox vs oracle, 2^16+3 paths")
    N = (1 << 16) + 3
    se = C.c_double()
    want = O.lib().orc_philox_mc_eur(100, 100, 0.05, 0.2, 1.0, N, 1.0, 7, C.byref(se))
    got = engine.mc_eur(100, 100, 0.05, 0.2, 1.0, N, "call", seed=7, precision="fp64")
    assert rel(got.price, want) < REL
    assert rel(got.std_err, se.value) < 1e-9


def test_eur_1e6_paths_fp32_vs_fp64_and_black_scholes(engine):
    N = 10**6
    bs = O.lib().orc_bs_price(100, 100, 0.05, 0.2, 1.0, 1.0)
    for seed in (1, 2, 3):
        say("mc_eur 1e6 paths fp32k vs fp64 seed", seed)
        r64 = engine.mc_eur(100, 100, 0.05, 0.2, 1.0, N, "call", seed=seed, precision="fp64")
        r32 = engine.mc_eur(100, 100, 0.05, 0.2, 1.0, N, "call", seed=seed, precision="fp32k")
        assert abs(r32.price - r64.price) < 0.05 * r64.std_err
        assert abs(r64.price - bs) < 4 * r64.std_err
        assert abs(r32.price - bs) < 4 * r32.std_err


def test_eur_deterministic_and_shape_invariant(engine):
    say("mc_eur determinism across launch shapes")
    a = engine.mc_eur(100, 100, 0.05, 0.2, 1.0, 123457, "call", seed=9)
    b = engine.mc_eur(100, 100, 0.05, 0.2, 1.0, 123457, "call", seed=9)
    c = engine.mc_eur(100, 100, 0.05, 0.2, 1.0, 123457, "call", seed=9, block_size=64, items_per_thread=7)
    assert a.price == b.price
    assert rel(a.price, c.price) < 1e-14  # same paths, different reduction tree


@pytest.mark.parametrize("nblk", [1, 2, 31, 32, 33, 63, 64, 65, 97, 1000])
def test_eur_grid_reduction_across_arrival_groups(engine, nblk):
    """Grids whose block counts straddle the 32 arrival groups of the grid
    reduction (block b counts on group b mod 32, the last of each group on a
    top counter): every block's partial must reach the total, so the FP64
    price equals the oracle's on the same Philox normals."""
    B = 64  # threads per block, one item (4 paths) per thread
    N = nblk * B * 4 - 3
    say("mc_eur grid reduction, blocks", nblk, "paths", N)
    se = C.c_double()
    want = O.lib().orc_philox_mc_eur(100, 100, 0.05, 0.2, 1.0, N, 1.0, 13, C.byref(se))
    got = engine.mc_eur(100, 100, 0.05, 0.2, 1.0, N, "call", seed=13, precision="fp64", block_size=B,
                        items_per_thread=1)
    assert got.grid_blocks == nblk
    assert rel(got.price, want) < REL
    assert rel(got.std_err, se.value) < 1e-9


def test_eur_invalid_arguments(engine):
    from parcompfin_amd import PcfError
    say("mc_eur invalid arguments")
    with pytest.raises(ValueError):
        engine.mc_eur(100, 100, 0.05, 0.2, 1.0, 1000, "straddle")
    with pytest.raises(PcfError):
        engine.mc_eur(100, 100, 0.05, 0.2, 1.0, 0, "call")
    with pytest.raises(PcfError):
        engine.mc_eur(100, 100, 0.05, 0.2, 1.0, 1000, "call", normals=np.zeros(10))


# ---------------------------------------------------------------- mc_asia
@pytest.mark.parametrize("N,M,pf,seed", [(1024, 252, "call", 42), (4096, 200, "put", 1), (333, 7, "call", 5)])
def test_asia_verification_vs_oracle(engine, N, M, pf, seed):
    say("mc_asia verify", N, M, pf, seed)
    z = O.normals_mt(seed, math.sqrt(1.0 / M), N * M)
    p = 1.0 if pf == "call" else -1.0
    want = O.lib().orc_mc_asia_z(100, 100, 0.05, 0.25, 1.0, N, M, p, O.ptr(z))
    got = engine.mc_asia(100, 100, 0.05, 0.25, 1.0, N, M, pf, normals=z)
    assert rel(got.price, want) < REL


def test_asia_verification_vs_reference_fixture(engine):
    rows = [r for r in csv.DictReader(open(os.path.join(HERE, "golden", "ref_kernels.csv")))
            if r["prog"] == "mc_asia"]
    for r in rows:
        N, M, seed = int(r["N"]), int(r["M"]), int(r["seed"])
        say("mc_asia fixture", N, M, seed, r["payoff"])
        T = float(r["T"])
        z = O.normals_mt(seed, math.sqrt(T / M), N * M)
        got = engine.mc_asia(float(r["S0"]), float(r["E"]), float(r["r"]), float(r["sigma"]), T, N, M,
                             r["payoff"], normals=z)
        assert rel(got.price, float(r["result"])) < REL


@pytest.mark.parametrize("S0,E,r,s,T,M,pf", [(100, 100, 0.05, 0.25, 1.0, 252, "call"),
                                             (100, 110, 0.02, 0.75, 2.0, 50, "call"),
                                             (90, 100, 0.08, 0.3, 0.5, 7, "put")])
def test_asia_production_arithmetic_on_reference_draws(engine, S0, E, r, s, T, M, pf):
    """The production FP64 step (kernel_variant 1 on host draws) prices the
    reference's own mt19937 draws as the reference does (src/mc_asia.cpp:27-37)
    and as the oracle's production restatement does, to 1e-12."""
    say("mc_asia production arithmetic on host draws", M, pf, T)
    N = 2048
    w = O.normals_mt(23, math.sqrt(T / M), N * M)
    p = 1.0 if pf == "call" else -1.0
    ref = O.lib().orc_mc_asia_z(S0, E, r, s, T, N, M, p, O.ptr(w))
    prod = O.lib().orc_mc_asia_prod_z(S0, E, r, s, T, N, M, p, O.ptr(w), None)
    got = engine.mc_asia(S0, E, r, s, T, N, M, pf, normals=w, kernel_variant=1)
    assert rel(got.price, prod) < REL
    assert rel(got.price, ref) < REL
    # kernel_variant 0 on the same draws is the reference's own arithmetic
    got0 = engine.mc_asia(S0, E, r, s, T, N, M, pf, normals=w)
    assert rel(got0.price, ref) < REL


@pytest.mark.parametrize("M", [252, 13])
def test_asia_fp64_philox_vs_oracle(engine, M):
    say("mc_asia fp64 philox vs oracle", M)
    N = 1500
    se, geo = C.c_double(), C.c_double()
    want = O.lib().orc_philox_mc_asia(100, 100, 0.05, 0.25, 1.0, N, M, 1.0, 11, C.byref(se), C.byref(geo))
    got = engine.mc_asia(100, 100, 0.05, 0.25, 1.0, N, M, "call", seed=11, precision="fp64", control=True)
    assert rel(got.price, want) < REL
    assert rel(got.control_price, geo.value) < 1e-11


def test_asia_2p22_paths_fp32_vs_fp64_and_geometric_control(engine):
    N, M = 1 << 22, 252
    say("mc_asia 2^22 x 252 fp32k/fp64 with control")
    r64 = engine.mc_asia(100, 100, 0.05, 0.25, 1.0, N, M, "call", seed=3, precision="fp64", control=True)
    r32 = engine.mc_asia(100, 100, 0.05, 0.25, 1.0, N, M, "call", seed=3, precision="fp32k", control=True)
    geo = O.lib().orc_geo_asian(100, 100, 0.05, 0.25, 1.0, M, 1.0)
    assert abs(r32.price - r64.price) < 0.05 * r64.std_err
    assert abs(r64.control_price - geo) < 4 * r64.control_se
    assert abs(r32.control_price - geo) < 4 * r32.control_se
    assert r64.price > r64.control_price


# ---------------------------------------------------------------- binom_embar
@pytest.mark.parametrize("N", [1, 2, 3, 100, 101, 1000, 4097, 10000, 30000])
def test_binom_chain_vs_oracle(engine, N):
    say("binom chain", N)
    want = O.lib().orc_binom_embar(100, 100, 0.05, 0.2, 1.0, N, 1.0, 1)
    got = engine.binom_embar(100, 100, 0.05, 0.2, 1.0, N, "call", ref_exact=True)
    assert rel(got.price, want) < REL


def test_binom_chain_vs_reference_fixtures(engine):
    rows = [r for r in csv.DictReader(open(os.path.join(HERE, "golden", "ref_kernels.csv")))
            if r["prog"] == "binom_embar"]
    # the reference driver's own N = 1e5 value is among them
    assert any(int(r["N"]) == 100000 for r in rows)
    for r in rows:
        N = int(r["N"])
        say("binom fixture", N, r["payoff"])
        got = engine.binom_embar(float(r["S0"]), float(r["E"]), float(r["r"]), float(r["sigma"]),
                                 float(r["T"]), N, r["payoff"], ref_exact=True)
        assert rel(got.price, float(r["result"])) < REL


def _csv(name):
    with open(os.path.join(HERE, "golden", name)) as f:
        return list(csv.DictReader(f))


@pytest.mark.parametrize("row", _csv("binom_large.csv"), ids=lambda r: f"{r['payoff']}-N{r['N']}")
def test_binom_chain_at_1e5_and_1e6(engine, row):
    """The reference-exact chain at N = 1e5 and 1e6 (call and put) against the
    prefix form's values, which equal the reference's chain bit for bit."""
    N = int(row["N"])
    say("binom chain large", N, row["payoff"])
    got = engine.binom_embar(100, 100, 0.05, 0.2, 1.0, N, row["payoff"], ref_exact=True)
    assert rel(got.price, float(row["result"])) < REL


@pytest.mark.parametrize("row", _csv("binom_exact.csv"), ids=lambda r: f"{r['payoff']}-N{r['N']}-E{r['E']}")
def test_binom_dd_production_vs_exact_sum(engine, row):
    """The double-double production form against the exact sum (mpmath, 40
    digits) to 1e-13, and against the oracle's restatement to 1e-13."""
    N = int(row["N"])
    S0, E, r, s, T = (float(row[k]) for k in ("S0", "E", "r", "sigma", "T"))
    say("binom dd", N, row["payoff"], E)
    got = engine.binom_embar(S0, E, r, s, T, N, row["payoff"], ref_exact=False)
    assert rel(got.price, float(row["exact"])) < 1e-13
    pf = 1.0 if row["payoff"] == "call" else -1.0
    assert rel(got.price, O.lib().orc_binom_dd(S0, E, r, s, T, N, pf)) < 1e-13


@pytest.mark.parametrize("N", [1, 2, 3, 100, 101, 4097, 30000])
@pytest.mark.parametrize("pf", ["call", "put"])
def test_binom_dd_production_vs_oracle(engine, N, pf):
    say("binom dd vs oracle", N, pf)
    p = 1.0 if pf == "call" else -1.0
    want = O.lib().orc_binom_dd(100, 110, 0.02, 0.75, 1.0, N, p)
    got = engine.binom_embar(100, 110, 0.02, 0.75, 1.0, N, pf, ref_exact=False)
    assert rel(got.price, want) < 1e-13
