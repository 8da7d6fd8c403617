"""GPU parity for the basket (mc_eur_multi) and the American LSM (mc_amer).

Bars as in test_gpu_parity.py: 1e-12 relative for verification / FP64
against the oracle on identical normals; FP32 within a fraction of a
standard error of FP64 on the same Philox paths; closed-form checks in SE.
The LSM decision can flip for a path whose payoff lies within rounding of
the regression value; the tests compare per-path exercise dates and require
no flips before asserting 1e-12.
"""
import csv
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


def equicorr(d, rho):
    return np.full((d, d), rho) + (1 - rho) * np.eye(d)


def chol(C):
    d = C.shape[0]
    L = np.empty((d, d))
    assert O.lib().orc_cholesky(d, O.ptr(np.ascontiguousarray(C)), O.ptr(L)) == 1
    return L


def random_corr(d, seed):
    rng = np.random.default_rng(seed)
    A = rng.normal(size=(d, d))
    C = A @ A.T + d * np.eye(d)
    Dm = 1 / np.sqrt(np.diag(C))
    return C * Dm[:, None] * Dm[None, :]


# ---------------------------------------------------------------- basket
@pytest.mark.parametrize("d,rho,pf", [(1, 0.0, "call"), (4, 0.5, "call"), (8, 0.3, "put"), (32, 0.3, "call"),
                                      (13, -0.05, "call")])
def test_basket_verification_vs_oracle(engine, d, rho, pf):
    say("basket verify", d, rho, pf)
    N = 20000
    Z = O.normals_mt(100 + d, 1.0, N * d)
    L = chol(equicorr(d, rho))
    p = 1.0 if pf == "call" else -1.0
    want = O.lib().orc_basket_z(100, 100, 0.1, 0.2, 1.0, N, p, d, O.ptr(L), O.ptr(Z), None, None, None)
    got = engine.mc_eur_multi(100, 100, 0.1, 0.2, 1.0, N, pf, d, rho, normals=Z, ref_exact=True)
    assert rel(got.price, want) < REL


def test_basket_extension_verification_vs_oracle(engine):
    """The basket extension: per-asset sigma in [0.15, 0.35] and a random SPD correlation."""
    d, N = 8, 20000
    rng = np.random.default_rng(2)
    sig = rng.uniform(0.15, 0.35, size=d)
    C = random_corr(d, 3)
    L = chol(C)
    Z = O.normals_mt(7, 1.0, N * d)
    say("basket ext verify")
    want = O.lib().orc_basket_z(100, 100, 0.05, 0.2, 1.0, N, 1.0, d, O.ptr(L), O.ptr(Z), O.ptr(sig), None, None)
    got = engine.mc_eur_multi(100, 100, 0.05, 0.2, 1.0, N, "call", d, 0.0, sigmas=sig, corr=C, normals=Z,
                              ref_exact=True)
    assert rel(got.price, want) < REL


@pytest.mark.parametrize("d", [8, 32])
def test_basket_fp64_philox_vs_oracle(engine, d):
    say("basket fp64 philox", d)
    N = 6000
    Z = np.empty(N * d)
    O.lib().orc_philox_basket_normals(21, N, d, O.ptr(Z))
    L = chol(equicorr(d, 0.3))
    want = O.lib().orc_basket_z(100, 100, 0.05, 0.2, 1.0, N, 1.0, d, O.ptr(L), O.ptr(Z), None, None, None)
    got = engine.mc_eur_multi(100, 100, 0.05, 0.2, 1.0, N, "call", d, 0.3, seed=21, precision="fp64")
    assert rel(got.price, want) < REL


@pytest.mark.parametrize("nblk", [1, 33, 64, 65, 130])
def test_basket_grid_reduction_across_arrival_groups(engine, nblk):
    """Basket grids across the arrival groups of the grid reduction (64
    threads per block, one path each): FP64 price equals the oracle's."""
    d, B = 4, 64
    N = nblk * B - 5
    say("basket grid reduction, blocks", nblk)
    Z = np.empty(N * d)
    O.lib().orc_philox_basket_normals(23, N, d, O.ptr(Z))
    L = chol(equicorr(d, 0.3))
    want = O.lib().orc_basket_z(100, 100, 0.05, 0.2, 1.0, N, 1.0, d, O.ptr(L), O.ptr(Z), None, None, None)
    got = engine.mc_eur_multi(100, 100, 0.05, 0.2, 1.0, N, "call", d, 0.3, seed=23, precision="fp64",
                              block_size=B, items_per_thread=1)
    assert got.grid_blocks == nblk
    assert rel(got.price, want) < REL


@pytest.mark.parametrize("d,variant", [(8, 0), (32, 1), (32, 2), (16, 2), (20, 2)])
def test_basket_fp32_vs_fp64(engine, d, variant):
    say("basket fp32 vs fp64", d, variant)
    N = 1 << 20
    r64 = engine.mc_eur_multi(100, 100, 0.05, 0.2, 1.0, N, "call", d, 0.3, seed=5, precision="fp64")
    r32 = engine.mc_eur_multi(100, 100, 0.05, 0.2, 1.0, N, "call", d, 0.3, seed=5, kernel_variant=variant)
    assert abs(r32.price - r64.price) < 0.05 * r64.std_err


@pytest.mark.parametrize("d", [24, 32])
def test_basket_mfma_full_factor(engine, d):
    """Assets i and i + d/2 perfectly correlated, the rest independent: the
    Cholesky factorisation fails at pivot d/2 and the eigen fallback's factor
    puts the pairs' common factor in columns d/2..d-1 of rows 0..d/2-1 (the
    Jacobi rotation of each pair leaves its eigenvalue 2 at the higher index).
    Rows below 16 then have terms at k >= 16, which the MFMA kernel's
    triangular form skips: FP32 MFMA must match FP32 FMA and FP64 on the
    same normals."""
    say("basket MFMA, full factor", d)
    N = 1 << 20
    h = d // 2
    corr = np.eye(d)
    for i in range(h):
        corr[i, i + h] = corr[i + h, i] = 1.0
    kw = dict(corr=corr, seed=7)
    r64 = engine.mc_eur_multi(100, 100, 0.05, 0.2, 1.0, N, "call", d, 0.0, precision="fp64", **kw)
    r32 = engine.mc_eur_multi(100, 100, 0.05, 0.2, 1.0, N, "call", d, 0.0, kernel_variant=2, **kw)
    rfma = engine.mc_eur_multi(100, 100, 0.05, 0.2, 1.0, N, "call", d, 0.0, kernel_variant=1, **kw)
    assert r64.lsm_fallback == 1 and r32.lsm_fallback == 1
    assert abs(r32.price - r64.price) < 0.05 * r64.std_err
    assert abs(r32.price - rfma.price) < 0.05 * r64.std_err


def test_basket_2p24_paths_full_size(engine):
    """d = 8 and 32, 2^24 paths, sigma_i ~ U[0.15,0.35], rho = 0.3 and a random SPD matrix."""
    N = 1 << 24
    for d in (8, 32):
        rng = np.random.default_rng(d)
        sig = rng.uniform(0.15, 0.35, size=d)
        for corr in (equicorr(d, 0.3), random_corr(d, 11)):
            say("basket 2^24 paths", d)
            r64 = engine.mc_eur_multi(100, 100, 0.05, 0.2, 1.0, N, "call", d, 0.3, sigmas=sig, corr=corr,
                                      seed=9, precision="fp64")
            r32 = engine.mc_eur_multi(100, 100, 0.05, 0.2, 1.0, N, "call", d, 0.3, sigmas=sig, corr=corr, seed=9)
            assert abs(r32.price - r64.price) < 0.05 * r64.std_err


def test_basket_perfect_correlation_is_black_scholes(engine):
    """rho = 1: the correlation is singular (eigen factor path, as mvnorm's
    fallback); every asset moves together, so the basket is one GBM."""
    say("basket rho=1")
    r = engine.mc_eur_multi(100, 100, 0.05, 0.2, 1.0, 1 << 20, "call", 8, 1.0, seed=4, precision="fp64")
    assert r.lsm_fallback == 1
    bs = O.lib().orc_bs_price(100, 100, 0.05, 0.2, 1.0, 1.0)
    assert abs(r.price - bs) < 4 * r.std_err


@pytest.mark.parametrize("T", [0.5, 2.0])
def test_basket_production_sqrt_T(engine, T):
    """Production (ref_exact 0) at T != 1, where the reference's missing
    sqrt(T) matters:
      * d = 1, FP64 Philox: the reference's mc_eur (src/mc_eur.cpp:15-27) on
        the same normals times sqrt(T), to 1e-12;
      * d = 8, FP64 Philox, two correlations: the oracle's production basket on
        the same normals, to 1e-12;
      * host normals (production arithmetic, reference order): the oracle's
        production basket to 1e-12;
      * rho = 1, FP64: the single-asset Black-Scholes price within 4 SE."""
    say("basket production sqrt(T)", T)
    N = 1 << 14
    z = np.empty(N)
    O.lib().orc_philox_basket_normals(31, N, 1, O.ptr(z))
    want = O.lib().orc_mc_eur_z(100, 100, 0.05, 0.2, T, N, 1.0, O.ptr(z * math.sqrt(T)))
    got = engine.mc_eur_multi(100, 100, 0.05, 0.2, T, N, "call", 1, 0.0, seed=31, precision="fp64")
    assert rel(got.price, want) < REL
    d = 8
    Z = np.empty(N * d)
    O.lib().orc_philox_basket_normals(32, N, d, O.ptr(Z))
    sig = np.random.default_rng(3).uniform(0.15, 0.35, size=d)
    for corr in (equicorr(d, 0.3), random_corr(d, 5)):
        L = chol(corr)
        for pf in ("call", "put"):
            p = 1.0 if pf == "call" else -1.0
            want = O.lib().orc_basket_prod_z(100, 100, 0.05, 0.2, T, N, p, d, O.ptr(L), O.ptr(Z), O.ptr(sig),
                                             None, None, None)
            got = engine.mc_eur_multi(100, 100, 0.05, 0.2, T, N, pf, d, 0.3, sigmas=sig, corr=corr, seed=32,
                                      precision="fp64")
            assert rel(got.price, want) < REL
            goth = engine.mc_eur_multi(100, 100, 0.05, 0.2, T, N, pf, d, 0.3, sigmas=sig, corr=corr, normals=Z)
            assert rel(goth.price, want) < REL
    r = engine.mc_eur_multi(100, 100, 0.05, 0.2, T, 1 << 20, "put", 4, 1.0, seed=6, precision="fp64")
    bs = O.lib().orc_bs_price(100, 100, 0.05, 0.2, T, -1.0)
    assert abs(r.price - bs) < 4 * r.std_err


def test_basket_factor_cache_collisions_never_mix_baskets(engine):
    """With the cache hash masked to a single value every basket collides;
    each price still equals its price with the full hash (the cache compares
    the full key on a hit), alone and interleaved, and in a served batch."""
    cases = [dict(assets=8, rho=0.3), dict(assets=8, rho=0.5), dict(assets=8, rho=0.3, sigmas=np.full(8, 0.25)),
             dict(assets=4, rho=0.3), dict(assets=8, rho=0.3, corr=random_corr(8, 2))]
    N = 1 << 14

    def price(kw, s):
        return engine.mc_eur_multi(100, 100, 0.05, 0.2, 1.0, N, "call", seed=s, precision="fp64", **kw).price

    from parcompfin_amd import _native
    L = _native.load()
    say("basket cache, full hash")
    want = [price(kw, 40 + i) for i, kw in enumerate(cases)]
    try:
        L.pcf_debug_basket_hash_mask(0)
        say("basket cache, every hash equal")
        for rnd in range(2):
            for i in ([0, 1, 2, 3, 4] if rnd == 0 else [4, 2, 0, 3, 1]):
                assert price(cases[i], 40 + i) == want[i], (rnd, i)
        from parcompfin_amd import Server
        with Server(engine, max_inflight=1) as srv:
            ts = [srv.submit("basket", 100, 100, 0.05, 0.2, 1.0, N, "call", seed=40 + i, precision="fp64", **kw)
                  for i, kw in enumerate(cases)]
            got = [srv.wait(t).price for t in ts]
        assert got == want
    finally:
        L.pcf_debug_basket_hash_mask(0xFFFFFFFFFFFFFFFF)


def test_basket_invalid(engine):
    from parcompfin_amd import PcfError
    say("basket invalid")
    with pytest.raises(PcfError):
        engine.mc_eur_multi(100, 100, 0.05, 0.2, 1.0, 1000, "call", 0, 0.3)
    with pytest.raises(PcfError):  # indefinite: rho < -1/(d-1)
        engine.mc_eur_multi(100, 100, 0.05, 0.2, 1.0, 1000, "call", 4, -0.5)
    with pytest.raises(PcfError):
        engine.mc_eur_multi(100, 100, 0.05, 0.2, 1.0, 1000, "call", 8, 0.3, kernel_variant=2)


def test_block_sizes_beyond_the_kernels_bound_are_refused(engine):
    """The basket kernels and the LSM's per-date kernels are built for at most
    256 threads per block: larger block sizes are refused before any launch,
    and the context keeps working."""
    from parcompfin_amd import PcfError
    say("block sizes beyond the launch bound")
    for B in (512, 1024):
        with pytest.raises(PcfError) as ei:
            engine.mc_eur_multi(100, 100, 0.05, 0.2, 1.0, 4096, "call", 8, 0.3, block_size=B)
        assert ei.value.status == 1 and "threads per block" in str(ei.value)
        with pytest.raises(PcfError) as ei:
            engine.mc_amer(100, 100, 0.06, 0.2, 1.0, 4096, 20, "put", kernel_variant=1, block_size=B)
        assert ei.value.status == 1 and "threads per block" in str(ei.value)
    # the persistent form takes up to 1024 threads per block
    r = engine.mc_amer(100, 100, 0.06, 0.2, 1.0, 4096, 20, "put", kernel_variant=2, block_size=1024,
                       items_per_thread=1)
    assert r.block_size == 1024
    got = engine.mc_eur_multi(100, 100, 0.05, 0.2, 1.0, 4096, "call", 4, 0.3, seed=23, precision="fp64",
                              block_size=256, items_per_thread=1)
    Z = np.empty(4096 * 4)
    O.lib().orc_philox_basket_normals(23, 4096, 4, O.ptr(Z))
    want = O.lib().orc_basket_z(100, 100, 0.05, 0.2, 1.0, 4096, 1.0, 4, O.ptr(chol(equicorr(4, 0.3))), O.ptr(Z),
                                None, None, None)
    assert rel(got.price, want) < REL


# ---------------------------------------------------------------- mc_amer
def _amer_ref(N, M, pf, w, quirk):
    when = np.zeros(N, dtype=np.int32)
    v = O.lib().orc_mc_amer_z(100, 100, 0.06, 0.2, 1.0, N, M, pf, O.ptr(w), quirk,
                              when.ctypes.data_as(O.C.POINTER(O.C.c_int32)))
    return v, when


@pytest.mark.parametrize("N,M,pf,seed", [(1024, 50, "put", 42), (16384, 50, "put", 1), (16384, 50, "call", 1),
                                         (2048, 7, "put", 3), (4096, 2, "put", 5), (64, 1, "put", 6)])
def test_amer_verification_vs_oracle(engine, N, M, pf, seed):
    say("amer verify", N, M, pf, seed)
    w = O.normals_mt(seed, math.sqrt(1.0 / M), (N // 2) * M)
    p = 1.0 if pf == "call" else -1.0
    want, wwhen = _amer_ref(N, M, p, w, 1)
    got = engine.mc_amer(100, 100, 0.06, 0.2, 1.0, N, M, pf, normals=w, ref_exact=True, return_when=True)
    flips = int((got.extra["when"] != wwhen).sum())
    say("  flips", flips)
    assert flips == 0
    assert rel(got.price, want) < REL


def test_amer_verification_vs_reference_fixture(engine):
    rows = [r for r in csv.DictReader(open(os.path.join(HERE, "golden", "ref_kernels.csv")))
            if r["prog"] == "mc_amer"]
    for r in rows:
        N, M, seed = int(r["N"]), int(r["M"]), int(r["seed"])
        say("amer fixture", N, M, seed, r["payoff"])
        T = float(r["T"])
        w = O.normals_mt(seed, math.sqrt(T / M), (N // 2) * M)
        got = engine.mc_amer(float(r["S0"]), float(r["E"]), float(r["r"]), float(r["sigma"]), T, N, M,
                             r["payoff"], normals=w, ref_exact=True)
        assert rel(got.price, float(r["result"])) < REL


def test_amer_production_arithmetic_on_host_normals(engine):
    """Corrected exercise test with the production basis S/E and Cholesky
    solve vs the oracle's corrected reference basis: the fitted values agree
    in exact arithmetic, so with no flipped decision the prices agree."""
    N, M = 8192, 50
    w = O.normals_mt(77, math.sqrt(1.0 / M), (N // 2) * M)
    want, wwhen = _amer_ref(N, M, -1.0, w, 0)
    say("amer production arithmetic, host normals")
    got = engine.mc_amer(100, 100, 0.06, 0.2, 1.0, N, M, "put", normals=w, return_when=True)
    flips = int((got.extra["when"] != wwhen).sum())
    say("  flips", flips)
    assert flips <= 2
    if flips == 0:
        assert rel(got.price, want) < 1e-11
    else:
        assert rel(got.price, want) < 1e-4


def test_amer_fp64_philox_vs_oracle(engine):
    N, M = 8192, 50
    w = np.empty((N // 2) * M)
    O.lib().orc_philox_amer_normals(13, N, M, 1.0, O.ptr(w))
    want, wwhen = _amer_ref(N, M, -1.0, w, 0)
    say("amer fp64 philox vs oracle")
    got = engine.mc_amer(100, 100, 0.06, 0.2, 1.0, N, M, "put", seed=13, precision="fp64", return_when=True)
    flips = int((got.extra["when"] != wwhen).sum())
    assert flips <= 2
    assert rel(got.price, want) < (1e-11 if flips == 0 else 1e-4)


def test_amer_2p20x50_fp32_vs_fp64_and_bounds(engine):
    N, M = 1 << 20, 50
    say("amer 2^20 x 50 fp32/fp64")
    r64 = engine.mc_amer(100, 100, 0.06, 0.2, 1.0, N, M, "put", seed=17, precision="fp64")
    r32 = engine.mc_amer(100, 100, 0.06, 0.2, 1.0, N, M, "put", seed=17)
    euro = O.lib().orc_bs_price(100, 100, 0.06, 0.2, 1.0, -1.0)
    lattice = 5.798877611783797  # reference binom_vanilla_amer, N = 2000 (tests/golden)
    say(f"  fp64 {r64.price:.6f} fp32 {r32.price:.6f} se {r64.std_err:.2e} euro-ctl {r64.control_price:.6f}")
    assert abs(r32.price - r64.price) < r64.std_err
    assert r64.price > euro
    assert abs(r64.control_price - euro) < 4 * r64.control_se
    assert abs(r64.price - lattice) < 0.05  # LSM with a quadratic basis sits slightly below


def test_amer_graph_replay_is_deterministic(engine):
    say("amer graph replay")
    a = engine.mc_amer(100, 100, 0.06, 0.2, 1.0, 1 << 16, 50, "put", seed=3)
    b = engine.mc_amer(100, 100, 0.06, 0.2, 1.0, 1 << 16, 50, "put", seed=3)
    c = engine.mc_amer(100, 100, 0.06, 0.2, 1.0, 1 << 16, 50, "put", seed=4)
    d = engine.mc_amer(100, 100, 0.06, 0.2, 1.0, 1 << 16, 50, "put", seed=3)
    assert a.price == b.price == d.price and c.price != a.price


def test_amer_errors_like_reference(engine):
    from parcompfin_amd import PcfError
    say("amer errors")
    with pytest.raises(PcfError):  # odd N: include/common.h:180
        engine.mc_amer(100, 100, 0.06, 0.2, 1.0, 1001, 50, "put")
    # sigma = 0: every in-the-money x is equal, the Gram matrix is singular and
    # the reference throws "Detereminant is not > 0".
    w = np.zeros((1024 // 2) * 10)
    with pytest.raises(PcfError) as ei:
        engine.mc_amer(100, 110, 0.01, 0.0, 1.0, 1024, 10, "put", normals=w, ref_exact=True)
    assert ei.value.status == 5
    # production falls back to a reduced basis instead
    r = engine.mc_amer(100, 110, 0.01, 0.0, 1.0, 1024, 10, "put", normals=w)
    assert r.lsm_fallback > 0


@pytest.mark.parametrize("N,M,prec,ref", [(1 << 16, 50, "fp32k", False), (1 << 16, 50, "fp64", False),
                                          (4096, 25, "fp64", True), (1 << 20, 50, "fp32k", False),
                                          (1 << 22, 50, "fp32k", False), (2, 3, "fp64", False),
                                          (300002, 30, "fp32k", False), (300002, 30, "fp64", True)])
def test_amer_persistent_matches_graph_form(engine, N, M, prec, ref):
    """The one-launch persistent sweep and the per-date graph give the same
    exercise dates (up to a flipped decision at the rounding boundary) and
    prices."""
    say("amer persistent vs graph", N, M, prec, ref)
    kw = dict(seed=29, precision=prec, ref_exact=ref, return_when=True)
    g = engine.mc_amer(100, 100, 0.06, 0.2, 1.0, N, M, "put", kernel_variant=1, **kw)
    p = engine.mc_amer(100, 100, 0.06, 0.2, 1.0, N, M, "put", kernel_variant=2, **kw)
    flips = int((g.extra["when"] != p.extra["when"]).sum())
    say("  flips", flips, "grid", p.grid_blocks, "kernel_ms graph/persistent", g.kernel_ms, p.kernel_ms)
    assert flips <= max(2, N // 100000)
    if flips == 0:
        assert abs(g.price - p.price) <= 1e-12 * g.price
    else:
        assert abs(g.price - p.price) <= 0.01 * g.std_err
    assert g.control_price == p.control_price or abs(g.control_price - p.control_price) < 1e-12 * g.control_price


def test_amer_small_grid_slots_across_launches(engine):
    """Grids of at most 64 blocks all-reduce through self-validating slots
    that every launch leaves zero: a 64-block sweep over 199 dates (each slot
    parity reused 66 times) matches the per-date graph, and repeats bit for
    bit after launches of other grid sizes (32 blocks, 256 blocks through the
    arrival counters) and another date count on the same context."""
    kw = dict(seed=37, return_when=True)
    first = engine.mc_amer(100, 100, 0.06, 0.2, 1.0, 1 << 14, 200, "put", kernel_variant=2, precision="fp32k", **kw)
    say("  small grid", first.grid_blocks, "x", first.block_size, "kernel_ms", first.kernel_ms)
    assert first.grid_blocks <= 64
    g = engine.mc_amer(100, 100, 0.06, 0.2, 1.0, 1 << 14, 200, "put", kernel_variant=1, precision="fp32k", **kw)
    flips = int((g.extra["when"] != first.extra["when"]).sum())
    assert flips <= 2
    if flips == 0:
        assert abs(g.price - first.price) <= 1e-12 * g.price
    else:
        assert abs(g.price - first.price) <= 0.01 * g.std_err
    others = [engine.mc_amer(100, 100, 0.06, 0.2, 1.0, 1 << 13, 37, "put", kernel_variant=2, precision="fp64", **kw),
              engine.mc_amer(100, 100, 0.06, 0.2, 1.0, 1 << 19, 50, "put", kernel_variant=2, precision="fp32k", **kw)]
    say("  between:", [(o.grid_blocks, o.block_size) for o in others])
    assert others[0].grid_blocks <= 64 < others[1].grid_blocks
    again = engine.mc_amer(100, 100, 0.06, 0.2, 1.0, 1 << 14, 200, "put", kernel_variant=2, precision="fp32k", **kw)
    assert (again.price, again.std_err) == (first.price, first.std_err)
    assert np.array_equal(again.extra["when"], first.extra["when"])


def test_amer_quadratic_persistent_grid_capped_at_256_blocks(engine):
    """The quadratic's persistent form gathers at most 256 blocks' sums per
    date through LDS: a forced shape of 512 blocks (64 threads x 1 pair at
    2^16 paths) is refused for the persistent form and runs as the per-date
    graph otherwise, with the same decisions as the default shape."""
    from parcompfin_amd import PcfError
    N, M = 1 << 16, 50
    kw = dict(seed=31, precision="fp32k", return_when=True)
    with pytest.raises(PcfError) as ei:
        engine.mc_amer(100, 100, 0.06, 0.2, 1.0, N, M, "put", kernel_variant=2, block_size=64, items_per_thread=1,
                       **kw)
    assert ei.value.status == 1  # PCF_ERR_INVALID_ARG
    forced = engine.mc_amer(100, 100, 0.06, 0.2, 1.0, N, M, "put", block_size=64, items_per_thread=1, **kw)
    default = engine.mc_amer(100, 100, 0.06, 0.2, 1.0, N, M, "put", **kw)
    say("  forced grid", forced.grid_blocks, "x", forced.block_size, "default", default.grid_blocks)
    assert default.grid_blocks <= 256
    flips = int((forced.extra["when"] != default.extra["when"]).sum())
    assert flips <= 2
    assert abs(forced.price - default.price) <= (1e-12 if flips == 0 else 1e-4) * default.price


# ---- K-term regression basis {1, x, .., x^(K-1)}, x = S/E - 1 -------------------

def _amer_poly_ref(N, M, pf, w, K):
    when = np.zeros(N, dtype=np.int32)
    flags = np.zeros(M + 1, dtype=np.int32)
    I32 = O.C.POINTER(O.C.c_int32)
    v = O.lib().orc_mc_amer_poly(100, 100, 0.06, 0.2, 1.0, N, M, pf, O.ptr(w), K, when.ctypes.data_as(I32),
                                 flags.ctypes.data_as(I32))
    return v, when, int((flags[1:M] == 4).sum())


@pytest.mark.parametrize("K,pf", [(2, "put"), (4, "put"), (5, "call"), (6, "put")])
def test_amer_basis_host_normals_vs_oracle(engine, K, pf):
    """FP64 on the host's normals: the device's K-term regression makes the
    oracle's decisions (up to a flip at the rounding boundary of a fit, from
    the different summation order) and its price."""
    N, M = 8192, 50
    w = O.normals_mt(70 + K, math.sqrt(1.0 / M), (N // 2) * M)
    p = 1.0 if pf == "call" else -1.0
    want, wwhen, wred = _amer_poly_ref(N, M, p, w, K)
    say("amer basis host normals", K, pf)
    got = engine.mc_amer(100, 100, 0.06, 0.2, 1.0, N, M, pf, normals=w, lsm_basis=K, return_when=True)
    flips = int((got.extra["when"] != wwhen).sum())
    say("  flips", flips, "reduced dates", got.lsm_fallback, wred)
    assert flips <= 2
    assert rel(got.price, want) < (1e-11 if flips == 0 else 1e-4)
    assert got.lsm_fallback == wred


@pytest.mark.parametrize("K", [2, 4, 6])
def test_amer_basis_fp64_philox_vs_oracle(engine, K):
    N, M = 16384, 30
    w = np.empty((N // 2) * M)
    O.lib().orc_philox_amer_normals(17, N, M, 1.0, O.ptr(w))
    want, wwhen, _ = _amer_poly_ref(N, M, -1.0, w, K)
    say("amer basis fp64 philox", K)
    got = engine.mc_amer(100, 100, 0.06, 0.2, 1.0, N, M, "put", seed=17, precision="fp64", lsm_basis=K,
                         return_when=True)
    flips = int((got.extra["when"] != wwhen).sum())
    assert flips <= 2
    assert rel(got.price, want) < (1e-11 if flips == 0 else 1e-4)


@pytest.mark.parametrize("K", [4, 5])
def test_amer_basis_fp32_at_full_size(engine, K):
    """2^20 paths, 50 dates (the American workload's size): FP32 paths within
    the standard error of FP64, both above the European put and near the
    reference lattice (binom_vanilla_amer N=2000: 5.7989)."""
    N, M = 1 << 20, 50
    say("amer basis fp32 full size", K)
    f32 = engine.mc_amer(100, 100, 0.06, 0.2, 1.0, N, M, "put", seed=3, lsm_basis=K)
    f64 = engine.mc_amer(100, 100, 0.06, 0.2, 1.0, N, M, "put", seed=3, precision="fp64", lsm_basis=K)
    q = engine.mc_amer(100, 100, 0.06, 0.2, 1.0, N, M, "put", seed=3)
    say("  K", K, f32.price, f64.price, "quadratic", q.price, "se", f64.std_err, "kernel_ms", f32.kernel_ms)
    assert abs(f32.price - f64.price) < f64.std_err
    eur = O.lib().orc_bs_price(100, 100, 0.06, 0.2, 1.0, -1.0)
    assert eur < f32.price and abs(f32.price - 5.798877611783797) < 0.03
    assert abs(f32.price - q.price) < 0.02  # more terms move the estimate by little


@pytest.mark.parametrize("K,N,prec", [(2, 1 << 16, "fp32k"), (4, 1 << 16, "fp64"), (6, 1 << 18, "fp32k"),
                                       (5, 1 << 20, "fp32k")])
def test_amer_basis_persistent_matches_graph_form(engine, K, N, prec):
    """The K-term regression in the per-date graph and in the persistent sweep:
    the same exercise dates up to a flip at the rounding boundary of a fit."""
    M = 50
    say("amer basis persistent vs graph", K, N, prec)
    kw = dict(seed=31, precision=prec, lsm_basis=K, return_when=True)
    g = engine.mc_amer(100, 100, 0.06, 0.2, 1.0, N, M, "put", kernel_variant=1, **kw)
    p = engine.mc_amer(100, 100, 0.06, 0.2, 1.0, N, M, "put", kernel_variant=2, **kw)
    flips = int((g.extra["when"] != p.extra["when"]).sum())
    say("  flips", flips, "kernel_ms graph/persistent", g.kernel_ms, p.kernel_ms)
    assert flips <= max(2, N // 100000)
    if flips == 0:
        assert abs(g.price - p.price) <= 1e-12 * g.price
    else:
        assert abs(g.price - p.price) <= 0.01 * g.std_err
    assert g.lsm_fallback == p.lsm_fallback


def test_amer_basis_graph_form_host_normals_vs_oracle(engine):
    N, M, K = 8192, 50, 4
    w = O.normals_mt(91, math.sqrt(1.0 / M), (N // 2) * M)
    want, wwhen, wred = _amer_poly_ref(N, M, -1.0, w, K)
    say("amer basis graph form, host normals")
    got = engine.mc_amer(100, 100, 0.06, 0.2, 1.0, N, M, "put", normals=w, lsm_basis=K, kernel_variant=1,
                         return_when=True)
    flips = int((got.extra["when"] != wwhen).sum())
    assert flips <= 2
    assert rel(got.price, want) < (1e-11 if flips == 0 else 1e-4)
    assert got.lsm_fallback == wred


def test_amer_basis_many_dates(engine):
    """M = 1100 exceeds the persistent form's date table: the graph form runs."""
    say("amer basis M=1100")
    r = engine.mc_amer(100, 100, 0.06, 0.2, 1.0, 1 << 16, 1100, "put", seed=6, lsm_basis=4)
    eur = O.lib().orc_bs_price(100, 100, 0.06, 0.2, 1.0, -1.0)
    assert eur < r.price and abs(r.price - 5.798877611783797) < 0.06
    assert r.n_steps == 1100


def test_amer_basis_errors(engine):
    from parcompfin_amd import PcfError
    say("amer basis errors")
    for kw in (dict(lsm_basis=7), dict(lsm_basis=1), dict(lsm_basis=-2), dict(lsm_basis=4, ref_exact=True)):
        with pytest.raises(PcfError):
            engine.mc_amer(100, 100, 0.06, 0.2, 1.0, 4096, 20, "put", **kw)
    with pytest.raises(PcfError):  # the persistent form asked for explicitly holds at most 1023 dates
        engine.mc_amer(100, 100, 0.06, 0.2, 1.0, 4096, 1100, "put", lsm_basis=4, kernel_variant=2)
    # 0 and 3 are the same quadratic; M = 1 has no regression date
    a = engine.mc_amer(100, 100, 0.06, 0.2, 1.0, 4096, 20, "put", seed=4)
    b = engine.mc_amer(100, 100, 0.06, 0.2, 1.0, 4096, 20, "put", seed=4, lsm_basis=3)
    assert a.price == b.price
    r = engine.mc_amer(100, 100, 0.06, 0.2, 1.0, 4096, 1, "put", seed=4, lsm_basis=5)
    assert abs(r.price - r.control_price) < 1e-12 * r.price  # no date to exercise on: the European value
