"""Binomial lattices (the reference's binom_vanilla programs) on the GPU.

The bar is bit equality: every node is formed from its two children with the
reference's operations (src/binom_vanilla_eur.cpp:32-36,
src/binom_vanilla_amer.cpp:31-37), so the tiled schedule must give the
oracle's bits, and the oracle's bits equal the reference drivers'
(tests/golden/ref_kernels.csv).  N runs across the tile edges: American
launches advance 32 levels with 32 exact nodes per wave (128 per block),
European ones 64 levels with 64 per wave (256 per block).
"""
import csv
import os

import pytest

import _oracle as O
from parcompfin_amd import PcfError

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))


def say(*a):
    print("[case]", *a, flush=True)


EDGES = [1, 2, 3, 31, 32, 33, 63, 64, 65, 95, 96, 97, 127, 128, 129, 191, 192, 193, 255, 256, 257, 511, 512, 513,
         639, 640, 641, 1000, 2000, 5000]


@pytest.mark.parametrize("american", [0, 1])
@pytest.mark.parametrize("pf", ["call", "put"])
def test_lattice_bitwise_vs_oracle_across_tile_edges(engine, american, pf):
    code = 1.0 if pf == "call" else -1.0
    for N in EDGES:
        say("lattice", pf, "american" if american else "european", N)
        got = engine.binom_vanilla(100, 100, 0.06, 0.2, 1.0, N, pf, american=bool(american)).price
        want = O.lib().orc_binom_vanilla(100, 100, 0.06, 0.2, 1.0, N, code, american)
        assert got == want, (N, got, want)


def test_lattice_other_parameters_bitwise(engine):
    # The published binom_embar parameters (results/results_binom_embar.csv) and a
    # deep in-the-money put.
    for args in ((100, 110, 0.02, 0.75, 1.0), (100, 140, 0.05, 0.3, 2.0), (80, 60, 0.01, 0.4, 0.5)):
        for american in (0, 1):
            for pf, code in (("call", 1.0), ("put", -1.0)):
                got = engine.binom_vanilla(*args, 777, pf, american=bool(american)).price
                want = O.lib().orc_binom_vanilla(*args, 777, code, american)
                assert got == want, (args, american, pf)


@pytest.mark.parametrize("american", [0, 1])
def test_lattice_tiny_values_bitwise(engine, american):
    """Far out-of-the-money calls whose values run down through the range
    where the division by R is rescaled (below 2^-900) and into subnormals
    (where it falls back to the division instruction): the root itself is
    1e-288 .. 1e-323, so any node's last bit would show in it."""
    for E in (3e6, 3.5e6, 4e6, 4.5e6, 5e6):
        say("lattice tiny values", E, "american" if american else "european")
        got = engine.binom_vanilla(100, E, 0.05, 0.3, 1.0, 2000, "call", american=bool(american)).price
        want = O.lib().orc_binom_vanilla(100, E, 0.05, 0.3, 1.0, 2000, 1.0, american)
        assert got == want, (E, got, want)


@pytest.mark.parametrize("american", [0, 1])
def test_lattice_n1e4_bitwise(engine, american):
    for args in ((100, 100, 0.06, 0.2, 1.0), (100, 110, 0.02, 0.75, 1.0)):
        for pf, code in (("call", 1.0), ("put", -1.0)):
            say("lattice N=1e4", args, pf, american)
            got = engine.binom_vanilla(*args, 10000, pf, american=bool(american)).price
            want = O.lib().orc_binom_vanilla(*args, 10000, code, american)
            assert got == want, (args, pf, got, want)


def test_lattice_vs_reference_fixtures(engine):
    rows = [r for r in csv.DictReader(open(os.path.join(HERE, "golden", "ref_kernels.csv")))
            if r["prog"].startswith("binom_vanilla")]
    assert len(rows) >= 4
    for r in rows:
        american = r["prog"] == "binom_vanilla_amer"
        got = engine.binom_vanilla(float(r["S0"]), float(r["E"]), float(r["r"]), float(r["sigma"]), float(r["T"]),
                                   int(r["N"]), r["payoff"], american=american).price
        assert got == float(r["result"]), (r, got)


def test_lattice_matches_closed_sum_and_bounds(engine):
    # European lattice vs the closed binomial sum on the same tree: the reference's
    # own results differ by ~1e-10 (results/results_binom_embar.csv:16-17).
    for N in (1000, 4096):
        lat = engine.binom_vanilla(100, 100, 0.05, 0.2, 1.0, N, "call").price
        emb = engine.binom_embar(100, 100, 0.05, 0.2, 1.0, N, "call", ref_exact=True).price
        assert abs(lat - emb) < 1e-9 * emb, (N, lat, emb)
    # N = 1e5 against the committed closed-sum fixture.
    rows = {(r["payoff"], int(r["N"])): float(r["result"])
            for r in csv.DictReader(open(os.path.join(HERE, "golden", "binom_large.csv")))}
    for pf in ("call", "put"):
        lat = engine.binom_vanilla(100, 100, 0.05, 0.2, 1.0, 100000, pf)
        want = rows[(pf, 100000)]
        say("lattice 1e5", pf, lat.price, want, "kernel_ms", lat.kernel_ms)
        assert abs(lat.price - want) < 1e-8 * want
    # American >= European; the call on a non-dividend stock is never exercised early.
    e = engine.binom_vanilla(100, 100, 0.06, 0.2, 1.0, 3000, "put").price
    a = engine.binom_vanilla(100, 100, 0.06, 0.2, 1.0, 3000, "put", american=True).price
    assert a > e
    ce = engine.binom_vanilla(100, 100, 0.06, 0.2, 1.0, 3000, "call").price
    ca = engine.binom_vanilla(100, 100, 0.06, 0.2, 1.0, 3000, "call", american=True).price
    assert abs(ca - ce) < 1e-12 * ce


def test_lattice_tables_follow_parameters(engine):
    # The resident u^i, d^k tables are keyed by (u, d, N): alternating calls must
    # not reuse stale tables.
    a1 = engine.binom_vanilla(100, 100, 0.06, 0.2, 1.0, 700, "put", american=True).price
    b1 = engine.binom_vanilla(100, 100, 0.06, 0.3, 1.0, 700, "put", american=True).price
    a2 = engine.binom_vanilla(100, 100, 0.06, 0.2, 1.0, 700, "put", american=True).price
    assert a1 == a2 and a1 != b1
    assert a1 == O.lib().orc_binom_vanilla(100, 100, 0.06, 0.2, 1.0, 700, -1.0, 1)


def test_lattice_errors(engine):
    with pytest.raises(PcfError):
        engine.binom_vanilla(100, 100, 0.06, 0.2, 1.0, 0, "put")
    with pytest.raises(PcfError):
        engine.binom_vanilla(100, 100, 0.06, 0.2, 1.0, (1 << 21) + 1, "put")
    with pytest.raises(PcfError):
        engine.binom_vanilla(100, 100, 0.06, 0.2, -1.0, 100, "put")


def test_lattice_repeated_requests_bitwise(engine):
    """Requests with the same N reuse the resident u^i, d^k tables and level
    rows; calls alternate styles and parameters, with other work and
    synchronisation between them, and each must still give the oracle's
    bits."""
    cases = [(100, 100, 0.06, 0.2, 1.0, "put", 1), (90, 100, 0.03, 0.35, 2.0, "call", 0),
             (120, 100, 0.06, 0.2, 0.5, "put", 1), (100, 110, 0.02, 0.75, 1.0, "call", 0),
             (100, 100, 0.06, 0.2, 1.0, "put", 1), (70, 100, 0.08, 0.5, 1.5, "put", 0)]
    for N in (700, 4099):
        for rep in range(2):
            for S0, E, r, sg, T, pf, am in cases:
                say("lattice repeat", N, rep, S0, E, pf, am)
                got = engine.binom_vanilla(S0, E, r, sg, T, N, pf, american=bool(am)).price
                want = O.lib().orc_binom_vanilla(S0, E, r, sg, T, N, 1.0 if pf == "call" else -1.0, am)
                assert got == want, (N, S0, E, pf, am, got, want)
                engine.mc_eur(100, 100, 0.05, 0.2, 1.0, 1 << 16, "call", seed=rep)  # other work between replays
                engine.synchronize()
