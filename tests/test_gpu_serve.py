"""Serving: continuous batching, memory admission, invariance to the mix.

The bar for "the same alone and in the mix" is bit equality: a batched
launch reduces each request over exactly the blocks its lone launch would
use, and the Philox stream is keyed by (seed, path, step).
"""
import math
import threading
import time

import numpy as np
import pytest

import _oracle as O
from parcompfin_amd import PcfError, Server, autotune, tune_reset

pytestmark = pytest.mark.gpu


def say(*a):
    print("[case]", *a, flush=True)


def mixed_requests(seed, n, lo=14, hi=17):
    rng = np.random.default_rng(seed)
    reqs = []
    for i in range(n):
        kind = ["eur", "basket8", "asia", "amer", "basket32"][i % 5]
        N = int(2 ** rng.uniform(lo, hi)) // 2 * 2
        s = int(rng.integers(1, 2**31))
        if kind == "eur":
            reqs.append(dict(kind="eur", args=(100, 100, 0.05, 0.2, 1.0, N, "call"), kw={}, seed=s))
        elif kind == "asia":
            reqs.append(dict(kind="asia", args=(100, 100, 0.05, 0.25, 1.0, N, "call"), kw=dict(M=252), seed=s))
        elif kind == "amer":
            reqs.append(dict(kind="amer", args=(100, 100, 0.06, 0.2, 1.0, N, "put"), kw=dict(M=50), seed=s))
        else:
            d = 8 if kind == "basket8" else 32
            reqs.append(dict(kind="basket", args=(100, 100, 0.05, 0.2, 1.0, N, "call"),
                             kw=dict(assets=d, rho=0.3), seed=s))
    return reqs


def price_alone(engine, q, precision="fp32k"):
    a, kw, s = q["args"], q["kw"], q["seed"]
    if q["kind"] == "eur":
        return engine.mc_eur(*a, seed=s, precision=precision)
    if q["kind"] == "asia":
        return engine.mc_asia(*a[:6], kw["M"], a[6], seed=s, precision=precision)
    if q["kind"] == "amer":
        return engine.mc_amer(*a[:6], kw["M"], a[6], seed=s, precision=precision)
    return engine.mc_eur_multi(*a, kw["assets"], kw["rho"], seed=s, precision=precision)


def test_results_identical_alone_and_in_the_mix(engine):
    reqs = mixed_requests(7, 60)
    say("alone", len(reqs))
    alone = [price_alone(engine, q) for q in reqs]
    say("served")
    with Server(engine, max_inflight=4) as srv:
        tickets = [srv.submit(q["kind"], *q["args"], seed=q["seed"], **q["kw"]) for q in reqs]
        served = [srv.wait(t) for t in tickets]
    for q, a, b in zip(reqs, alone, served):
        assert a.price == b.price, (q, a.price, b.price)
        assert a.std_err == b.std_err
    batched = [b.batch_size for b in served]
    say("batch sizes", batched)
    assert max(batched) > 1  # requests submitted together shared launches


def test_concurrent_submitters_and_poisson_arrivals(engine):
    """Requests arriving while others run are admitted and batched; every
    result still equals its lone price."""
    reqs = mixed_requests(11, 40, 14, 18)
    alone = [price_alone(engine, q) for q in reqs]
    rng = np.random.default_rng(3)
    gaps = rng.exponential(0.0005, size=len(reqs))
    results = [None] * len(reqs)
    with Server(engine, max_inflight=6) as srv:
        def client(idx):
            for i in idx:
                time.sleep(gaps[i])
                t = srv.submit(reqs[i]["kind"], *reqs[i]["args"], seed=reqs[i]["seed"], **reqs[i]["kw"])
                results[i] = srv.wait(t)
        threads = [threading.Thread(target=client, args=(list(range(k, len(reqs), 4)),)) for k in range(4)]
        say("4 client threads")
        for th in threads:
            th.start()
        for th in threads:
            th.join(timeout=120)
    for a, b in zip(alone, results):
        assert b is not None and a.price == b.price
        assert b.extra["timing"]["latency_ms"] > 0


def test_memory_budget_admission(engine):
    N, M = 1 << 16, 50
    # FP32 store 50 x 2^16 x 4 B = 13 MB per request; budget fits one at a time.
    with Server(engine, mem_budget_bytes=20 << 20, max_inflight=4) as srv:
        say("budget admission")
        tickets = [srv.submit("amer", 100, 100, 0.06, 0.2, 1.0, N, "put", M=M, seed=s) for s in range(6)]
        res = [srv.wait(t) for t in tickets]
        for s, r in zip(range(6), res):
            assert r.price == engine.mc_amer(100, 100, 0.06, 0.2, 1.0, N, M, "put", seed=s).price
        # overlapping is impossible here: each admission waited for the previous release
        spans = sorted((r.extra["timing"]["start_ms"], r.extra["timing"]["done_ms"]) for r in res)
        for (s0, e0), (s1, e1) in zip(spans, spans[1:]):
            assert s1 >= e0 - 1e-6
        with pytest.raises(PcfError) as ei:  # can never fit
            srv.submit("amer", 100, 100, 0.06, 0.2, 1.0, 1 << 20, "put", M=50)
        assert ei.value.status == 4


def test_bad_request_is_isolated(engine):
    with Server(engine, max_inflight=2) as srv:
        say("error isolation")
        good = srv.submit("eur", 100, 100, 0.05, 0.2, 1.0, 1 << 15, "call", seed=1)
        with pytest.raises(PcfError):
            srv.submit("eur", 100, 100, 0.05, -0.2, 1.0, 1 << 15, "call")
        with pytest.raises(PcfError):
            srv.submit("amer", 100, 100, 0.06, 0.2, 1.0, 1001, "put", M=50)
        with pytest.raises(PcfError):
            srv.submit("basket", 100, 100, 0.05, 0.2, 1.0, 1000, "call", assets=4, rho=-0.9)
        r = srv.wait(good)
        assert r.price == engine.mc_eur(100, 100, 0.05, 0.2, 1.0, 1 << 15, "call", seed=1).price


@pytest.mark.skip(reason="batched persistent American launches are switched off: the first GPU run of that "
                         "code passed every test, then the GPU was found faulted; the cause was not found, "
                         "so the code was taken out and American requests run one launch each")
def test_american_requests_share_persistent_launches(engine):
    """American requests with one block size run in one persistent launch, each
    on its own blocks, store and date barriers; the results are bitwise those of
    the lone launches.  Eight 2^18-path requests exceed one co-resident grid,
    so that unit is split over consecutive launches."""
    cases = []
    for s in range(6):  # 2^16 and 2^17 paths share a block size
        cases.append(((100, 100, 0.06, 0.2, 1.0, (1 << 16) + 2 * s * 1000, "put"), dict(M=50, seed=100 + s)))
        cases.append(((100, 100, 0.06, 0.2, 1.0, 1 << 17, "put"), dict(M=30, seed=200 + s)))
    for s in range(4):
        cases.append(((100, 110, 0.05, 0.3, 1.0, 1 << 16, "put"), dict(M=50, seed=300 + s, precision="fp64")))
    for s in range(8):
        cases.append(((100, 100, 0.06, 0.2, 1.0, 1 << 18, "put"), dict(M=20, seed=400 + s)))
    say("american batch", len(cases))
    alone = [engine.mc_amer(*a[:6], kw["M"], a[6], **{k: v for k, v in kw.items() if k != "M"}) for a, kw in cases]
    with Server(engine, max_inflight=1) as srv:
        tickets = [srv.submit("amer", *a, **kw) for a, kw in cases]
        served = [srv.wait(t) for t in tickets]
    for (a, kw), x, y in zip(cases, alone, served):
        assert x.price == y.price and x.std_err == y.std_err, (a, kw, x.price, y.price)
        assert x.control_price == y.control_price
        assert (x.lsm_small_itm, x.lsm_fallback) == (y.lsm_small_itm, y.lsm_fallback)
        assert (x.grid_blocks, x.block_size) == (y.grid_blocks, y.block_size)
    sizes = [y.batch_size for y in served]
    say("american batch sizes", sizes)
    assert max(sizes) >= 4


def test_served_american_basis_requests(engine):
    """K-term regressions are served like any American request: same result as
    the direct call, alongside quadratic ones."""
    cases = [(K, N, s) for K, N, s in ((4, 1 << 16, 3), (0, 1 << 16, 3), (5, 1 << 15, 4), (2, 1 << 17, 5))]
    alone = [engine.mc_amer(100, 100, 0.06, 0.2, 1.0, N, 50, "put", seed=s, lsm_basis=K) for K, N, s in cases]
    with Server(engine, max_inflight=3) as srv:
        say("served basis", [c[0] for c in cases])
        tickets = [srv.submit("amer", 100, 100, 0.06, 0.2, 1.0, N, "put", M=50, seed=s, lsm_basis=K)
                   for K, N, s in cases]
        served = [srv.wait(t) for t in tickets]
    for c, a, b in zip(cases, alone, served):
        assert a.price == b.price and a.std_err == b.std_err, (c, a.price, b.price)
    assert alone[0].price != alone[1].price  # the 4-term fit is a different estimator


def test_served_fp32_within_se_of_fp64(engine):
    reqs = mixed_requests(5, 20, 16, 18)
    with Server(engine, max_inflight=4) as srv:
        say("served fp32 vs alone fp64")
        t32 = [srv.submit(q["kind"], *q["args"], seed=q["seed"], **q["kw"]) for q in reqs]
        r32 = [srv.wait(t) for t in t32]
    for q, a in zip(reqs, r32):
        b = price_alone(engine, q, precision="fp64")
        assert abs(a.price - b.price) < b.std_err, (q["kind"], a.price, b.price, b.std_err)


def test_autotune_then_invariance(engine):
    say("autotune")
    try:
        entries = autotune(engine, max_log2_paths=18)
        assert entries and all(e["best_ms"] > 0 for e in entries)
        for e in entries:
            assert e["best_ms"] <= e["default_ms"] * 1.0001
        reqs = mixed_requests(9, 25)
        alone = [price_alone(engine, q) for q in reqs]
        with Server(engine, max_inflight=4) as srv:
            tickets = [srv.submit(q["kind"], *q["args"], seed=q["seed"], **q["kw"]) for q in reqs]
            served = [srv.wait(t) for t in tickets]
        for a, b in zip(alone, served):
            assert a.price == b.price
        # Tuned shapes move only the reduction tree: last-bit differences, except
        # that an American exercise decision within rounding of its regression
        # value may flip (bounded by a small fraction of a standard error).
        tune_reset()
        for q, a in zip(reqs, alone):
            b = price_alone(engine, q)
            if q["kind"] == "amer":
                assert abs(a.price - b.price) <= 0.05 * b.std_err
            else:
                assert abs(a.price - b.price) <= 1e-12 * abs(b.price) + 1e-15
    finally:
        tune_reset()


def test_binom_served(engine):
    with Server(engine, max_inflight=2) as srv:
        say("binom served")
        t = srv.submit("binom", 100, 100, 0.05, 0.2, 1.0, 10000, "call", ref_exact=True)
        r = srv.wait(t)
    want = O.lib().orc_binom_embar(100, 100, 0.05, 0.2, 1.0, 10000, 1.0, 1)
    assert abs(r.price - want) / want < 1e-12


def test_kept_path_stores_bounded_by_the_budget(engine):
    """Slots keep their American path stores between requests only while the
    stores kept by all slots fit the budget; past it a slot frees its store.
    Results do not depend on which happened: each equals its lone call bit
    for bit, in two rounds of the same requests."""
    # An FP64 2^16 x 50 store is 2^16 * 50 * 8 B = 26 MB: with 64 MB and 4 slots
    # at most two stay kept, so later requests free theirs.
    N, M = 1 << 16, 50
    seeds = [11, 12, 13, 14, 15, 16, 17, 18]
    want = {s: engine.mc_amer(100, 100, 0.06, 0.2, 1.0, N, M, "put", seed=s, precision="fp64").price
            for s in seeds}
    say("kept path stores, 64 MB budget, 4 slots")
    with Server(engine, mem_budget_bytes=64 << 20, max_inflight=4) as srv:
        for _ in range(2):
            ts = {s: srv.submit("amer", 100, 100, 0.06, 0.2, 1.0, N, "put", M=M, seed=s, precision="fp64")
                  for s in seeds}
            for s, t in ts.items():
                assert srv.wait(t).price == want[s], s
