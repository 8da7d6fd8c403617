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


def test_concurrent_persistent_launches_match_lone_calls(engine):
    """Persistent LSM launches from several host threads (one context each),
    of 64-, 128- and 256-block grids, hand over to the device's persistent
    stream in whatever order the threads reach it, each after its own inputs;
    every result equals its lone call bit for bit, twice over."""
    from parcompfin_amd import Engine
    cases = []
    for s in range(6):
        cases.append((1 << 14, 50, 300 + s, "fp32k"))   # 64 blocks
        cases.append((1 << 16, 50, 400 + s, "fp64"))    # 128 blocks
    cases += [(1 << 20, 50, 500, "fp32k"), (1 << 18, 30, 501, "fp64"), (3 << 15, 20, 502, "fp32k")]
    want = [engine.mc_amer(100, 100, 0.06, 0.2, 1.0, N, M, "put", seed=s, precision=p) for N, M, s, p in cases]
    got = [None] * len(cases)
    engines = [Engine(0) for _ in range(4)]

    def work(k):
        for i in range(k, len(cases), 4):
            N, M, s, p = cases[i]
            got[i] = engines[k].mc_amer(100, 100, 0.06, 0.2, 1.0, N, M, "put", seed=s, precision=p)

    say("concurrent persistent launches, 4 threads")
    for _ in range(2):
        threads = [threading.Thread(target=work, args=(k,)) for k in range(4)]
        for th in threads:
            th.start()
        for th in threads:
            th.join(timeout=120)
        for c, a, b in zip(cases, want, got):
            assert b is not None and a.price == b.price and a.std_err == b.std_err, (c, a.price, b and b.price)
            assert a.control_price == b.control_price and a.grid_blocks == b.grid_blocks
    for e in engines:
        e.close()


def test_persistent_launches_ordered_against_heavy_kernels(engine):
    """Persistent LSM launches wait for the kernels other contexts had
    enqueued, and heavy Asian/basket launches wait for the last persistent
    launch.  Threads on four contexts interleave both kinds (FP32 and FP64)
    twice over: every call completes and equals its lone call bit for bit."""
    from parcompfin_amd import Engine
    amer = [(1 << 18, 50, 600 + s, "fp32k" if s % 2 else "fp64") for s in range(4)] + [(1 << 20, 50, 610, "fp32k")]
    heavy = [("asia", 1 << 20, 700 + s, "fp32k" if s % 2 else "fp64") for s in range(3)]
    heavy += [("basket", 1 << 21, 710 + s, "fp32k" if s % 2 else "fp64") for s in range(2)]
    heavy += [("asia", 1 << 12, 720, "fp32k")]  # below the ordering threshold

    def run_heavy(e, c):
        kind, N, s, p = c
        if kind == "asia":
            return e.mc_asia(100, 100, 0.05, 0.25, 1.0, N, 252, "call", seed=s, precision=p)
        return e.mc_eur_multi(100, 100, 0.05, 0.2, 1.0, N, "call", 8, 0.3, seed=s, precision=p)

    want_a = [engine.mc_amer(100, 100, 0.06, 0.2, 1.0, N, M, "put", seed=s, precision=p) for N, M, s, p in amer]
    want_h = [run_heavy(engine, c) for c in heavy]
    engines = [Engine(0) for _ in range(4)]
    got_a, got_h = [None] * len(amer), [None] * len(heavy)

    def work(k):
        e = engines[k]
        if k < 2:
            for i in range(k, len(amer), 2):
                N, M, s, p = amer[i]
                got_a[i] = e.mc_amer(100, 100, 0.06, 0.2, 1.0, N, M, "put", seed=s, precision=p)
        else:
            for i in range(k - 2, len(heavy), 2):
                got_h[i] = run_heavy(e, heavy[i])

    say("persistent launches beside heavy kernels, 4 contexts")
    for _ in range(2):
        threads = [threading.Thread(target=work, args=(k,)) for k in range(4)]
        for th in threads:
            th.start()
        for th in threads:
            th.join(timeout=120)
        for c, a, b in zip(amer, want_a, got_a):
            assert b is not None and a.price == b.price and a.std_err == b.std_err, (c, a.price, b and b.price)
        for c, a, b in zip(heavy, want_h, got_h):
            assert b is not None and a.price == b.price and a.std_err == b.std_err, (c, a.price, b and b.price)
        say("  American kernel ms, lone vs beside:",
            [(round(a.kernel_ms, 3), round(b.kernel_ms, 3)) for a, b in zip(want_a, got_a)])
        got_a, got_h = [None] * len(amer), [None] * len(heavy)
    for e in engines:
        e.close()


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


def test_fp64_error_per_result(engine):
    """opts.fp64_check: every FP32/Kahan result carries |FP32/Kahan - FP64| on
    the same Philox paths, direct and served; it equals the difference of the
    two lone calls exactly, and it is a small fraction of a standard error."""
    reqs = mixed_requests(17, 10, 14, 17)
    say("fp64 check direct")
    for q in reqs:
        a = price_alone_kw(engine, q, fp64_check=True)
        b = price_alone(engine, q, precision="fp64")
        assert a.fp64_checked == 1 and a.fp64_err == abs(a.price - b.price), q["kind"]
        assert a.fp64_err < 0.05 * b.std_err
        c = price_alone(engine, q)
        assert c.fp64_checked == 0 and c.price == a.price
    say("fp64 check served")
    with Server(engine, max_inflight=3) as srv:
        ts = [srv.submit(q["kind"], *q["args"], seed=q["seed"], fp64_check=True, **q["kw"]) for q in reqs]
        served = [srv.wait(t) for t in ts]
    for q, s in zip(reqs, served):
        b = price_alone(engine, q, precision="fp64")
        assert s.fp64_checked == 1 and s.fp64_err == abs(s.price - b.price), q["kind"]


def price_alone_kw(engine, q, **extra):
    a, kw, s = q["args"], q["kw"], q["seed"]
    if q["kind"] == "eur":
        return engine.mc_eur(*a, seed=s, **extra)
    if q["kind"] == "asia":
        return engine.mc_asia(*a[:6], kw["M"], a[6], seed=s, **extra)
    if q["kind"] == "amer":
        return engine.mc_amer(*a[:6], kw["M"], a[6], seed=s, **extra)
    return engine.mc_eur_multi(*a, kw["assets"], kw["rho"], seed=s, **extra)


def test_binom_served(engine):
    with Server(engine, max_inflight=2) as srv:
        say("binom served")
        t = srv.submit("binom", 100, 100, 0.05, 0.2, 1.0, 10000, "call", ref_exact=True)
        r = srv.wait(t)
    want = O.lib().orc_binom_embar(100, 100, 0.05, 0.2, 1.0, 10000, 1.0, 1)
    assert abs(r.price - want) / want < 1e-12


def test_held_device_memory_never_exceeds_the_budget(engine):
    """Kept path stores and in-flight requests share one budget: with FP64
    2^22 x 50 American stores (1.7 GB each) and a 4 GiB budget on 4 slots, at
    most two stores exist at a time; a slot reuses the store it keeps; a
    request with a 3.4 GB store (M = 100) then needs the others' kept stores
    freed.  The charged bytes never pass the budget, no unit holds more than it
    was charged, and every result equals its lone call bit for bit."""
    N, M = 1 << 22, 50
    seeds = [21, 22, 23, 24, 25]
    want = {s: engine.mc_amer(100, 100, 0.06, 0.2, 1.0, N, M, "put", seed=s, precision="fp64").price
            for s in seeds}
    want_big = engine.mc_amer(100, 100, 0.06, 0.2, 1.0, N, 100, "put", seed=26, precision="fp64").price
    small = [engine.mc_eur(100, 100, 0.05, 0.2, 1.0, 1 << 20, "call", seed=s).price for s in seeds]
    budget = 4 << 30
    say("held memory, 4 GiB budget, 4 slots, FP64 2^22 stores")
    with Server(engine, mem_budget_bytes=budget, max_inflight=4) as srv:
        for rnd in range(2):
            ts = {s: srv.submit("amer", 100, 100, 0.06, 0.2, 1.0, N, "put", M=M, seed=s, precision="fp64")
                  for s in seeds}
            te = [srv.submit("eur", 100, 100, 0.05, 0.2, 1.0, 1 << 20, "call", seed=s) for s in seeds]
            for s, t in ts.items():
                assert srv.wait(t).price == want[s], (rnd, s)
            assert [srv.wait(t).price for t in te] == small
        tb = srv.submit("amer", 100, 100, 0.06, 0.2, 1.0, N, "put", M=100, seed=26, precision="fp64")
        assert srv.wait(tb).price == want_big
        st = srv.stats()
    say("server stats", st)
    assert st["budget_bytes"] == budget
    assert st["peak_charged_bytes"] <= budget
    assert st["peak_held_bytes"] <= budget
    assert st["underestimates"] == 0
    assert st["completed"] == 4 * len(seeds) + 1
    # the M = 100 store and a kept 2^22 x 50 store do not fit beside a second
    # kept one: kept stores were freed while the server ran
    assert st["evictions"] >= 1
    assert st["held_bytes"] <= budget


def test_binomial_tables_are_charged(engine):
    """Binomial requests hold resident tables (log table and chain prefix, or
    the double-double log k! table): they are charged, so a budget just above
    one request's tables serves them one at a time and frees what is kept."""
    N = 1 << 20
    want_chain = engine.binom_embar(100, 100, 0.05, 0.2, 1.0, N, "call", ref_exact=True).price
    want_dd = engine.binom_embar(100, 100, 0.05, 0.2, 1.0, N, "call", ref_exact=False).price
    budget = 2 * (N + 1) * 8 + (4 << 20)
    say("binomial under a tight budget")
    with Server(engine, mem_budget_bytes=budget, max_inflight=3) as srv:
        ts = [srv.submit("binom", 100, 100, 0.05, 0.2, 1.0, N, "call", ref_exact=(i % 2 == 0)) for i in range(6)]
        got = [srv.wait(t).price for t in ts]
        st = srv.stats()
    assert got == [want_chain if i % 2 == 0 else want_dd for i in range(6)]
    assert st["peak_charged_bytes"] <= budget and st["underestimates"] == 0
    assert st["evictions"] >= 1


@pytest.mark.parametrize("precision", ["fp32k", "fp64"])
def test_full_size_range_served_mix_identical_alone(engine, precision):
    """The served mix over the whole request range, 2^14..2^22 paths, in both
    precisions, with a budget small enough that FP64 2^22 path stores are
    freed and re-made: every result equals its lone call bit for bit."""
    reqs = mixed_requests(13, 40, 14, 22)
    for q in reqs:
        q["kw"] = dict(q["kw"], precision=precision)
    say("full-range mix alone", precision, [q["args"][5] for q in reqs])
    alone = [price_alone(engine, q, precision=precision) for q in reqs]
    budget = 3 << 30
    with Server(engine, mem_budget_bytes=budget, max_inflight=6) as srv:
        say("full-range mix served", precision)
        tickets = [srv.submit(q["kind"], *q["args"], seed=q["seed"], **q["kw"]) for q in reqs]
        served = [srv.wait(t) for t in tickets]
        st = srv.stats()
    for q, a, b in zip(reqs, alone, served):
        assert a.price == b.price and a.std_err == b.std_err, (q["kind"], q["args"][5], a.price, b.price)
    assert max(q["args"][5] for q in reqs) > (1 << 21)
    assert st["peak_charged_bytes"] <= budget and st["underestimates"] == 0
