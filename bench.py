#!/usr/bin/env python3
"""bench.py -- the pricing hot path served on one MI355X.

Prints ONE JSON line on stdout (diagnostics go to stderr).

Workload (config.workload = "served mix"): a seeded request set mixing the
single-GPU configurations -- European call (S0=K=100, r=0.05, sigma=0.2),
basket call on 8 equicorrelated assets (rho=0.3, sigma_i ~ U[0.15,0.35]),
arithmetic Asian call (sigma=0.25, M=252) and American put by Longstaff-
Schwartz (r=0.06, sigma=0.2, M=50 exercise dates, path store on the device)
-- with path counts log-uniform in [2^14, 2^22].  One step is the whole set
served by the continuous-batching server (pcf_server_*); value = sum over
requests of paths x steps (1 for single-step
products, M otherwise) / step time, in the production FP32/Kahan form.  The
K timed steps run under sustained load: their K sets are submitted back to
back and every result is waited for, as a server sees a steady stream; the
closed-loop time (one set at a time, waiting for each) is reported beside
it.  The same steps in the plain FP64 form give the FP32-vs-FP64 ratio.  A second
phase feeds a Poisson stream at ~75% of the measured capacity and reports
p50/p99 latency overall and for 2^16-path requests.

Every configuration is also priced at its full size in FP32/Kahan and FP64
with the error in standard errors and the closed-form checks.

roofline: the dominant kernel of the served mix by device time (the kind
with the largest summed kernel time in the FP32/Kahan steps), measured at its
full configuration with HIP events recorded by the library on the stream it
launches on.  `achieved` is the algorithmic count per launch (bytes for the
American path store, flops otherwise) / the mean launch time; `traffic` is
the HBM bytes per launch from rocprofv3 FETCH_SIZE and WRITE_SIZE runs of the
same kernel, each scaled by what the counter reads for a copy of known size
(profiles/pmc_traffic.json, written from those runs).  Every kernel's
roofline object is also listed under "rooflines".

cpu_baseline: the reference's own OpenMP programs (src/mc_eur_omp.cpp,
src/mc_asia_omp.cpp, src/mc_amer_omp.cpp, src/binom_embar_omp.cpp built from
source at -O2 -fopenmp into oracle/_ref; the serial lattice programs) on this
host's cores (at most 16, the thread count and CPU model stated), each at its
configuration's own size, best of 3 T_calculation (binom_embar at N = 10^6
once, with --cpu-full).  value = the served mix's European/Asian/American
requests (paths x steps) over the time those per-kind rates give; the
reference basket program needs Eigen/Boost, absent here, so basket requests
are left out.  Each leg is also attached to its configuration's object in
per_config with the GPU/CPU ratio at equal size.
"""
import argparse
import json
import math
import os
import subprocess
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK = {"fp32": 157.3, "fp64": 78.6, "hbm": 8000.0}


def log(*a):
    print(*a, file=sys.stderr, flush=True)


METRIC = "paths·steps/s FP32-Kahan vs FP64 at ≤1 SE; p50/p99 latency of served mix"


def metric_name():
    return METRIC


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


# ---------------------------------------------------------------- the mix
def make_mix(seed, n, lo=14.0, hi=22.0):
    import numpy as np
    rng = np.random.default_rng(seed)
    kinds = ["eur", "basket", "asia", "amer"]
    sig8 = np.random.default_rng(8).uniform(0.15, 0.35, size=8)
    mix = []
    for i in range(n):
        k = kinds[int(rng.integers(0, 4))]
        N = max(2, int(round(2 ** rng.uniform(lo, hi))) // 2 * 2)
        s = int(rng.integers(1, 2**62))
        if k == "eur":
            mix.append(dict(kind="eur", args=(100.0, 100.0, 0.05, 0.2, 1.0, N, "call"), kw={}, steps=1, seed=s))
        elif k == "basket":
            mix.append(dict(kind="basket", args=(100.0, 100.0, 0.05, 0.2, 1.0, N, "call"),
                            kw=dict(assets=8, rho=0.3, sigmas=sig8), steps=1, seed=s))
        elif k == "asia":
            mix.append(dict(kind="asia", args=(100.0, 100.0, 0.05, 0.25, 1.0, N, "call"), kw=dict(M=252),
                            steps=252, seed=s))
        else:
            mix.append(dict(kind="amer", args=(100.0, 100.0, 0.06, 0.2, 1.0, N, "put"), kw=dict(M=50),
                            steps=50, seed=s))
    return mix


def path_steps(mix):
    return sum(q["args"][5] * q["steps"] for q in mix)


def serve_once(srv, mix, precision):
    """One closed-loop step: submit the set, wait for every result."""
    t0 = time.perf_counter()
    tickets = [srv.submit(q["kind"], *q["args"], seed=q["seed"], precision=precision, **q["kw"]) for q in mix]
    res = [srv.wait(t) for t in tickets]
    return time.perf_counter() - t0, res


def serve_steps(srv, mix, precision, K):
    """K steps under sustained load: the K sets are submitted back to back
    (the server keeps working while later sets arrive), then every result is
    waited for.  Returns the elapsed time and every set's results."""
    t0 = time.perf_counter()
    tickets = [[srv.submit(q["kind"], *q["args"], seed=q["seed"], precision=precision, **q["kw"]) for q in mix]
               for _ in range(K)]
    res = [[srv.wait(t) for t in ts] for ts in tickets]
    return time.perf_counter() - t0, res


def kernel_spread(mix, sets, kind="amer"):
    """Spread of each `kind` request's kernel time over the sustained sets:
    max/min per request (the worst and the median over requests), and the
    summed kernel time per set against the sum of the per-request minima."""
    import numpy as np
    idx = [i for i, q in enumerate(mix) if q["kind"] == kind]
    if not idx or len(sets) < 2:
        return None
    kt = np.array([[s[i].kernel_ms for i in idx] for s in sets])
    ratio = kt.max(axis=0) / kt.min(axis=0)
    return {"requests": len(idx), "sets": len(sets), "worst_max_over_min": float(ratio.max()),
            "median_max_over_min": float(np.median(ratio)), "kernel_ms_per_set": float(kt.sum(axis=1).mean()),
            "sum_of_min_kernel_ms": float(kt.min(axis=0).sum())}


def poisson_phase(srv, mix, rate, precision="fp32k"):
    """Open-loop arrivals at `rate` req/s from a submitter thread; returns latencies."""
    import threading
    import numpy as np
    rng = np.random.default_rng(123)
    gaps = rng.exponential(1.0 / rate, size=len(mix))
    tickets = [None] * len(mix)
    t_start = time.perf_counter()

    def submitter():
        t = time.perf_counter()
        for i, q in enumerate(mix):
            t += gaps[i]
            while True:
                d = t - time.perf_counter()
                if d <= 0:
                    break
                time.sleep(min(d, 0.002))
            tickets[i] = srv.submit(q["kind"], *q["args"], seed=q["seed"], precision=precision, **q["kw"])

    th = threading.Thread(target=submitter)
    th.start()
    res = [None] * len(mix)
    for i in range(len(mix)):
        while tickets[i] is None:
            time.sleep(0.0002)
        res[i] = srv.wait(tickets[i])
    th.join()
    return time.perf_counter() - t_start, res


def pct(v, p):
    import numpy as np
    return float(np.percentile(np.asarray(v), p)) if len(v) else None


# ---------------------------------------------------------------- per config
def per_config(eng):
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import numpy as np
    out = {}

    def pair(fn):
        fn("fp32k")
        fn("fp64")  # warm both
        r32 = fn("fp32k")
        r64 = fn("fp64")
        return r32, r64

    def summary(name, r32, r64, steps, check=None):
        d = {"price_fp32k": r32.price, "price_fp64": r64.price, "std_err": r64.std_err,
             "fp32k_minus_fp64_in_se": (r32.price - r64.price) / r64.std_err if r64.std_err else None,
             "kernel_ms_fp32k": r32.kernel_ms, "kernel_ms_fp64": r64.kernel_ms,
             "paths_steps_per_s_fp32k": r32.n_paths * steps / (r32.kernel_ms * 1e-3),
             "paths_steps_per_s_fp64": r64.n_paths * steps / (r64.kernel_ms * 1e-3),
             "fp32k_over_fp64": r64.kernel_ms / r32.kernel_ms}
        if check:
            d.update(check)
        out[name] = d
        log(name, json.dumps({k: (round(v, 6) if isinstance(v, float) else v) for k, v in d.items()}))

    bs = 10.450583572185565
    r32, r64 = pair(lambda p: eng.mc_eur(100, 100, 0.05, 0.2, 1.0, 10**6, "call", seed=1, precision=p))
    summary("mc_eur_1e6", r32, r64, 1, {"black_scholes": bs, "fp64_minus_bs_in_se": (r64.price - bs) / r64.std_err})
    # exact sums (mpmath, 40 digits) for the production form's accuracy
    import csv
    with open(os.path.join(ROOT, "tests", "golden", "binom_exact.csv")) as f:
        exact = {int(r["N"]): float(r["exact"]) for r in csv.DictReader(f)
                 if r["payoff"] == "call" and r["E"] == "100.0"}
    for N in (10**5, 10**6):
        eng.binom_embar(100, 100, 0.05, 0.2, 1.0, N, "call", ref_exact=True)
        eng.binom_embar(100, 100, 0.05, 0.2, 1.0, N, "call", ref_exact=False)
        rc = eng.binom_embar(100, 100, 0.05, 0.2, 1.0, N, "call", ref_exact=True)
        rd = eng.binom_embar(100, 100, 0.05, 0.2, 1.0, N, "call", ref_exact=False)
        d = {"price_chain": rc.price, "kernel_ms_chain": rc.kernel_ms,
             "price_dd": rd.price, "kernel_ms_dd": rd.kernel_ms,
             "log_adds_per_s_chain": rc.flops / (rc.kernel_ms * 1e-3)}
        if N in exact:
            d["exact"] = exact[N]
            d["dd_minus_exact_rel"] = (rd.price - exact[N]) / exact[N]
            d["chain_minus_exact_rel"] = (rc.price - exact[N]) / exact[N]
        out[f"binom_embar_N{N}"] = d
        log(f"binom_embar N={N}", d)
    for d in (8, 32):
        sig = np.random.default_rng(d).uniform(0.15, 0.35, size=d)
        r32, r64 = pair(lambda p: eng.mc_eur_multi(100, 100, 0.05, 0.2, 1.0, 1 << 24, "call", d, 0.3,
                                                   sigmas=sig, seed=2, precision=p))
        summary(f"basket_d{d}_2^24", r32, r64, 1, {"tflops_fp32k": r32.flops / (r32.kernel_ms * 1e-3) / 1e12})
    r32, r64 = pair(lambda p: eng.mc_asia(100, 100, 0.05, 0.25, 1.0, 1 << 22, 252, "call", seed=3, precision=p,
                                          control=True))
    geo = 6.551716578
    summary("mc_asia_2^22x252", r32, r64, 252,
            {"geometric_control_fp32k": r32.control_price, "geometric_closed_form": geo,
             "control_minus_closed_in_se": (r32.control_price - geo) / r32.control_se})
    lat = eng.binom_vanilla(100, 100, 0.06, 0.2, 1.0, 2000, "put", american=True)
    r32, r64 = pair(lambda p: eng.mc_amer(100, 100, 0.06, 0.2, 1.0, 1 << 20, 50, "put", seed=4, precision=p))
    summary("mc_amer_put_2^20x50", r32, r64, 50,
            {"european_put": 5.166002511, "european_control_fp32k": r32.control_price,
             "lattice_binom_vanilla_amer_N2000": lat.price,
             "fp64_minus_lattice_in_se": (r64.price - lat.price) / r64.std_err,
             "hbm_gbs_fp32k": r32.bytes / (r32.kernel_ms * 1e-3) / 1e9})
    # The reference's lattices on the GPU (src/binom_vanilla_{eur,amer}.cpp), at the
    # sizes its runscripts use for the comparison column (runscript_mc_eur.sh:22,
    # runscript_mc_amer.sh:23), each checked against the reference's own value.
    for name, N, pf, r_, am, want in (("lattice_eur_call_N1e5", 100000, "call", 0.05, False, None),
                                      ("lattice_eur_call_N1e4", 10000, "call", 0.05, False, None),
                                      ("lattice_amer_put_N1e4", 10000, "put", 0.06, True, 5.7989254457596333),
                                      ("lattice_amer_put_N2000", 2000, "put", 0.06, True, 5.798877611783797)):
        eng.binom_vanilla(100, 100, r_, 0.2, 1.0, N, pf, american=am)
        lr = eng.binom_vanilla(100, 100, r_, 0.2, 1.0, N, pf, american=am)
        d = {"price": lr.price, "kernel_ms": lr.kernel_ms, "launches": lr.grid_blocks,
             "node_updates_per_s": 0.5 * (N + 1) * (N + 2) / (lr.kernel_ms * 1e-3)}
        if want is not None:
            d["reference"] = want
            d["equal_to_reference"] = lr.price == want
        out[name] = d
        log(name, d)
    return out


def load_traffic():
    try:
        with open(os.path.join(ROOT, "profiles", "pmc_traffic.json")) as f:
            return json.load(f)
    except (OSError, ValueError):
        return {}


def traffic_of(kernel):
    t = load_traffic().get(kernel)
    return t["bytes_per_launch"] if t else None


def counters_of(kernel):
    """VALU/MFMA counter figures of a kernel from its committed rocprofv3 --pmc
    summary (profiles/pmc_valu.json), or None."""
    try:
        with open(os.path.join(ROOT, "profiles", "pmc_valu.json")) as f:
            c = json.load(f).get(kernel)
    except (OSError, ValueError):
        return None
    if c:
        c = dict(c, source="profiles/k_asia_2p22x252_pmc_valu.md" if kernel.startswith("k_asia")
                 else "profiles/k_basket_2p24_paths_pmc_valu.md")
    return c


def asia_roofline(eng, K=10):
    """k_asia<fp32k> at 2^22 x 252: algorithmic FP32 flops / HIP-event time."""
    N, M = 1 << 22, 252
    rs = [eng.mc_asia(100, 100, 0.05, 0.25, 1.0, N, M, "call", seed=100 + k) for k in range(K + 1)][1:]
    kms = sum(r.kernel_ms for r in rs) / K
    fl = rs[0].flops
    ach = fl / (kms * 1e-3) / 1e12
    return {"bound": "valu",
            "pipe": "VALU FP32 (non-packed FMA, 32x32->64 integer multiplies of Philox, transcendentals); no MFMA; "
                    "157.3 TFLOP/s is MI355X's FP32 vector peak",
            "kernel": "k_asia<fp32k> (mc_asia, 2^22 paths x 252 steps)", "achieved": ach,
            "peak": PEAK["fp32"], "unit": "TFLOP/s", "frac": ach / PEAK["fp32"], "traffic": traffic_of("k_asia<0>"),
            "kernel_ms": kms, "flops_per_launch": fl,
            "flops_per_unit": "14 per path-step (normal 6, I update 4, S update 4)",
            "counters": counters_of("k_asia<0, false>")}


def lsm_roofline(eng, K=10):
    """k_lsm_persistent<float> at 2^20 paths x 50 dates: algorithmic bytes / HIP-event time."""
    N, M = 1 << 20, 50
    rs = [eng.mc_amer(100, 100, 0.06, 0.2, 1.0, N, M, "put", seed=200 + k) for k in range(K + 1)][1:]
    kms = sum(r.kernel_ms for r in rs) / K
    by = rs[0].bytes
    ach = by / (kms * 1e-3) / 1e9
    traffic = traffic_of("k_lsm_persistent<float>")
    return {"bound": "latency",
            "limiter": "per date: each SIMD's issue of the decide, y-terms and x-moments passes over its four "
                       "waves, then the grid all-reduce (publish, wait for every block, gather, solve); "
                       "~5.5 us per date at 2^20 paths against ~0.5 us of row traffic",
            "kernel": f"k_lsm_persistent<float> (mc_amer put, 2^20 paths x 50 dates, grid "
                      f"{rs[0].grid_blocks}x{rs[0].block_size})",
            "achieved": ach, "peak": PEAK["hbm"], "unit": "GB/s", "frac": ach / PEAK["hbm"],
            "frac_of_hbm_from_counters": (traffic / (kms * 1e-3) / 1e9 / PEAK["hbm"]) if traffic else None,
            "traffic": traffic, "kernel_ms": kms, "bytes_per_launch": by,
            "bytes_per_unit": "24 per path-date in FP32: store write 4 + backward read 4 + per-path state "
                              "{cash flow 4, date 4} read and written (the state is register-resident here, "
                              "so traffic is below this count)"}


def basket_roofline(eng, K=10):
    """k_basket_mfma at d=32, 2^24 paths: algorithmic FP32 flops / HIP-event time."""
    import numpy as np
    sig = np.random.default_rng(32).uniform(0.15, 0.35, size=32)
    rs = [eng.mc_eur_multi(100, 100, 0.05, 0.2, 1.0, 1 << 24, "call", 32, 0.3, sigmas=sig, seed=300 + k)
          for k in range(K + 1)][1:]
    kms = sum(r.kernel_ms for r in rs) / K
    fl = rs[0].flops
    ach = fl / (kms * 1e-3) / 1e12
    return {"bound": "valu", "pipe": "exponents c0 + (c1 L) Z^T on MFMA (v_mfma_f32_32x32x2f32); the SIMD "
                                     "cycles split between VALU issue (Philox normals, per-asset exp) and the "
                                     "MFMA pipe, whose sum bounds the kernel",
            "kernel": "k_basket_mfma (mc_eur_multi d=32, 2^24 paths)", "achieved": ach,
            "peak": PEAK["fp32"], "unit": "TFLOP/s", "frac": ach / PEAK["fp32"],
            "traffic": traffic_of("k_basket_mfma"), "kernel_ms": kms, "flops_per_launch": fl,
            "flops_per_unit": "d^2 + 12 d + 3 per path (normals 6d, Z L^T d(d+1), asset 5d, payoff 3)",
            "counters": counters_of("k_basket_mfma")}


def profile_one(eng, which, K, W, slots=4):
    """One configuration at full size, W warm-up + K timed calls per precision."""
    import numpy as np
    sig8 = np.random.default_rng(8).uniform(0.15, 0.35, size=8)
    sig32 = np.random.default_rng(32).uniform(0.15, 0.35, size=32)
    calls = {
        "eur": lambda p, s: eng.mc_eur(100, 100, 0.05, 0.2, 1.0, 10**6, "call", seed=s, precision=p),
        "binom": lambda p, s: eng.binom_embar(100, 100, 0.05, 0.2, 1.0, 10**5, "call", ref_exact=(p == "fp64")),
        "basket": lambda p, s: [eng.mc_eur_multi(100, 100, 0.05, 0.2, 1.0, 1 << 24, "call", d, 0.3, sigmas=sg,
                                               seed=s, precision=p) for d, sg in ((8, sig8), (32, sig32))][-1],
        "asia": lambda p, s: eng.mc_asia(100, 100, 0.05, 0.25, 1.0, 1 << 22, 252, "call", seed=s, precision=p),
        "amer": lambda p, s: eng.mc_amer(100, 100, 0.06, 0.2, 1.0, 1 << 20, 50, "put", seed=s, precision=p),
        # 2^23 paths do not fit a co-resident grid: the per-date kernels
        "amer_graph": lambda p, s: eng.mc_amer(100, 100, 0.06, 0.2, 1.0, 1 << 23, 50, "put", seed=s, precision=p),
    }
    out = {"profile": which}
    if which == "lattice":
        # binom_vanilla_eur call N=1e5 and binom_vanilla_amer put N=1e4 (FP64, reference bits)
        legs = {"european_call_N1e5": lambda: eng.binom_vanilla(100, 100, 0.05, 0.2, 1.0, 10**5, "call"),
                "american_put_N1e4": lambda: eng.binom_vanilla(100, 100, 0.06, 0.2, 1.0, 10**4, "put",
                                                               american=True)}
        for name, f in legs.items():
            for _ in range(W):
                f()
            rs = [f() for _ in range(K)]
            out[name] = {"kernel_ms": [r.kernel_ms for r in rs], "launches": rs[0].grid_blocks}
        return out
    if which == "mix":
        from parcompfin_amd import Server
        mix = make_mix(2026, 64)
        with Server(eng, mem_budget_bytes=16 << 30, max_inflight=slots) as srv:
            for p in ("fp32k", "fp64"):
                for _ in range(W):
                    serve_once(srv, mix, p)
                ts = [serve_once(srv, mix, p)[0] for _ in range(K)]
                out[p] = {"ms_per_step": 1e3 * sum(ts) / K}
        return out
    for p in ("fp32k", "fp64"):
        for k in range(W):
            calls[which](p, 1 + k)
        rs = [calls[which](p, 100 + k) for k in range(K)]
        out[p] = {"kernel_ms": [r.kernel_ms for r in rs], "total_ms": [r.total_ms for r in rs]}
    return out


# ---------------------------------------------------------------- CPU baseline
def cpu_threads():
    """Threads for the reference's OpenMP programs: every core this process may
    run on, at most 16 (one GPU's share of the host on the box)."""
    try:
        n = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        n = os.cpu_count() or 1
    return max(1, min(16, n))


def _ref_row(ref, env, argv):
    exe = os.path.join(ref, argv[0])
    row = subprocess.run([exe] + [str(x) for x in argv[1:]], env=env, check=True, capture_output=True,
                         text=True, timeout=900).stdout.strip().split(",")
    return float(row[12]), float(row[13])  # T_calculation, Result


# (name, argv, work units, unit) -- each configuration at its own size.  The
# work unit is the one the GPU figure of the same configuration uses.
def cpu_legs(threads, full):
    legs = [
        ("mc_eur_1e6", ["mc_eur_omp", "call", 100, 100, 0.05, 0.2, 1.0, 10**6, threads], 10**6, "paths·steps"),
        ("mc_asia_2^22x252", ["mc_asia_omp", "call", 100, 100, 0.05, 0.25, 1.0, 1 << 22, 252, threads],
         (1 << 22) * 252, "paths·steps"),
        ("mc_amer_put_2^20x50", ["mc_amer_omp", "put", 100, 100, 0.06, 0.2, 1.0, 1 << 20, 50, threads],
         (1 << 20) * 50, "paths·steps"),
        ("binom_embar_N100000", ["binom_embar_omp", "call", 100, 100, 0.05, 0.2, 1.0, 100000, threads],
         0.75 * 1e10, "reference log-adds (0.75 N^2)"),
        ("lattice_amer_put_N1e4", ["binom_vanilla_amer", "put", 100, 100, 0.06, 0.2, 1.0, 10000],
         0.5 * 10001 * 10002, "node updates"),
        ("lattice_eur_call_N1e4", ["binom_vanilla_eur", "call", 100, 100, 0.05, 0.2, 1.0, 10000],
         0.5 * 10001 * 10002, "node updates"),
    ]
    if full:
        legs.append(("binom_embar_N1000000", ["binom_embar_omp", "call", 100, 100, 0.05, 0.2, 1.0, 10**6, threads],
                     0.75 * 1e12, "reference log-adds (0.75 N^2)"))
    return legs


def cpu_reference(full=False, repeats=3):
    """The reference's own programs (src/*_omp.cpp at -O2 -fopenmp, the serial
    lattices as shipped at -O0; built from source into oracle/_ref), each at its
    configuration's own size: best of `repeats` T_calculation (the O(N^2)
    binom_embar at N = 1e6, only with full=True, is timed once).  The basket
    program is not buildable here (Eigen and Boost are absent)."""
    threads = cpu_threads()
    ref = os.path.join(ROOT, "oracle", "_ref")
    env = dict(os.environ, OMP_PROC_BIND="close", OMP_NUM_THREADS=str(threads))
    out = {"threads": threads, "nproc": os.cpu_count(), "cpu_model": cpu_model(),
           "compiler": "g++ -O2 -fopenmp (the reference ships -O0, Makefile:9); lattices -O0 as shipped",
           "legs": {}}
    for name, argv, work, unit in cpu_legs(threads, full):
        if not os.path.exists(os.path.join(ref, argv[0])):
            out["legs"][name] = None
            continue
        reps = 1 if name == "binom_embar_N1000000" else repeats
        runs = [_ref_row(ref, env, argv) for _ in range(reps)]
        t = min(r[0] for r in runs)
        out["legs"][name] = {"program": argv[0], "argv": " ".join(str(x) for x in argv[1:]),
                             "seconds_best": t, "runs": reps, "result": runs[0][1], "work": work, "unit": unit,
                             "per_s": work / t, "threads": threads if argv[0].endswith("_omp") else 1}
        log("cpu", name, out["legs"][name])
    return out


def attach_cpu(configs, cpu):
    """Puts each CPU leg beside the GPU figure of the same configuration and
    size, with the GPU/CPU ratio in the leg's unit."""
    gpu_rate = {
        "mc_eur_1e6": ("paths_steps_per_s_fp32k", "paths_steps_per_s_fp64"),
        "mc_asia_2^22x252": ("paths_steps_per_s_fp32k", "paths_steps_per_s_fp64"),
        "mc_amer_put_2^20x50": ("paths_steps_per_s_fp32k", "paths_steps_per_s_fp64"),
        "binom_embar_N100000": ("log_adds_per_s_chain", None),
        "binom_embar_N1000000": ("log_adds_per_s_chain", None),
        "lattice_amer_put_N1e4": ("node_updates_per_s", None),
        "lattice_eur_call_N1e4": ("node_updates_per_s", None),
    }
    for name, leg in cpu["legs"].items():
        if not leg or name not in configs:
            continue
        c = configs[name]
        c["cpu_reference"] = leg
        k32, k64 = gpu_rate[name]
        c["gpu_over_cpu"] = {"gpu_fp32k" if k64 else "gpu_fp64": c[k32] / leg["per_s"]}
        if k64:
            c["gpu_over_cpu"]["gpu_fp64"] = c[k64] / leg["per_s"]


def cpu_baseline_line(cpu, mix):
    """The headline CPU figure in the metric's unit: the served mix's
    paths·steps over the time the reference's per-configuration rates imply
    for it (European, Asian, American requests; the basket needs Eigen/Boost)."""
    rate = {}
    for kind, name in (("eur", "mc_eur_1e6"), ("asia", "mc_asia_2^22x252"), ("amer", "mc_amer_put_2^20x50")):
        leg = cpu["legs"].get(name)
        if leg:
            rate[kind] = leg["per_s"]
    sample = [q for q in mix if q["kind"] in rate]
    if not sample:
        return {"value": None, "unit": "paths·steps/s", "cores": cpu["threads"], "kind": None,
                "sample": "oracle/_ref reference programs not built"}
    ps = path_steps(sample)
    t = sum(q["args"][5] * q["steps"] / rate[q["kind"]] for q in sample)
    total_s = sum(leg["seconds_best"] * leg["runs"] for leg in cpu["legs"].values() if leg)
    return {"value": ps / t, "unit": "paths·steps/s", "cores": cpu["threads"], "kind": "reference",
            "nproc": cpu["nproc"], "cpu_model": cpu["cpu_model"],
            "sample": f"the reference's mc_eur_omp (N=1e6), mc_asia_omp (N=2^22, M=252) and mc_amer_omp "
                      f"(N=2^20, M=50), each at its configuration's own size, best of 3 T_calculation, "
                      f"{cpu['threads']} threads (OMP_PROC_BIND=close) on '{cpu['cpu_model']}' "
                      f"(nproc {cpu['nproc']}); value = the {len(sample)} European/Asian/American requests "
                      f"of the served mix ({ps:.3g} paths·steps) over the time those rates give "
                      f"({t:.2f} s); basket requests excluded (the reference basket needs Eigen/Boost); "
                      f"{total_s:.1f} s of CPU runs in all",
            "legs": cpu["legs"]}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--requests", type=int, default=64)
    ap.add_argument("--slots", type=int, default=4)
    ap.add_argument("--latency-requests", type=int, default=400)
    ap.add_argument("--cpu-full", action="store_true",
                    help="also time the reference's binom_embar_omp at N = 1e6 once (O(N^2): minutes)")
    ap.add_argument("--cpu-only", action="store_true",
                    help="run only the CPU reference timings and print them as one JSON line")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-autotune", action="store_true")
    ap.add_argument("--no-configs", action="store_true")
    ap.add_argument("--profile", choices=["eur", "binom", "basket", "asia", "amer", "amer_graph", "mix", "lattice"],
                    help="run only one configuration --steps times in FP32/Kahan then FP64 "
                         "(for rocprofv3 kernel traces), print its timings and exit")
    args = ap.parse_args()
    if args.gpus != 1:
        log("bench.py measures one GPU; the multi-GPU sharding of the path is a design only, with no "
            "collective code. Use --gpus 1.")
        sys.exit(2)
    K, W = args.steps, args.warmup
    if args.cpu_only:
        print(json.dumps(cpu_reference(full=args.cpu_full)), flush=True)
        return

    from parcompfin_amd import Engine, Server, autotune
    eng = Engine(0)
    info = eng.device_info()
    log("device:", info)
    if args.profile:
        print(json.dumps(profile_one(eng, args.profile, K, W, args.slots)), flush=True)
        eng.close()
        return
    tuned = []
    if not args.no_autotune:
        t0 = time.perf_counter()
        tuned = autotune(eng, max_log2_paths=22)
        log(f"autotune: {len(tuned)} entries in {time.perf_counter() - t0:.1f} s")

    configs = {} if args.no_configs else per_config(eng)
    rooflines = {"asia": asia_roofline(eng), "amer": lsm_roofline(eng), "basket": basket_roofline(eng)}
    for k, v in rooflines.items():
        log(f"roofline {k}:", v)

    mix = make_mix(2026, args.requests)
    ps = path_steps(mix)
    srv = Server(eng, mem_budget_bytes=16 << 30, max_inflight=args.slots)
    results, sstats, closed, spread = {}, {}, {}, {}
    for prec in ("fp32k", "fp64"):
        for _ in range(W):
            serve_once(srv, mix, prec)
        # closed loop: one set at a time (reported beside the sustained figure)
        closed[prec] = 1e3 * sum(serve_once(srv, mix, prec)[0] for _ in range(K)) / K
        # sustained load: the timed K steps
        t, sets = serve_steps(srv, mix, prec, K)
        rs = sets[-1]
        results[prec] = ([t], rs)
        spread[prec] = kernel_spread(mix, sets)
        # cumulative server counters after this precision's steps
        sstats[prec] = {k: v for k, v in srv.stats().items()
                        if k in ("evictions", "underestimates", "peak_held_bytes", "peak_charged_bytes",
                                 "budget_bytes", "completed")}
        log(f"mix {prec}: {1e3 * t / K:.2f} ms/step sustained ({closed[prec]:.2f} closed loop), "
            f"{ps * K / t:.4g} paths·steps/s; American kernel spread {spread[prec]}")
    t32, r32 = results["fp32k"]
    t64, r64 = results["fp64"]
    value = ps * K / sum(t32)
    value64 = ps * K / sum(t64)
    dse = [abs(a.price - b.price) / b.std_err for a, b in zip(r32, r64) if b.std_err > 0]
    by_kind = {}
    for q, a in zip(mix, r32):
        k = by_kind.setdefault(q["kind"], {"requests": 0, "paths_steps": 0, "device_ms": 0.0})
        k["requests"] += 1
        k["paths_steps"] += q["args"][5] * q["steps"]
        k["device_ms"] += a.kernel_ms / max(a.batch_size, 1)

    dominant = max(by_kind, key=lambda k: by_kind[k]["device_ms"])
    roof = dict(rooflines.get(dominant, rooflines["amer"]), dominant_kind=dominant)
    log("roofline (dominant kernel of the mix):", roof)

    # Poisson latency phase at ~75% of capacity.
    req_rate = len(mix) * K / sum(t32)
    lat_mix = make_mix(77, args.latency_requests)
    for q in lat_mix[: len(lat_mix) // 8]:  # a share of exactly-2^16 requests
        q["args"] = q["args"][:5] + (1 << 16,) + q["args"][6:]
    rate = 0.75 * req_rate * (ps / len(mix)) / (path_steps(lat_mix) / len(lat_mix))
    wall, lres = poisson_phase(srv, lat_mix, rate)
    lat = [r.extra["timing"]["latency_ms"] for r in lres]
    lat16 = [r.extra["timing"]["latency_ms"] for q, r in zip(lat_mix, lres) if q["args"][5] == 1 << 16]
    lat_by = {}
    for q, r in zip(lat_mix, lres):
        lat_by.setdefault(q["kind"], []).append(r.extra["timing"]["latency_ms"])
    latency = {"arrival_rate_req_per_s": rate, "requests": len(lat_mix), "wall_s": wall,
               "p50_ms": pct(lat, 50), "p99_ms": pct(lat, 99),
               "p50_ms_2^16_paths": pct(lat16, 50), "p99_ms_2^16_paths": pct(lat16, 99), "n_2^16": len(lat16),
               "p99_ms_by_kind": {k: pct(v, 99) for k, v in lat_by.items()},
               "served_paths_steps_per_s": path_steps(lat_mix) / wall}
    log("latency:", latency)
    srv.close()

    cpu = None
    if not args.no_cpu:
        try:
            ref = cpu_reference(full=args.cpu_full)
            cpu = cpu_baseline_line(ref, mix)
            log("cpu baseline:", cpu)
            attach_cpu(configs, ref)
        except Exception as e:  # reported, never replaced by something else
            log("cpu baseline failed:", repr(e))
            cpu = {"value": None, "unit": "paths·steps/s", "cores": None, "kind": None, "sample": f"failed: {e!r}"}

    line = {
        "metric": metric_name(),
        "value": value,
        "unit": "paths·steps/s",
        "n_gpus": 1,
        "steps": K,
        "warmup": W,
        "ms_per_step": 1e3 * sum(t32) / K,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32",
        "data": "synthetic (seeded request mix; Philox4x32-10 normals generated on the device)",
        "config": {"workload": "served mix: European call, basket call d=8, Asian call M=252, American "
                               "put M=50, paths log-uniform in 2^14..2^22, FP32 paths + Kahan/FP64 sums",
                   "requests_per_step": len(mix), "paths_steps_per_step": ps, "slots": args.slots,
                   "parallelism": "single GPU"},
        "roofline": roof,
        "rooflines": rooflines,
        "cpu_baseline": cpu,
        "fp64": {"value": value64, "ms_per_step": 1e3 * sum(t64) / K, "fp32k_over_fp64": value / value64},
        "closed_loop_ms_per_step": closed,
        "accuracy": {"max_abs_fp32k_minus_fp64_in_se": max(dse) if dse else None,
                     "mean_abs_fp32k_minus_fp64_in_se": sum(dse) / len(dse) if dse else None},
        "latency": latency,
        "mix_by_kind": by_kind,
        "server_stats": sstats,
        "american_kernel_spread": spread,
        "per_config": configs,
        "autotune": tuned,
        "device": info["name"],
    }
    print(json.dumps(line), flush=True)
    eng.close()


if __name__ == "__main__":
    main()
