"""Python face of the engine, mirroring the reference kernel functions.

Each method keeps the reference function's name, argument order and meaning
(S0, E, r, sigma, T, N[, M][, assets, rho], payoff_fun) and returns a Result
whose .price is the reference function's return value:

  Engine.mc_eur        <- double mc_eur(...)        src/mc_eur.cpp:5-27
  Engine.binom_embar   <- double binom(...)         src/binom_embar.cpp:5-50
  Engine.mc_asia       <- double mc_asia(...)       src/mc_asia.cpp:5-40
  Engine.mc_amer       <- double mc_amer(...)       src/mc_amer.cpp:5-114
  Engine.mc_eur_multi  <- double mc_eur(...assets, rho)  src/mc_eur_multi.cpp:6-35

payoff_fun is "call"/"put" or +1/-1; anything else raises ValueError
("Unknown payoff function", as src/mc_eur.cpp:42 throws).  All arithmetic
runs in libpcf.so on the GPU; there is no CPU path here.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field

import numpy as np

from . import _native as N_

PRECISIONS = {"fp32k": N_.PCF_FP32K, "fp64": N_.PCF_FP64}


def payoff_code(payoff_fun) -> float:
    if isinstance(payoff_fun, str):
        if payoff_fun == "call":
            return 1.0
        if payoff_fun == "put":
            return -1.0
        raise ValueError("Unknown payoff function")
    v = float(payoff_fun)
    if v not in (1.0, -1.0):
        raise ValueError("Unknown payoff function")
    return v


@dataclass
class Result:
    price: float
    std_err: float
    mean: float
    variance: float
    control_price: float
    control_se: float
    n_paths: int
    n_steps: int
    kernel_ms: float
    total_ms: float
    flops: float
    bytes: float
    lsm_small_itm: int
    lsm_fallback: int
    grid_blocks: int
    block_size: int
    batch_size: int
    fp64_checked: int
    fp64_err: float
    extra: dict = field(default_factory=dict)

    @classmethod
    def from_c(cls, r: N_.PcfResult, **extra) -> "Result":
        d = r.as_dict()
        return cls(**d, extra=extra)


class Engine:
    """One device context (one HIP stream, resident buffers)."""

    def __init__(self, device: int = 0):
        self.lib = N_.load()
        h = C.c_void_p()
        st = self.lib.pcf_ctx_create(device, C.byref(h))
        if st != N_.PCF_OK:
            raise N_.PcfError(st, f"pcf_ctx_create(device={device}) failed")
        self.ctx = h

    # -- plumbing -----------------------------------------------------------
    def close(self):
        if getattr(self, "ctx", None):
            self.lib.pcf_ctx_destroy(self.ctx)
            self.ctx = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def last_error(self) -> str:
        return (self.lib.pcf_ctx_last_error(self.ctx) or b"").decode()

    def _check(self, st: int):
        if st != N_.PCF_OK:
            raise N_.PcfError(st, self.last_error())

    def device_info(self) -> dict:
        buf = C.create_string_buffer(256)
        cus = C.c_int()
        mem = C.c_int64()
        self._check(self.lib.pcf_device_info(self.ctx, buf, 256, C.byref(cus), C.byref(mem)))
        return {"name": buf.value.decode(), "compute_units": cus.value, "total_mem": mem.value}

    def synchronize(self):
        self._check(self.lib.pcf_ctx_synchronize(self.ctx))

    @staticmethod
    def make_opts(seed=0x5EED, precision="fp32k", normals=None, ref_exact=False, control=False,
                  block_size=0, items_per_thread=0, when_out=None, kernel_variant=0, lsm_basis=0,
                  fp64_check=False):
        o = N_.PcfOpts()
        N_.load().pcf_opts_init(C.byref(o))
        o.seed = int(seed) & 0xFFFFFFFFFFFFFFFF
        o.precision = PRECISIONS[precision] if isinstance(precision, str) else int(precision)
        keep = []
        if normals is not None:
            arr = np.ascontiguousarray(normals, dtype=np.float64)
            keep.append(arr)
            o.rng = N_.PCF_RNG_HOST
            o.host_normals = arr.ctypes.data_as(C.POINTER(C.c_double))
            o.n_host_normals = arr.size
        o.ref_exact = 1 if ref_exact else 0
        o.control = 1 if control else 0
        o.block_size = int(block_size)
        o.items_per_thread = int(items_per_thread)
        o.kernel_variant = int(kernel_variant)
        o.lsm_basis = int(lsm_basis)
        o.fp64_check = 1 if fp64_check else 0
        if when_out is not None:
            keep.append(when_out)
            o.when_out = when_out.ctypes.data_as(C.POINTER(C.c_int32))
        return o, keep

    # -- products ------------------------------------------------------------
    def mc_eur(self, S0, E, r, sigma, T, N, payoff_fun, **kw) -> Result:
        o, keep = self.make_opts(**kw)
        res = N_.PcfResult()
        st = self.lib.pcf_mc_eur(self.ctx, S0, E, r, sigma, T, int(N), payoff_code(payoff_fun),
                                 C.byref(o), C.byref(res))
        self._check(st)
        del keep
        return Result.from_c(res)

    def binom_embar(self, S0, E, r, sigma, T, N, payoff_fun, ref_exact=True, **kw) -> Result:
        o, keep = self.make_opts(ref_exact=ref_exact, **kw)
        res = N_.PcfResult()
        st = self.lib.pcf_binom_embar(self.ctx, S0, E, r, sigma, T, int(N), payoff_code(payoff_fun),
                                      C.byref(o), C.byref(res))
        self._check(st)
        return Result.from_c(res)

    def binom_vanilla(self, S0, E, r, sigma, T, N, payoff_fun, american=False, **kw) -> Result:
        """The reference's lattice programs (src/binom_vanilla_eur.cpp:5-39,
        src/binom_vanilla_amer.cpp:5-40): backward induction, FP64, bit-equal."""
        o, keep = self.make_opts(**kw)
        res = N_.PcfResult()
        st = self.lib.pcf_binom_vanilla(self.ctx, S0, E, r, sigma, T, int(N), payoff_code(payoff_fun),
                                        1 if american else 0, C.byref(o), C.byref(res))
        self._check(st)
        return Result.from_c(res)

    def mc_asia(self, S0, E, r, sigma, T, N, M, payoff_fun, **kw) -> Result:
        o, keep = self.make_opts(**kw)
        res = N_.PcfResult()
        st = self.lib.pcf_mc_asia(self.ctx, S0, E, r, sigma, T, int(N), int(M), payoff_code(payoff_fun),
                                  C.byref(o), C.byref(res))
        self._check(st)
        del keep
        return Result.from_c(res)

    def mc_amer(self, S0, E, r, sigma, T, N, M, payoff_fun, return_when=False, **kw) -> Result:
        when = np.zeros(int(N), dtype=np.int32) if return_when else None
        o, keep = self.make_opts(when_out=when, **kw)
        res = N_.PcfResult()
        st = self.lib.pcf_mc_amer(self.ctx, S0, E, r, sigma, T, int(N), int(M), payoff_code(payoff_fun),
                                  C.byref(o), C.byref(res))
        self._check(st)
        out = Result.from_c(res)
        if return_when:
            out.extra["when"] = when
        del keep
        return out

    def mc_eur_multi(self, S0, E, r, sigma, T, N, payoff_fun, assets, rho, sigmas=None, spots=None,
                     weights=None, corr=None, **kw) -> Result:
        o, keep = self.make_opts(**kw)
        ext = None
        arrs = []
        if any(v is not None for v in (sigmas, spots, weights, corr)):
            ext = N_.PcfBasketExt()
            for name, v in (("sigma", sigmas), ("S0", spots), ("weights", weights), ("corr", corr)):
                if v is not None:
                    a = np.ascontiguousarray(v, dtype=np.float64).ravel()
                    arrs.append(a)
                    setattr(ext, name, a.ctypes.data_as(C.POINTER(C.c_double)))
        res = N_.PcfResult()
        st = self.lib.pcf_mc_eur_multi(self.ctx, S0, E, r, sigma, T, int(N), payoff_code(payoff_fun),
                                       int(assets), rho, C.byref(ext) if ext is not None else None,
                                       C.byref(o), C.byref(res))
        self._check(st)
        del keep, arrs
        return Result.from_c(res)

    def normals(self, seed, stream, path0, n_paths, K, precision="fp64") -> np.ndarray:
        out = np.empty((int(n_paths), int(K)), dtype=np.float64)
        prec = PRECISIONS[precision] if isinstance(precision, str) else int(precision)
        st = self.lib.pcf_normals(self.ctx, int(seed), int(stream), int(path0), int(n_paths), int(K),
                                  prec, out.ctypes.data_as(C.POINTER(C.c_double)))
        self._check(st)
        return out


KINDS = {"eur": N_.PCF_KIND_EUR, "asia": N_.PCF_KIND_ASIA, "amer": N_.PCF_KIND_AMER,
         "basket": N_.PCF_KIND_BASKET, "binom": N_.PCF_KIND_BINOM}


def autotune(engine: "Engine", kinds=("eur", "asia", "basket", "amer"), max_log2_paths=22) -> list:
    """Runs the launch-shape autotuner (process-wide table) and returns its entries."""
    mask = 0
    for k in kinds:
        mask |= 1 << KINDS[k]
    out = (N_.PcfTuneEntry * 128)()
    n = C.c_int32()
    engine._check(engine.lib.pcf_autotune(engine.ctx, mask, int(max_log2_paths), out, 128, C.byref(n)))
    return [{f: getattr(out[i], f) for f, _ in N_.PcfTuneEntry._fields_} for i in range(n.value)]


def tune_reset():
    N_.load().pcf_tune_reset()


class Server:
    """Continuous-batching request loop (libpcf pcf_server_*)."""

    def __init__(self, engine: Engine, mem_budget_bytes: int = 0, max_inflight: int = 0, use_graphs: int = 1):
        self.lib = engine.lib
        cfg = N_.PcfServerConfig()
        cfg.mem_budget_bytes = int(mem_budget_bytes)
        cfg.max_inflight = int(max_inflight)
        cfg.use_graphs = int(use_graphs)
        h = C.c_void_p()
        st = self.lib.pcf_server_create(engine.ctx, C.byref(cfg), C.byref(h))
        if st != N_.PCF_OK:
            raise N_.PcfError(st, "pcf_server_create failed")
        self.srv = h
        self._keep = {}

    def last_error(self) -> str:
        return (self.lib.pcf_server_last_error(self.srv) or b"").decode()

    def submit(self, kind, S0, E, r, sigma, T, N, payoff_fun, M=0, assets=1, rho=0.0, sigmas=None,
               spots=None, weights=None, corr=None, **opts) -> int:
        q = N_.PcfRequest()
        q.kind = KINDS[kind] if isinstance(kind, str) else int(kind)
        q.M, q.assets = int(M), int(assets)
        q.S0, q.E, q.r, q.sigma, q.T = S0, E, r, sigma, T
        q.payoff_fun = payoff_code(payoff_fun)
        q.rho = rho
        q.N = int(N)
        o, keep = Engine.make_opts(**opts)
        q.opts = o
        if any(v is not None for v in (sigmas, spots, weights, corr)):
            ext = N_.PcfBasketExt()
            for name, v in (("sigma", sigmas), ("S0", spots), ("weights", weights), ("corr", corr)):
                if v is not None:
                    a = np.ascontiguousarray(v, dtype=np.float64).ravel()
                    keep.append(a)
                    setattr(ext, name, a.ctypes.data_as(C.POINTER(C.c_double)))
            keep.append(ext)
            q.basket = C.pointer(ext)
        t = C.c_uint64()
        st = self.lib.pcf_server_submit(self.srv, C.byref(q), C.byref(t))
        if st != N_.PCF_OK:
            raise N_.PcfError(st, self.last_error())
        self._keep[t.value] = keep
        return t.value

    def wait(self, ticket: int):
        res = N_.PcfResult()
        tm = N_.PcfRequestTiming()
        st = self.lib.pcf_server_wait(self.srv, int(ticket), C.byref(res), C.byref(tm))
        self._keep.pop(int(ticket), None)
        if st != N_.PCF_OK:
            raise N_.PcfError(st, self.last_error())
        timing = {f: getattr(tm, f) for f, _ in N_.PcfRequestTiming._fields_}
        return Result.from_c(res, timing=timing)

    def stats(self) -> dict:
        """Device-memory accounting of the server (pcf_server_get_stats)."""
        s = N_.PcfServerStats()
        st = self.lib.pcf_server_get_stats(self.srv, C.byref(s))
        if st != N_.PCF_OK:
            raise N_.PcfError(st, self.last_error())
        return s.as_dict()

    def close(self):
        if getattr(self, "srv", None):
            self.lib.pcf_server_destroy(self.srv)
            self.srv = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
