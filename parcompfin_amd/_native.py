"""ctypes mirror of include/pcf.h (the libpcf C-ABI).

This is the binding a Python caller of the engine uses; it adds nothing to
the arithmetic.  The library is loaded from the package tree
(parcompfin_amd/lib/libpcf.so, built in place by __graft_entry__.build());
there is no fallback: if it is missing or a call fails, an exception says so.
"""
from __future__ import annotations

import ctypes as C
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "lib", "libpcf.so")

# ---- enums (include/pcf.h) ------------------------------------------------
PCF_OK = 0
PCF_ERR_INVALID_ARG = 1
PCF_ERR_NO_DEVICE = 2
PCF_ERR_HIP = 3
PCF_ERR_OUT_OF_MEMORY = 4
PCF_ERR_SINGULAR = 5
PCF_ERR_NOT_FOUND = 6
PCF_ERR_SHUTDOWN = 7
PCF_ERR_TIMEOUT = 8

PCF_FP32K = 0
PCF_FP64 = 1
PCF_RNG_PHILOX = 0
PCF_RNG_HOST = 1
PCF_STREAM_EUR = 1
PCF_STREAM_ASIA = 2
PCF_STREAM_AMER = 3
PCF_STREAM_BASKET = 4

PCF_KIND_EUR = 0
PCF_KIND_ASIA = 1
PCF_KIND_AMER = 2
PCF_KIND_BASKET = 3
PCF_KIND_BINOM = 4

_STATUS_NAMES = {
    0: "PCF_OK", 1: "PCF_ERR_INVALID_ARG", 2: "PCF_ERR_NO_DEVICE", 3: "PCF_ERR_HIP",
    4: "PCF_ERR_OUT_OF_MEMORY", 5: "PCF_ERR_SINGULAR", 6: "PCF_ERR_NOT_FOUND",
    7: "PCF_ERR_SHUTDOWN", 8: "PCF_ERR_TIMEOUT",
}


class PcfOpts(C.Structure):
    _fields_ = [
        ("seed", C.c_uint64),
        ("precision", C.c_int32),
        ("rng", C.c_int32),
        ("host_normals", C.POINTER(C.c_double)),
        ("n_host_normals", C.c_int64),
        ("ref_exact", C.c_int32),
        ("control", C.c_int32),
        ("block_size", C.c_int32),
        ("items_per_thread", C.c_int32),
        ("when_out", C.POINTER(C.c_int32)),
        ("kernel_variant", C.c_int32),
        ("lsm_basis", C.c_int32),
        ("fp64_check", C.c_int32),
        ("reserved", C.c_int32 * 3),
    ]


class PcfResult(C.Structure):
    _fields_ = [
        ("price", C.c_double),
        ("std_err", C.c_double),
        ("mean", C.c_double),
        ("variance", C.c_double),
        ("control_price", C.c_double),
        ("control_se", C.c_double),
        ("n_paths", C.c_int64),
        ("n_steps", C.c_int64),
        ("kernel_ms", C.c_double),
        ("total_ms", C.c_double),
        ("flops", C.c_double),
        ("bytes", C.c_double),
        ("lsm_small_itm", C.c_int32),
        ("lsm_fallback", C.c_int32),
        ("grid_blocks", C.c_int32),
        ("block_size", C.c_int32),
        ("batch_size", C.c_int32),
        ("fp64_checked", C.c_int32),
        ("fp64_err", C.c_double),
    ]

    def as_dict(self) -> dict:
        return {k: getattr(self, k) for k, _ in self._fields_ if k != "reserved"}


class PcfBasketExt(C.Structure):
    _fields_ = [
        ("sigma", C.POINTER(C.c_double)),
        ("S0", C.POINTER(C.c_double)),
        ("weights", C.POINTER(C.c_double)),
        ("corr", C.POINTER(C.c_double)),
    ]


class PcfRequest(C.Structure):
    _fields_ = [
        ("kind", C.c_int32),
        ("M", C.c_int32),
        ("assets", C.c_int32),
        ("pad", C.c_int32),
        ("S0", C.c_double),
        ("E", C.c_double),
        ("r", C.c_double),
        ("sigma", C.c_double),
        ("T", C.c_double),
        ("payoff_fun", C.c_double),
        ("rho", C.c_double),
        ("N", C.c_int64),
        ("opts", PcfOpts),
        ("basket", C.POINTER(PcfBasketExt)),
    ]


class PcfTuneEntry(C.Structure):
    _fields_ = [
        ("kind", C.c_int32),
        ("mode", C.c_int32),
        ("size_class", C.c_int32),
        ("block_size", C.c_int32),
        ("items_per_thread", C.c_int32),
        ("candidates", C.c_int32),
        ("best_ms", C.c_double),
        ("default_ms", C.c_double),
    ]


class PcfServerConfig(C.Structure):
    _fields_ = [
        ("mem_budget_bytes", C.c_int64),
        ("max_inflight", C.c_int32),
        ("use_graphs", C.c_int32),
        ("reserved", C.c_int32 * 8),
    ]


class PcfRequestTiming(C.Structure):
    _fields_ = [
        ("submit_ms", C.c_double),
        ("start_ms", C.c_double),
        ("done_ms", C.c_double),
        ("latency_ms", C.c_double),
    ]


class PcfServerStats(C.Structure):
    _fields_ = [
        ("budget_bytes", C.c_int64),
        ("charged_bytes", C.c_int64),
        ("peak_charged_bytes", C.c_int64),
        ("held_bytes", C.c_int64),
        ("peak_held_bytes", C.c_int64),
        ("fixed_bytes_per_slot", C.c_int64),
        ("evictions", C.c_int64),
        ("underestimates", C.c_int64),
        ("completed", C.c_int64),
        ("slots", C.c_int32),
        ("reserved", C.c_int32 * 5),
    ]

    def as_dict(self) -> dict:
        return {n: getattr(self, n) for n, _ in self._fields_ if n != "reserved"}


_d = C.c_double
_i64 = C.c_int64
_i32 = C.c_int32
_ctx = C.c_void_p

# name -> (restype, argtypes); every symbol include/pcf.h declares.
SIGNATURES = {
    "pcf_ctx_create": (C.c_int, [C.c_int, C.POINTER(C.c_void_p)]),
    "pcf_ctx_destroy": (C.c_int, [_ctx]),
    "pcf_ctx_last_error": (C.c_char_p, [_ctx]),
    "pcf_ctx_synchronize": (C.c_int, [_ctx]),
    "pcf_device_info": (C.c_int, [_ctx, C.c_char_p, C.c_int, C.POINTER(C.c_int), C.POINTER(_i64)]),
    "pcf_opts_init": (None, [C.POINTER(PcfOpts)]),
    "pcf_abi_version": (C.c_int, []),
    "pcf_struct_size": (C.c_int, [C.c_char_p]),
    "pcf_mc_eur": (C.c_int, [_ctx, _d, _d, _d, _d, _d, _i64, _d, C.POINTER(PcfOpts), C.POINTER(PcfResult)]),
    "pcf_binom_embar": (C.c_int, [_ctx, _d, _d, _d, _d, _d, _i64, _d, C.POINTER(PcfOpts), C.POINTER(PcfResult)]),
    "pcf_binom_vanilla": (C.c_int, [_ctx, _d, _d, _d, _d, _d, _i64, _d, C.c_int32, C.POINTER(PcfOpts),
                                    C.POINTER(PcfResult)]),
    "pcf_mc_asia": (C.c_int, [_ctx, _d, _d, _d, _d, _d, _i64, _i32, _d, C.POINTER(PcfOpts), C.POINTER(PcfResult)]),
    "pcf_mc_asia_noalloc": (C.c_int, [_ctx, _d, _d, _d, _d, _d, _i64, _i32, _d, C.POINTER(PcfOpts),
                                      C.POINTER(PcfResult)]),
    "pcf_mc_amer": (C.c_int, [_ctx, _d, _d, _d, _d, _d, _i64, _i32, _d, C.POINTER(PcfOpts), C.POINTER(PcfResult)]),
    "pcf_mc_eur_multi": (C.c_int, [_ctx, _d, _d, _d, _d, _d, _i64, _d, _i32, _d, C.POINTER(PcfBasketExt),
                                   C.POINTER(PcfOpts), C.POINTER(PcfResult)]),
    "pcf_normals": (C.c_int, [_ctx, C.c_uint64, C.c_uint32, C.c_uint64, _i64, _i32, _i32, C.POINTER(_d)]),
    "pcf_server_create": (C.c_int, [_ctx, C.POINTER(PcfServerConfig), C.POINTER(C.c_void_p)]),
    "pcf_server_submit": (C.c_int, [C.c_void_p, C.POINTER(PcfRequest), C.POINTER(C.c_uint64)]),
    "pcf_server_wait": (C.c_int, [C.c_void_p, C.c_uint64, C.POINTER(PcfResult), C.POINTER(PcfRequestTiming)]),
    "pcf_server_destroy": (C.c_int, [C.c_void_p]),
    "pcf_server_last_error": (C.c_char_p, [C.c_void_p]),
    "pcf_server_get_stats": (C.c_int, [C.c_void_p, C.POINTER(PcfServerStats)]),
    "pcf_autotune": (C.c_int, [_ctx, _i32, _i32, C.POINTER(PcfTuneEntry), _i32, C.POINTER(_i32)]),
    "pcf_tune_reset": (None, []),
    "pcf_debug_basket_hash_mask": (None, [C.c_uint64]),
}

STRUCTS = {
    "pcf_opts": PcfOpts,
    "pcf_result": PcfResult,
    "pcf_basket_ext": PcfBasketExt,
    "pcf_request": PcfRequest,
    "pcf_server_config": PcfServerConfig,
    "pcf_request_timing": PcfRequestTiming,
    "pcf_tune_entry": PcfTuneEntry,
    "pcf_server_stats": PcfServerStats,
}

_lib = None


class PcfError(RuntimeError):
    def __init__(self, status: int, message: str):
        self.status = status
        super().__init__(f"{_STATUS_NAMES.get(status, status)}: {message}")


def load(path: str | None = None) -> C.CDLL:
    """Loads libpcf.so (once).  Raises if it has not been built."""
    global _lib
    if _lib is not None and path is None:
        return _lib
    p = path or LIB_PATH
    if not os.path.exists(p):
        raise RuntimeError(
            f"libpcf.so not found at {p}: build it first (python -c 'import __graft_entry__ as g; g.build()')")
    lib = C.CDLL(p)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name, None)
        if fn is None:
            continue
        fn.restype = res
        fn.argtypes = args
    if path is None:
        _lib = lib
    return lib


def exported_symbols() -> list[str]:
    lib = load()
    return [n for n in SIGNATURES if getattr(lib, n, None) is not None]
