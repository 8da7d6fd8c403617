"""ctypes access to oracle/liboracle.so -- the CPU checker (tests only)."""
import ctypes as C
import os

C = C  # re-exported for tests that pass pointers

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_SO = os.path.join(ROOT, "oracle", "liboracle.so")
REF_DIR = os.path.join(ROOT, "oracle", "_ref")

_d, _i64, _i32, _u64, _u32 = C.c_double, C.c_int64, C.c_int, C.c_uint64, C.c_uint32
_pd = C.POINTER(C.c_double)
_SIG = {
    "orc_normals_mt": (None, [_u64, _d, _i64, _pd]),
    "orc_mc_eur_z": (_d, [_d, _d, _d, _d, _d, _i64, _d, _pd]),
    "orc_mc_eur_seed": (_d, [_d, _d, _d, _d, _d, _i64, _d, _u64]),
    "orc_mc_asia_z": (_d, [_d, _d, _d, _d, _d, _i64, _i32, _d, _pd]),
    "orc_mc_asia_seed": (_d, [_d, _d, _d, _d, _d, _i64, _i32, _d, _u64]),
    "orc_mc_asia_prod_z": (_d, [_d, _d, _d, _d, _d, _i64, _i32, _d, _pd, _pd]),
    "orc_comb": (_d, [_i32, _i32]),
    "orc_beta_tree": (None, [_d, _d, _d, _i32, _pd]),
    "orc_binom_embar": (_d, [_d, _d, _d, _d, _d, _i32, _d, _i32]),
    "orc_binom_embar_prefix": (_d, [_d, _d, _d, _d, _d, _i32, _d, _i32]),
    "orc_logfact_dd": (None, [_i32, _pd]),
    "orc_binom_dd": (_d, [_d, _d, _d, _d, _d, _i32, _d]),
    "orc_binom_vanilla": (_d, [_d, _d, _d, _d, _d, _i32, _d, _i32]),
    "orc_inverse3_solve": (_i32, [_pd, _pd, _pd]),
    "orc_mc_amer_z": (_d, [_d, _d, _d, _d, _d, _i64, _i32, _d, _pd, _i32, C.POINTER(C.c_int32)]),
    "orc_mc_amer_seed": (_d, [_d, _d, _d, _d, _d, _i64, _i32, _d, _u64, _i32]),
    "orc_lsm_solve_poly": (_i32, [_pd, _i32, _pd]),
    "orc_mc_amer_poly": (_d, [_d, _d, _d, _d, _d, _i64, _i32, _d, _pd, _i32, C.POINTER(C.c_int32),
                              C.POINTER(C.c_int32)]),
    "orc_basket_z": (_d, [_d, _d, _d, _d, _d, _i64, _d, _i32, _pd, _pd, _pd, _pd, _pd]),
    "orc_basket_prod_z": (_d, [_d, _d, _d, _d, _d, _i64, _d, _i32, _pd, _pd, _pd, _pd, _pd, _pd]),
    "orc_cholesky": (_i32, [_i32, _pd, _pd]),
    "orc_philox4x32_10": (None, [C.POINTER(_u32), C.POINTER(_u32), C.POINTER(_u32)]),
    "orc_philox_normals4": (None, [_u64, _u32, _u64, _u32, _pd]),
    "orc_philox_mc_eur": (_d, [_d, _d, _d, _d, _d, _i64, _d, _u64, _pd]),
    "orc_philox_mc_asia": (_d, [_d, _d, _d, _d, _d, _i64, _i32, _d, _u64, _pd, _pd]),
    "orc_philox_amer_normals": (None, [_u64, _i64, _i32, _d, _pd]),
    "orc_philox_basket_normals": (None, [_u64, _i64, _i32, _pd]),
    "orc_philox_normals": (None, [_u64, _u32, _u64, _i64, _i32, _pd]),
    "orc_bs_price": (_d, [_d, _d, _d, _d, _d, _d]),
    "orc_geo_asian": (_d, [_d, _d, _d, _d, _d, _i32, _d]),
    "orc_omp_max_threads": (_i32, []),
}

_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(ORACLE_SO):
            raise RuntimeError("oracle/liboracle.so missing: run `make -C oracle`")
        L = C.CDLL(ORACLE_SO)
        for n, (r, a) in _SIG.items():
            f = getattr(L, n)
            f.restype = r
            f.argtypes = a
        _lib = L
    return _lib


def ptr(a):
    return a.ctypes.data_as(_pd) if a is not None else None


def normals_mt(seed, sd, count):
    out = np.empty(int(count), dtype=np.float64)
    lib().orc_normals_mt(seed, sd, int(count), ptr(out))
    return out


def philox_normals(seed, stream, p0, n, K):
    out = np.empty((int(n), int(K)), dtype=np.float64)
    lib().orc_philox_normals(seed, stream, p0, int(n), int(K), ptr(out))
    return out


def philox_block(ctr, key):
    U4 = C.c_uint32 * 4
    K2 = C.c_uint32 * 2
    o = U4()
    lib().orc_philox4x32_10(U4(*ctr), K2(*key), o)
    return tuple(o)


def ref_bin(name):
    p = os.path.join(REF_DIR, name)
    return p if os.path.exists(p) else None
