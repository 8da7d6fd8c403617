"""parcompfin_amd -- MI355X-native Monte Carlo / binomial option pricing.

The engine is libpcf.so (HIP kernels for gfx950 behind the C-ABI of
include/pcf.h).  This package holds its sources (csrc/), the drop-in CLIs
(host/), and a thin ctypes binding (api.Engine) with the reference's
function names.
"""
from ._native import LIB_PATH, PcfError, load  # noqa: F401
from .api import Engine, Result, Server, autotune, payoff_code, tune_reset  # noqa: F401

__all__ = ["Engine", "Result", "Server", "PcfError", "load", "LIB_PATH", "payoff_code", "autotune",
           "tune_reset"]
