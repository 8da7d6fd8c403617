"""The C-ABI library: loads, exports every symbol include/pcf.h declares, and
its record layouts match the ctypes mirror.  No compute calls (CPU only)."""
import ctypes as C
import os
import re

import pytest

from parcompfin_amd import _native as N

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    src = open(os.path.join(ROOT, "include", "pcf.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(pcf_[a-z0-9_]+)\s*\(", src)))


def test_library_built_and_loads():
    assert os.path.exists(N.LIB_PATH), "run __graft_entry__.build() first"
    lib = N.load()
    assert lib.pcf_abi_version() == 2


def test_every_declared_symbol_is_exported_and_bound():
    lib = N.load()
    names = _declared()
    assert len(names) >= 18
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    unbound = [n for n in names if n not in N.SIGNATURES]
    assert not unbound, unbound


@pytest.mark.parametrize("name", sorted(N.STRUCTS))
def test_struct_layout_matches(name):
    lib = N.load()
    assert lib.pcf_struct_size(name.encode()) == C.sizeof(N.STRUCTS[name])


def test_opts_defaults():
    o = N.PcfOpts()
    N.load().pcf_opts_init(C.byref(o))
    assert o.precision == N.PCF_FP32K and o.rng == N.PCF_RNG_PHILOX and o.seed != 0


def test_missing_library_fails_loudly(tmp_path):
    with pytest.raises(RuntimeError):
        N.load(str(tmp_path / "nope.so"))


def test_payoff_strings_like_reference():
    from parcompfin_amd import payoff_code
    assert payoff_code("call") == 1.0 and payoff_code("put") == -1.0
    with pytest.raises(ValueError):
        payoff_code("straddle")
    with pytest.raises(ValueError):
        payoff_code(0.5)


def test_result_record_roundtrip():
    """Every C result field lands in the Python Result (no GPU needed)."""
    from parcompfin_amd.api import Result
    r = N.PcfResult()
    r.price, r.batch_size = 1.5, 3
    out = Result.from_c(r, timing={"latency_ms": 1.0})
    assert out.price == 1.5 and out.batch_size == 3 and out.extra["timing"]["latency_ms"] == 1.0


def test_request_record_builds_without_device():
    import ctypes as C
    import numpy as np
    from parcompfin_amd.api import Engine
    o, keep = Engine.make_opts(seed=3, precision="fp64", normals=np.zeros(4), kernel_variant=1)
    assert o.rng == N.PCF_RNG_HOST and o.n_host_normals == 4 and o.kernel_variant == 1
    q = N.PcfRequest()
    q.opts = o
    ext = N.PcfBasketExt()
    q.basket = C.pointer(ext)
    assert bool(q.basket)
