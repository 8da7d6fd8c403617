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
    assert lib.pcf_abi_version() == 1


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
