"""The C-ABI library loads and exports every symbol include/dips_hip.h declares
(no compute call: these run without a GPU)."""
import ctypes
import os
import re

from conftest import ROOT


def declared_symbols():
    txt = open(os.path.join(ROOT, "include", "dips_hip.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(dips_[a-z0-9_]+)\s*\(", txt)))


def test_library_exports_declared_symbols():
    from jax_dips import _native

    lib = _native.lib()
    syms = declared_symbols()
    assert len(syms) >= 6
    missing = [s for s in syms if not hasattr(lib, s)]
    assert not missing, missing
    assert set(syms) == set(_native.SIGNATURES), "ctypes signature table out of sync with the header"


def test_host_only_queries():
    from jax_dips import _native

    lib = _native.lib()
    net = _native.Net(1, 1, 2, 10)
    assert lib.dips_param_count(ctypes.byref(net)) == 167
    assert lib.dips_net_supported(ctypes.byref(net)) == 1
    assert lib.dips_net_supported(ctypes.byref(_native.Net(7, 5, 3, 9))) == 0
    ps = lib.dips_partial_size(ctypes.byref(net), 256)
    assert ps > 0 and ps % 168 == 0  # whole partial rows of P + 1 floats
    assert lib.dips_version() >= 1
