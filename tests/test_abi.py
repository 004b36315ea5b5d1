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


STRUCTS = {"Grid": "dips_grid_t", "Records": "dips_records_t", "Net": "dips_net_t", "Opt": "dips_opt_t",
           "Tiles": "dips_tiles_t", "MfmaBatch": "dips_mfma_batch_t", "Assemble": "dips_assemble_t",
           "BatchProblem": "dips_batch_problem_t", "HashGrid": "dips_hashgrid_t", "BatchCfg": "dips_batch_cfg_t",
           "BatchPlan": "dips_batch_plan_t"}


def check_layout(classes, tmp_path):
    """Size and every field offset of ctypes structures ({C struct name: class})
    equal the C header's (a C program built with gcc prints offsetof/sizeof)."""
    import shutil
    import subprocess

    import pytest

    gcc = shutil.which("gcc")
    if gcc is None:
        pytest.skip("gcc not available")
    lines = ["#include <stdio.h>", "#include <stddef.h>", '#include "dips_hip.h"', "int main(void) {"]
    expect = []
    for c, cls in classes.items():
        lines.append(f'  printf("%zu\\n", sizeof({c}));')
        expect.append(ctypes.sizeof(cls))
        for name, _ in cls._fields_:
            lines.append(f'  printf("%zu\\n", offsetof({c}, {name}));')
            expect.append(getattr(cls, name).offset)
    lines += ["  return 0;", "}"]
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines) + "\n")
    exe = tmp_path / "layout"
    subprocess.run([gcc, "-std=c99", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)], check=True)
    got = [int(v) for v in subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.split()]
    assert got == expect


def test_ctypes_structs_match_header(tmp_path):
    """The package's ctypes mirrors (jax_dips._native) have the C layout."""
    from jax_dips import _native

    check_layout({c: getattr(_native, py) for py, c in STRUCTS.items()}, tmp_path)


def test_integration_stub_matches_header(tmp_path):
    """The ctypes stub that INTEGRATION.md shows a maintainer (its python block
    with the Structure classes) declares every field of each C struct at the
    header's offsets and sizes."""
    import ast

    txt = open(os.path.join(ROOT, "INTEGRATION.md")).read()
    blocks = re.findall(r"```python\n(.*?)```", txt, flags=re.S)
    stub = [b for b in blocks if "ctypes.Structure" in b]
    assert len(stub) == 1
    tree = ast.parse(stub[0])
    # the imports, type aliases and Structure classes (not the CDLL load or the calls)
    keep = [n for n in tree.body if isinstance(n, (ast.Import, ast.ClassDef))
            or (isinstance(n, ast.Assign) and not isinstance(n.value, ast.Call)
                and all(isinstance(t, (ast.Name, ast.Tuple)) for t in n.targets))]
    ns = {}
    exec(compile(ast.Module(body=keep, type_ignores=[]), "INTEGRATION.md", "exec"), ns)
    names = {"Grid": "dips_grid_t", "Records": "dips_records_t", "Net": "dips_net_t", "Opt": "dips_opt_t",
             "Tiles": "dips_tiles_t", "MfmaBatch": "dips_mfma_batch_t"}
    assert set(names) <= set(ns)
    check_layout({c: ns[py] for py, c in names.items()}, tmp_path)


def test_precond_shapes_and_counts():
    from jax_dips import _native

    lib = _native.lib()
    net = _native.Net(1, 1, 2, 10, 8, 4, 1.0)
    # networks 167 + preconditioner 26*8+8 + 8*4+4 + 4+1
    assert lib.dips_param_count(ctypes.byref(net)) == 167 + 216 + 36 + 5
    assert lib.dips_tile_supported(ctypes.byref(net)) == 1
    assert lib.dips_net_supported(ctypes.byref(net)) == 0  # only the tile kernel applies it
    assert lib.dips_mfma_supported(ctypes.byref(net)) == 0
    assert lib.dips_tile_supported(ctypes.byref(_native.Net(1, 1, 2, 10, 7, 4, 1.0))) == 0
    one = _native.Net(1, 1, 2, 10, 8, 0, 0.5)
    assert lib.dips_param_count(ctypes.byref(one)) == 167 + 216 + 9
    assert lib.dips_tile_supported(ctypes.byref(one)) == 1
    bad = _native.Net(1, 1, 2, 10, 8, 4, float("nan"))
    assert lib.dips_param_count(ctypes.byref(bad)) < 0
    # the LPBE model in full: CELU coordinate networks 3 -> [16] -> 1 and the [8, 4] preconditioner
    lpbe = _native.Net(1, 16, 1, 16, 8, 4, 1.0, 1)
    assert lib.dips_param_count(ctypes.byref(lpbe)) == 2 * (4 * 16 + 17) + 216 + 36 + 5
    assert lib.dips_tile_supported(ctypes.byref(lpbe)) == 1
    assert lib.dips_tile_supported(ctypes.byref(_native.Net(1, 16, 1, 16, 8, 0, 1.0, 1))) == 0


def test_hashgrid_table_and_param_counts():
    """Table rows and parameter count of the encoder equal the oracle's level
    layout (dense levels res^3 rows, hashed levels table_size rows)."""
    from jax_dips import _native
    from oracle.hashgrid import HashGrid, mlp_size

    lib = _native.lib()
    for L, T, nmin, nmax, bound in ((16, 2 ** 19, 16, 2048, 1.0), (4, 2 ** 10, 4, 32, 1.5), (8, 2 ** 14, 16, 512, 1.0)):
        hg = _native.HashGrid(L, 2, T, nmin, int(nmax * bound), bound)
        ref = HashGrid(L, T, 2, nmin, nmax, bound)
        assert lib.dips_hash_table_rows(ctypes.byref(hg)) == ref.rows
        assert lib.dips_hash_param_count(ctypes.byref(hg)) == 2 * mlp_size(2 * L, [16]) + 2 * ref.rows
        dbl = ctypes.c_int64(0)
        assert lib.dips_hash_workspace(ctypes.byref(hg), 1000, 5000, ctypes.byref(dbl)) > 5000
        assert dbl.value > 0
    for bad in (_native.HashGrid(1, 2, 1024, 16, 64, 1.0), _native.HashGrid(17, 2, 1024, 16, 64, 1.0),
                _native.HashGrid(4, 4, 1024, 16, 64, 1.0), _native.HashGrid(4, 2, 1024, 16, 8, 1.0),
                _native.HashGrid(4, 2, 1024, 16, 64, 0.0)):
        assert lib.dips_hash_table_rows(ctypes.byref(bad)) < 0
        assert lib.dips_hash_param_count(ctypes.byref(bad)) < 0
