"""The C-ABI library loads and exports every symbol include/ewpt.h declares
(no device calls)."""
import ctypes
import os
import re

import ewpt
from ewpt import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    txt = open(os.path.join(ROOT, "include", "ewpt.h")).read()
    return sorted(set(re.findall(r"\b(ewpt_[a-z_]+)\s*\(", txt)))


def test_exports_match_header():
    L = ewpt.load()
    decl = declared_symbols()
    assert decl, "no declarations parsed"
    for name in decl:
        assert hasattr(L, name), name
    assert set(decl) == set(_lib.EXPORTS)


def test_strerror_and_version():
    L = ewpt.load()
    assert L.ewpt_strerror(0) == b"ok"
    assert L.ewpt_strerror(_lib.EWPT_ERR_NONFINITE).startswith(b"non-finite")
    assert b"gfx950" in L.ewpt_version()


def test_invalid_arguments_rejected_without_device():
    L = ewpt.load()
    h = ctypes.c_void_p()
    bad = _lib.EwptLattice(4, 64, 64, 0.25)          # nx < 8
    assert L.ewpt_create(ctypes.byref(h), ctypes.byref(bad), 0, 0) == _lib.EWPT_ERR_INVAL
    good = _lib.EwptLattice(16, 16, 16, 0.25)
    assert L.ewpt_create(ctypes.byref(h), ctypes.byref(good), 7, 0) == _lib.EWPT_ERR_INVAL
    assert L.ewpt_create(None, ctypes.byref(good), 0, 0) == _lib.EWPT_ERR_INVAL
    assert L.ewpt_destroy(None) == _lib.EWPT_ERR_INVAL
    assert L.ewpt_step(None, None, 1) == _lib.EWPT_ERR_INVAL


def test_struct_layouts_match_c(tmp_path):
    """ewpt_lattice and ewpt_params as a C compiler lays them out (offsetof /
    sizeof from include/ewpt.h) equal the ctypes structures the host code
    passes, and the Julia structs of the integration guide (Cint x 3 then a
    Cdouble: dx at byte 16; ten Cdoubles)."""
    import shutil
    import subprocess

    import pytest
    cc = shutil.which("gcc") or shutil.which("cc")
    if cc is None:
        pytest.skip("no C compiler")
    src = tmp_path / "layout.c"
    src.write_text(
        '#include <stddef.h>\n#include <stdio.h>\n#include "ewpt.h"\n'
        "int main(void) {\n"
        '  printf("lattice %zu %zu %zu %zu %zu\\n", sizeof(ewpt_lattice), offsetof(ewpt_lattice, nx),\n'
        "         offsetof(ewpt_lattice, ny), offsetof(ewpt_lattice, nz), offsetof(ewpt_lattice, dx));\n"
        '  printf("params %zu %zu %zu %zu %zu %zu %zu %zu %zu %zu %zu\\n", sizeof(ewpt_params),\n'
        "         offsetof(ewpt_params, gw), offsetof(ewpt_params, gy), offsetof(ewpt_params, gp2),\n"
        "         offsetof(ewpt_params, beta_w), offsetof(ewpt_params, beta_y), offsetof(ewpt_params, vev),\n"
        "         offsetof(ewpt_params, lambda), offsetof(ewpt_params, dt), offsetof(ewpt_params, jex),\n"
        "         offsetof(ewpt_params, gamma));\n"
        "  return 0;\n}\n")
    exe = tmp_path / "layout"
    subprocess.run([cc, "-I", os.path.join(ROOT, "include"), "-o", str(exe), str(src)], check=True)
    lines = dict(l.split(" ", 1) for l in subprocess.run([str(exe)], check=True, capture_output=True,
                                                           text=True).stdout.splitlines())
    lat = [int(v) for v in lines["lattice"].split()]
    L = _lib.EwptLattice
    assert lat == [ctypes.sizeof(L), L.nx.offset, L.ny.offset, L.nz.offset, L.dx.offset]
    assert lat[4] == 16 and lat[0] == 24
    par = [int(v) for v in lines["params"].split()]
    P = _lib.EwptParams
    names = ("gw", "gy", "gp2", "beta_w", "beta_y", "vev", "lambda_", "dt", "jex", "gamma")
    assert par == [ctypes.sizeof(P)] + [getattr(P, n).offset for n in names]
    assert par == [80] + [8 * k for k in range(10)]
