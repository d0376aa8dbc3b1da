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
