"""ewpt_energies_device: the energy table with no allocation by the library
(partials in a fixed device array), for callers that cannot allocate.

Called through the C ABI on a state in a device buffer, twice in a row (a
lattice of 3072 bricks, more than the 2048 partial rows, so blocks stride
over several bricks; then a smaller one, which must not see the first
call's partials), and checked against the existing entry point
ewpt_energies on the same state: the same per-site densities summed in a
different grouping, so the columns agree to rounding (1e-12 relative).
"""
import ctypes

import numpy as np
import pytest

import ewpt
from ewpt import _lib
from ewpt.lattice import as_params
from oracle import fields

pytestmark = pytest.mark.gpu

DX = 0.25


def _hip():
    """The HIP runtime libewpt is linked against, opened by soname after the
    library (the same loaded instance)."""
    ewpt.load()
    hip = ctypes.CDLL("libamdhip64.so.7")
    hip.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
    hip.hipFree.argtypes = [ctypes.c_void_p]
    hip.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    hip.hipDeviceSynchronize.argtypes = []
    return hip


def _device_energies(hip, shape, s, p):
    L = ewpt.load()
    nbytes = s.nbytes
    dstate, dout = ctypes.c_void_p(), ctypes.c_void_p()
    assert hip.hipMalloc(ctypes.byref(dstate), nbytes) == 0
    assert hip.hipMalloc(ctypes.byref(dout), 13 * 8) == 0
    try:
        assert hip.hipMemcpy(dstate, s.ctypes.data, nbytes, 1) == 0          # host to device
        lat = _lib.EwptLattice(shape[0], shape[1], shape[2], DX)
        P = as_params(p)
        rc = L.ewpt_energies_device(ctypes.byref(lat), ctypes.byref(P), dstate, dout, None)
        assert rc == 0, rc
        assert hip.hipDeviceSynchronize() == 0
        out = np.zeros(13)
        assert hip.hipMemcpy(out.ctypes.data, dout, 13 * 8, 2) == 0          # device to host
    finally:
        hip.hipFree(dstate)
        hip.hipFree(dout)
    return out


def test_energies_device_matches_energies_twice():
    hip = _hip()
    p = fields.params(gamma=0.1, jex=0.2)
    for k, shape in enumerate(((128, 64, 96), (32, 16, 16))):
        print(f"energies_device call {k + 1}: lattice {shape}", flush=True)
        s = fields.smooth_state(shape, seed=40 + k, amp=0.4)
        with ewpt.Lattice(shape, DX) as lat:
            lat.set_state(s)
            want = lat.energies(p)
        got = _device_energies(hip, shape, s, p)
        np.testing.assert_allclose(got, want, rtol=1e-12, atol=1e-14)
        assert got[9] > 0 and np.all(got[[7, 8, 10, 12]] == 0.0)


def test_energies_device_rejects():
    L = ewpt.load()
    P = as_params(fields.params())
    bad = _lib.EwptLattice(20, 16, 16, DX)                  # 20 % 16: not a 16x4x4 tiling
    assert L.ewpt_energies_device(ctypes.byref(bad), ctypes.byref(P), ctypes.c_void_p(8), ctypes.c_void_p(8),
                                  None) == _lib.EWPT_ERR_UNSUPPORTED
    small = _lib.EwptLattice(4, 16, 16, DX)                 # below the minimum size
    assert L.ewpt_energies_device(ctypes.byref(small), ctypes.byref(P), ctypes.c_void_p(8), ctypes.c_void_p(8),
                                  None) == _lib.EWPT_ERR_INVAL
    ok = _lib.EwptLattice(32, 16, 16, DX)
    assert L.ewpt_energies_device(ctypes.byref(ok), ctypes.byref(P), None, None, None) == _lib.EWPT_ERR_INVAL
