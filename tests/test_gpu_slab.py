"""z-slab contexts: one lattice split over several contexts with halo planes
exchanged before each stage (the per-GPU share of a multi-GPU run, with
device copies standing in for the transfers between GPUs).

The slab stage kernels read the z halo from ghost planes instead of wrapping
the index, so stepping the slabs must reproduce stepping the stacked lattice
in one context with the same variant bit for bit: a wrong plane offset, a
swapped neighbour or a missing exchange changes the result.
"""
import ctypes

import numpy as np
import pytest

import ewpt
from ewpt import _lib
from oracle import fields

pytestmark = pytest.mark.gpu

DX = 0.25
P = fields.params(beta_w=0.1, beta_y=0.05, gamma=0.2, jex=0.3)


def _state(shape, seed):
    return fields.smooth_state(shape, seed=seed, amp=0.5)


def _split(s, n):
    return np.split(s, n, axis=1)          # (36, nz, ny, nx): z is axis 1


@pytest.mark.parametrize("precision,variant", [(_lib.EWPT_FP64, 320), (_lib.EWPT_FP64, 10), (_lib.EWPT_MIXED, 311),
                                               (_lib.EWPT_FP64, 334), (_lib.EWPT_FP64, 324), (_lib.EWPT_MIXED, 333)])
def test_slab_alone_is_periodic(precision, variant):
    shape = (32, 32, 16)
    print(f"slab alone {shape} precision {precision} variant {variant}", flush=True)
    s = _state(shape, 31)
    with ewpt.Lattice(shape, DX, precision) as a, ewpt.Lattice(shape, DX, precision, slab=True) as b:
        assert b.active_variant == (334 if precision == _lib.EWPT_FP64 else 333)   # the automatic choice
        a.variant = variant
        b.variant = variant
        assert a.active_variant == b.active_variant
        a.set_state(s)
        b.set_state(s)
        assert np.array_equal(b.get_state(), a.get_state())
        a.step(P, 3)
        b.step(P, 3)
        assert np.array_equal(b.get_state(), a.get_state())


@pytest.mark.parametrize("nslab", [2, 4])
@pytest.mark.parametrize("precision,variant", [(_lib.EWPT_FP64, 311), (_lib.EWPT_FP64, 2), (_lib.EWPT_FP64, 10),
                                               (_lib.EWPT_MIXED, 311), (_lib.EWPT_MIXED, 210),
                                               (_lib.EWPT_FP64, 334), (_lib.EWPT_FP64, 332), (_lib.EWPT_MIXED, 333),
                                               (_lib.EWPT_MIXED, 315)])
def test_slabs_equal_one_lattice(nslab, precision, variant):
    full = (32, 32, 32)
    if variant % 100 == 2 and 32 // nslab % 8:
        pytest.skip("8x8x8 bricks need 8 planes per slab")
    print(f"{nslab} slabs of {full} precision {precision} variant {variant}", flush=True)
    s = _state(full, 32)
    local = (32, 32, 32 // nslab)
    slabs = [ewpt.Lattice(local, DX, precision, slab=True) for _ in range(nslab)]
    try:
        with ewpt.Lattice(full, DX, precision) as one:
            one.variant = variant
            one.set_state(s)
            one.step(P, 3)
            want = one.get_state()
        for lat, part in zip(slabs, _split(s, nslab)):
            lat.variant = variant
            lat.set_state(part)
        ewpt.slab_step(slabs, P, 3)
        got = np.concatenate([lat.get_state() for lat in slabs], axis=1)
    finally:
        for lat in slabs:
            lat.close()
    assert np.array_equal(got, want), np.abs(got - want).max()


@pytest.mark.parametrize("variant", [0, 325])
def test_exchange_primitives_as_a_multi_gpu_driver(variant):
    """pack -> hand over -> unpack -> stage through the public calls, the
    way a driver with one rank per GPU would, equals ewpt_slab_step (the
    automatic variant exchanges the stored stage inputs, 325 the k's)."""
    full = (32, 32, 32)
    s = _state(full, 33)
    parts = _split(s, 2)
    drv = [ewpt.Lattice((32, 32, 16), DX, slab=True) for _ in range(2)]
    ref = [ewpt.Lattice((32, 32, 16), DX, slab=True) for _ in range(2)]
    try:
        for lat, part in zip(drv + ref, parts + parts):
            lat.variant = variant
            lat.set_state(part)
        nk = 36 if variant == 0 else 20
        assert drv[0].halo_bytes(1) == nk * 3 * 32 * 32 * 8
        bufs = [lat.slab_buffers() for lat in drv]      # (send_lo, send_hi) each
        for _ in range(2):
            for stage in (1, 2, 3, 4):
                what = stage - 1
                for lat, (lo, hi) in zip(drv, bufs):
                    lat.slab_pack(what, lo, hi)
                for lat in drv:
                    lat.sync()
                for i, lat in enumerate(drv):
                    other = bufs[1 - i]       # two slabs: each is the other's lower and upper neighbour
                    lat.slab_unpack(what, other[1], other[0])
                for lat in drv:
                    lat.sync()
                for lat in drv:
                    lat.stage(P, stage)
                for lat in drv:
                    lat.sync()
        ewpt.slab_step(ref, P, 2)
        for a, b in zip(drv, ref):
            assert np.array_equal(a.get_state(), b.get_state())
    finally:
        for lat in drv + ref:
            lat.close()


def _hip():
    """The HIP runtime libewpt is linked against: opened by its soname after
    the library, so the loader hands back that same instance (opening the
    file by path would load it again against whichever HSA runtime the
    process brought up first)."""
    ewpt.load()
    hip = ctypes.CDLL("libamdhip64.so.7")
    hip.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
    hip.hipFree.argtypes = [ctypes.c_void_p]
    hip.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    return hip


@pytest.mark.parametrize("precision", [_lib.EWPT_FP64, _lib.EWPT_MIXED])
def test_pack_holds_the_boundary_planes(precision):
    shape = (16, 8, 12)
    s = _state(shape, 34)
    word = 8 if precision == _lib.EWPT_FP64 else 4
    hip = _hip()
    with ewpt.Lattice(shape, DX, precision, slab=True) as lat:
        lat.set_state(s)
        nb = lat.halo_bytes(0)
        assert nb == 36 * 3 * 16 * 8 * word
        assert lat.halo_bytes(1) == 36 * 3 * 16 * 8 * word     # the stored stage input u2
        lat.variant = 325                                       # 16x4x4, inputs formed on load
        assert lat.halo_bytes(1) == 20 * 3 * 16 * 8 * word     # k_1
        lat.variant = 0
        lo, hi = lat.slab_buffers()
        lat.slab_pack(0, lo, hi)
        lat.sync()
        dt = np.float64 if word == 8 else np.float32
        got = []
        for src in (lo, hi):
            h = np.zeros(nb // word, dtype=dt)
            assert hip.hipMemcpy(h.ctypes.data, src, nb, 2) == 0      # hipMemcpyDeviceToHost
            got.append(h.reshape(36, 3, 8, 16))
    want_lo, want_hi = s[:, :3], s[:, -3:]
    if word == 4:      # the mixed state's high word
        want_lo, want_hi = want_lo.astype(np.float32), want_hi.astype(np.float32)
    assert np.array_equal(got[0], want_lo)
    assert np.array_equal(got[1], want_hi)


def test_slab_rejects_whole_lattice_calls():
    with ewpt.Lattice((16, 16, 8), DX, slab=True) as lat:
        lat.set_state(_state((16, 16, 8), 35))
        for call in (lambda: lat.energies(P), lambda: lat.gauss(P), lambda: lat.init_thermal(0.5, 1.0, 1, "config")):
            with pytest.raises(ewpt.EwptError) as ei:
                call()
            assert ei.value.code == _lib.EWPT_ERR_UNSUPPORTED
        with pytest.raises(ewpt.EwptError) as ei:
            lat.variant = 9                                  # the gather kernel wraps z
        assert ei.value.code == _lib.EWPT_ERR_UNSUPPORTED
        with ewpt.Lattice((16, 16, 8), DX) as periodic:
            periodic.set_state(_state((16, 16, 8), 35))
            with pytest.raises(ewpt.EwptError):
                ewpt.slab_step([lat, periodic], P, 1)
            # state copies work across the two layouts
            lat.copy_from(periodic)
            assert np.array_equal(lat.get_state(), periodic.get_state())
