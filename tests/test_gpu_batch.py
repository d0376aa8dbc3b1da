"""Batched stepping (ewpt_batch_*): every lattice of a batch must come out
bitwise equal to the same lattice stepped alone with the same stage variant
(the batched kernels run the same brick code), through the captured hipGraph
(pairs of steps, replayed several times with eager work in between) and the
eager odd last step, for both launch forms and both precisions."""
import math

import numpy as np
import pytest

import ewpt
from ewpt import ic
from oracle import fields

pytestmark = pytest.mark.gpu

DX = 0.25
MH = 2.0 * math.sqrt(0.129)
MEFF = 8119.43 / 174.0 ** 2


def _init(lat, k):
    """Alternate thermal and bubble initial conditions, distinct seeds."""
    n = lat.nx
    if k % 2 == 0:
        lat.init_thermal(0.3 + 0.1 * k, MEFF, 1000 + k, "config")
    else:
        lat.init_bubbles(ic.bubble_table((lat.nx, lat.ny, lat.nz), 3 + k, 2000 + k, "config"), "config", MH)
    return n


def _params(k):
    # distinct per-lattice scalars: thermal ones get the external mass term
    return fields.params(gy=0.35, jex=(MH ** 2 if k % 2 == 0 else 0.0), gamma=0.05 * k)


@pytest.mark.parametrize("precision,variant", [(ewpt.EWPT_MIXED, 210), (ewpt.EWPT_MIXED, 1010),
                                               (ewpt.EWPT_FP64, 210), (ewpt.EWPT_FP64, 1002),
                                               (ewpt.EWPT_FP64, 332), (ewpt.EWPT_FP64, 1332),
                                               (ewpt.EWPT_FP64, 334), (ewpt.EWPT_FP64, 1334)])
def test_batch_bitwise_equals_single(precision, variant):
    shapes = [(32, 32, 32), (48, 48, 48), (64, 32, 16), (32, 48, 64)]
    print(f"batch precision={precision} variant={variant}", flush=True)
    lats = [ewpt.Lattice(s, DX, precision=precision) for s in shapes]
    solo = [ewpt.Lattice(s, DX, precision=precision) for s in shapes]
    try:
        for k, (a, b) in enumerate(zip(lats, solo)):
            _init(a, k)
            _init(b, k)
            b.variant = variant % 1000
        ps = [_params(k) for k in range(len(shapes))]
        with ewpt.Batch(lats) as bt:
            bt.variant = variant
            assert bt.variant == variant
            bt.step(ps, 4)                       # two graph replays
            mid = [a.get_state() for a in lats]  # eager work between replays
            bt.step(ps, 5)                       # two replays + one eager step
        for k, b in enumerate(solo):
            b.step(ps[k], 4)
            assert np.array_equal(mid[k], b.get_state(), equal_nan=True), (k, "after 4")
            b.step(ps[k], 5)
            assert np.array_equal(lats[k].get_state(), b.get_state(), equal_nan=True), (k, "after 9")
            e = lats[k].energies(ps[k])
            assert np.all(np.isfinite(e[:7]))
    finally:
        for x in lats + solo:
            x.close()


@pytest.mark.parametrize("variant", [333, 331, 315, 1333])
def test_batch_32_wide_mixed_forms_bitwise(variant):
    # the 32x4x4 stored-input forms (early y, lean stage-4 stash) on lattices they tile
    shapes = [(32, 32, 32), (64, 32, 16), (32, 64, 32)]
    lats = [ewpt.Lattice(s, DX, precision=ewpt.EWPT_MIXED) for s in shapes]
    solo = [ewpt.Lattice(s, DX, precision=ewpt.EWPT_MIXED) for s in shapes]
    try:
        for k, (a, b) in enumerate(zip(lats, solo)):
            _init(a, k)
            _init(b, k)
            b.variant = variant % 1000
        ps = [_params(k) for k in range(len(shapes))]
        with ewpt.Batch(lats) as bt:
            bt.variant = variant
            bt.step(ps, 5)
        for k, b in enumerate(solo):
            b.step(ps[k], 5)
            assert np.array_equal(lats[k].get_state(), b.get_state(), equal_nan=True), k
    finally:
        for x in lats + solo:
            x.close()


def test_batch_tune_leaves_states_and_caches():
    shapes = [(32, 32, 32), (48, 48, 48)]
    lats = [ewpt.Lattice(s, DX, precision=ewpt.EWPT_MIXED) for s in shapes]
    try:
        for k, a in enumerate(lats):
            _init(a, k)
        with ewpt.Batch(lats) as bt:
            # the FP64 view of a fresh mixed state is the uploaded FP64 value; after
            # any step it is hi + lo of the FP32 pair, so compare from there on
            bt.step(fields.params(dt=0.0), 1)
            before = [a.get_state() for a in lats]
            v, ms = bt.tune(_params(0), 2)
            assert v % 100 in (10, 11, 15, 31, 33) and ms > 0
            assert bt.variant == v
            for a, s in zip(lats, before):
                assert np.array_equal(a.get_state(), s)      # dt = 0 timing steps change nothing
        with ewpt.Batch(lats[::-1]) as bt2:                  # same shapes: served from the cache
            v2, ms2 = bt2.tune(_params(0), 2)
            assert v2 == v and ms2 == -1.0
    finally:
        for x in lats:
            x.close()


def test_batch_errors():
    a = ewpt.Lattice((40, 40, 40), DX)
    b = ewpt.Lattice((32, 32, 32), DX)
    m = ewpt.Lattice((32, 32, 32), DX, precision=ewpt.EWPT_MIXED)
    try:
        with pytest.raises(ewpt.EwptError):
            ewpt.Batch([b, b])                               # the same context twice
        with pytest.raises(ewpt.EwptError):
            ewpt.Batch([b, m])                               # mixed precisions
        with ewpt.Batch([a, b]) as bt:
            assert bt.variant == 2                           # only 8x8x8 tiles 40^3
            with pytest.raises(ewpt.EwptError):
                bt.variant = 210
            with pytest.raises(ewpt.EwptError):
                bt.step(fields.params(), 1)                  # no state yet
            a.set_state(fields.smooth_state((40, 40, 40), seed=1))
            b.set_state(fields.smooth_state((32, 32, 32), seed=2))
            with pytest.raises(ValueError):
                bt.step([fields.params()], 1)                # one parameter set per lattice
            bt.step(fields.params(), 1)
    finally:
        for x in (a, b, m):
            x.close()
