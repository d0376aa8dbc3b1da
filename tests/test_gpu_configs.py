"""The precision and batching configurations at their own workloads.

* Precision (512^3 thermal IC, T = 0.5): the mixed path (FP32 storage and
  stencils, compensated update, FP64 reductions) against the FP64 path from
  the same initial state, 100 steps at dt = dx/10 with the same scalars:
  per-field relative L2 <= 1e-5, energy drifts equal to 1e-6, Gauss-norm
  ratio within 1e-4 of 1.  Both states cannot sit in one GPU at once (the two
  contexts need ~310 GB), so the FP64 result is parked in a device buffer
  after its context is gone.
* Batching: 48 realisations (16 x 64^3, 16 x 96^3, 16 x 128^3; even members
  thermal with T uniform in [0.3, 0.7] and J_ext = mH^2, odd members 6-20
  tanh-wall bubbles; distinct seeds), mixed precision, tile shape from the
  autotuner, hipGraph-captured steps, 300 steps.  Every member equals its
  single-lattice run with the same variant bit for bit, and one 64^3 thermal
  member run alone for 50 steps stays within 1e-5 per field of the FP64
  oracle (main_rk4.jl:6146-6441 as restated by oracle/ewpt_oracle.c).  96^3
  is a size the reference cannot run (its 64-thread blocks need N % 64 == 0,
  parameters.jl:90-92).
"""
import numpy as np
import pytest

import ewpt
from ewpt import _lib, ic
from oracle import cref

pytestmark = pytest.mark.gpu

DX = 0.25


def test_precision_512_thermal_100_steps():
    import torch

    n, T, steps = 512, 0.5, 100
    P = ewpt.Parameters(N=n, dx=DX, gy=0.35, T=T)
    p = P.step_params()
    res, final = {}, {}
    for prec in (_lib.EWPT_FP64, _lib.EWPT_MIXED):
        print(f"512^3 thermal precision={prec} x {steps}", flush=True)
        with ewpt.Lattice(n, DX, precision=prec) as lat:
            lat.init_thermal(T, P.meff_sq, P.base_seed, "config")
            e0 = lat.energies(p)
            lat.step(p, steps)
            e1, g1 = lat.energies(p), lat.gauss(p)
            buf = torch.empty(lat.shape, dtype=torch.float64, device="cuda")
            lat.get_state_device(buf.data_ptr())
        final[prec] = buf
        res[prec] = ((sum(e1[:7]) - sum(e0[:7])) / sum(e0[:7]), float(np.linalg.norm(g1[:4])))
    ref, mix = final[_lib.EWPT_FP64], final[_lib.EWPT_MIXED]
    err = [float(torch.linalg.vector_norm(mix[c] - ref[c]) / torch.linalg.vector_norm(ref[c])) for c in range(16)]
    del final, ref, mix, buf
    torch.cuda.empty_cache()
    (d64, g64), (d32, g32) = res[_lib.EWPT_FP64], res[_lib.EWPT_MIXED]
    print(f"max field rel-L2 {max(err):.3e}; drift fp64 {d64:.6e} mixed {d32:.6e}; gauss {g64:.6e} {g32:.6e}",
          flush=True)
    assert max(err) <= 1e-5, err
    assert abs(d32 - d64) <= 1e-6, (d32, d64)
    assert g64 > 0 and abs(g32 / g64 - 1.0) <= 1e-4, (g32, g64)


SHAPES = [64] * 16 + [96] * 16 + [128] * 16


def _batch_inits():
    rng = np.random.default_rng(2024)
    out = []
    for k in range(len(SHAPES)):
        if k % 2 == 0:
            out.append(("thermal", float(rng.uniform(0.3, 0.7)), 5000 + k))
        else:
            out.append(("bubbles", int(rng.integers(6, 21)), 7000 + k))
    return out


def test_batch_48_realisations_300_steps():
    par = ewpt.Parameters(N=64, dx=DX, gy=0.35)
    inits = _batch_inits()
    ps = [par.step_params(jex=par.mH ** 2) if kind == "thermal" else par.step_params() for kind, _, _ in inits]
    assert sum(k == "thermal" for k, _, _ in inits) == 24
    assert all(6 <= a <= 20 for k, a, _ in inits if k == "bubbles")
    assert all(0.3 <= a <= 0.7 for k, a, _ in inits if k == "thermal")

    def init(lat, k):
        kind, a, seed = inits[k]
        if kind == "thermal":
            lat.init_thermal(a, par.meff_sq, seed, "config")
        else:
            lat.init_bubbles(ic.bubble_table(lat.nx, a, seed, "config"), "config", par.mH)

    lats = []
    try:
        for k, n in enumerate(SHAPES):
            lats.append(ewpt.Lattice(n, DX, precision=_lib.EWPT_MIXED))
            init(lats[-1], k)
        # the 64^3 thermal member 0 against the FP64 oracle, run alone
        s0 = lats[0].get_state()
        with ewpt.Lattice(64, DX, precision=_lib.EWPT_MIXED) as one:
            one.set_state(s0)
            one.step(ps[0], 50)
            y = one.get_state()
        print("48-lattice batch: oracle run of member 0 (64^3 x 50)", flush=True)
        r = cref.rk4(s0, ps[0], DX, 50)
        num = np.linalg.norm((y - r)[:16].reshape(16, -1), axis=1)
        den = np.linalg.norm(r[:16].reshape(16, -1), axis=1)
        assert np.max(num / den) < 1e-5, num / den
        with ewpt.Batch(lats) as bt:
            v, ms = bt.tune(ps, 2)
            print(f"48-lattice batch: tuned variant {v} ({ms:.2f} ms/step)", flush=True)
            for k, lat in enumerate(lats):          # restart from the initial states
                init(lat, k)
            bt.step(ps, 300)
        for k, n in enumerate(SHAPES):
            with ewpt.Lattice(n, DX, precision=_lib.EWPT_MIXED) as one:
                one.variant = v % 1000
                init(one, k)
                one.step(ps[k], 300)
                assert np.array_equal(one.get_state(), lats[k].get_state(), equal_nan=True), k
            e = lats[k].energies(ps[k])
            assert np.all(np.isfinite(e[:7])), k
        print("48-lattice batch: every member equals its solo run", flush=True)
    finally:
        for lat in lats:
            lat.close()
