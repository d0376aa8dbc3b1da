"""Size-independent properties at the benchmark configurations' full sizes
(256^3 thermal state), where the CPU oracle is too slow to run.

* RK4 order: the final state of the same physical time integrated with
  dt, dt/2 and dt/4 converges with the Richardson ratio
  |y_dt - y_dt/2| / |y_dt/2 - y_dt/4| = 2^p; classical RK4 has p = 4
  (ratio 16).  A dropped stage, a wrong stage weight or a stage input formed
  with the wrong k lowers the order (ratio 4 or 2).
* Mixed precision: 500 steps at dt = dx/10 of the mixed path stay within
  1e-5 relative L2 per field channel of the FP64 path, with the same energy
  and Gauss-residual history.
"""
import numpy as np
import pytest

import ewpt

pytestmark = pytest.mark.gpu

N, DX, T = 256, 0.25, 0.5


def _par():
    return ewpt.Parameters(N=N, dx=DX, gy=0.35, T=T)


def _fields(lat):
    return lat.get_state()[:16]


def test_rk4_order_at_256():
    P = _par()
    out = []
    with ewpt.Lattice(N, DX) as src, ewpt.Lattice(N, DX) as lat:
        src.init_thermal(T, P.meff_sq, 77, "config")
        for dt, nst in ((0.1, 40), (0.05, 80), (0.025, 160)):
            print(f"256^3 FP64 dt={dt} x {nst}", flush=True)
            p = P.step_params()
            p.dt = dt
            lat.copy_from(src)
            lat.step(p, nst)
            out.append(_fields(lat))
    d1 = np.linalg.norm(out[0] - out[1])
    d2 = np.linalg.norm(out[1] - out[2])
    ratio = d1 / d2
    print(f"Richardson ratio {ratio:.2f} (order {np.log2(ratio):.2f})", flush=True)
    print(f"|y_dt/2 - y_dt/4| / |y| = {d2 / np.linalg.norm(out[2]):.2e}", flush=True)
    assert 12.0 < ratio < 24.0, ratio


def test_mixed_tracks_fp64_at_256_over_500_steps():
    P = _par()
    p = P.step_params()          # dt = dx/10
    res = {}
    for prec in (ewpt.EWPT_FP64, ewpt.EWPT_MIXED):
        print(f"256^3 precision {prec} x 500", flush=True)
        with ewpt.Lattice(N, DX, precision=prec) as lat:
            lat.init_thermal(T, P.meff_sq, 78, "config")
            e0 = lat.energies(p)
            lat.step(p, 500)
            res[prec] = (_fields(lat), e0, lat.energies(p), lat.gauss(p))
    y64, e0_64, e1_64, g64 = res[ewpt.EWPT_FP64]
    y32, e0_32, e1_32, g32 = res[ewpt.EWPT_MIXED]
    err = np.linalg.norm((y32 - y64).reshape(16, -1), axis=1) / np.linalg.norm(y64.reshape(16, -1), axis=1)
    print(f"max field rel-L2 {err.max():.3e}", flush=True)
    assert err.max() < 1e-5, err
    drift64 = (sum(e1_64[:7]) - sum(e0_64[:7])) / sum(e0_64[:7])
    drift32 = (sum(e1_32[:7]) - sum(e0_32[:7])) / sum(e0_32[:7])
    assert abs(drift32 - drift64) < 1e-6, (drift32, drift64)
    np.testing.assert_allclose(g32, g64, rtol=1e-4)
