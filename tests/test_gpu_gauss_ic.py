"""The initial Gauss-law solve (ewpt_init_thermal_gauss) against the numpy
restatement in oracle/gauss_np.py, and its residual measured by the
independent residual code of the C oracle (oracle/ewpt_oracle.c, from the
k_Gamma / k_Sigma charges of main_rk4.jl:603-657) and of the library
(ewpt_gauss).

* Small lattices (one with an odd side, where only the modes 0 and pi of
  the even sides are annihilated by the d1 stencil): the device state equals
  the oracle's to 1e-12 relative L2 per channel, and the Gauss residual
  after the projection is below 1e-12 of the one before it.
* The thermal configuration's size, 256^3: the residual drops below 1e-12 of
  its start; over the configuration's 500 FP64 steps the residual grows to
  the level the equations produce from the reference's own pi = E = 0 start
  (within 2x of it), far below the residual of the unprojected draw (below
  1/10 of it after 200 steps, 1/5 after 500).
"""
import numpy as np
import pytest

import ewpt
from oracle import cref, fields, gauss_np

pytestmark = pytest.mark.gpu

DX = 0.25
MEFF = 8119.43 / 174.0 ** 2
GW, GY = 0.65, 0.35


def _rel(a, b):
    num = np.sqrt(np.sum((a - b) ** 2, axis=(1, 2, 3)))
    den = np.sqrt(np.sum(b ** 2, axis=(1, 2, 3)))
    return np.where(den > 0, num / np.where(den > 0, den, 1), num)


@pytest.mark.parametrize("shape,seed", [((32, 32, 32), 11), ((24, 16, 20), 12), ((20, 18, 15), 13)])
def test_gauss_solve_matches_oracle(shape, seed):
    print(f"Gauss-law IC {shape} seed {seed}", flush=True)
    p = fields.params(gw=GW, gy=GY)
    with ewpt.Lattice(shape, DX) as lat:
        st = lat.init_thermal_gauss(0.5, MEFF, seed, p)
        got = lat.get_state()
        g_dev = lat.gauss(p)
    want, info = gauss_np.thermal_gauss_state(shape, DX, 0.5, MEFF, seed, GW, GY)
    err = _rel(got, want)
    print(f"max rel-L2 per channel {err.max():.3e}; device {st}; oracle {info}", flush=True)
    assert err.max() < 1e-12, (err.argmax(), err.max())
    assert st["residual_after"] < 1e-12 * st["residual_before"], st
    assert st["residual_before"] > 0.1
    # the residual by the C oracle's own charge code and by ewpt_gauss
    g_ref = cref.gauss(got, p, DX)
    before = st["residual_before"]
    assert np.sqrt(np.sum(g_ref[:4] ** 2)) < 1e-12 * before, g_ref
    assert np.sqrt(np.sum(g_dev[:4] ** 2)) < 1e-12 * before, g_dev
    assert np.max(g_ref[4:]) < 1e-12 * before            # Gamma = div W, Sigma = div Y
    assert np.abs(got[20:36]).max() > 1e-3                 # the momenta are drawn


def test_gauss_solve_mixed_context():
    """The mixed path solves in FP64 and keeps the result as its FP32 pair."""
    shape = (32, 32, 32)
    p = fields.params(gw=GW, gy=GY)
    with ewpt.Lattice(shape, DX) as a, ewpt.Lattice(shape, DX, precision=ewpt.EWPT_MIXED) as b:
        sa = a.init_thermal_gauss(0.5, MEFF, 21, p)
        sb = b.init_thermal_gauss(0.5, MEFF, 21, p)
        ya, yb = a.get_state(), b.get_state()
    assert abs(sa["residual_after"] - sb["residual_after"]) < 1e-12 * sa["residual_before"]
    assert np.abs(ya - yb).max() <= 1e-13 * np.abs(ya).max()


def test_gauss_ic_config_256():
    n, T = 256, 0.5
    P = ewpt.Parameters(N=n, dx=DX, gy=GY, T=T)
    p = P.step_params()
    print("Gauss-law IC at 256^3", flush=True)
    with ewpt.Lattice(n, DX) as lat:
        st = lat.init_thermal_gauss(T, P.meff_sq, P.base_seed, p)
        g0 = lat.gauss(p)
        e0 = lat.energies(p)
        print(f"256^3 solve: {st}", flush=True)
        assert st["residual_after"] < 1e-12 * st["residual_before"], st
        lat.step(p, 200)
        g200 = lat.gauss(p)
        lat.step(p, 300)
        g1 = lat.gauss(p)
        e1 = lat.energies(p)
    with ewpt.Lattice(n, DX) as lat:
        lat.init_thermal(T, P.meff_sq, P.base_seed, "config")   # pi = E = 0
        lat.step(p, 200)
        gz200 = lat.gauss(p)
        lat.step(p, 300)
        g_zero = lat.gauss(p)
    G0, G200, G1, Gz200, Gz = (float(np.sqrt(np.sum(np.asarray(g[:4]) ** 2))) for g in (g0, g200, g1, gz200, g_zero))
    drift = (sum(e1[:7]) - sum(e0[:7])) / sum(e0[:7])
    print(f"Gauss residual: start {G0:.3e}, after 200 / 500 steps {G200:.3e} / {G1:.3e} "
          f"(from pi = E = 0: {Gz200:.3e} / {Gz:.3e}); energy drift {drift:.3e}", flush=True)
    assert G0 < 1e-12 * st["residual_before"]
    # the residual the equations build up (the k_Gamma / k_Sigma terms do not
    # keep div E = rho) is that of the reference's own pi = E = 0 start, far
    # below the unprojected draw's
    assert G200 < 0.1 * st["residual_before"]
    assert G1 < 2.0 * Gz and G1 < 0.2 * st["residual_before"]
    assert abs(drift) < 1e-3
