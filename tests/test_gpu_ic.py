"""Device initial conditions (ewpt_init_bubbles, ewpt_init_thermal) against
the oracle (oracle/ic_np.py), and a reference-configuration run from a bubble
IC evolved on the device and in the C oracle.

Tolerances: the bubble profiles are elementwise transcendental functions,
compared to 1e-13 absolute (|phi| <= O(1)); the thermal state goes through
an FFT of N^3 terms in a different order than numpy's, compared to 1e-12 of
the field's maximum.  The mixed context holds the state as an FP32 pair, whose
FP64 view reproduces the FP64 value to about 2^-48 relative.
"""
import math

import numpy as np
import pytest

import ewpt
from ewpt import ic
from oracle import cref, fields, ic_np

pytestmark = pytest.mark.gpu

MH = 2.0 * math.sqrt(0.129)
MEFF = 8119.43 / 174.0 ** 2
SHAPE = (32, 24, 16)     # nx, ny, nz: distinct so an axis mix-up shows


@pytest.mark.parametrize("precision", [ewpt.EWPT_FP64, ewpt.EWPT_MIXED])
@pytest.mark.parametrize("profile", ["ref", "config"])
def test_bubbles_match_oracle(profile, precision):
    print(f"bubbles profile={profile} precision={precision} shape={SHAPE}", flush=True)
    dx = 0.25
    if profile == "ref":
        t = ic.bubble_table(SHAPE, 6, 123456, "ref", bub_diam=4)
    else:
        t = ic.bubble_table(SHAPE, 6, 123456, "config")
    with ewpt.Lattice(SHAPE, dx, precision=precision) as lat:
        lat.init_bubbles(t, profile, MH)
        got = lat.get_state()
    want = ic_np.bubbles_state(SHAPE, dx, t, ic.profile_id(profile), MH)
    assert np.max(np.abs(got - want)) < 1e-13
    assert np.all(got[4:] == 0.0)


@pytest.mark.parametrize("precision", [ewpt.EWPT_FP64, ewpt.EWPT_MIXED])
@pytest.mark.parametrize("profile", ["ref", "config"])
def test_thermal_matches_oracle(profile, precision):
    print(f"thermal profile={profile} precision={precision} shape={SHAPE}", flush=True)
    dx, T, seed = 0.25, 0.5, 987654321
    with ewpt.Lattice(SHAPE, dx, precision=precision) as lat:
        lat.init_thermal(T, MEFF, seed, profile)
        got = lat.get_state()
    want = ic_np.thermal_state(SHAPE, dx, T, MEFF, seed, ic.profile_id(profile))
    scale = np.max(np.abs(want[:20]))
    assert scale > 0
    assert np.max(np.abs(got - want)) < 1e-12 * scale
    assert np.all(got[20:] == 0.0)


def test_thermal_massless_zero_mode_and_errors():
    with ewpt.Lattice((16, 16, 16), 0.5) as lat:
        lat.init_thermal(1.0, 0.0, 3, "config")         # omega = 0 mode left empty
        s = lat.get_state()
        assert np.all(np.isfinite(s))
        want = ic_np.thermal_state((16, 16, 16), 0.5, 1.0, 0.0, 3, 1)
        assert np.max(np.abs(s - want)) < 1e-12 * np.max(np.abs(want))
        for bad in [dict(T=0.0), dict(T=float("nan")), dict(meff_sq=float("inf")), dict(profile=2)]:
            kw = dict(T=1.0, meff_sq=0.0, seed=1, profile=1)
            kw.update(bad)
            with pytest.raises((ewpt.EwptError, ValueError)):
                lat.init_thermal(**kw)
        row = np.array([[1.0, 2.0, 3.0, 0.0, 0.0, 0.0, 1.0]])
        with pytest.raises(ewpt.EwptError):
            lat.init_bubbles(row, "ref", -1.0)
        with pytest.raises(ewpt.EwptError):
            lat.init_bubbles(np.where(row == 2.0, np.nan, row), "ref", MH)
        lat.init_bubbles(np.zeros((0, 7)), "ref", MH)   # no bubbles: zero state
        assert not np.any(lat.get_state())


def test_thermal_constraints_256():
    """At 256^3 (profile 1) the auxiliary constraints Gamma = div W,
    Sigma = div Y hold to rounding and, with E = pi = 0, so does the Gauss law."""
    n, dx = 256, 0.25
    print(f"thermal constraints n={n}", flush=True)
    p = fields.params()
    with ewpt.Lattice(n, dx) as lat:
        lat.init_thermal(0.5, MEFF, 42, "ref")
        g0 = lat.gauss(p)          # Gamma = Sigma = 0: slots 4..7 are ||div W^a||, ||div Y||
        lat.init_thermal(0.5, MEFF, 42, "config")
        g = lat.gauss(p)
        e = lat.energies(p)
    assert np.all(g[:4] == 0.0) and np.all(g0[:4] == 0.0)
    assert np.min(g0[4:]) > 0
    assert np.max(g[4:]) < 1e-12 * np.max(g0[4:])
    assert np.all(np.isfinite(e)) and e[2] > 0 and e[0] > 0


@pytest.mark.parametrize("profile", ["ref", "config"])
def test_reference_bubble_run_64_200_steps(profile):
    """The CPU-runnable reference configuration: 64^3, dx = 0.25, dt = dx/10,
    10 seeded bubbles, gauge fields and momenta zero, 200 RK4 steps; with the
    reference's profile (sub-lattice centres, its sech-like profile) and with
    the configuration's literal one (uniform centres, uniform SU(2)
    orientations, (1 - tanh)/2 wall of radius 3/mH).  The device run from the
    device-built IC matches the C oracle's run from the same state to 1e-11
    relative L2 per field, with matching energies and Gauss residuals."""
    n, dx = 64, 0.25
    P = ewpt.Parameters(N=n, dx=dx, gy=0.35)
    t = ic.bubble_table(n, 10, P.base_seed, profile, bub_diam=P.bub_diam)
    p = fields.params(gy=0.35, dt=P.dt)
    print(f"bubble run 64^3 x 200, profile {profile}", flush=True)
    with ewpt.Lattice(n, dx) as lat:
        lat.init_bubbles(t, profile, P.mH)
        s0 = lat.get_state()
        lat.step(p, 200)
        got = lat.get_state()
        e_dev, g_dev = lat.energies(p), lat.gauss(p)
    np.testing.assert_allclose(s0, ic_np.bubbles_state((n, n, n), dx, t, ic.profile_id(profile), P.mH),
                               atol=1e-13, rtol=0)
    want = cref.rk4(s0, p, dx, 200)
    num = np.sqrt(((got - want) ** 2).sum(axis=(1, 2, 3)))
    den = np.sqrt((want ** 2).sum(axis=(1, 2, 3)))
    moving = den > 1e-6 * den.max()
    assert np.max(num[moving] / den[moving]) < 1e-11
    assert np.all(num[~moving] < 1e-11 * den.max())
    e_ref, g_ref = cref.energies(want, p, dx), cref.gauss(want, p, dx)
    np.testing.assert_allclose(e_dev[:7], e_ref[:7], rtol=1e-11, atol=1e-14)
    np.testing.assert_allclose(g_dev, g_ref, rtol=1e-8, atol=1e-12)
