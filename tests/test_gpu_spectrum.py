"""Device EM diagnostics and shell spectrum against the oracle.

Tolerances: the EM fields are pointwise FP64 expressions of stencil values,
compared to 1e-12 of each channel's maximum; the energies are sums over the
lattice (1e-12 relative); the spectrum passes through a D2Z transform and an
atomic histogram, whose summation order differs from numpy's (1e-10 relative
per shell, or 1e-14 of the largest shell where the shell power is tiny).
"""
import math

import numpy as np
import pytest

import ewpt
from oracle import cref, fields, spec_np

pytestmark = pytest.mark.gpu

TH = math.asin(math.sqrt(0.22))
DX = 0.25
SHAPE = (32, 24, 16)


def _state(seed, shape=SHAPE):
    s = fields.noise_state(shape, seed=seed)
    s[3] += 1.0          # keep |phi| away from 0 so B_2 is well scaled
    return s


@pytest.mark.parametrize("shape", [SHAPE, (20, 12, 12)])
def test_em_fields_and_energies_match_oracle(shape):
    """SHAPE takes the brick reductions, (20, 12, 12) the gather kernels."""
    s = _state(21, shape)
    p = fields.params()
    print(f"em fields shape={shape}", flush=True)
    with ewpt.Lattice(shape, DX) as lat:
        lat.set_state(s)
        got = lat.em_fields(p, TH)
        e = lat.em_energies(p, TH)
        row = lat.energy_row(p, TH)
    want = cref.em_fields(s, p, DX, TH)
    for c in range(9):
        assert np.max(np.abs(got[c] - want[c])) < 1e-12 * np.max(np.abs(want[c])), c
    e_ref = spec_np.em_energies(want, DX)
    np.testing.assert_allclose(e, e_ref, rtol=1e-12)
    np.testing.assert_allclose(row[[7, 8, 12]], e_ref, rtol=1e-12)


@pytest.mark.parametrize("field", ["B", "B2"])
@pytest.mark.parametrize("cut", [None, (32, 24, 16)])
def test_spectrum_matches_oracle(field, cut):
    s = _state(22)
    p = fields.params()
    print(f"spectrum field={field} cut={cut}", flush=True)
    with ewpt.Lattice(SHAPE, DX) as lat:
        lat.set_state(s)
        k, pw = lat.spectrum(p, TH, field=field, cut=cut)
    em = cref.em_fields(s, p, DX, TH)
    B = em[0:3] if field == "B" else em[3:6]
    c = (SHAPE[0] // 4, SHAPE[1] // 4, SHAPE[2] // 4) if cut is None else cut
    kr, pr = spec_np.spectrum(B.copy(), c)
    assert k.shape == kr.shape
    np.testing.assert_array_equal(k, kr)
    np.testing.assert_allclose(pw, pr, rtol=1e-10, atol=1e-14 * pr.max())
    if cut == SHAPE:   # every shell: Parseval
        k2 = spec_np.k2_signed(SHAPE[::-1])
        cnt = np.array([(k2 == round(v * v)).sum() for v in k])
        assert np.sum(pw * cnt) == pytest.approx(np.prod(SHAPE) * np.sum(B ** 2), rel=1e-11)


def test_spectrum_single_mode_128():
    """Photon-limit state with one Fourier mode: B_y = cos(th) d_x Y_z, so the
    shell |K| = m holds (N^3/2)^2 (cos(th) a k_eff)^2 * 2 / (points on the shell)."""
    n, m, a = 128, 5, 0.3
    p = fields.params()
    s = np.zeros((36, n, n, n))
    s[3] = 1.0
    x = np.arange(n)
    s[15] = a * np.cos(2 * np.pi * m * x / n)[None, None, :]       # Y_z(x)
    t = 2 * np.pi * m / n
    keff = (45 * math.sin(t) - 9 * math.sin(2 * t) + math.sin(3 * t)) / (30 * DX)
    print(f"spectrum single mode n={n}", flush=True)
    with ewpt.Lattice(n, DX) as lat:
        lat.set_state(s)
        k, pw = lat.spectrum(p, TH)
        e = lat.em_energies(p, TH)
    k2 = spec_np.k2_signed((n, n, n))
    shell = int((k2 == m * m).sum())
    i = int(np.argmin(np.abs(k - m)))
    want = 2 * (n ** 3 / 2) ** 2 * (math.cos(TH) * a * keff) ** 2 / shell
    assert pw[i] == pytest.approx(want, rel=1e-12)
    assert np.max(np.delete(pw, i)) < 1e-20 * want
    # 0.5 sum B^2 dx^3 with <cos^2> = 1/2
    assert e[0] == pytest.approx(0.5 * (math.cos(TH) * a * keff) ** 2 * 0.5 * n ** 3 * DX ** 3, rel=1e-12)


def test_mixed_context_em_matches_fp64():
    s = _state(23)
    p = fields.params()
    with ewpt.Lattice(SHAPE, DX) as lat:
        lat.set_state(s)
        e64 = lat.em_energies(p, TH)
        k64, p64 = lat.spectrum(p, TH)
    with ewpt.Lattice(SHAPE, DX, precision=ewpt.EWPT_MIXED) as lat:
        lat.set_state(s)
        em = lat.em_energies(p, TH)
        km, pm = lat.spectrum(p, TH)
    np.testing.assert_allclose(em, e64, rtol=1e-12)
    np.testing.assert_array_equal(km, k64)
    np.testing.assert_allclose(pm, p64, rtol=1e-10, atol=1e-14 * p64.max())


def test_spectrum_argument_errors():
    p = fields.params()
    with ewpt.Lattice((16, 16, 16), DX) as lat:
        with pytest.raises(ewpt.EwptError):
            lat.em_energies(p, TH)                          # no state yet
        lat.set_state(_state_small())
        with pytest.raises(ewpt.EwptError):
            lat.spectrum(p, TH, cut=(64, 64, 64))           # more shells than the lattice has
        with pytest.raises(ValueError):
            lat.spectrum(p, TH, field="E")
        k, pw = lat.spectrum(p, TH, cut=(4, 4, 4))
        assert len(k) == spec_np.kc_bin_nums(4, 4, 4)


def _state_small():
    s = fields.noise_state((16, 16, 16), seed=5)
    s[3] += 1.0
    return s
