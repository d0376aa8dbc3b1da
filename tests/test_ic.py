"""Initial-condition oracle pins and host-side IC logic (CPU only).

The oracle's IC functions (oracle/ic_np.py) are pinned here by facts that do
not restate their formulas: the bubble profile against its sech closed form,
the distance quirk by its symmetries, the orientations by the moments of the
uniform distribution on S^3, the counter-based normal draws by their
moments, and the thermal field by recovering the Bose-Einstein spectrum and
the transversality of the gauge modes from an FFT of the generated field.
"""
import math

import numpy as np
import pytest

import ewpt
from ewpt import ic
from oracle import cref, fields, ic_np

MH = 2.0 * math.sqrt(0.129)
MEFF = 8119.43 / 174.0 ** 2


# ------------------------------------------------------------------ RNG

def test_gauss_draw_moments():
    q = np.arange(2_000_000, dtype=np.uint64)
    x = ic_np.gauss_sqrt2(123456, 3, q)
    assert abs(x.mean()) < 0.005
    assert abs(x.var() - 2.0) < 0.01                 # Normal(0, sqrt2)
    assert abs(np.mean(x ** 4) / 4.0 - 3.0) < 0.03   # Gaussian kurtosis
    y = ic_np.gauss_sqrt2(123456, 4, q)
    assert abs(np.corrcoef(x, y)[0, 1]) < 0.003      # streams independent
    z = ic_np.gauss_sqrt2(123457, 3, q)
    assert abs(np.corrcoef(x, z)[0, 1]) < 0.003      # seeds independent
    assert np.array_equal(x[:1000], ic_np.gauss_sqrt2(123456, 3, q[:1000]))


# ------------------------------------------------------------------ bubbles

def test_bubble_profile_closed_form():
    r = np.linspace(0.0, 20.0, 2001)
    for mH in (MH, 1.0, 2.5):
        closed = math.sqrt(2.0) / np.cosh(mH * r / math.sqrt(2.0) + math.asinh(1.0))
        np.testing.assert_allclose(ic_np.rmag_ref(r, mH), closed, rtol=1e-13, atol=1e-300)
    assert ic_np.rmag_ref(0.0, MH) == pytest.approx(1.0, abs=1e-15)


def test_tanh_profile_values():
    assert ic_np.rmag_tanh(3.0 / MH, MH) == pytest.approx(0.5, abs=1e-15)
    assert ic_np.rmag_tanh(0.0, MH) == pytest.approx(1.0 / (1.0 + math.exp(-3.0)), rel=1e-14)
    assert ic_np.rmag_tanh(100.0, MH) < 1e-10


def test_reference_distance_quirks():
    n, dx = 64, 0.25
    ib = 20.0
    r = ic_np.dist_ref((n, n, n), dx, ib, ib, ib)
    # the centre sits half a cell below ib: sites ib-1 and ib (0-based) tie
    c = int(ib)
    assert r[c, c, c] == pytest.approx(r[c - 1, c - 1, c - 1], rel=1e-14)
    assert r[c, c, c] == pytest.approx(math.sqrt(3) * 0.5 * dx, rel=1e-3)
    # near the centre the chord distance is the Euclidean one, shortened by
    # at most sin x >= x (1 - x^2/6) per axis, x = rk * distance
    e = ic_np.dist_min_image((n, n, n), dx, ib - 0.5, ib - 0.5, ib - 0.5)
    rk = math.pi / ((n - 1) * dx)
    near = e < 2.0
    assert np.all(r[near] <= e[near] * (1 + 1e-14))
    assert np.all(r[near] >= e[near] * (1 - (rk * e[near]) ** 2 / 6) - 1e-14)
    # period N-1 rather than N (rk = pi/((N-1) dx)): sites 0 and N-1 tie
    assert r[0, c, c] == pytest.approx(r[n - 1, c, c], rel=1e-13)


def test_min_image_distance():
    n, dx = 16, 0.5
    r = ic_np.dist_min_image((n, n, n), dx, n - 1.0, 0.0, 0.0)
    assert r[0, 0, 0] == pytest.approx(dx)           # wraps across the boundary
    assert r[0, 0, n - 1] == 0.0


def test_single_bubble_state():
    shape, dx = (24, 20, 16), 0.25
    p = ic.gen_phi((0.3, 0.6, 0.1), -1.0)
    row = np.array([[10, 8, 6, *p]], dtype=np.float64)
    s = ic_np.bubbles_state(shape, dx, row, 0, MH)
    mod = np.sqrt((s[:4] ** 2).sum(0))
    np.testing.assert_allclose(s[:4] / mod, np.broadcast_to(p[:, None, None, None], s[:4].shape), atol=1e-14)
    assert np.all(s[4:] == 0.0)
    closed = math.sqrt(2.0) / np.cosh(MH * ic_np.dist_ref(shape, dx, 10, 8, 6) / math.sqrt(2.0) + math.asinh(1.0))
    np.testing.assert_allclose(mod, closed, rtol=1e-13)
    two = np.vstack([row, [[3, 15, 12, *ic.gen_phi((0.9, 0.2, 0.7), 1.0)]]])
    both = ic_np.bubbles_state(shape, dx, two, 0, MH)
    np.testing.assert_allclose(both, s + ic_np.bubbles_state(shape, dx, two[1:], 0, MH), atol=1e-15)


# ------------------------------------------------------------------ orientations and host logic

def test_gen_phi_uniform_on_s3():
    rng = np.random.default_rng(7)
    abc = rng.uniform(size=(40000, 3))
    sg = rng.choice([-1.0, 1.0], size=40000)
    p = np.array([ic.gen_phi(a, s) for a, s in zip(abc, sg)])
    np.testing.assert_allclose(np.linalg.norm(p, axis=1), 1.0, atol=1e-15)
    np.testing.assert_allclose((p ** 2).mean(0), 0.25, atol=0.006)
    np.testing.assert_allclose((p ** 4).mean(0), 1.0 / 8.0, atol=0.005)   # E x^4 = 3/(n(n+2)), n = 4
    assert abs(np.mean(p[:, 0] * p[:, 2])) < 0.005
    # the oracle's complex form agrees with the host's real form
    for a, s in zip(abc[:50], sg[:50]):
        np.testing.assert_allclose(ic.gen_phi(a, s), ic_np.gen_phi(a, s), atol=1e-15)


def test_bubble_table_reference_profile():
    P = ewpt.Parameters(N=64, dx=0.25)
    assert P.bub_diam == 12
    t = ic.bubble_table(64, 10, P.base_seed, "ref", bub_diam=P.bub_diam)
    assert t.shape == (10, 7)
    assert np.all(t[:, :3] % 12 == 0) and np.all(t[:, :3] >= 12) and np.all(t[:, :3] <= 64)
    assert len({tuple(r) for r in t[:, :3]}) == 10                 # without replacement
    np.testing.assert_allclose(np.linalg.norm(t[:, 3:], axis=1), 1.0, atol=1e-15)
    assert np.array_equal(t, ic.bubble_table(64, 10, P.base_seed, "ref", bub_diam=12))
    two = ic.bubble_table(64, 2, P.base_seed, "ref", bub_diam=12)
    assert two[:, :3].tolist() == [[8, 32, 32], [56, 32, 32]]     # N/2 -+ 2 bub_diam
    with pytest.raises(ValueError):
        ic.bubble_table(64, 126, 1, "ref", bub_diam=12)            # only 5^3 sub-lattice sites
    with pytest.raises(ValueError):
        ic.bubble_table(64, 3, 1, "ref")
    # the Hopf draws behind the orientations: gen_phi of each row with one of
    # the two signs gives that row's orientation (uniforms in [0, 1))
    t2, hoft = ic.bubble_table(64, 10, P.base_seed, "ref", bub_diam=12, with_angles=True)
    assert np.array_equal(t2, t) and hoft.shape == (10, 3)
    assert np.all((hoft >= 0) & (hoft < 1))
    for row, h in zip(t, hoft):
        assert min(np.abs(ic.gen_phi(h, sg) - row[3:]).max() for sg in (-1.0, 1.0)) < 1e-15


def test_bubble_table_config_profile():
    t = ic.bubble_table((64, 48, 32), 20, 99, "config")
    assert t.shape == (20, 7)
    assert np.all(t[:, :3] >= 0) and np.all(t[:, :3] < np.array([64, 48, 32]))
    np.testing.assert_allclose(np.linalg.norm(t[:, 3:], axis=1), 1.0, atol=1e-15)
    assert ic.bubble_table((64, 48, 32), 20, 99, "config", with_angles=True)[1] is None
    with pytest.raises(ValueError):
        ic.bubble_table(64, 3, 1, "tanh")


# ------------------------------------------------------------------ thermal

def _shell_ratio(power, expect, modk, nbins=8, min_count=400):
    """Mean of power/expect over |k| shells; returns the list of shell means."""
    keep = (expect > 0) & np.isfinite(expect)
    edges = np.quantile(modk[keep], np.linspace(0, 1, nbins + 1))
    out = []
    for lo, hi in zip(edges[:-1], edges[1:]):
        m = keep & (modk >= lo) & (modk <= hi)
        if m.sum() >= min_count:
            out.append((power[m] / expect[m]).mean())
    return np.array(out)


@pytest.mark.parametrize("T", [0.5, 2.0])
def test_thermal_gauge_spectrum_and_transversality(T):
    shape, dx = (32, 32, 32), 0.25
    n3 = 32 ** 3
    s = ic_np.thermal_state(shape, dx, T, MEFF, 2024, 1)
    kx, ky, kz, modk = ic_np.lattice_k(shape, dx, 1)
    amp2 = ic_np.be_amplitude(modk, T, n3 * dx ** 3) ** 2
    for grp, ch in enumerate((4, 7, 10, 13)):
        F = np.fft.fftn(s[ch:ch + 3], axes=(-3, -2, -1))
        # transverse to the lattice momentum: sum_i k_i F_i = 0 mode by mode,
        # away from the Nyquist planes (there k(-q) != -k(q) on the lattice)
        div = kx * F[0] + ky * F[1] + kz * F[2]
        scale = modk.max() * np.sqrt((np.abs(F) ** 2).sum(0)).max()
        nyq = np.zeros(shape[::-1], dtype=bool)
        nyq[16, :, :] = nyq[:, 16, :] = nyq[:, :, 16] = True
        assert np.max(np.abs(div)[~nyq]) < 1e-13 * scale, grp
        # F(q) = (N^3/2)(A(q) + conj A(-q)) and <sum_i |A_i|^2> = 8 amp^2 (four
        # N(0, sqrt2) draws on two unit polarisations), so
        # <sum_i |F_i|^2> = (N^3/2)^2 * 16 amp^2
        power = (np.abs(F) ** 2).sum(0) / ((n3 / 2.0) ** 2 * 16.0)
        r = _shell_ratio(power, amp2, modk)
        assert len(r) >= 6
        np.testing.assert_allclose(r, 1.0, atol=0.12)


def test_thermal_higgs_spectrum():
    shape, dx, T = (32, 32, 32), 0.25, 0.5
    n3 = 32 ** 3
    s = ic_np.thermal_state(shape, dx, T, MEFF, 77, 1)
    modk = ic_np.lattice_k(shape, dx, 1)[3]
    amp2 = ic_np.be_amplitude(np.sqrt(modk ** 2 + MEFF), T, n3 * dx ** 3) ** 2
    for b in range(4):
        F = np.fft.fftn(s[b])
        power = np.abs(F) ** 2 / ((n3 / 2.0) ** 2 * 4.0)   # <|F|^2> = (N^3/2)^2 (2 + 2) amp^2, real draws
        r = _shell_ratio(power, amp2, modk)
        np.testing.assert_allclose(r, 1.0, atol=0.12)
    assert np.all(s[20:] == 0.0)


def test_thermal_reference_index_offset():
    """Profile 0 evaluates array slot m at mode number m+1 (the reference's
    1-based thread index used as the frequency), i.e. the momenta of
    profile 1 rolled by one slot."""
    shape, dx = (12, 10, 8), 0.5
    k0 = ic_np.lattice_k(shape, dx, 0)
    k1 = ic_np.lattice_k(shape, dx, 1)
    for a, b in zip(k0, k1):
        np.testing.assert_allclose(np.abs(a), np.abs(np.roll(b, (-1, -1, -1), axis=(0, 1, 2))), atol=1e-12)
    # 2 sin(pi m/N)/dx with m = 1..N never turns negative: the reference's
    # polarisations are built against a momentum with non-negative components
    assert all(np.all(a >= 0) for a in k0[:3])
    # the k = 0 mode lands on the last slot
    assert k0[3][-1, -1, -1] < 1e-7 and k1[3][0, 0, 0] == 0.0


def test_thermal_constraints_config_profile():
    shape, dx = (16, 12, 20), 0.25
    s = ic_np.thermal_state(shape, dx, 0.5, MEFF, 5, 1)
    g = cref.gauss(np.ascontiguousarray(s), fields.params(), dx)
    scale = np.sqrt((s[4:16] ** 2).sum() * dx ** 3) / dx
    assert np.all(g[:4] == 0.0)                      # E = pi = 0: Gauss law exact
    assert np.max(g[4:]) < 1e-13 * scale             # Gamma = div W, Sigma = div Y
    s0 = ic_np.thermal_state(shape, dx, 0.5, MEFF, 5, 0)
    assert np.all(s0[16:20] == 0.0)                  # reference leaves Gamma, Sigma at 0


@pytest.mark.parametrize("plant", ["no_vol", "omega_sq", "no_offset_shift"])
def test_thermal_spectrum_catches_planted_errors(plant, monkeypatch):
    """A planted mistake in the oracle's amplitude or mode numbering moves the
    recovered spectrum off the Bose-Einstein curve."""
    shape, dx, T = (32, 32, 32), 0.25, 0.5
    n3 = 32 ** 3
    good = ic_np.be_amplitude
    if plant == "no_vol":
        monkeypatch.setattr(ic_np, "be_amplitude", lambda w, T_, vol: good(w, T_, vol) * math.sqrt(vol))
    elif plant == "omega_sq":
        monkeypatch.setattr(ic_np, "be_amplitude", lambda w, T_, vol: good(w, T_, vol) / np.sqrt(np.where(w > 0, w, 1)))
    else:   # signed-index momenta replaced by the unsigned slot index
        real = ic_np.mode_numbers
        monkeypatch.setattr(ic_np, "mode_numbers", lambda sh, prof: real(sh, 0) if prof == 1 else real(sh, prof))
    s = ic_np.thermal_state(shape, dx, T, MEFF, 2024, 1)
    monkeypatch.undo()
    modk = ic_np.lattice_k(shape, dx, 1)[3]
    amp2 = ic_np.be_amplitude(modk, T, n3 * dx ** 3) ** 2
    F = np.fft.fftn(s[4:7], axes=(-3, -2, -1))
    r = _shell_ratio((np.abs(F) ** 2).sum(0) / ((n3 / 2.0) ** 2 * 16.0), amp2, modk)
    assert np.max(np.abs(r - 1.0)) > 0.3
