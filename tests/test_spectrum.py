"""EM-field and spectrum oracle pins (CPU only).

The C transcription (oracle_em_fields) and the numpy vector form
(oracle/spec_np.py) check each other; both are then pinned by limits the
physics fixes: in the unitary vacuum with only W^3 and Y the magnetic field is
minus the lattice curl of the photon field A = sin(th) W^3 + cos(th) Y, a pure
Higgs configuration gives the X_i term alone, and n^a is a unit vector.  The
shell spectrum is pinned by Parseval, a single Fourier mode, and a literal
count of K_c_mag values.
"""
import math

import numpy as np
import pytest

from oracle import cref, fields, spec_np

TH = math.asin(math.sqrt(0.22))
DX = 0.25


def keff(m, n, dx=DX):
    """Symbol of the 6th-order first derivative on exp(i 2 pi m x/n)."""
    t = 2.0 * math.pi * m / n
    return (45.0 * math.sin(t) - 9.0 * math.sin(2 * t) + math.sin(3 * t)) / (30.0 * dx)


def test_c_and_numpy_em_fields_agree():
    s = fields.noise_state((16, 12, 20), seed=9)
    p = fields.params()
    a = cref.em_fields(s, p, DX, TH)
    b = spec_np.em_fields(s, p, DX, TH)
    for c in range(9):
        scale = np.max(np.abs(b[c]))
        assert scale > 0
        assert np.max(np.abs(a[c] - b[c])) < 1e-12 * scale, c


def test_photon_limit_is_minus_curl_of_A():
    n = 16
    p = fields.params()
    s = np.zeros((36, n, n, n))
    s[3] = 1.0                                      # phi = (0, 0, 0, v), v = 1
    x = np.arange(n)
    X = x[None, None, :]
    Yc = x[None, :, None]
    Z = x[:, None, None]
    s[4 + 6 + 1] = 0.3 * np.cos(2 * np.pi * 2 * X / n)        # W^3_y(x)
    s[13 + 2] = 0.2 * np.cos(2 * np.pi * 1 * Yc / n)          # Y_z(y)
    s[13 + 0] = 0.1 * np.cos(2 * np.pi * 3 * Z / n + 0.4)     # Y_x(z)
    em = cref.em_fields(s, p, DX, TH)
    sw, cw = math.sin(TH), math.cos(TH)
    # A = sin th W^3 + cos th Y (n = (0, 0, 1))
    np.testing.assert_allclose(em[6 + 1], sw * s[11] + 0 * X, atol=1e-15)
    np.testing.assert_allclose(em[6 + 2], cw * s[15] + 0 * X, atol=1e-15)
    # B = -curl A, each derivative through the stencil symbol
    Bx = -(cw * (-0.2 * keff(1, n) * np.sin(2 * np.pi * Yc / n)))                  # -d_y A_z
    By = -(cw * (-0.1 * keff(3, n) * np.sin(2 * np.pi * 3 * Z / n + 0.4)))         # -d_z A_x
    Bz = -(sw * (-0.3 * keff(2, n) * np.sin(2 * np.pi * 2 * X / n)))               # -d_x A_y
    for c, want in enumerate((Bx, By, Bz)):
        np.testing.assert_allclose(em[c], np.broadcast_to(want, em[c].shape), atol=1e-13)
        np.testing.assert_allclose(em[3 + c], em[c], atol=1e-13)     # |phi| = v: B_2 = B
    b = spec_np.em_fields(s, p, DX, TH)
    np.testing.assert_allclose(b, em, atol=1e-13)


def test_pure_higgs_gives_the_x_term():
    n = 16
    p = fields.params()
    s = np.zeros((36, n, n, n))
    x = np.arange(n)
    a, b_ = 0.7, 0.5
    s[0] = a * np.cos(2 * np.pi * 1 * x / n)[None, None, :]     # phi_1(x)
    s[1] = b_ * np.cos(2 * np.pi * 2 * x / n)[None, :, None]    # phi_2(y)
    s[3] = 0.9                                                   # keeps |phi| > 0
    em = cref.em_fields(s, p, DX, TH)
    c1 = 4.0 * math.sin(TH) / (p["gw"] * p["vev"] ** 2)
    sx = np.sin(2 * np.pi * x / n)[None, None, :]
    sy = np.sin(2 * np.pi * 2 * x / n)[None, :, None]
    Xz = (-a * keff(1, n) * sx) * (-b_ * keff(2, n) * sy)       # D_x phi_1 D_y phi_2
    np.testing.assert_allclose(em[0], 0.0, atol=1e-14)
    np.testing.assert_allclose(em[1], 0.0, atol=1e-14)
    np.testing.assert_allclose(em[2], np.broadcast_to(-c1 * Xz, em[2].shape), atol=1e-13)
    mag = (s[:4] ** 2).sum(0)
    c2 = 4.0 * math.sin(TH) / (p["gw"] * mag ** 2)
    np.testing.assert_allclose(em[5], -c2 * Xz, atol=1e-12)


def test_n_is_a_unit_vector():
    """With W^a_x = delta_ab and Y = 0, A_x = sin(th) n^b."""
    n = 8
    p = fields.params()
    s = fields.noise_state((n, n, n), seed=3)
    s[4:16] = 0.0
    comps = []
    for b in range(3):
        t = s.copy()
        t[4 + 3 * b] = 1.0
        comps.append(cref.em_fields(t, p, DX, TH)[6] / math.sin(TH))
    nn = np.sqrt(sum(c ** 2 for c in comps))
    np.testing.assert_allclose(nn, 1.0, atol=1e-14)


def test_kc_bin_nums_literal():
    """Kc_bin_nums counts distinct K_c_mag values (spec_routines.jl:5-43)."""
    def k_c_mag(x, y, z, nx, ny, nz):        # 1-based, as the reference
        ks = []
        for v, nn in ((x, nx), (y, ny), (z, nz)):
            v = v - 1
            ks.append(v if v <= nn // 2 else v - nn)
        return math.sqrt(ks[0] ** 2 + ks[1] ** 2 + ks[2] ** 2)
    for cut in ((4, 4, 4), (8, 6, 5), (16, 16, 16)):
        vals = {k_c_mag(i, j, k, *cut) for i in range(1, cut[0] + 1) for j in range(1, cut[1] + 1)
                for k in range(1, cut[2] + 1)}
        assert spec_np.kc_bin_nums(*cut) == len(vals)


def test_spectrum_parseval_and_single_mode():
    n = 32
    rng = np.random.default_rng(4)
    B = rng.standard_normal((3, n, n, n))
    k, pw = spec_np.spectrum(B, (n, n, n))                 # every shell
    k2 = spec_np.k2_signed((n, n, n))
    cnt = np.array([(k2 == round(v * v)).sum() for v in k])
    assert cnt.sum() == n ** 3
    assert np.sum(pw * cnt) == pytest.approx(n ** 3 * np.sum(B ** 2), rel=1e-12)
    assert np.all(np.diff(k) > 0) and k[0] == 0.0
    # one mode: B_y = cos(2 pi 3 x/n) -> power (n^3/2)^2 at K = (+-3, 0, 0)
    x = np.arange(n)
    B1 = np.zeros((3, n, n, n))
    B1[1] = np.cos(2 * np.pi * 3 * x / n)[None, None, :]
    k, pw = spec_np.spectrum(B1, (n // 4,) * 3)
    i3 = int(np.argmin(np.abs(k - 3.0)))
    shell = (k2 == 9).sum()
    assert pw[i3] == pytest.approx(2 * (n ** 3 / 2) ** 2 / shell, rel=1e-12)
    others = np.delete(pw, i3)
    assert np.max(others) < 1e-20 * pw[i3]
