"""Pins of the initial Gauss-law solve's CPU restatement (oracle/gauss_np.py),
CPU only, by facts that do not restate its formulas:

* the projected state satisfies the Gauss law as the C oracle's independent
  residual code measures it (the k_Gamma / k_Sigma charges of
  main_rk4.jl:603-657): a wrong sign or a dropped term in the restatement's
  Gauss operator or charges leaves a residual there;
* the Gauss operator and the covariant gradient are adjoint up to sign,
  <G E, chi> = -<E, D chi> (a planted sign slip in the W-term fails it);
* the U(1) correction is a lattice gradient (its lattice curl vanishes) and
  the projection leaves divergence-free parts of E untouched;
* the momentum draws carry the Bose-Einstein spectrum n*omega/vol shell by
  shell (gauge modes transverse), and the pi correction removes the U(1)
  charge's component on the modes the d1 stencil annihilates.
"""
import math

import numpy as np
import pytest

from oracle import cref, fields, gauss_np, ic_np

MEFF = 8119.43 / 174.0 ** 2
GW, GY = 0.65, 0.35


def test_projected_state_satisfies_c_oracle_gauss_law():
    shape, dx = (16, 16, 16), 0.25
    s, info = gauss_np.thermal_gauss_state(shape, dx, 0.5, MEFF, 3, GW, GY)
    g = cref.gauss(np.ascontiguousarray(s), fields.params(gw=GW, gy=GY), dx)
    assert info["before"] > 0.1
    assert np.sqrt(np.sum(g[:4] ** 2)) < 1e-12 * info["before"], (g, info)
    assert info["after"] < 1e-12 * info["before"]
    assert np.abs(s[20:36]).max() > 1e-3


@pytest.mark.parametrize("plant", ["su2_charge_sign", "u1_charge_sign", "no_w_term"])
def test_planted_mistakes_leave_a_c_oracle_residual(plant, monkeypatch):
    shape, dx = (12, 12, 12), 0.25
    good_q, good_g = gauss_np.higgs_charges, gauss_np.gauss_op
    if plant == "su2_charge_sign":
        monkeypatch.setattr(gauss_np, "higgs_charges", lambda f, q: (-good_q(f, q)[0], good_q(f, q)[1]))
    elif plant == "u1_charge_sign":
        monkeypatch.setattr(gauss_np, "higgs_charges", lambda f, q: (good_q(f, q)[0], -good_q(f, q)[1]))
    else:
        monkeypatch.setattr(gauss_np, "gauss_op", lambda E, W, gw, dx_: good_g(E, 0.0 * W, gw, dx_))
    s, info = gauss_np.thermal_gauss_state(shape, dx, 0.5, MEFF, 8, GW, GY)
    monkeypatch.undo()
    g = cref.gauss(np.ascontiguousarray(s), fields.params(gw=GW, gy=GY), dx)
    assert np.sqrt(np.sum(g[:4] ** 2)) > 1e-3 * info["before"], (plant, g)


def test_gauss_operator_is_minus_adjoint_of_covariant_gradient():
    rng = np.random.default_rng(4)
    grid = (6, 5, 7)
    dx = 0.3
    W = rng.standard_normal((3, 3) + grid)
    E = rng.standard_normal((3, 3) + grid)
    chi = rng.standard_normal((3,) + grid)
    lhs = np.sum(gauss_np.gauss_op(E, W, GW, dx) * chi)
    rhs = -np.sum(E * gauss_np.cov_grad(chi, W, GW, dx))
    assert abs(lhs - rhs) < 1e-12 * abs(lhs)
    # a planted sign slip in the W term of the gradient breaks the identity
    bad = gauss_np.cov_grad(chi, -W, GW, dx)
    assert abs(lhs + np.sum(E * bad)) > 1e-3 * abs(lhs)


def test_u1_correction_is_a_lattice_gradient():
    shape, dx = (12, 10, 8), 0.25
    rng = np.random.default_rng(5)
    EY0 = rng.standard_normal((3,) + shape[::-1])
    QY = rng.standard_normal(shape[::-1])
    for m in gauss_np.null_modes(shape):          # a charge the lattice divergence can reach
        e = gauss_np.sign_pattern(shape, m)
        QY -= e * np.sum(QY * e) / np.sum(e * e)
    EY = gauss_np.solve_u1(EY0, QY, GY, dx, shape)
    d = lambda A, i: gauss_np.d1(A, gauss_np._ax(i), dx)
    div = sum(d(EY[i], i) for i in range(3))
    np.testing.assert_allclose(div, GY * QY, atol=1e-11 * np.abs(GY * QY).max())
    dE = EY - EY0
    for i, j in ((0, 1), (1, 2), (0, 2)):
        curl = d(dE[j], i) - d(dE[i], j)
        assert np.abs(curl).max() < 1e-11 * np.abs(dE).max()
    # a divergence-free field with zero charge is left as it is
    psi = rng.standard_normal(shape[::-1])
    Ecurl = np.stack([d(psi, 1), -d(psi, 0), np.zeros_like(psi)])      # d_x E_x + d_y E_y = 0
    np.testing.assert_allclose(gauss_np.solve_u1(Ecurl, np.zeros_like(psi), GY, dx, shape), Ecurl,
                               atol=1e-12 * np.abs(Ecurl).max())


def test_pi_correction_clears_the_unreachable_charge():
    shape = (8, 6, 10)
    rng = np.random.default_rng(6)
    phi = rng.standard_normal((4,) + shape[::-1])
    pi = rng.standard_normal((4,) + shape[::-1])
    pi2 = gauss_np.correct_pi(phi, pi, GY, shape)
    _, QY = gauss_np.higgs_charges(phi, pi2)
    _, QY0 = gauss_np.higgs_charges(phi, pi)
    for m in gauss_np.null_modes(shape):
        e = gauss_np.sign_pattern(shape, m)
        assert abs(np.sum(QY * e)) < 1e-12 * np.abs(QY0).sum()
    assert len(gauss_np.null_modes(shape)) == 8
    assert len(gauss_np.null_modes((7, 6, 10))) == 4       # odd side: only 0 on that axis


def _shell_means(power, expect, modk, nbins=8, min_count=400):
    keep = (expect > 0) & np.isfinite(expect)
    edges = np.quantile(modk[keep], np.linspace(0, 1, nbins + 1))
    out = []
    for lo, hi in zip(edges[:-1], edges[1:]):
        m = keep & (modk >= lo) & (modk <= hi)
        if m.sum() >= min_count:
            out.append((power[m] / expect[m]).mean())
    return np.array(out)


def test_momenta_carry_bose_einstein_spectrum():
    shape, dx, T = (32, 32, 32), 0.25, 0.5
    n3 = 32 ** 3
    vol = n3 * dx ** 3
    kx, ky, kz, modk = ic_np.lattice_k(shape, dx, 1)
    kpos = np.where(modk > 0, modk, 1.0)
    expect_g = np.where(modk > 0, kpos / np.expm1(kpos / T) / vol, 0.0)   # n omega / vol, omega = |k|
    for grp in (0, 3):
        E = ic_np.bfft(gauss_np.momentum_modes_gauge(shape, dx, T, 17, grp)).real
        F = np.fft.fftn(E, axes=(-3, -2, -1))
        div = kx * F[0] + ky * F[1] + kz * F[2]
        nyq = np.zeros(shape[::-1], dtype=bool)
        nyq[16, :, :] = nyq[:, 16, :] = nyq[:, :, 16] = True
        assert np.max(np.abs(div)[~nyq]) < 1e-12 * modk.max() * np.abs(F).max()
        power = (np.abs(F) ** 2).sum(0) / ((n3 / 2.0) ** 2 * 16.0)
        r = _shell_means(power, expect_g, modk)
        assert len(r) >= 6
        np.testing.assert_allclose(r, 1.0, atol=0.12)
    w = np.sqrt(modk ** 2 + MEFF)
    expect_h = w / np.expm1(w / T) / vol
    pi = ic_np.bfft(gauss_np.momentum_modes_higgs(shape, dx, T, MEFF, 17, 2)).real
    power = np.abs(np.fft.fftn(pi)) ** 2 / ((n3 / 2.0) ** 2 * 4.0)
    np.testing.assert_allclose(_shell_means(power, expect_h, modk), 1.0, atol=0.12)
