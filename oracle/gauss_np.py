"""CPU restatement of the thermal momenta and the initial Gauss-law solve
(test infrastructure only; tests/ use it as the checker of
``ewpt_init_thermal`` profile 2).

The reference starts thermal runs with pi = E = 0 and drives the Higgs with
an external mass term instead (gpu-kernels-rk4/main_rk4.jl:5154-5162,
5349-5638).  The benchmark configuration asks for thermal electric fields
that satisfy the Gauss law.  Profile 2 therefore adds to the profile-1
fields (signed mode index, Gamma = div W, Sigma = div Y):

1. Bose-Einstein momenta: the conjugate amplitudes sqrt(n omega / vol) of
   the field modes of ``ic_np`` (a free mode of occupation n has
   <|A_k|^2> = n/(omega vol) and <|E_k|^2> = n omega/vol): transverse
   gauge modes for E^a, E^Y (streams 100 + 10 grp ...), Higgs modes for pi
   (streams 140 + b), real part of the unnormalised inverse DFT.
2. The charges the reference's k_Gamma / k_Sigma weight by gp2
   (main_rk4.jl:603-657): rho^a = gw (-(W x E)^a + Q^a(phi, pi)),
   rho^Y = gy Q^Y(phi, pi).  The Gauss operator on the electric fields is
   (G E)^a = d_i E^a_i + gw eps^{abc} W^b_i E^c_i with the 6th-order d1 of
   diff_scheme.jl:157-233, so the law reads G E = gw Q (SU(2)) and
   d_i E^Y_i = gy Q^Y (U(1)).
3. The lattice d1 annihilates the modes whose wave numbers are all 0 or pi
   (the zero mode and, on even sides, the Nyquist modes), so the U(1) charge
   must have no component there.  pi is corrected by
   delta pi = c(x) J^T phi (phi^T J delta pi = |phi|^2 c, J the U(1)
   generator's matrix of Q^Y = phi^T J pi), c a combination of those modes
   fixed by an 8 x 8 (or smaller) linear system.
4. U(1): E^Y += d chi_Y with d.d chi_Y = gy Q^Y - d.E^Y, solved exactly by FFT
   with the d1 symbol K(theta) = (90 sin t - 18 sin 2t + 2 sin 3t)/(60 dx).
5. SU(2): E^a += (D chi)^a, (D chi)^a_i = d_i chi^a + gw eps^{abc} W^b_i chi^c.
   Since G = -D^T, the law becomes D^T D chi = -(gw Q - G E0), a symmetric
   positive definite system solved by conjugate gradients preconditioned
   with the flat operator (sum_i K_i^2 + mu)^-1 applied by FFT,
   mu = (2/3) gw^2 <|W|^2>.
The correction E - E0 is a (covariant) gradient, the smallest change of E
that satisfies the law.
"""
from __future__ import annotations

import numpy as np

from . import ic_np

NCH = 36
EPS = np.zeros((3, 3, 3))
for _a, _b, _c in ((0, 1, 2), (1, 2, 0), (2, 0, 1)):
    EPS[_a, _b, _c] = 1.0
    EPS[_a, _c, _b] = -1.0


def d1(A, axis, dx):
    c = lambda s: np.roll(A, -s, axis=axis)
    return (45.0 * (c(1) - c(-1)) - 9.0 * (c(2) - c(-2)) + (c(3) - c(-3))) / (60.0 * dx)


def _ax(i):
    """numpy axis of lattice direction i (0 = x) in a (..., nz, ny, nx) array."""
    return -1 - i


def higgs_charges(phi, pi):
    """Q^1..3 and Q^Y of k_Gamma_1..3 / k_Sigma (main_rk4.jl:612-616, 631-635,
    650-654, 660-664), 0-based components."""
    p, q = phi, pi
    Q1 = p[0] * q[3] - p[3] * q[0] + p[2] * q[1] - p[1] * q[2]
    Q2 = p[2] * q[0] - p[0] * q[2] + p[3] * q[1] - p[1] * q[3]
    Q3 = p[0] * q[1] - p[1] * q[0] + p[3] * q[2] - p[2] * q[3]
    QY = p[0] * q[1] - p[1] * q[0] + p[2] * q[3] - p[3] * q[2]
    return np.stack([Q1, Q2, Q3]), QY


def gauss_op(E, W, gw, dx):
    """(G E)^a = d_i E^a_i + gw eps^{abc} W^b_i E^c_i; E, W: (3 colour, 3 dir, grid)."""
    div = np.stack([sum(d1(E[a, i], _ax(i), dx) for i in range(3)) for a in range(3)])
    return div + gw * np.einsum("abc,bi...,ci...->a...", EPS, W, E)


def cov_grad(chi, W, gw, dx):
    """(D chi)^a_i = d_i chi^a + gw eps^{abc} W^b_i chi^c -> (3, 3, grid)."""
    g = np.stack([np.stack([d1(chi[a], _ax(i), dx) for i in range(3)]) for a in range(3)])
    return g + gw * np.einsum("abc,bi...,c...->ai...", EPS, W, chi)


def k_symbol(shape, dx):
    """sum_i K(theta_i)^2 on the FFT grid of a (nz, ny, nx) array,
    K(theta) = (90 sin t - 18 sin 2t + 2 sin 3t)/(60 dx) the d1 symbol."""
    nx, ny, nz = shape
    tot = 0.0
    for n, ax in ((nx, 2), (ny, 1), (nz, 0)):
        th = 2.0 * np.pi * np.arange(n) / n
        K = (90.0 * np.sin(th) - 18.0 * np.sin(2 * th) + 2.0 * np.sin(3 * th)) / (60.0 * dx)
        shp = [1, 1, 1]
        shp[ax] = n
        tot = tot + (K ** 2).reshape(shp)
    return tot


def null_modes(shape):
    """Mode numbers (mx, my, mz) in {0, 1}^3 whose d1 symbol vanishes on every
    axis: 0, or pi on an even side."""
    nx, ny, nz = shape
    out = []
    for mz in (0, 1):
        for my in (0, 1):
            for mx in (0, 1):
                if (mx and nx % 2) or (my and ny % 2) or (mz and nz % 2):
                    continue
                out.append((mx, my, mz))
    return out


def sign_pattern(shape, m):
    nx, ny, nz = shape
    k, j, i = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    return (-1.0) ** ((m[0] * i + m[1] * j + m[2] * k) % 2)


def correct_pi(phi, pi, gy, shape):
    """pi + c(x) J^T phi with c in the span of the d1 null modes, chosen so
    that gy Q^Y has no component on those modes."""
    modes = null_modes(shape)
    e = [sign_pattern(shape, m) for m in modes]
    r2 = np.sum(phi ** 2, axis=0)
    _, QY = higgs_charges(phi, pi)
    P = np.array([np.sum(gy * QY * em) for em in e])
    M = np.array([[gy * np.sum(r2 * em * en) for en in e] for em in e])
    alpha = np.linalg.solve(M, -P)
    c = sum(a * em for a, em in zip(alpha, e))
    JTphi = np.stack([-phi[1], phi[0], -phi[3], phi[2]])
    return pi + c * JTphi


def precond(r, shape, dx, mu):
    """(sum_i K_i^2 + mu)^-1 per colour component, by FFT."""
    sym = k_symbol(shape, dx) + mu
    return np.stack([np.fft.ifftn(np.fft.fftn(r[a]) / sym).real for a in range(3)])


def solve_su2(E0, W, Q, gw, dx, shape, tol=1e-13, maxit=2000):
    """E = E0 + D chi with G E = gw Q; returns (E, iterations, relative residual)."""
    mu = (2.0 / 3.0) * gw ** 2 * np.mean(np.sum(W ** 2, axis=(0, 1)))
    A = lambda v: -gauss_op(cov_grad(v, W, gw, dx), W, gw, dx)
    b = -(gw * Q - gauss_op(E0, W, gw, dx))
    x = np.zeros_like(b)
    r = b.copy()
    z = precond(r, shape, dx, mu)
    p = z.copy()
    rz = np.sum(r * z)
    bn = np.sqrt(np.sum(b * b))
    it = 0
    while it < maxit and np.sqrt(np.sum(r * r)) > tol * bn:
        q = A(p)
        alpha = rz / np.sum(p * q)
        x += alpha * p
        r -= alpha * q
        z = precond(r, shape, dx, mu)
        rz_new = np.sum(r * z)
        p = z + (rz_new / rz) * p
        rz = rz_new
        it += 1
    E = E0 + cov_grad(x, W, gw, dx)
    res = gauss_op(E, W, gw, dx) - gw * Q
    return E, it, float(np.sqrt(np.sum(res * res)) / bn)


def solve_u1(EY0, QY, gy, dx, shape):
    """E^Y + d chi with d.E^Y = gy Q^Y (QY free of the d1 null modes)."""
    div0 = sum(d1(EY0[i], _ax(i), dx) for i in range(3))
    rhs = gy * QY - div0
    sym = k_symbol(shape, dx)
    F = np.fft.fftn(rhs)
    with np.errstate(divide="ignore", invalid="ignore"):
        chiF = np.where(sym > 1e-12 * sym.max(), -F / sym, 0.0)
    chi = np.fft.ifftn(chiF).real
    return np.stack([EY0[i] + d1(chi, _ax(i), dx) for i in range(3)])


def momentum_modes_gauge(shape, dx, T, seed, grp):
    """Complex E modes of gauge group grp: gauge_modes' construction with
    streams 100 + 10 grp ... and amplitude sqrt(n omega / vol)."""
    nx, ny, nz = shape
    n3 = nx * ny * nz
    q = np.arange(n3, dtype=np.uint64)
    kx, ky, kz, modk = (a.reshape(-1) for a in ic_np.lattice_k(shape, dx, 1))
    s0 = 100 + 10 * grp
    r1, r2, r3 = (ic_np.gauss_sqrt2(seed, s0 + t, q) for t in range(3))
    x1, x2, x3, x4 = (ic_np.gauss_sqrt2(seed, s0 + 6 + t, q) for t in range(4))
    k = np.stack([kx, ky, kz])
    e1 = np.cross(np.stack([r1, r2, r3]), k, axis=0)
    with np.errstate(invalid="ignore", divide="ignore"):
        e1 = e1 / np.linalg.norm(e1, axis=0)
        e2 = np.cross(e1, k, axis=0)
        e2 = e2 / np.linalg.norm(e2, axis=0)
    amp = ic_np.be_amplitude(modk, T, n3 * dx ** 3) * modk
    out = np.zeros((3, n3), dtype=np.complex128)
    nz_ = modk != 0
    out[:, nz_] = amp[nz_] * (e1[:, nz_] * x1[nz_] + e2[:, nz_] * x2[nz_]) + \
        1j * amp[nz_] * (e1[:, nz_] * x3[nz_] + e2[:, nz_] * x4[nz_])
    return out.reshape(3, nz, ny, nx)


def momentum_modes_higgs(shape, dx, T, meff_sq, seed, b):
    nx, ny, nz = shape
    n3 = nx * ny * nz
    q = np.arange(n3, dtype=np.uint64)
    modk = ic_np.lattice_k(shape, dx, 1)[3].reshape(-1)
    w = np.sqrt(modk ** 2 + meff_sq)
    amp = ic_np.be_amplitude(w, T, n3 * dx ** 3) * w
    return (amp * ic_np.gauss_sqrt2(seed, 140 + b, q)).reshape(nz, ny, nx).astype(np.complex128)


def thermal_gauss_state(shape, dx, T, meff_sq, seed, gw, gy, tol=1e-13):
    """Profile-2 thermal state (36, nz, ny, nx) and a dict of solve figures."""
    s = ic_np.thermal_state(shape, dx, T, meff_sq, seed, 1)
    for grp in range(4):
        ch = 24 + 3 * grp if grp < 3 else 33
        s[ch:ch + 3] = ic_np.bfft(momentum_modes_gauge(shape, dx, T, seed, grp)).real
    for b in range(4):
        s[20 + b] = ic_np.bfft(momentum_modes_higgs(shape, dx, T, meff_sq, seed, b)).real
    phi, W = s[0:4], s[4:13].reshape((3, 3) + s.shape[1:])
    s[20:24] = correct_pi(phi, s[20:24], gy, shape)
    Q, QY = higgs_charges(phi, s[20:24])
    E0 = s[24:33].reshape((3, 3) + s.shape[1:]).copy()
    G0 = gauss_op(E0, W, gw, dx) - gw * Q
    div0Y = sum(d1(s[33 + i], _ax(i), dx) for i in range(3)) - gy * QY
    E, it, rel = solve_su2(E0, W, Q, gw, dx, shape, tol=tol)
    s[24:33] = E.reshape((9,) + s.shape[1:])
    s[33:36] = solve_u1(s[33:36], QY, gy, dx, shape)
    G1 = gauss_op(E, W, gw, dx) - gw * Q
    G1Y = sum(d1(s[33 + i], _ax(i), dx) for i in range(3)) - gy * QY
    before = np.sqrt(np.sum(G0 ** 2) + np.sum(div0Y ** 2))
    after = np.sqrt(np.sum(G1 ** 2) + np.sum(G1Y ** 2))
    return s, {"iterations": it, "cg_rel": rel, "before": float(before), "after": float(after)}
