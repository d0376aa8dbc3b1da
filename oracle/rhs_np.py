"""Independent numpy restatement of the reference RHS (test infrastructure only).

The literal C oracle (``ewpt_oracle.c``) copies the reference's scalar
expressions term by term.  This module instead writes the same equations in
generator form, read off the reference once:

* covariant derivative (cov_derivs_temporal.jl:24-129):
  ``D_i phi = d_i phi + A_i phi``,  ``A_i = (gw/2) W^a_i T_a + (gy/2) Y_i T_Y``
  with the 4x4 real antisymmetric generators ``T_a``, ``T_Y`` below;
* field strength (field_strengths.jl:35-311):
  ``F^a_ij = d_i W^a_j - d_j W^a_i + gw eps^abc W^b_i W^c_j``;
* Higgs EOM (main_rk4.jl:337-407), SU(2) EOM (409-551), U(1) EOM (553-581),
  Gauss auxiliaries (603-657) as matrix/einsum expressions.

Agreement with the C transcription is to rounding (the sums are associated
differently), which is what makes the pair a check on each other.
"""
from __future__ import annotations

import numpy as np

# generators acting on (phi_1..phi_4); read off cov_derivs_temporal.jl:24-57
T = np.zeros((3, 4, 4))
T[0][0, 3], T[0][3, 0], T[0][1, 2], T[0][2, 1] = 1, -1, -1, 1
T[1][0, 2], T[1][2, 0], T[1][1, 3], T[1][3, 1] = -1, 1, -1, 1
T[2][0, 1], T[2][1, 0], T[2][2, 3], T[2][3, 2] = 1, -1, -1, 1
TY = np.zeros((4, 4))
TY[0, 1], TY[1, 0], TY[2, 3], TY[3, 2] = 1, -1, 1, -1

EPS = np.zeros((3, 3, 3))
for a, b, c in ((0, 1, 2), (1, 2, 0), (2, 0, 1)):
    EPS[a, b, c], EPS[a, c, b] = 1, -1

# lattice axis of spatial direction i for a (nz, ny, nx) channel array
_AX = {0: -1, 1: -2, 2: -3}


def d1(A: np.ndarray, i: int, dx: float) -> np.ndarray:
    """6th-order central first derivative, periodic (diff_scheme.jl:157-233)."""
    ax = _AX[i]
    s = lambda k: np.roll(A, -k, axis=ax)  # noqa: E731  A[x+k]
    return (s(3) - 9.0 * s(2) + 45.0 * s(1) - 45.0 * s(-1) + 9.0 * s(-2) - s(-3)) / (60.0 * dx)


def d2(A: np.ndarray, i: int, dx: float) -> np.ndarray:
    """6th-order central second derivative, periodic (diff_scheme.jl:235-311)."""
    ax = _AX[i]
    s = lambda k: np.roll(A, -k, axis=ax)  # noqa: E731
    return (s(3) / 90.0 - 3.0 * s(2) / 20.0 + 3.0 * s(1) / 2.0 - 49.0 * A / 18.0
            + 3.0 * s(-1) / 2.0 - 3.0 * s(-2) / 20.0 + s(-3) / 90.0) / (dx * dx)


def lap(A, dx):
    return d2(A, 0, dx) + d2(A, 1, dx) + d2(A, 2, dx)


def unpack(s: np.ndarray):
    """36-channel state -> named views (main_rk4.jl:52-74 argument order)."""
    return dict(phi=s[0:4], W=s[4:13].reshape((3, 3) + s.shape[1:]), Y=s[13:16],
                Gam=s[16:19], Sig=s[19], pi=s[20:24],
                E=s[24:33].reshape((3, 3) + s.shape[1:]), EY=s[33:36])


def derived(s: np.ndarray, p, dx: float):
    """Gradients, covariant derivatives and field strengths of a state."""
    f = unpack(s)
    gw, gy = p["gw"], p["gy"]
    phi, W, Y = f["phi"], f["W"], f["Y"]
    dphi = np.stack([np.stack([d1(phi[b], j, dx) for j in range(3)]) for b in range(4)])
    dW = np.stack([np.stack([np.stack([d1(W[a, i], j, dx) for j in range(3)])
                             for i in range(3)]) for a in range(3)])      # [a,i,j] = d_j W^a_i
    dY = np.stack([np.stack([d1(Y[i], j, dx) for j in range(3)]) for i in range(3)])
    # connection A_i[b,c]
    Acon =0.5 * gw * np.einsum("ai...,abc->ibc...", W, T) + 0.5 * gy * np.einsum("i...,bc->ibc...", Y, TY)
    D = dphi + np.einsum("ibc...,c...->bi...", Acon, phi)                 # [b,i]
    F = (np.transpose(dW, (0, 2, 1) + tuple(range(3, dW.ndim))) - dW
         + gw * np.einsum("abc,bi...,cj...->aij...", EPS, W, W))          # [a,i,j]
    FY = np.transpose(dY, (1, 0) + tuple(range(2, dY.ndim))) - dY          # [i,j]
    Bw = np.stack([F[:, 1, 2], F[:, 2, 0], F[:, 0, 1]], axis=1)             # [a,i]
    BY = np.stack([FY[1, 2], FY[2, 0], FY[0, 1]])
    return f, dict(dphi=dphi, dW=dW, dY=dY, A=Acon, D=D, F=F, FY=FY, B=Bw, BY=BY)


def rhs(s: np.ndarray, p, dx: float) -> np.ndarray:
    """36 -> 36 channel RHS, the rk4_step! equations in generator form."""
    f, g = derived(s, p, dx)
    gw, gy, gp2 = p["gw"], p["gy"], p["gp2"]
    bW, bY, vev, lam, Jex, gam = p["beta_w"], p["beta_y"], p["vev"], p["lambda_"], p["jex"], p["gamma"]
    phi, W, Gam, Sig, pi, E, EY = f["phi"], f["W"], f["Gam"], f["Sig"], f["pi"], f["E"], f["EY"]
    dphi, dW, D, F, Bw, BY, Acon = g["dphi"], g["dW"], g["D"], g["F"], g["B"], g["BY"], g["A"]
    out = np.empty_like(s)
    r2 = np.sum(phi * phi, axis=0)

    # Higgs, main_rk4.jl:337-407.  beta_W bracket keeps the reference's
    # E^1_y, E^1_z in the a=2,3 rows (main_rk4.jl:350-352).
    Ewq = E.copy()
    Ewq[1, 1], Ewq[1, 2], Ewq[2, 1], Ewq[2, 2] = E[0, 1], E[0, 2], E[0, 1], E[0, 2]
    QW = np.sum(Ewq * Bw, axis=(0, 1))
    QY = np.sum(EY * BY, axis=0)
    Mgam = 0.5 * gw * np.einsum("a...,abc->bc...", Gam, T) + 0.5 * gy * np.einsum("...,bc->bc...", Sig, TY)
    ktphi = (np.stack([lap(phi[b], dx) for b in range(4)])
             + np.einsum("ibc...,ci...->b...", Acon, dphi + D)
             + np.einsum("bc...,c...->b...", Mgam, phi)
             - 2.0 * lam * (r2 - vev * vev) * phi
             - 0.5 * Jex * phi
             - gam * phi * (phi * pi) / r2
             + bW * QW * phi + bY * QY * phi)

    # SU(2), main_rk4.jl:409-551
    u = np.einsum("b...,bj...->j...", phi, dphi)            # u_j = phi . d_j phi
    phipi = np.sum(phi * pi, axis=0)

    def cross(a, b):
        return np.stack([a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2], a[0] * b[1] - a[1] * b[0]])

    Wlap = np.stack([np.stack([lap(W[a, i], dx) for i in range(3)]) for a in range(3)])
    dGam = np.stack([np.stack([d1(Gam[a], i, dx) for i in range(3)]) for a in range(3)])
    ktW = (Wlap
           - gw * np.einsum("abc,bij...,cj...->ai...", EPS, dW, W)
           - gw * np.einsum("abc,bj...,cij...->ai...", EPS, W, F)
           + gw * np.einsum("b...,abc,ci...->ai...", phi, T, D)
           - dGam
           - gw * np.einsum("abc,bi...,c...->ai...", EPS, W, Gam)
           - 2.0 * bW * phipi * Bw
           + 2.0 * bW * np.stack([cross(u, E[a]) for a in range(3)]))

    # U(1), main_rk4.jl:553-581
    ktY = (np.stack([lap(f["Y"][i], dx) for i in range(3)])
           + gy * np.einsum("b...,bc,ci...->i...", phi, TY, D)
           - np.stack([d1(Sig, i, dx) for i in range(3)])
           - 2.0 * bY * phipi * BY + 2.0 * bY * cross(u, EY))

    # Gauss auxiliaries, main_rk4.jl:603-657
    divE = np.stack([sum(d1(E[a, i], i, dx) for i in range(3)) for a in range(3)])
    divEY = sum(d1(EY[i], i, dx) for i in range(3))
    rho = (-np.einsum("abc,bi...,ci...->a...", EPS, W, E)
           + np.einsum("b...,abc,c...->a...", phi, T, pi))
    kGam = (1.0 - gp2) * divE + gp2 * gw * rho
    kSig = (1.0 - gp2) * divEY + gp2 * gy * np.einsum("b...,bc,c...->...", phi, TY, pi)

    out[0:16] = s[20:36]
    out[16:19] = kGam
    out[19] = kSig
    out[20:24] = ktphi
    out[24:33] = ktW.reshape((9,) + s.shape[1:])
    out[33:36] = ktY
    return out


def rk4(s: np.ndarray, p, dx: float, nsteps: int, rhs_fn=None) -> np.ndarray:
    """RK4 loop in the reference's stage/accumulation order (main_rk4.jl:6146-6441)."""
    f = rhs_fn or (lambda u: rhs(u, p, dx))
    y = s.copy()
    dt = p["dt"]
    for _ in range(nsteps):
        k = f(y)
        n = k * (dt / 6.0) + y
        k = f(y + k * (dt / 2.0))
        n = k * (dt / 3.0) + n
        k = f(y + k * (dt / 2.0))
        n = k * (dt / 3.0) + n
        k = f(y + k * dt)
        y = k * (dt / 6.0) + n
    return y


def energies(s: np.ndarray, p, dx: float) -> np.ndarray:
    """Energy table columns 1-7, 10, 12 (main_rk4.jl:1334-2301, 6104-6116)."""
    f, g = derived(s, p, dx)
    phi = f["phi"]
    r2 = np.sum(phi * phi, axis=0)
    vol = dx ** 3
    out = np.zeros(13)
    out[0] = np.sum(p["lambda_"] * (r2 - p["vev"] ** 2) ** 2) * vol
    out[1] = np.sum(f["pi"] ** 2) * vol
    out[2] = np.sum(g["D"] ** 2) * vol
    out[3] = 0.5 * np.sum(f["E"] ** 2) * vol
    out[4] = 0.5 * np.sum(g["B"] ** 2) * vol
    out[5] = 0.5 * np.sum(f["EY"] ** 2) * vol
    out[6] = 0.5 * np.sum(g["BY"] ** 2) * vol
    out[9] = np.min(np.sqrt(r2))
    out[11] = np.mean(np.sqrt(r2))
    return out


def gauss(s: np.ndarray, p, dx: float) -> np.ndarray:
    """L2 norms of the Gauss residuals G^a, G^Y and consistency C^a, C^Y."""
    f = unpack(s)
    gw, gy = p["gw"], p["gy"]
    phi, W, E, EY, pi = f["phi"], f["W"], f["E"], f["EY"], f["pi"]
    divE = np.stack([sum(d1(E[a, i], i, dx) for i in range(3)) for a in range(3)])
    divEY = sum(d1(EY[i], i, dx) for i in range(3))
    divW = np.stack([sum(d1(W[a, i], i, dx) for i in range(3)) for a in range(3)])
    divY = sum(d1(f["Y"][i], i, dx) for i in range(3))
    rho = gw * (-np.einsum("abc,bi...,ci...->a...", EPS, W, E) + np.einsum("b...,abc,c...->a...", phi, T, pi))
    rhoY = gy * np.einsum("b...,bc,c...->...", phi, TY, pi)
    res = [divE[0] - rho[0], divE[1] - rho[1], divE[2] - rho[2], divEY - rhoY,
           f["Gam"][0] - divW[0], f["Gam"][1] - divW[1], f["Gam"][2] - divW[2], f["Sig"] - divY]
    vol = dx ** 3
    return np.array([np.sqrt(np.sum(r * r) * vol) for r in res])
