// ewpt_point.h -- pointwise part of the RHS of the Higgs-gauge equations.
//
// Given the centre values of the 36 stage-state channels and their spatial
// derivatives at one site, evaluates the 20 non-trivial outputs of one
// rk4_step! evaluation (gpu-kernels-rk4/main_rk4.jl:52-661):
//   out[0..2] k_Gamma^a, out[3] k_Sigma, out[4..7] kt_phi, out[8..16] kt_W,
//   out[17..19] kt_Y.
// (The 16 field k's are copies of the stage momenta and never computed.)
//
// The equations are written in generator form instead of the reference's
// expanded scalar expressions, which saves arithmetic:
//   D_i phi = d_i phi + A_i phi,  A_i = gw/2 W^a_i T_a + gy/2 Y_i T_Y
//     (cov_derivs_temporal.jl:24-129; T_a, T_Y are the 4x4 real
//      antisymmetric matrices those functions apply, see apply_gen below)
//   F^a_ij = d_i W^a_j - d_j W^a_i + gw eps^abc W^b_i W^c_j
//     (field_strengths.jl:35-311)
//   kt_phi  = lap phi + sum_i A_i (d_i phi + D_i phi) + M_Gamma phi
//             - (2 lambda (|phi|^2 - v^2) + Jex/2) phi
//             - gamma phi_a (phi_a pi_a)/|phi|^2 + (beta terms) phi
//             (main_rk4.jl:337-407)
//   kt_W^a_i = lap W^a_i - gw sum_j [ (d_j W_i x W_j)^a + (W_j x F_ij)^a ]
//             + gw phi.T_a D_i phi - d_i Gamma^a - gw (W_i x Gamma)^a + beta terms
//             (main_rk4.jl:409-551)
//   kt_Y_i  = lap Y_i + gy phi.T_Y D_i phi - d_i Sigma + beta terms (553-581)
//   k_Gamma^a = (1-gp2) div E^a + gp2 gw [-(W_i x E_i)^a + phi.T_a pi] (603-648)
//   k_Sigma   = (1-gp2) div E^Y + gp2 gy phi.T_Y pi                     (651-657)
// The beta_W bracket of kt_phi keeps the reference's use of dW^1_y/dt and
// dW^1_z/dt in its a=2,3 rows (main_rk4.jl:350-352).
#pragma once

#include <hip/hip_runtime.h>

namespace ewpt {

// channel numbers of the 36-channel state
enum : int {
    CH_PHI = 0,   // 0..3
    CH_W = 4,     // 4..12, W^a_i at CH_W + 3a + i (a,i = 0..2)
    CH_Y = 13,    // 13..15
    CH_GAM = 16,  // 16..18
    CH_SIG = 19,
    CH_PI = 20,   // 20..23
    CH_E = 24,    // 24..32, E^a_i at CH_E + 3a + i
    CH_EY = 33,   // 33..35
    NCH = 36,
    NK = 20,      // stored derivative channels per stage (state channels 16..35)
};

struct Phys {
    double gw, gy, gp2, bw, by, vev, lam, jex, gam;
};

// Centre values and derivatives of one site (registers).
template <typename R>
struct SiteIn {
    R c[NCH];        // centre values of the stage state
    R dphi[4][3];    // d_j phi_b
    R dW[3][3][3];   // [a][i][j] = d_j W^a_i
    R dY[3][3];      // [i][j] = d_j Y_i
    R dG[3][3];      // [a][j] = d_j Gamma^a
    R dS[3];         // d_j Sigma
    R lap[16];       // Laplacians of channels 0..15
    R divE[3];       // d_i E^a_i
    R divEY;         // d_i E^Y_i
};

// o = (h1 W1 T_1 + h1 W2 T_2 + h1 W3 T_3 + hy Yv T_Y) v, T's as read off
// cov_derivs_temporal.jl:24-57 (row b of D_2phi_b).
template <typename R>
__device__ __forceinline__ void apply_gen(R h1, R W1, R W2, R W3, R hy, R Yv, const R v[4], R o[4])
{
    o[0] = h1 * (W1 * v[3] - W2 * v[2] + W3 * v[1]) + hy * Yv * v[1];
    o[1] = -h1 * (W1 * v[2] + W2 * v[3] + W3 * v[0]) - hy * Yv * v[0];
    o[2] = h1 * (W1 * v[1] + W2 * v[0] - W3 * v[3]) + hy * Yv * v[3];
    o[3] = -h1 * (W1 * v[0] - W2 * v[1] - W3 * v[2]) - hy * Yv * v[2];
}

// phi . T_a v  (a = 0,1,2) and phi . T_Y v
template <typename R>
__device__ __forceinline__ R phiT(int a, const R p[4], const R v[4])
{
    if (a == 0) return p[0] * v[3] - p[1] * v[2] + p[2] * v[1] - p[3] * v[0];
    if (a == 1) return -p[0] * v[2] - p[1] * v[3] + p[2] * v[0] + p[3] * v[1];
    return p[0] * v[1] - p[1] * v[0] - p[2] * v[3] + p[3] * v[2];
}
template <typename R>
__device__ __forceinline__ R phiTY(const R p[4], const R v[4])
{
    return p[0] * v[1] - p[1] * v[0] + p[2] * v[3] - p[3] * v[2];
}

// colour cross product (X x Y)^a = eps^abc X^b Y^c
template <typename R>
__device__ __forceinline__ R xprod(int a, R X0, R X1, R X2, R Y0, R Y1, R Y2)
{
    if (a == 0) return X1 * Y2 - X2 * Y1;
    if (a == 1) return X2 * Y0 - X0 * Y2;
    return X0 * Y1 - X1 * Y0;
}

// Covariant derivatives D[b][i] of phi.
template <typename R>
__device__ __forceinline__ void cov_derivs(const SiteIn<R> &s, R gw, R gy, R D[4][3])
{
    const R h1 = R(0.5) * gw, hy = R(0.5) * gy;
    const R p[4] = {s.c[0], s.c[1], s.c[2], s.c[3]};
#pragma unroll
    for (int i = 0; i < 3; i++) {
        R Ap[4];
        apply_gen(h1, s.c[CH_W + i], s.c[CH_W + 3 + i], s.c[CH_W + 6 + i], hy, s.c[CH_Y + i], p, Ap);
#pragma unroll
        for (int b = 0; b < 4; b++) D[b][i] = s.dphi[b][i] + Ap[b];
    }
}

// Field strengths as magnetic vectors B[a][i] = (F_yz, F_zx, F_xy) and BY.
template <typename R>
__device__ __forceinline__ void field_strengths(const SiteIn<R> &s, R gw, R B[3][3], R BY[3])
{
#pragma unroll
    for (int a = 0; a < 3; a++) {
#pragma unroll
        for (int i = 0; i < 3; i++) {
            const int j = (i + 1) % 3, k = (i + 2) % 3;  // B_i = F_jk
            const R nl = xprod(a, s.c[CH_W + j], s.c[CH_W + 3 + j], s.c[CH_W + 6 + j],
                               s.c[CH_W + k], s.c[CH_W + 3 + k], s.c[CH_W + 6 + k]);
            B[a][i] = s.dW[a][k][j] - s.dW[a][j][k] + gw * nl;
        }
    }
#pragma unroll
    for (int i = 0; i < 3; i++) {
        const int j = (i + 1) % 3, k = (i + 2) % 3;
        BY[i] = s.dY[k][j] - s.dY[j][k];
    }
}

template <typename R>
__device__ __forceinline__ void rhs_point(const SiteIn<R> &s, const Phys &P, R out[NK])
{
    const R gw = R(P.gw), gy = R(P.gy), gp2 = R(P.gp2);
    const R h1 = R(0.5) * gw, hy = R(0.5) * gy;
    const R p[4] = {s.c[0], s.c[1], s.c[2], s.c[3]};
    const R pi[4] = {s.c[CH_PI], s.c[CH_PI + 1], s.c[CH_PI + 2], s.c[CH_PI + 3]};
    auto W = [&](int a, int i) { return s.c[CH_W + 3 * a + i]; };
    auto E = [&](int a, int i) { return s.c[CH_E + 3 * a + i]; };

    R D[4][3];
    cov_derivs(s, gw, gy, D);
    R B[3][3], BY[3];
    field_strengths(s, gw, B, BY);
    // F_ij for i != j from B: F_xy = B_z, F_yz = B_x, F_zx = B_y, F_ji = -F_ij
    auto F = [&](int a, int i, int j) -> R {
        if (i == j) return R(0);
        const int k = 3 - i - j;
        const R v = B[a][k];
        return ((j - i + 3) % 3 == 1) ? v : -v;
    };

    const R r2 = p[0] * p[0] + p[1] * p[1] + p[2] * p[2] + p[3] * p[3];
    const R phipi = pi[0] * p[0] + pi[1] * p[1] + pi[2] * p[2] + pi[3] * p[3];
    R u[3];  // u_j = phi . d_j phi
#pragma unroll
    for (int j = 0; j < 3; j++) u[j] = s.dphi[0][j] * p[0] + s.dphi[1][j] * p[1] + s.dphi[2][j] * p[2] + s.dphi[3][j] * p[3];

    // ---- k_Gamma, k_Sigma (main_rk4.jl:603-657)
#pragma unroll
    for (int a = 0; a < 3; a++) {
        R na = R(0);
#pragma unroll
        for (int i = 0; i < 3; i++) na -= xprod(a, W(0, i), W(1, i), W(2, i), E(0, i), E(1, i), E(2, i));
        out[a] = (R(1) - gp2) * s.divE[a] + gp2 * gw * (na + phiT(a, p, pi));
    }
    out[3] = (R(1) - gp2) * s.divEY + gp2 * gy * phiTY(p, pi);

    // ---- kt_phi (main_rk4.jl:337-407)
    {
        R QW = R(0);
#pragma unroll
        for (int a = 0; a < 3; a++) {
            QW += E(a, 0) * B[a][0];
            QW += E(0, 1) * B[a][1] + E(0, 2) * B[a][2];   // reference uses E^1_y, E^1_z for every a
        }
        const R QY = s.c[CH_EY] * BY[0] + s.c[CH_EY + 1] * BY[1] + s.c[CH_EY + 2] * BY[2];
        const R pot = R(2) * R(P.lam) * (r2 - R(P.vev) * R(P.vev)) + R(0.5) * R(P.jex);
        const R beta = R(P.bw) * QW + R(P.by) * QY;
        R acc[4];
#pragma unroll
        for (int b = 0; b < 4; b++) acc[b] = s.lap[b] + (beta - pot) * p[b];
#pragma unroll
        for (int i = 0; i < 3; i++) {
            const R v[4] = {s.dphi[0][i] + D[0][i], s.dphi[1][i] + D[1][i], s.dphi[2][i] + D[2][i], s.dphi[3][i] + D[3][i]};
            R Av[4];
            apply_gen(h1, W(0, i), W(1, i), W(2, i), hy, s.c[CH_Y + i], v, Av);
#pragma unroll
            for (int b = 0; b < 4; b++) acc[b] += Av[b];
        }
        R Gp[4];
        apply_gen(h1, s.c[CH_GAM], s.c[CH_GAM + 1], s.c[CH_GAM + 2], hy, s.c[CH_SIG], p, Gp);
        const R gam = R(P.gam);
        // the reference's 0/0 at |phi| = 0 is kept (NaN), also for gamma = 0
        const R inv_r2 = R(1) / r2;
#pragma unroll
        for (int b = 0; b < 4; b++) out[4 + b] = acc[b] + Gp[b] - gam * p[b] * (p[b] * pi[b]) * inv_r2;
    }

    // ---- kt_W (main_rk4.jl:409-551)
#pragma unroll
    for (int a = 0; a < 3; a++) {
#pragma unroll
        for (int i = 0; i < 3; i++) {
            R acc = s.lap[CH_W + 3 * a + i] - s.dG[a][i];
            R nl = R(0);
#pragma unroll
            for (int j = 0; j < 3; j++) {
                nl += xprod(a, s.dW[0][i][j], s.dW[1][i][j], s.dW[2][i][j], W(0, j), W(1, j), W(2, j));
                if (j != i) nl += xprod(a, W(0, j), W(1, j), W(2, j), F(0, i, j), F(1, i, j), F(2, i, j));
            }
            nl += xprod(a, W(0, i), W(1, i), W(2, i), s.c[CH_GAM], s.c[CH_GAM + 1], s.c[CH_GAM + 2]);
            const R Di[4] = {D[0][i], D[1][i], D[2][i], D[3][i]};
            acc += gw * (phiT(a, p, Di) - nl);
            // beta_W: -2 beta (pi.phi) B^a_i + 2 beta (u x E^a)_i
            const int j = (i + 1) % 3, k = (i + 2) % 3;
            acc += R(2) * R(P.bw) * (u[j] * E(a, k) - u[k] * E(a, j) - phipi * B[a][i]);
            out[8 + 3 * a + i] = acc;
        }
    }

    // ---- kt_Y (main_rk4.jl:553-581)
#pragma unroll
    for (int i = 0; i < 3; i++) {
        const R Di[4] = {D[0][i], D[1][i], D[2][i], D[3][i]};
        const int j = (i + 1) % 3, k = (i + 2) % 3;
        out[17 + i] = s.lap[CH_Y + i] + gy * phiTY(p, Di) - s.dS[i]
                      + R(2) * R(P.by) * (u[j] * s.c[CH_EY + k] - u[k] * s.c[CH_EY + j] - phipi * BY[i]);
    }
}

}  // namespace ewpt

namespace ewpt {

// Electromagnetic fields at one site (compute_Bx!/By!/Bz!, compute_B_2_*!,
// compute_Ax!/Ay!/Az!, main_rk4.jl:2303-2822, 3012-3528, 3718-4250):
//   n^a   = (-2(p1 p3 + p2 p4), 2(p2 p3 - p1 p4), -p1^2 - p2^2 + p3^2 + p4^2)/|phi|^2,
//           0 where |phi| = 0 (2459-2467)
//   B_i   = -(sin th n^a F^a_jk + cos th FY_jk + 4 sin th/(gw v^2) X_i)
//   B2_i  = the same with |phi|^4 in place of v^2; 0 where |phi| = 0
//   A_i   = sin th n^a W^a_i + cos th Y_i
//   X_i   = D_j p1 D_k p2 - D_j p2 D_k p1 + D_j p3 D_k p4 - D_j p4 D_k p3,
// (i, j, k) cyclic, F^a_jk = B[a][i] and FY_jk = BY[i] of field_strengths.
template <typename R>
__device__ __forceinline__ void em_point(const SiteIn<R> &s, const Phys &P, R sw, R cw, R Bem[3], R B2[3], R A[3])
{
    R D[4][3];
    cov_derivs(s, R(P.gw), R(P.gy), D);
    R B[3][3], BY[3];
    field_strengths(s, R(P.gw), B, BY);
    const R p1 = s.c[0], p2 = s.c[1], p3 = s.c[2], p4 = s.c[3];
    const R mag = p1 * p1 + p2 * p2 + p3 * p3 + p4 * p4;
    R n[3] = {R(0), R(0), R(0)};
    if (mag != R(0)) {
        n[0] = R(-2) * (p1 * p3 + p2 * p4) / mag;
        n[1] = R(2) * (p2 * p3 - p1 * p4) / mag;
        n[2] = (-p1 * p1 - p2 * p2 + p3 * p3 + p4 * p4) / mag;
    }
    const R c1 = R(4) * sw / (R(P.gw) * R(P.vev) * R(P.vev));
    const R c2 = mag != R(0) ? R(4) * sw / (R(P.gw) * mag * mag) : R(0);
#pragma unroll
    for (int i = 0; i < 3; i++) {
        const int j = (i + 1) % 3, k = (i + 2) % 3;
        const R X = D[0][j] * D[1][k] - D[1][j] * D[0][k] + D[2][j] * D[3][k] - D[3][j] * D[2][k];
        const R F = sw * (n[0] * B[0][i] + n[1] * B[1][i] + n[2] * B[2][i]) + cw * BY[i];
        Bem[i] = -(F + c1 * X);
        B2[i] = mag != R(0) ? -(F + c2 * X) : R(0);
        A[i] = sw * (n[0] * s.c[CH_W + i] + n[1] * s.c[CH_W + 3 + i] + n[2] * s.c[CH_W + 6 + i]) + cw * s.c[CH_Y + i];
    }
}

}  // namespace ewpt
