/*
 * ewpt_oracle.c -- literal FP64 CPU restatement of the reference RK4 path.
 *
 * TEST INFRASTRUCTURE ONLY (the checker, never the product).  Built with
 * -O2 -ffp-contract=off so that every + - * / is one IEEE operation in the
 * order the Julia source writes it (left to right, as Julia parses
 * a+b+c and a*b*c).
 *
 * Followed line by line:
 *   stencils            gpu-kernels-rk4/diff_scheme.jl:157-311
 *   covariant derivs    gpu-kernels-rk4/cov_derivs_temporal.jl:24-129
 *   field strengths     gpu-kernels-rk4/field_strengths.jl:35-311
 *   RHS (rk4_step!)     gpu-kernels-rk4/main_rk4.jl:52-661
 *   RK4 loop            gpu-kernels-rk4/main_rk4.jl:6146-6441
 *   energy densities    gpu-kernels-rk4/main_rk4.jl:1334-2301, totals 6104-6116
 *
 * The periodic branches of dfdx & co. (diff_scheme.jl:161-178) all evaluate
 * the same formula with wrapped indices, so a modulo index is identical.
 */
#include "ewpt_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#define NCH 36

static inline int wrapi(int i, int n) { i %= n; return i < 0 ? i + n : i; }

typedef struct { int nx, ny, nz; double dx; } lat_t;

static inline long sidx(const lat_t *L, int i, int j, int k)
{
    return (long)i + (long)L->nx * ((long)j + (long)L->ny * (long)k);
}

/* value of A at (i,j,k) shifted by s along axis ax, periodic */
static inline double at(const lat_t *L, const double *A, int i, int j, int k, int ax, int s)
{
    if (ax == 0) i = wrapi(i + s, L->nx);
    else if (ax == 1) j = wrapi(j + s, L->ny);
    else k = wrapi(k + s, L->nz);
    return A[sidx(L, i, j, k)];
}

/* diff_scheme.jl:157-233 */
static double d1(const lat_t *L, const double *A, int i, int j, int k, int ax)
{
    double a3 = at(L, A, i, j, k, ax, 3), a2 = at(L, A, i, j, k, ax, 2);
    double a1 = at(L, A, i, j, k, ax, 1), m1 = at(L, A, i, j, k, ax, -1);
    double m2 = at(L, A, i, j, k, ax, -2), m3 = at(L, A, i, j, k, ax, -3);
    return (a3 - 9. * a2 + 45. * a1 - 45. * m1 + 9. * m2 - m3) / (60 * L->dx);
}

/* diff_scheme.jl:235-311 */
static double d2(const lat_t *L, const double *A, int i, int j, int k, int ax)
{
    double a3 = at(L, A, i, j, k, ax, 3), a2 = at(L, A, i, j, k, ax, 2);
    double a1 = at(L, A, i, j, k, ax, 1), a0 = A[sidx(L, i, j, k)];
    double m1 = at(L, A, i, j, k, ax, -1), m2 = at(L, A, i, j, k, ax, -2);
    double m3 = at(L, A, i, j, k, ax, -3);
    return (a3 / 90. - 3. * a2 / 20. + 3. * a1 / 2. - 49. * a0 / 18. + 3. * m1 / 2.
            - 3. * m2 / 20. + m3 / 90.) / (L->dx * L->dx);
}

/* channel numbers */
enum {
    C_P1, C_P2, C_P3, C_P4,
    C_W12, C_W13, C_W14, C_W22, C_W23, C_W24, C_W32, C_W33, C_W34,
    C_Y2, C_Y3, C_Y4, C_G1, C_G2, C_G3, C_SG,
    C_PI1, C_PI2, C_PI3, C_PI4,
    C_E12, C_E13, C_E14, C_E22, C_E23, C_E24, C_E32, C_E33, C_E34,
    C_EY2, C_EY3, C_EY4
};

#define CH(c) (in + (long)(c) * N3)

/* cov_derivs_temporal.jl:24-129: D_{2,3,4}phi_{1..4}; dir-generic since the
 * twelve functions only differ in which W^a_i, Y_i they are handed. */
static inline double Dphi1(double d, double p2, double p3, double p4,
                           double W1, double W2, double W3, double Y, double gw, double gy)
{ return (d + 0.5 * gw * ((W1) * (p4) - (W2) * (p3) + (W3) * (p2)) + 0.5 * gy * ((Y) * (p2))); }
static inline double Dphi2(double d, double p1, double p3, double p4,
                           double W1, double W2, double W3, double Y, double gw, double gy)
{ return (d - 0.5 * gw * ((W1) * (p3) + (W2) * (p4) + (W3) * (p1)) - 0.5 * gy * ((Y) * (p1))); }
static inline double Dphi3(double d, double p1, double p2, double p4,
                           double W1, double W2, double W3, double Y, double gw, double gy)
{ return (d + 0.5 * gw * ((W1) * (p2) + (W2) * (p1) - (W3) * (p4)) + 0.5 * gy * ((Y) * (p4))); }
static inline double Dphi4(double d, double p1, double p2, double p3,
                           double W1, double W2, double W3, double Y, double gw, double gy)
{ return (d - 0.5 * gw * ((W1) * (p1) - (W2) * (p2) - (W3) * (p3)) - 0.5 * gy * ((Y) * (p3))); }

/* everything the RHS and the energy densities need at one site */
typedef struct {
    /* centre values */
    double p1, p2, p3, p4;
    double W12, W13, W14, W22, W23, W24, W32, W33, W34;
    double Y2, Y3, Y4, G1, G2, G3, Sg;
    double pi1, pi2, pi3, pi4;
    double E12, E13, E14, E22, E23, E24, E32, E33, E34, EY2, EY3, EY4;
    /* first derivatives d<field>d<x|y|z> */
    double dp[4][3], dW[9][3], dY[3][3], dG[3][3], dS[3];
    /* covariant derivatives Dphi[a][dir] */
    double D[4][3];
    /* field strengths */
    double W_1_23, W_1_24, W_1_34, W_2_23, W_2_24, W_2_34, W_3_23, W_3_24, W_3_34;
    double Y_2_3, Y_2_4, Y_3_4;
} site_t;

static void load_site(const lat_t *L, const oracle_params *P, const double *in, long N3,
                      int i, int j, int k, site_t *s)
{
    long c = sidx(L, i, j, k);
    const double gw = P->gw, gy = P->gy;
    s->p1 = CH(C_P1)[c]; s->p2 = CH(C_P2)[c]; s->p3 = CH(C_P3)[c]; s->p4 = CH(C_P4)[c];
    s->W12 = CH(C_W12)[c]; s->W13 = CH(C_W13)[c]; s->W14 = CH(C_W14)[c];
    s->W22 = CH(C_W22)[c]; s->W23 = CH(C_W23)[c]; s->W24 = CH(C_W24)[c];
    s->W32 = CH(C_W32)[c]; s->W33 = CH(C_W33)[c]; s->W34 = CH(C_W34)[c];
    s->Y2 = CH(C_Y2)[c]; s->Y3 = CH(C_Y3)[c]; s->Y4 = CH(C_Y4)[c];
    s->G1 = CH(C_G1)[c]; s->G2 = CH(C_G2)[c]; s->G3 = CH(C_G3)[c]; s->Sg = CH(C_SG)[c];
    s->pi1 = CH(C_PI1)[c]; s->pi2 = CH(C_PI2)[c]; s->pi3 = CH(C_PI3)[c]; s->pi4 = CH(C_PI4)[c];
    s->E12 = CH(C_E12)[c]; s->E13 = CH(C_E13)[c]; s->E14 = CH(C_E14)[c];
    s->E22 = CH(C_E22)[c]; s->E23 = CH(C_E23)[c]; s->E24 = CH(C_E24)[c];
    s->E32 = CH(C_E32)[c]; s->E33 = CH(C_E33)[c]; s->E34 = CH(C_E34)[c];
    s->EY2 = CH(C_EY2)[c]; s->EY3 = CH(C_EY3)[c]; s->EY4 = CH(C_EY4)[c];

    /* main_rk4.jl:84-173 */
    for (int a = 0; a < 4; a++)
        for (int d = 0; d < 3; d++) s->dp[a][d] = d1(L, CH(C_P1 + a), i, j, k, d);
    for (int a = 0; a < 9; a++)
        for (int d = 0; d < 3; d++) s->dW[a][d] = d1(L, CH(C_W12 + a), i, j, k, d);
    for (int a = 0; a < 3; a++)
        for (int d = 0; d < 3; d++) s->dY[a][d] = d1(L, CH(C_Y2 + a), i, j, k, d);
    for (int a = 0; a < 3; a++)
        for (int d = 0; d < 3; d++) s->dG[a][d] = d1(L, CH(C_G1 + a), i, j, k, d);
    for (int d = 0; d < 3; d++) s->dS[d] = d1(L, CH(C_SG), i, j, k, d);

    /* main_rk4.jl:256-283: direction d uses W^a_{d} and Y_{d} */
    const double Wd[3][3] = {{s->W12, s->W22, s->W32}, {s->W13, s->W23, s->W33}, {s->W14, s->W24, s->W34}};
    const double Yd[3] = {s->Y2, s->Y3, s->Y4};
    for (int d = 0; d < 3; d++) {
        s->D[0][d] = Dphi1(s->dp[0][d], s->p2, s->p3, s->p4, Wd[d][0], Wd[d][1], Wd[d][2], Yd[d], gw, gy);
        s->D[1][d] = Dphi2(s->dp[1][d], s->p1, s->p3, s->p4, Wd[d][0], Wd[d][1], Wd[d][2], Yd[d], gw, gy);
        s->D[2][d] = Dphi3(s->dp[2][d], s->p1, s->p2, s->p4, Wd[d][0], Wd[d][1], Wd[d][2], Yd[d], gw, gy);
        s->D[3][d] = Dphi4(s->dp[3][d], s->p1, s->p2, s->p3, Wd[d][0], Wd[d][1], Wd[d][2], Yd[d], gw, gy);
    }

    /* field_strengths.jl:35-311 as called at main_rk4.jl:297-330.
     * dW[3*(a-1)+(i-2)][dir]: W^a_i derivative */
#define DW(a, i, d) s->dW[3 * ((a) - 1) + ((i) - 2)][d]
    s->W_1_23 = DW(1, 3, 0) - DW(1, 2, 1) + gw * (s->W22 * s->W33 - s->W23 * s->W32);
    s->W_1_24 = DW(1, 4, 0) - DW(1, 2, 2) + gw * (s->W22 * s->W34 - s->W24 * s->W32);
    s->W_1_34 = DW(1, 4, 1) - DW(1, 3, 2) + gw * (s->W23 * s->W34 - s->W24 * s->W33);
    s->W_2_23 = DW(2, 3, 0) - DW(2, 2, 1) + gw * (s->W32 * s->W13 - s->W33 * s->W12);
    s->W_2_24 = DW(2, 4, 0) - DW(2, 2, 2) + gw * (s->W32 * s->W14 - s->W34 * s->W12);
    s->W_2_34 = DW(2, 4, 1) - DW(2, 3, 2) + gw * (s->W33 * s->W14 - s->W34 * s->W13);
    s->W_3_23 = DW(3, 3, 0) - DW(3, 2, 1) + gw * (s->W12 * s->W23 - s->W13 * s->W22);
    s->W_3_24 = DW(3, 4, 0) - DW(3, 2, 2) + gw * (s->W12 * s->W24 - s->W14 * s->W22);
    s->W_3_34 = DW(3, 4, 1) - DW(3, 3, 2) + gw * (s->W13 * s->W24 - s->W14 * s->W23);
#undef DW
    s->Y_2_3 = s->dY[1][0] - s->dY[0][1];
    s->Y_2_4 = s->dY[2][0] - s->dY[0][2];
    s->Y_3_4 = s->dY[2][1] - s->dY[1][2];
}

static void rhs_site(const lat_t *L, const oracle_params *P, const double *in, double *out,
                     long N3, int i, int j, int k)
{
    site_t S;
    load_site(L, P, in, N3, i, j, k, &S);
    const long c = sidx(L, i, j, k);
    const double gw = P->gw, gy = P->gy, gp2 = P->gp2, beta_W = P->beta_w, beta_Y = P->beta_y;
    const double vev = P->vev, lambda = P->lambda, Jex = P->jex, gam = P->gamma;

    /* names as in main_rk4.jl */
    const double p1 = S.p1, p2 = S.p2, p3 = S.p3, p4 = S.p4;
    const double W_1_2 = S.W12, W_1_3 = S.W13, W_1_4 = S.W14;
    const double W_2_2 = S.W22, W_2_3 = S.W23, W_2_4 = S.W24;
    const double W_3_2 = S.W32, W_3_3 = S.W33, W_3_4 = S.W34;
    const double Y_2 = S.Y2, Y_3 = S.Y3, Y_4 = S.Y4;
    const double G_1 = S.G1, G_2 = S.G2, G_3 = S.G3, Sig = S.Sg;
    const double pi1 = S.pi1, pi2 = S.pi2, pi3 = S.pi3, pi4 = S.pi4;
    const double dW_1_2_dt = S.E12, dW_1_3_dt = S.E13, dW_1_4_dt = S.E14;
    const double dW_2_2_dt = S.E22, dW_2_3_dt = S.E23, dW_2_4_dt = S.E24;
    const double dW_3_2_dt = S.E32, dW_3_3_dt = S.E33, dW_3_4_dt = S.E34;
    const double dY_2_dt = S.EY2, dY_3_dt = S.EY3, dY_4_dt = S.EY4;

    const double dp1dx = S.dp[0][0], dp1dy = S.dp[0][1], dp1dz = S.dp[0][2];
    const double dp2dx = S.dp[1][0], dp2dy = S.dp[1][1], dp2dz = S.dp[1][2];
    const double dp3dx = S.dp[2][0], dp3dy = S.dp[2][1], dp3dz = S.dp[2][2];
    const double dp4dx = S.dp[3][0], dp4dy = S.dp[3][1], dp4dz = S.dp[3][2];
#define DWd(a, i, d) S.dW[3 * ((a) - 1) + ((i) - 2)][d]
    const double dW_1_2_dx = DWd(1, 2, 0), dW_1_2_dy = DWd(1, 2, 1), dW_1_2_dz = DWd(1, 2, 2);
    const double dW_1_3_dx = DWd(1, 3, 0), dW_1_3_dy = DWd(1, 3, 1), dW_1_3_dz = DWd(1, 3, 2);
    const double dW_1_4_dx = DWd(1, 4, 0), dW_1_4_dy = DWd(1, 4, 1), dW_1_4_dz = DWd(1, 4, 2);
    const double dW_2_2_dx = DWd(2, 2, 0), dW_2_2_dy = DWd(2, 2, 1), dW_2_2_dz = DWd(2, 2, 2);
    const double dW_2_3_dx = DWd(2, 3, 0), dW_2_3_dy = DWd(2, 3, 1), dW_2_3_dz = DWd(2, 3, 2);
    const double dW_2_4_dx = DWd(2, 4, 0), dW_2_4_dy = DWd(2, 4, 1), dW_2_4_dz = DWd(2, 4, 2);
    const double dW_3_2_dx = DWd(3, 2, 0), dW_3_2_dy = DWd(3, 2, 1), dW_3_2_dz = DWd(3, 2, 2);
    const double dW_3_3_dx = DWd(3, 3, 0), dW_3_3_dy = DWd(3, 3, 1), dW_3_3_dz = DWd(3, 3, 2);
    const double dW_3_4_dx = DWd(3, 4, 0), dW_3_4_dy = DWd(3, 4, 1), dW_3_4_dz = DWd(3, 4, 2);
#undef DWd
    const double dG_1_dx = S.dG[0][0], dG_1_dy = S.dG[0][1], dG_1_dz = S.dG[0][2];
    const double dG_2_dx = S.dG[1][0], dG_2_dy = S.dG[1][1], dG_2_dz = S.dG[1][2];
    const double dG_3_dx = S.dG[2][0], dG_3_dy = S.dG[2][1], dG_3_dz = S.dG[2][2];
    const double dS_dx = S.dS[0], dS_dy = S.dS[1], dS_dz = S.dS[2];

    const double Dx_1 = S.D[0][0], Dy_1 = S.D[0][1], Dz_1 = S.D[0][2];
    const double Dx_2 = S.D[1][0], Dy_2 = S.D[1][1], Dz_2 = S.D[1][2];
    const double Dx_3 = S.D[2][0], Dy_3 = S.D[2][1], Dz_3 = S.D[2][2];
    const double Dx_4 = S.D[3][0], Dy_4 = S.D[3][1], Dz_4 = S.D[3][2];

    const double W_1_23 = S.W_1_23, W_1_24 = S.W_1_24, W_1_34 = S.W_1_34;
    const double W_2_23 = S.W_2_23, W_2_24 = S.W_2_24, W_2_34 = S.W_2_34;
    const double W_3_23 = S.W_3_23, W_3_24 = S.W_3_24, W_3_34 = S.W_3_34;
    const double Y_2_3 = S.Y_2_3, Y_2_4 = S.Y_2_4, Y_3_4 = S.Y_3_4;

    /* Laplacian pieces (main_rk4.jl:175-248) */
    double lap[16][3];
    for (int a = 0; a < 16; a++)
        for (int d = 0; d < 3; d++) lap[a][d] = d2(L, CH(a), i, j, k, d);
#define LAP(ch) (lap[ch][0] + lap[ch][1] + lap[ch][2])

    double *o = out;
#define OUT(ch) o[(long)(ch) * N3 + c]

    /* main_rk4.jl:337-355 */
    OUT(C_PI1) = (LAP(C_P1) -
        0.5 * gw * (((-W_1_2 * dp4dx) - (W_1_3 * dp4dy) - (W_1_4 * dp4dz)) -
        ((-W_2_2 * dp3dx) - (W_2_3 * dp3dy) - (W_2_4 * dp3dz)) +
        ((-W_3_2 * dp2dx) - (W_3_3 * dp2dy) - (W_3_4 * dp2dz))) -
        0.5 * gy * (-Y_2 * dp2dx - Y_3 * dp2dy - Y_4 * dp2dz) -
        0.5 * gw * ((-W_1_2 * Dx_4 - W_1_3 * Dy_4 - W_1_4 * Dz_4) -
        (-W_2_2 * Dx_3 - W_2_3 * Dy_3 - W_2_4 * Dz_3) +
        (-W_3_2 * Dx_2 - W_3_3 * Dy_2 - W_3_4 * Dz_2)) -
        0.5 * gy * (-Y_2 * Dx_2 - Y_3 * Dy_2 - Y_4 * Dz_2) -
        2.0 * lambda * (p1 * p1 + p2 * p2 + p3 * p3 + p4 * p4 - vev * vev) * p1 +
        0.5 * ((gw * G_3 + gy * Sig) * p2 - gw * G_2 * p3 + gw * G_1 * p4) -
        0.5 * Jex * p1 -
        gam * p1 * (p1 * pi1) / (p1 * p1 + p2 * p2 + p3 * p3 + p4 * p4) +
        beta_W * (dW_1_2_dt * W_1_34 + dW_1_3_dt * (-W_1_24) + dW_1_4_dt * W_1_23 +
                  dW_2_2_dt * W_2_34 + dW_1_3_dt * (-W_2_24) + dW_1_4_dt * W_2_23 +
                  dW_3_2_dt * W_3_34 + dW_1_3_dt * (-W_3_24) + dW_1_4_dt * W_3_23) * p1 +
        beta_Y * (dY_2_dt * Y_3_4 + dY_3_dt * (-Y_2_4) + dY_4_dt * Y_2_3) * p1);

    /* main_rk4.jl:355-371 */
    OUT(C_PI2) = (LAP(C_P2) +
        0.5 * gw * ((-W_1_2 * dp3dx - W_1_3 * dp3dy - W_1_4 * dp3dz) +
        (-W_2_2 * dp4dx - W_2_3 * dp4dy - W_2_4 * dp4dz) +
        (-W_3_2 * dp1dx - W_3_3 * dp1dy - W_3_4 * dp1dz)) +
        0.5 * gy * (-Y_2 * dp1dx - Y_3 * dp1dy - Y_4 * dp1dz) +
        0.5 * gw * ((-W_1_2 * Dx_3 - W_1_3 * Dy_3 - W_1_4 * Dz_3) +
        (-W_2_2 * Dx_4 - W_2_3 * Dy_4 - W_2_4 * Dz_4) +
        (-W_3_2 * Dx_1 - W_3_3 * Dy_1 - W_3_4 * Dz_1)) +
        0.5 * gy * (-Y_2 * Dx_1 - Y_3 * Dy_1 - Y_4 * Dz_1) -
        2.0 * lambda * (p1 * p1 + p2 * p2 + p3 * p3 + p4 * p4 - vev * vev) * p2 -
        0.5 * ((gw * G_3 + gy * Sig) * p1 + gw * G_1 * p3 + gw * G_2 * p4) -
        0.5 * Jex * p2 -
        gam * p2 * (p2 * pi2) / (p1 * p1 + p2 * p2 + p3 * p3 + p4 * p4) +
        beta_W * (dW_1_2_dt * W_1_34 + dW_1_3_dt * (-W_1_24) + dW_1_4_dt * W_1_23 +
                  dW_2_2_dt * W_2_34 + dW_1_3_dt * (-W_2_24) + dW_1_4_dt * W_2_23 +
                  dW_3_2_dt * W_3_34 + dW_1_3_dt * (-W_3_24) + dW_1_4_dt * W_3_23) * p2 +
        beta_Y * (dY_2_dt * Y_3_4 + dY_3_dt * (-Y_2_4) + dY_4_dt * Y_2_3) * p2);

    /* main_rk4.jl:373-389 */
    OUT(C_PI3) = (LAP(C_P3) -
        0.5 * gw * ((-W_1_2 * dp2dx - W_1_3 * dp2dy - W_1_4 * dp2dz) +
        (-W_2_2 * dp1dx - W_2_3 * dp1dy - W_2_4 * dp1dz) -
        (-W_3_2 * dp4dx - W_3_3 * dp4dy - W_3_4 * dp4dz)) -
        0.5 * gy * (-Y_2 * dp4dx - Y_3 * dp4dy - Y_4 * dp4dz) -
        0.5 * gw * ((-W_1_2 * Dx_2 - W_1_3 * Dy_2 - W_1_4 * Dz_2) +
        (-W_2_2 * Dx_1 - W_2_3 * Dy_1 - W_2_4 * Dz_1) -
        (-W_3_2 * Dx_4 - W_3_3 * Dy_4 - W_3_4 * Dz_4)) -
        0.5 * gy * (-Y_2 * Dx_4 - Y_3 * Dy_4 - Y_4 * Dz_4) -
        2.0 * lambda * (p1 * p1 + p2 * p2 + p3 * p3 + p4 * p4 - vev * vev) * p3 +
        0.5 * ((-gw * G_3 + gy * Sig) * p4 + gw * G_2 * p1 + gw * G_1 * p2) -
        0.5 * Jex * p3 -
        gam * p3 * (p3 * pi3) / (p1 * p1 + p2 * p2 + p3 * p3 + p4 * p4) +
        beta_W * (dW_1_2_dt * W_1_34 + dW_1_3_dt * (-W_1_24) + dW_1_4_dt * W_1_23 +
                  dW_2_2_dt * W_2_34 + dW_1_3_dt * (-W_2_24) + dW_1_4_dt * W_2_23 +
                  dW_3_2_dt * W_3_34 + dW_1_3_dt * (-W_3_24) + dW_1_4_dt * W_3_23) * p3 +
        beta_Y * (dY_2_dt * Y_3_4 + dY_3_dt * (-Y_2_4) + dY_4_dt * Y_2_3) * p3);

    /* main_rk4.jl:391-407 (dY_4_dt[i,1j,k] is dY_4_dt[i,j,k]) */
    OUT(C_PI4) = (LAP(C_P4) +
        0.5 * gw * ((-W_1_2 * dp1dx - W_1_3 * dp1dy - W_1_4 * dp1dz) -
        (-W_2_2 * dp2dx - W_2_3 * dp2dy - W_2_4 * dp2dz) -
        (-W_3_2 * dp3dx - W_3_3 * dp3dy - W_3_4 * dp3dz)) +
        0.5 * gy * (-Y_2 * dp3dx - Y_3 * dp3dy - Y_4 * dp3dz) +
        0.5 * gw * ((-W_1_2 * Dx_1 - W_1_3 * Dy_1 - W_1_4 * Dz_1) -
        (-W_2_2 * Dx_2 - W_2_3 * Dy_2 - W_2_4 * Dz_2) -
        (-W_3_2 * Dx_3 - W_3_3 * Dy_3 - W_3_4 * Dz_3)) +
        0.5 * gy * (-Y_2 * Dx_3 - Y_3 * Dy_3 - Y_4 * Dz_3) -
        2.0 * lambda * (p1 * p1 + p2 * p2 + p3 * p3 + p4 * p4 - vev * vev) * p4 -
        0.5 * ((-gw * G_3 + gy * Sig) * p3 + gw * G_1 * p1 - gw * G_2 * p2) -
        0.5 * Jex * p4 -
        gam * p4 * (p4 * pi4) / (p1 * p1 + p2 * p2 + p3 * p3 + p4 * p4) +
        beta_W * (dW_1_2_dt * W_1_34 + dW_1_3_dt * (-W_1_24) + dW_1_4_dt * W_1_23 +
                  dW_2_2_dt * W_2_34 + dW_1_3_dt * (-W_2_24) + dW_1_4_dt * W_2_23 +
                  dW_3_2_dt * W_3_34 + dW_1_3_dt * (-W_3_24) + dW_1_4_dt * W_3_23) * p4 +
        beta_Y * (dY_2_dt * Y_3_4 + dY_3_dt * (-Y_2_4) + dY_4_dt * Y_2_3) * p4);

    /* phi.pi and phi.dphi/d<dir> appear in every beta term */
#define PHIPI (pi1 * p1 + pi2 * p2 + pi3 * p3 + pi4 * p4)
#define PHID(d) (dp1##d * p1 + dp2##d * p2 + dp3##d * p3 + dp4##d * p4)

    /* main_rk4.jl:409-423 */
    OUT(C_E12) = (LAP(C_W12) +
        gw * (-(dW_2_2_dx * W_3_2 - dW_3_2_dx * W_2_2) -
        (dW_2_2_dy * W_3_3 - dW_3_2_dy * W_2_3) -
        (dW_2_2_dz * W_3_4 - dW_3_2_dz * W_2_4) -
        (W_2_3 * W_3_23 - W_3_3 * W_2_23) -
        (W_2_4 * W_3_24 - W_3_4 * W_2_24)) +
        gw * (p1 * Dx_4 - p2 * Dx_3 + p3 * Dx_2 - p4 * Dx_1) -
        dG_1_dx - gw * (W_2_2 * G_3 - W_3_2 * G_2) -
        beta_W * PHIPI * 2.0 * W_1_34 +
        2.0 * beta_W * (PHID(dy) * dW_1_4_dt - PHID(dz) * dW_1_3_dt));

    /* main_rk4.jl:425-439 */
    OUT(C_E13) = (LAP(C_W13) +
        gw * (-(dW_2_3_dx * W_3_2 - dW_3_3_dx * W_2_2) -
        (dW_2_3_dy * W_3_3 - dW_3_3_dy * W_2_3) -
        (dW_2_3_dz * W_3_4 - dW_3_3_dz * W_2_4) -
        (W_2_2 * (-W_3_23) - W_3_2 * (-W_2_23)) -
        (W_2_4 * W_3_34 - W_3_4 * W_2_34)) +
        gw * (p1 * Dy_4 - p2 * Dy_3 + p3 * Dy_2 - p4 * Dy_1) -
        dG_1_dy - gw * (W_2_3 * G_3 - W_3_3 * G_2) -
        beta_W * PHIPI * 2.0 * (-W_1_24) +
        2.0 * beta_W * (PHID(dz) * dW_1_2_dt - PHID(dx) * dW_1_4_dt));

    /* main_rk4.jl:441-455 */
    OUT(C_E14) = (LAP(C_W14) +
        gw * (-(dW_2_4_dx * W_3_2 - dW_3_4_dx * W_2_2) -
        (dW_2_4_dy * W_3_3 - dW_3_4_dy * W_2_3) -
        (dW_2_4_dz * W_3_4 - dW_3_4_dz * W_2_4) -
        (W_2_2 * (-W_3_24) - W_3_2 * (-W_2_24)) -
        (W_2_3 * (-W_3_34) - W_3_3 * (-W_2_34))) +
        gw * (p1 * Dz_4 - p2 * Dz_3 + p3 * Dz_2 - p4 * Dz_1) -
        dG_1_dz - gw * (W_2_4 * G_3 - W_3_4 * G_2) -
        beta_W * PHIPI * 2.0 * W_1_23 +
        2.0 * beta_W * (PHID(dx) * dW_1_3_dt - PHID(dy) * dW_1_2_dt));

    /* main_rk4.jl:457-471 */
    OUT(C_E22) = (LAP(C_W22) +
        gw * (-(dW_3_2_dx * W_1_2 - dW_1_2_dx * W_3_2) -
        (dW_3_2_dy * W_1_3 - dW_1_2_dy * W_3_3) -
        (dW_3_2_dz * W_1_4 - dW_1_2_dz * W_3_4) -
        (W_3_3 * W_1_23 - W_1_3 * W_3_23) -
        (W_3_4 * W_1_24 - W_1_4 * W_3_24)) +
        gw * (-p1 * Dx_3 - p2 * Dx_4 + p3 * Dx_1 + p4 * Dx_2) -
        dG_2_dx - gw * (W_3_2 * G_1 - W_1_2 * G_3) -
        beta_W * PHIPI * 2.0 * W_2_34 +
        2.0 * beta_W * (PHID(dy) * dW_2_4_dt - PHID(dz) * dW_2_3_dt));

    /* main_rk4.jl:473-487 */
    OUT(C_E23) = (LAP(C_W23) +
        gw * (-(dW_3_3_dx * W_1_2 - dW_1_3_dx * W_3_2) -
        (dW_3_3_dy * W_1_3 - dW_1_3_dy * W_3_3) -
        (dW_3_3_dz * W_1_4 - dW_1_3_dz * W_3_4) -
        (W_3_2 * (-W_1_23) - W_1_2 * (-W_3_23)) -
        (W_3_4 * (W_1_34) - W_1_4 * W_3_34)) +
        gw * (-p1 * Dy_3 - p2 * Dy_4 + p3 * Dy_1 + p4 * Dy_2) -
        dG_2_dy - gw * (W_3_3 * G_1 - W_1_3 * G_3) -
        beta_W * PHIPI * 2.0 * (-W_2_24) +
        2.0 * beta_W * (PHID(dz) * dW_2_2_dt - PHID(dx) * dW_2_4_dt));

    /* main_rk4.jl:489-503 */
    OUT(C_E24) = (LAP(C_W24) +
        gw * (-(dW_3_4_dx * W_1_2 - dW_1_4_dx * W_3_2) -
        (dW_3_4_dy * W_1_3 - dW_1_4_dy * W_3_3) -
        (dW_3_4_dz * W_1_4 - dW_1_4_dz * W_3_4) -
        (W_3_2 * (-W_1_24) - W_1_2 * (-W_3_24)) -
        (W_3_3 * (-W_1_34) - W_1_3 * (-W_3_34))) +
        gw * (-p1 * Dz_3 - p2 * Dz_4 + p3 * Dz_1 + p4 * Dz_2) -
        dG_2_dz - gw * (W_3_4 * G_1 - W_1_4 * G_3) -
        beta_W * PHIPI * 2.0 * W_2_23 +
        2.0 * beta_W * (PHID(dx) * dW_2_3_dt - PHID(dy) * dW_2_2_dt));

    /* main_rk4.jl:505-519 */
    OUT(C_E32) = (LAP(C_W32) +
        gw * (-(dW_1_2_dx * W_2_2 - dW_2_2_dx * W_1_2) -
        (dW_1_2_dy * W_2_3 - dW_2_2_dy * W_1_3) -
        (dW_1_2_dz * W_2_4 - dW_2_2_dz * W_1_4) -
        (W_1_3 * W_2_23 - W_2_3 * W_1_23) -
        (W_1_4 * W_2_24 - W_2_4 * W_1_24)) +
        gw * (p1 * Dx_2 - p2 * Dx_1 - p3 * Dx_4 + p4 * Dx_3) -
        dG_3_dx - gw * (W_1_2 * G_2 - W_2_2 * G_1) -
        beta_W * PHIPI * 2.0 * W_3_34 +
        2.0 * beta_W * (PHID(dy) * dW_3_4_dt - PHID(dz) * dW_3_3_dt));

    /* main_rk4.jl:521-535 */
    OUT(C_E33) = (LAP(C_W33) +
        gw * (-(dW_1_3_dx * W_2_2 - dW_2_3_dx * W_1_2) -
        (dW_1_3_dy * W_2_3 - dW_2_3_dy * W_1_3) -
        (dW_1_3_dz * W_2_4 - dW_2_3_dz * W_1_4) -
        (W_1_2 * (-W_2_23) - W_2_2 * (-W_1_23)) -
        (W_1_4 * W_2_34 - W_2_4 * (W_1_34))) +
        gw * (p1 * Dy_2 - p2 * Dy_1 - p3 * Dy_4 + p4 * Dy_3) -
        dG_3_dy - gw * (W_1_3 * G_2 - W_2_3 * G_1) -
        beta_W * PHIPI * 2.0 * (-W_3_24) +
        2.0 * beta_W * (PHID(dz) * dW_3_2_dt - PHID(dx) * dW_3_4_dt));

    /* main_rk4.jl:537-551 */
    OUT(C_E34) = (LAP(C_W34) +
        gw * (-(dW_1_4_dx * W_2_2 - dW_2_4_dx * W_1_2) -
        (dW_1_4_dy * W_2_3 - dW_2_4_dy * W_1_3) -
        (dW_1_4_dz * W_2_4 - dW_2_4_dz * W_1_4) -
        (W_1_2 * (-W_2_24) - W_2_2 * (-W_1_24)) -
        (W_1_3 * (-W_2_34) - W_2_3 * (-W_1_34))) +
        gw * (p1 * Dz_2 - p2 * Dz_1 - p3 * Dz_4 + p4 * Dz_3) -
        dG_3_dz - gw * (W_1_4 * G_2 - W_2_4 * G_1) -
        beta_W * PHIPI * 2.0 * W_3_23 +
        2.0 * beta_W * (PHID(dx) * dW_3_3_dt - PHID(dy) * dW_3_2_dt));

    /* kt_Y_2..4, main_rk4.jl:553-581 */
    OUT(C_EY2) = ((LAP(C_Y2) +
        gy * (p1 * Dx_2 - p2 * Dx_1 + p3 * Dx_4 - p4 * Dx_3) - dS_dx) -
        beta_Y * PHIPI * 2.0 * Y_3_4 +
        2.0 * beta_Y * (PHID(dy) * dY_4_dt - PHID(dz) * dY_3_dt));
    OUT(C_EY3) = ((LAP(C_Y3) +
        gy * (p1 * Dy_2 - p2 * Dy_1 + p3 * Dy_4 - p4 * Dy_3) - dS_dy) -
        beta_Y * PHIPI * 2.0 * (-Y_2_4) +
        2.0 * beta_Y * (PHID(dz) * dY_2_dt - PHID(dx) * dY_4_dt));
    OUT(C_EY4) = ((LAP(C_Y4) +
        gy * (p1 * Dz_2 - p2 * Dz_1 + p3 * Dz_4 - p4 * Dz_3) - dS_dz) -
        beta_Y * PHIPI * 2.0 * Y_2_3 +
        2.0 * beta_Y * (PHID(dx) * dY_3_dt - PHID(dy) * dY_2_dt));
#undef PHIPI
#undef PHID
#undef LAP

    /* k_f = pi copies, main_rk4.jl:585-600 */
    for (int a = 0; a < 16; a++) OUT(a) = in[(long)(C_PI1 + a) * N3 + c];

    /* k_Gamma, k_Sigma, main_rk4.jl:603-657 */
    const double divE1 = d1(L, CH(C_E12), i, j, k, 0) + d1(L, CH(C_E13), i, j, k, 1) + d1(L, CH(C_E14), i, j, k, 2);
    const double divE2 = d1(L, CH(C_E22), i, j, k, 0) + d1(L, CH(C_E23), i, j, k, 1) + d1(L, CH(C_E24), i, j, k, 2);
    const double divE3 = d1(L, CH(C_E32), i, j, k, 0) + d1(L, CH(C_E33), i, j, k, 1) + d1(L, CH(C_E34), i, j, k, 2);
    const double divEY = d1(L, CH(C_EY2), i, j, k, 0) + d1(L, CH(C_EY3), i, j, k, 1) + d1(L, CH(C_EY4), i, j, k, 2);

    OUT(C_G1) = ((1.0 - gp2) * (divE1) +
        gp2 * gw * (
        -((W_2_2) * (dW_3_2_dt) - (W_3_2) * (dW_2_2_dt)) -
        ((W_2_3) * (dW_3_3_dt) - (W_3_3) * (dW_2_3_dt)) -
        ((W_2_4) * (dW_3_4_dt) - (W_3_4) * (dW_2_4_dt))) +
        gp2 * gw * ((p1) * (pi4) - (p4) * (pi1) + (p3) * (pi2) - (p2) * (pi3)));

    OUT(C_G2) = ((1.0 - gp2) * (divE2) +
        gp2 * gw * (
        -((W_3_2) * (dW_1_2_dt) - (W_1_2) * (dW_3_2_dt)) -
        ((W_3_3) * (dW_1_3_dt) - (W_1_3) * (dW_3_3_dt)) -
        ((W_3_4) * (dW_1_4_dt) - (W_1_4) * (dW_3_4_dt))) +
        gp2 * gw * ((p3) * (pi1) - (p1) * (pi3) + (p4) * (pi2) - (p2) * (pi4)));

    OUT(C_G3) = ((1.0 - gp2) * (divE3) +
        gp2 * gw * (
        -((W_1_2) * (dW_2_2_dt) - (W_2_2) * (dW_1_2_dt)) -
        ((W_1_3) * (dW_2_3_dt) - (W_2_3) * (dW_1_3_dt)) -
        ((W_1_4) * (dW_2_4_dt) - (W_2_4) * (dW_1_4_dt))) +
        gp2 * gw * ((p1) * (pi2) - (p2) * (pi1) + (p4) * (pi3) - (p3) * (pi4)));

    OUT(C_SG) = ((1.0 - gp2) * (divEY) +
        gp2 * gy * ((p1) * (pi2) - (p2) * (pi1) + (p3) * (pi4) - (p4) * (pi3)));
#undef OUT
}

void oracle_rhs(int nx, int ny, int nz, double dx, const oracle_params *p,
                const double *in, double *out)
{
    lat_t L = {nx, ny, nz, dx};
    const long N3 = (long)nx * ny * nz;
#pragma omp parallel for schedule(static)
    for (int k = 0; k < nz; k++)
        for (int j = 0; j < ny; j++)
            for (int i = 0; i < nx; i++) rhs_site(&L, p, in, out, N3, i, j, k);
}

/* main_rk4.jl:6146-6441: stage inputs y + k*(c*dt) as temporaries, the
 * accumulator _n = k*(dt/6) + y, += k*(dt/3) twice, y = k*(dt/6) + _n. */
void oracle_rk4(int nx, int ny, int nz, double dx, const oracle_params *p,
                double *y, int nsteps)
{
    const long N = (long)nx * ny * nz * NCH;
    double *kk = (double *)malloc(sizeof(double) * N);
    double *u = (double *)malloc(sizeof(double) * N);
    double *n = (double *)malloc(sizeof(double) * N);
    const double dt = p->dt;
    for (int it = 0; it < nsteps; it++) {
        oracle_rhs(nx, ny, nz, dx, p, y, kk);
#pragma omp parallel for schedule(static)
        for (long q = 0; q < N; q++) { n[q] = kk[q] * (dt / 6.) + y[q]; u[q] = y[q] + kk[q] * (dt / 2.); }
        oracle_rhs(nx, ny, nz, dx, p, u, kk);
#pragma omp parallel for schedule(static)
        for (long q = 0; q < N; q++) { n[q] = kk[q] * (dt / 3.) + n[q]; u[q] = y[q] + kk[q] * (dt / 2.); }
        oracle_rhs(nx, ny, nz, dx, p, u, kk);
#pragma omp parallel for schedule(static)
        for (long q = 0; q < N; q++) { n[q] = kk[q] * (dt / 3.) + n[q]; u[q] = y[q] + kk[q] * dt; }
        oracle_rhs(nx, ny, nz, dx, p, u, kk);
#pragma omp parallel for schedule(static)
        for (long q = 0; q < N; q++) y[q] = kk[q] * (dt / 6.) + n[q];
    }
    free(kk); free(u); free(n);
}

/* Sums are taken per z-plane, then over planes in order, so the result does
 * not depend on the thread count. */
void oracle_energies(int nx, int ny, int nz, double dx, const oracle_params *p,
                     const double *y, double out[13])
{
    lat_t L = {nx, ny, nz, dx};
    const long N3 = (long)nx * ny * nz;
    double *pl = (double *)calloc((size_t)nz * 9, sizeof(double));
    const double *in = y;
#pragma omp parallel for schedule(static)
    for (int k = 0; k < nz; k++) {
        double acc[7] = {0}, mn = INFINITY, msum = 0.0;
        for (int j = 0; j < ny; j++)
            for (int i = 0; i < nx; i++) {
                site_t s;
                load_site(&L, p, in, N3, i, j, k, &s);
                const double S2 = s.p1 * s.p1 + s.p2 * s.p2 + s.p3 * s.p3 + s.p4 * s.p4;
                const double X = S2 - p->vev * p->vev;
                acc[0] += p->lambda * (X * X);                                        /* 1334-1343 */
                acc[1] += s.pi1 * s.pi1 + s.pi2 * s.pi2 + s.pi3 * s.pi3 + s.pi4 * s.pi4; /* 1659 */
                double ge = 0.0;                                                        /* 1498-1501 */
                for (int a = 0; a < 4; a++)
                    ge = ge + s.D[a][0] * s.D[a][0] + s.D[a][1] * s.D[a][1] + s.D[a][2] * s.D[a][2];
                acc[2] += ge;
                acc[3] += 0.5 * (s.E12 * s.E12 + s.E13 * s.E13 + s.E14 * s.E14 + s.E22 * s.E22 +
                                 s.E23 * s.E23 + s.E24 * s.E24 + s.E32 * s.E32 + s.E33 * s.E33 +
                                 s.E34 * s.E34);                                         /* 1817-1820 */
                acc[4] += 0.5 * (s.W_1_23 * s.W_1_23 + s.W_1_24 * s.W_1_24 + s.W_1_34 * s.W_1_34 +
                                 s.W_2_23 * s.W_2_23 + s.W_2_24 * s.W_2_24 + s.W_2_34 * s.W_2_34 +
                                 s.W_3_23 * s.W_3_23 + s.W_3_24 * s.W_3_24 + s.W_3_34 * s.W_3_34); /* 1978-1981 */
                acc[5] += 0.5 * (s.EY2 * s.EY2 + s.EY3 * s.EY3 + s.EY4 * s.EY4);        /* 2139-2140 */
                acc[6] += 0.5 * (s.Y_2_3 * s.Y_2_3 + s.Y_2_4 * s.Y_2_4 + s.Y_3_4 * s.Y_3_4); /* 2298 */
                const double m = sqrt(S2);
                if (m < mn) mn = m;
                msum += m;
            }
        for (int q = 0; q < 7; q++) pl[(long)k * 9 + q] = acc[q];
        pl[(long)k * 9 + 7] = mn;
        pl[(long)k * 9 + 8] = msum;
    }
    const double Vol = dx * dx * dx;
    double tot[7] = {0}, mn = INFINITY, msum = 0.0;
    for (int k = 0; k < nz; k++) {
        for (int q = 0; q < 7; q++) tot[q] += pl[(long)k * 9 + q];
        if (pl[(long)k * 9 + 7] < mn) mn = pl[(long)k * 9 + 7];
        msum += pl[(long)k * 9 + 8];
    }
    for (int q = 0; q < 7; q++) out[q] = tot[q] * Vol;
    out[7] = 0.0; out[8] = 0.0;
    out[9] = mn;
    out[10] = 0.0;
    out[11] = msum / (double)N3;
    out[12] = 0.0;
    free(pl);
}

/* Gauss law.  k_Gamma^a (main_rk4.jl:603-648) is (1-gp2)*div E^a + gp2*rho^a
 * with rho^a = gw*(non-abelian + Higgs charge); k_Sigma (651-657) likewise
 * with rho^Y = gy*(Higgs U(1) charge).  Residuals:
 *   G^a = div E^a - rho^a,  G^Y = div E^Y - rho^Y,
 *   C^a = Gamma^a - div W^a, C^Y = Sigma - div Y.
 * Each output is sqrt(sum_x r^2 * dx^3). */
void oracle_gauss(int nx, int ny, int nz, double dx, const oracle_params *p,
                  const double *y, double out[8])
{
    lat_t L = {nx, ny, nz, dx};
    const long N3 = (long)nx * ny * nz;
    const double *in = y;
    const double gw = p->gw, gy = p->gy;
    double *pl = (double *)calloc((size_t)nz * 8, sizeof(double));
#pragma omp parallel for schedule(static)
    for (int k = 0; k < nz; k++) {
        double acc[8] = {0};
        for (int j = 0; j < ny; j++)
            for (int i = 0; i < nx; i++) {
                const long c = sidx(&L, i, j, k);
#define V(ch) CH(ch)[c]
#define DIV(c0) (d1(&L, CH(c0), i, j, k, 0) + d1(&L, CH((c0) + 1), i, j, k, 1) + d1(&L, CH((c0) + 2), i, j, k, 2))
                const double p1 = V(C_P1), p2 = V(C_P2), p3 = V(C_P3), p4 = V(C_P4);
                const double pi1 = V(C_PI1), pi2 = V(C_PI2), pi3 = V(C_PI3), pi4 = V(C_PI4);
                const double rho1 = gw * (-(V(C_W22) * V(C_E32) - V(C_W32) * V(C_E22)) -
                                          (V(C_W23) * V(C_E33) - V(C_W33) * V(C_E23)) -
                                          (V(C_W24) * V(C_E34) - V(C_W34) * V(C_E24))) +
                                    gw * (p1 * pi4 - p4 * pi1 + p3 * pi2 - p2 * pi3);
                const double rho2 = gw * (-(V(C_W32) * V(C_E12) - V(C_W12) * V(C_E32)) -
                                          (V(C_W33) * V(C_E13) - V(C_W13) * V(C_E33)) -
                                          (V(C_W34) * V(C_E14) - V(C_W14) * V(C_E34))) +
                                    gw * (p3 * pi1 - p1 * pi3 + p4 * pi2 - p2 * pi4);
                const double rho3 = gw * (-(V(C_W12) * V(C_E22) - V(C_W22) * V(C_E12)) -
                                          (V(C_W13) * V(C_E23) - V(C_W23) * V(C_E13)) -
                                          (V(C_W14) * V(C_E24) - V(C_W24) * V(C_E14))) +
                                    gw * (p1 * pi2 - p2 * pi1 + p4 * pi3 - p3 * pi4);
                const double rhoY = gy * (p1 * pi2 - p2 * pi1 + p3 * pi4 - p4 * pi3);
                const double r[8] = {
                    DIV(C_E12) - rho1, DIV(C_E22) - rho2, DIV(C_E32) - rho3, DIV(C_EY2) - rhoY,
                    V(C_G1) - DIV(C_W12), V(C_G2) - DIV(C_W22), V(C_G3) - DIV(C_W32), V(C_SG) - DIV(C_Y2)};
#undef V
#undef DIV
                for (int q = 0; q < 8; q++) acc[q] += r[q] * r[q];
            }
        for (int q = 0; q < 8; q++) pl[(long)k * 8 + q] = acc[q];
    }
    const double Vol = dx * dx * dx;
    for (int q = 0; q < 8; q++) {
        double t = 0.0;
        for (int k = 0; k < nz; k++) t += pl[(long)k * 8 + q];
        out[q] = sqrt(t * Vol);
    }
    free(pl);
}

/* Electromagnetic fields, literal from the reference kernels:
 *   compute_Bx!/By!/Bz! (main_rk4.jl:2303-2822): n^a at 2459-2467, B_x at 2469-2471,
 *     B_y at 2643-2645, B_z at 2817-2819;
 *   compute_B_2_x!/y!/z! (3718-4250): the same with gw*phi_mag^2 for gw*vev^2
 *     (phi_mag = |phi|^2) and 0 where phi_mag == 0;
 *   compute_Ax!/Ay!/Az! (3012-3528; A_x at 3179, A_y at 3352, A_z at 3525):
 *     A_i = sin(th) n^a W^a_i + cos(th) Y_i.
 * out: 9 channels Bx, By, Bz, B2x, B2y, B2z, Ax, Ay, Az, each nx*ny*nz. */
void oracle_em_fields(int nx, int ny, int nz, double dx, const oracle_params *p, double theta_w,
                      const double *y, double *out)
{
    lat_t L = {nx, ny, nz, dx};
    const long N3 = (long)nx * ny * nz;
    const double *in = y;
    const double sw = sin(theta_w), cw = cos(theta_w), gw = p->gw, vev = p->vev;
#pragma omp parallel for schedule(static)
    for (int k = 0; k < nz; k++)
        for (int j = 0; j < ny; j++)
            for (int i = 0; i < nx; i++) {
                site_t s;
                load_site(&L, p, in, N3, i, j, k, &s);
                const long c = sidx(&L, i, j, k);
                const double mag = s.p1 * s.p1 + s.p2 * s.p2 + s.p3 * s.p3 + s.p4 * s.p4;
                double n1, n2, n3;
                if (mag == 0.0) {
                    n1 = 0.0; n2 = 0.0; n3 = 0.0;
                } else {
                    n1 = -2. * (s.p1 * s.p3 + s.p2 * s.p4) / mag;
                    n2 = +2. * (s.p2 * s.p3 - s.p1 * s.p4) / mag;
                    n3 = (-s.p1 * s.p1 - s.p2 * s.p2 + s.p3 * s.p3 + s.p4 * s.p4) / mag;
                }
                /* D[b][dir]: dir 0 = x, 1 = y, 2 = z */
                const double Xx = s.D[0][1] * s.D[1][2] - s.D[1][1] * s.D[0][2] + s.D[2][1] * s.D[3][2] - s.D[3][1] * s.D[2][2];
                const double Xy = s.D[0][2] * s.D[1][0] - s.D[1][2] * s.D[0][0] + s.D[2][2] * s.D[3][0] - s.D[3][2] * s.D[2][0];
                const double Xz = s.D[0][0] * s.D[1][1] - s.D[1][0] * s.D[0][1] + s.D[2][0] * s.D[3][1] - s.D[3][0] * s.D[2][1];
                const double Fx = sw * (n1 * s.W_1_34 + n2 * s.W_2_34 + n3 * s.W_3_34) + cw * s.Y_3_4;
                const double Fy = sw * (n1 * (-s.W_1_24) + n2 * (-s.W_2_24) + n3 * (-s.W_3_24)) + cw * (-s.Y_2_4);
                const double Fz = sw * (n1 * (s.W_1_23) + n2 * s.W_2_23 + n3 * s.W_3_23) + cw * s.Y_2_3;
                const double c1 = 4. * sw / (gw * vev * vev);
                out[0 * N3 + c] = -(Fx + c1 * Xx);
                out[1 * N3 + c] = -(Fy + c1 * Xy);
                out[2 * N3 + c] = -(Fz + c1 * Xz);
                if (mag == 0.0) {
                    out[3 * N3 + c] = 0.0; out[4 * N3 + c] = 0.0; out[5 * N3 + c] = 0.0;
                } else {
                    const double c2 = 4. * sw / (gw * mag * mag);
                    out[3 * N3 + c] = -(Fx + c2 * Xx);
                    out[4 * N3 + c] = -(Fy + c2 * Xy);
                    out[5 * N3 + c] = -(Fz + c2 * Xz);
                }
                out[6 * N3 + c] = sw * (n1 * s.W12 + n2 * s.W22 + n3 * s.W32) + cw * s.Y2;
                out[7 * N3 + c] = sw * (n1 * s.W13 + n2 * s.W23 + n3 * s.W33) + cw * s.Y3;
                out[8 * N3 + c] = sw * (n1 * s.W14 + n2 * s.W24 + n3 * s.W34) + cw * s.Y4;
            }
}
