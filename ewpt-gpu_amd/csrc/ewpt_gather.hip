// ewpt_gather.hip -- FP64 stage kernel, one site per thread, stencil taps
// gathered straight from global memory through the caches.
//
// This is the simple form of the rk4_step! launch (main_rk4.jl:52-661 with
// the grid (N/64,N,N) x (64,1,1) of parameters.jl:74,90-92), extended with
// the RK4 bookkeeping of ewpt_stage.h: the stage input y + c*k is formed on
// load and stage 4 writes y' directly.  It works for any lattice size >= 4.
// It also provides the energy and Gauss-law reductions.
#include <algorithm>
#include <mutex>
#include <utility>

#include "ewpt_internal.h"

namespace ewpt {

namespace {

constexpr int BX = 64, BY = 4;

struct Geo {
    int nx, ny, nz;
    double inv60dx, invdx2;
};

// Gather centre values and all derivatives of the stage-S input at (i,j,k).
template <int S>
__device__ __forceinline__ void gather_site(const StageBufs<double> &b, const Geo &g, int i, int j, int k,
                                            double dt, SiteIn<double> &s)
{
    const long nx = g.nx, nxy = (long)g.nx * g.ny;
    const long q = i + nx * j + nxy * k;
    long ox[7], oy[7], oz[7];
#pragma unroll
    for (int t = 0; t < 7; t++) {
        ox[t] = wrap_idx(i + t - 3, g.nx) - i;
        oy[t] = (wrap_idx(j + t - 3, g.ny) - j) * nx;
        oz[t] = (wrap_idx(k + t - 3, g.nz) - k) * nxy;
    }
#pragma unroll
    for (int c = 0; c < NCH; c++) s.c[c] = stage_in<S>(b, c, q, dt);

    auto line = [&](int c, const long *o, double v[7]) {
#pragma unroll
        for (int t = 0; t < 7; t++) v[t] = (t == 3) ? s.c[c] : stage_in<S>(b, c, q + o[t], dt);
    };
    // channels 0..15: gradient and Laplacian
#pragma unroll 1
    for (int c = 0; c < 16; c++) {
        double lap = 0.0, d[3];
#pragma unroll
        for (int ax = 0; ax < 3; ax++) {
            double v[7];
            line(c, ax == 0 ? ox : ax == 1 ? oy : oz, v);
            d[ax] = sten_d1(v[0], v[1], v[2], v[4], v[5], v[6], g.inv60dx);
            lap += sten_d2(v[0], v[1], v[2], v[3], v[4], v[5], v[6], g.invdx2);
        }
        s.lap[c] = lap;
#pragma unroll
        for (int ax = 0; ax < 3; ax++) {
            if (c < 4) s.dphi[c][ax] = d[ax];
            else if (c < 13) s.dW[(c - 4) / 3][(c - 4) % 3][ax] = d[ax];
            else s.dY[c - 13][ax] = d[ax];
        }
    }
    // Gamma, Sigma: gradient
#pragma unroll
    for (int c = 16; c < 20; c++) {
#pragma unroll
        for (int ax = 0; ax < 3; ax++) {
            double v[7];
            line(c, ax == 0 ? ox : ax == 1 ? oy : oz, v);
            const double d = sten_d1(v[0], v[1], v[2], v[4], v[5], v[6], g.inv60dx);
            if (c < 19) s.dG[c - 16][ax] = d; else s.dS[ax] = d;
        }
    }
    // divergences of E^a, E^Y: one direction per channel
#pragma unroll
    for (int a = 0; a < 4; a++) {
        double dv = 0.0;
#pragma unroll
        for (int ax = 0; ax < 3; ax++) {
            const int c = (a < 3) ? CH_E + 3 * a + ax : CH_EY + ax;
            double v[7];
            line(c, ax == 0 ? ox : ax == 1 ? oy : oz, v);
            dv += sten_d1(v[0], v[1], v[2], v[4], v[5], v[6], g.inv60dx);
        }
        if (a < 3) s.divE[a] = dv; else s.divEY = dv;
    }
}

template <int S>
__global__ void __launch_bounds__(BX *BY) stage_gather(StageBufs<double> b, Geo g, Phys P, double dt)
{
    const int i = blockIdx.x * BX + threadIdx.x;
    const int j = blockIdx.y * BY + threadIdx.y;
    const int k = blockIdx.z;
    if (i >= g.nx || j >= g.ny) return;
    SiteIn<double> s;
    gather_site<S>(b, g, i, j, k, dt, s);
    double o[NK];
    rhs_point(s, P, o);
    const long q = i + (long)g.nx * (j + (long)g.ny * k);
    if constexpr (S < 4) {
#pragma unroll
        for (int m = 0; m < NK; m++) b.kout[(long)m * b.n3 + q] = o[m];
    } else {
        // y' = k4*(dt/6) + (((k1*(dt/6) + y) + k2*(dt/3)) + k3*(dt/3))
        const double d6 = dt / 6.0, d3 = dt / 3.0;
#pragma unroll
        for (int c = 0; c < NCH; c++) {
            const double yv = b.y[(long)c * b.n3 + q];
            double k1, k2, k3, k4;
            if (c < 16) {
                const double piv = b.y[(long)(c + 20) * b.n3 + q];
                k1 = piv;
                k2 = piv + b.k1[(long)(c + 4) * b.n3 + q] * (dt * 0.5);
                k3 = piv + b.k2[(long)(c + 4) * b.n3 + q] * (dt * 0.5);
                k4 = piv + b.k3[(long)(c + 4) * b.n3 + q] * dt;
            } else {
                const long m = (long)(c - 16) * b.n3 + q;
                k1 = b.k1[m]; k2 = b.k2[m]; k3 = b.k3[m]; k4 = o[c - 16];
            }
            double n = k1 * d6 + yv;
            n = k2 * d3 + n;
            n = k3 * d3 + n;
            b.ynew[(long)c * b.n3 + q] = k4 * d6 + n;
        }
    }
}

// RHS only: 36 in -> 36 out (the 16 field k's are the momenta copies).
__global__ void __launch_bounds__(BX *BY) rhs_gather(const double *in, double *out, long n3, Geo g, Phys P)
{
    const int i = blockIdx.x * BX + threadIdx.x;
    const int j = blockIdx.y * BY + threadIdx.y;
    const int k = blockIdx.z;
    if (i >= g.nx || j >= g.ny) return;
    StageBufs<double> b{};
    b.y = in;
    b.n3 = n3;
    SiteIn<double> s;
    gather_site<1>(b, g, i, j, k, 0.0, s);
    double o[NK];
    rhs_point(s, P, o);
    const long q = i + (long)g.nx * (j + (long)g.ny * k);
#pragma unroll
    for (int c = 0; c < 16; c++) out[(long)c * n3 + q] = s.c[CH_PI + c];
#pragma unroll
    for (int m = 0; m < NK; m++) out[(long)(16 + m) * n3 + q] = o[m];
}

// ---------------------------------------------------------------- reductions

constexpr int NRED = 16;  // partial slots per block

__device__ __forceinline__ double wave_sum(double v)
{
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}
__device__ __forceinline__ double wave_min(double v)
{
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v = fmin(v, __shfl_xor(v, off, 64));
    return v;
}

// Block-reduce nq values (slot q uses min if minmask bit q set), write to part[blk*NRED + q].
template <int NQ>
__device__ void block_reduce(double (&v)[NQ], unsigned minmask, double *part)
{
    __shared__ double red[(BX * BY) / 64][NQ];
    const int tid = threadIdx.x + BX * threadIdx.y;
    const int w = tid >> 6, lane = tid & 63;
#pragma unroll
    for (int q = 0; q < NQ; q++) {
        const double r = (minmask >> q & 1u) ? wave_min(v[q]) : wave_sum(v[q]);
        if (lane == 0) red[w][q] = r;
    }
    __syncthreads();
    if (tid < NQ) {
        double r = red[0][tid];
        for (int ww = 1; ww < (BX * BY) / 64; ww++)
            r = (minmask >> tid & 1u) ? fmin(r, red[ww][tid]) : r + red[ww][tid];
        const long blk = blockIdx.x + (long)gridDim.x * (blockIdx.y + (long)gridDim.y * blockIdx.z);
        part[blk * NRED + tid] = r;
    }
}

// Energy densities at one site, the reference's compute_* kernels:
// e[0] PE (main_rk4.jl:1341-1342), e[1] KE (1659), e[2] GE (1498-1501),
// e[3] EEW (1817-1820), e[4] MEW (1978-1981), e[5] EEY (2139-2140), e[6] MEY (2298).
// The terms that need only centre values: e[0], e[1], e[3], e[5].
__device__ __forceinline__ void energy_centre_terms(const SiteIn<double> &s, const Phys &P, double e[7])
{
    const double r2 = s.c[0] * s.c[0] + s.c[1] * s.c[1] + s.c[2] * s.c[2] + s.c[3] * s.c[3];
    const double X = r2 - P.vev * P.vev;
    e[0] = P.lam * (X * X);
    e[1] = s.c[CH_PI] * s.c[CH_PI] + s.c[CH_PI + 1] * s.c[CH_PI + 1] + s.c[CH_PI + 2] * s.c[CH_PI + 2] +
           s.c[CH_PI + 3] * s.c[CH_PI + 3];
    double ee = 0.0;
#pragma unroll
    for (int q = 0; q < 9; q++) ee += s.c[CH_E + q] * s.c[CH_E + q];
    e[3] = 0.5 * ee;
    e[5] = 0.5 * (s.c[CH_EY] * s.c[CH_EY] + s.c[CH_EY + 1] * s.c[CH_EY + 1] + s.c[CH_EY + 2] * s.c[CH_EY + 2]);
}

// e[2] = sum |D_i phi|^2: centres of phi, W, Y and the phi gradients.
__device__ __forceinline__ double energy_gradient_term(const SiteIn<double> &s, const Phys &P)
{
    double D[4][3];
    cov_derivs(s, P.gw, P.gy, D);
    double ge = 0.0;
#pragma unroll
    for (int a = 0; a < 4; a++) ge += D[a][0] * D[a][0] + D[a][1] * D[a][1] + D[a][2] * D[a][2];
    return ge;
}

// e[4], e[6]: centres of W and the W, Y gradients.
__device__ __forceinline__ void energy_magnetic_terms(const SiteIn<double> &s, const Phys &P, double e[7])
{
    double B[3][3], BYv[3];
    field_strengths(s, P.gw, B, BYv);
    double me = 0.0;
#pragma unroll
    for (int a = 0; a < 3; a++) me += B[a][0] * B[a][0] + B[a][1] * B[a][1] + B[a][2] * B[a][2];
    e[4] = 0.5 * me;
    e[6] = 0.5 * (BYv[0] * BYv[0] + BYv[1] * BYv[1] + BYv[2] * BYv[2]);
}

__device__ __forceinline__ void energy_terms(const SiteIn<double> &s, const Phys &P, double e[7])
{
    energy_centre_terms(s, P, e);
    e[2] = energy_gradient_term(s, P);
    energy_magnetic_terms(s, P, e);
}

// slots: 0 PE, 1 KE, 2 GE, 3 EEW, 4 MEW, 5 EEY, 6 MEY, 7 sum|phi|, 8 min|phi|,
// 9 count of non-finite densities.
__global__ void __launch_bounds__(BX *BY) energy_density(const double *y, long n3, Geo g, Phys P, double *part)
{
    const int i = blockIdx.x * BX + threadIdx.x;
    const int j = blockIdx.y * BY + threadIdx.y;
    const int k = blockIdx.z;
    double v[10] = {0, 0, 0, 0, 0, 0, 0, 0, INFINITY, 0};
    if (i < g.nx && j < g.ny) {
        StageBufs<double> b{};
        b.y = y;
        b.n3 = n3;
        SiteIn<double> s;
        gather_site<1>(b, g, i, j, k, 0.0, s);
        energy_terms(s, P, v);
        const double m = sqrt(s.c[0] * s.c[0] + s.c[1] * s.c[1] + s.c[2] * s.c[2] + s.c[3] * s.c[3]);
        v[7] = m;
        v[8] = m;
        int bad = 0;
#pragma unroll
        for (int q = 0; q < 8; q++) bad += !isfinite(v[q]);
        v[9] = (double)bad;
    }
    block_reduce<10>(v, 1u << 8, part);
}

// The same densities on 16x4x4 bricks of 256 threads: each of the 16
// stencilled channels (phi, W, Y) is staged once with its 3-deep star halo
// in LDS and differentiated there, so a value is read from memory once per
// brick instead of once per tap.  The derivatives use sten_d1 on the same
// taps as gather_site, so every site's densities equal energy_density's bit
// for bit; only the grouping of the block sums differs.
constexpr int EBX = 16, EBY = 4, EBZ = 4, EBT = EBX * EBY * EBZ;
static_assert(EBT == BX * BY, "block_reduce assumes BX*BY threads");
constexpr int EHX = EBX + 6, EHY = EBY + 6, EHZ = EBZ + 6, EBOX = EHX * EHY * EHZ;

// Stage channel yc of brick (i0, j0, k0) with its star halo into the tile
// and return the three first derivatives at thread t's site.  Not inlined:
// inlined into one body, the 16 passes kept ~500 values live at once.  One
// copy per calling kernel (Tag): a shared copy cost the energy kernel spills.
// DIR >= 0: only the faces across axis DIR are staged and only d[DIR] is
// formed (the one-direction derivatives of the Gauss residuals).
template <int Tag, int DIR = -1>
__device__ __noinline__ void energy_pass(const __attribute__((address_space(1))) double *yc, int nx, int ny, int nz,
                                         double inv60dx, int i0, int j0, int k0, int t, double *tile, double *d)
{
    // all of this thread's loads in flight at once, then into the tile
    constexpr int NIT = (EBOX + EBT - 1) / EBT;
    double val[NIT];
    bool use[NIT];
#pragma unroll
    for (int it = 0; it < NIT; it++) {
        const int e = t + it * EBT;
        const int lx = e % EHX, ly = (e / EHX) % EHY, lz = e / (EHX * EHY);
        const int out = (lx < 3 || lx >= EBX + 3) + (ly < 3 || ly >= EBY + 3) + (lz < 3 || lz >= EBZ + 3);
        const bool face_ok = DIR < 0 || (DIR == 0 ? !(ly < 3 || ly >= EBY + 3) && !(lz < 3 || lz >= EBZ + 3)
                                       : DIR == 1 ? !(lx < 3 || lx >= EBX + 3) && !(lz < 3 || lz >= EBZ + 3)
                                                  : !(lx < 3 || lx >= EBX + 3) && !(ly < 3 || ly >= EBY + 3));
        use[it] = e < EBOX && out <= 1 && face_ok;   // edges and corners of the box are never read
        if (use[it]) {
            const int gi = wrap_idx(i0 + lx - 3, nx), gj = wrap_idx(j0 + ly - 3, ny),
                      gk = wrap_idx(k0 + lz - 3, nz);
            val[it] = yc[gi + (long)nx * (gj + (long)ny * gk)];
        }
    }
    __syncthreads();   // the previous channel's reads of the tile are done
#pragma unroll
    for (int it = 0; it < NIT; it++)
        if (use[it]) tile[t + it * EBT] = val[it];
    __syncthreads();
    const int tx = t % EBX, ty = (t / EBX) % EBY, tz = t / (EBX * EBY);
    const int c0 = (tx + 3) + EHX * ((ty + 3) + EHY * (tz + 3));
    constexpr int sx = 1, sy = EHX, sz = EHX * EHY;
    if constexpr (DIR < 0 || DIR == 0)
        d[0] = sten_d1(tile[c0 - 3 * sx], tile[c0 - 2 * sx], tile[c0 - sx], tile[c0 + sx], tile[c0 + 2 * sx],
                       tile[c0 + 3 * sx], inv60dx);
    if constexpr (DIR < 0 || DIR == 1)
        d[1] = sten_d1(tile[c0 - 3 * sy], tile[c0 - 2 * sy], tile[c0 - sy], tile[c0 + sy], tile[c0 + 2 * sy],
                       tile[c0 + 3 * sy], inv60dx);
    if constexpr (DIR < 0 || DIR == 2)
        d[2] = sten_d1(tile[c0 - 3 * sz], tile[c0 - 2 * sz], tile[c0 - sz], tile[c0 + sz], tile[c0 + 2 * sz],
                       tile[c0 + 3 * sz], inv60dx);
}

template <int Tag, int C>
__device__ __forceinline__ void energy_stage_channel(const double *y, long n3, const Geo &g, int i0, int j0, int k0,
                                                     int t, double *tile, SiteIn<double> &s)
{
    // global address space: generic (FLAT) loads would also count against
    // the LDS wait of the barriers
    const __attribute__((address_space(1))) double *yc =
        (const __attribute__((address_space(1))) double *)(y + (long)C * n3);
    double d[3];
    energy_pass<Tag>(yc, g.nx, g.ny, g.nz, g.inv60dx, i0, j0, k0, t, tile, d);
#pragma unroll
    for (int ax = 0; ax < 3; ax++) {
        if constexpr (C < 4) s.dphi[C][ax] = d[ax];
        else if constexpr (C < 13) s.dW[(C - 4) / 3][(C - 4) % 3][ax] = d[ax];
        else s.dY[C - 13][ax] = d[ax];
    }
}

template <int Tag, int... Cs>
__device__ __forceinline__ void energy_stage_all(const double *y, long n3, const Geo &g, int i0, int j0, int k0,
                                                 int t, double *tile, SiteIn<double> &s,
                                                 std::integer_sequence<int, Cs...>)
{
    (energy_stage_channel<Tag, Cs>(y, n3, g, i0, j0, k0, t, tile, s), ...);
}

// The ten reduction slots of brick b at thread t's site (energy_density's
// per-site values, computed on the LDS-staged brick).
template <int Tag>
__device__ __forceinline__ void energy_brick_site(const double *y, long n3, const Geo &g, const Phys &P, long b, int t,
                                                  double *tile, double (&v)[10])
{
    const int nbx = g.nx / EBX, nby = g.ny / EBY;
    const int bx = (int)(b % nbx), by = (int)((b / nbx) % nby), bz = (int)(b / ((long)nbx * nby));
    const int i0 = bx * EBX, j0 = by * EBY, k0 = bz * EBZ;
    const int tx = t % EBX, ty = (t / EBX) % EBY, tz = t / (EBX * EBY);
    const long q = (i0 + tx) + (long)g.nx * ((j0 + ty) + (long)g.ny * (k0 + tz));
    SiteIn<double> s;
#pragma unroll
    for (int qq = 0; qq < 10; qq++) v[qq] = 0.0;
#pragma unroll
    for (int c = 0; c < NCH; c++)
        if (c < 16 || c >= CH_PI) s.c[c] = y[(long)c * n3 + q];
    energy_centre_terms(s, P, v);   // the momenta are not needed after this
    // phi passes, then the gradient term (the phi gradients die), then W and Y
    energy_stage_all<Tag>(y, n3, g, i0, j0, k0, t, tile, s, std::integer_sequence<int, 0, 1, 2, 3>{});
    v[2] = energy_gradient_term(s, P);
    energy_stage_all<Tag>(y, n3, g, i0, j0, k0, t, tile, s,
                          std::integer_sequence<int, 4, 5, 6, 7, 8, 9, 10, 11, 12, 13, 14, 15>{});
    energy_magnetic_terms(s, P, v);
    const double m = sqrt(s.c[0] * s.c[0] + s.c[1] * s.c[1] + s.c[2] * s.c[2] + s.c[3] * s.c[3]);
    v[7] = m;
    v[8] = m;
    int bad = 0;
#pragma unroll
    for (int qq = 0; qq < 8; qq++) bad += !isfinite(v[qq]);
    v[9] = (double)bad;
}

__global__ void __launch_bounds__(EBT) __attribute__((amdgpu_waves_per_eu(3))) energy_brick(const double *y, long n3, Geo g, Phys P, double *part)
{
    __shared__ double tile[EBOX];
    double v[10];
    energy_brick_site<0>(y, n3, g, P, blockIdx.x, (int)threadIdx.x, tile, v);
    block_reduce<10>(v, 1u << 8, part);
}

// The same densities with a bounded number of partials: block b sums the
// bricks b, b + gridDim.x, ... in that order before its block reduction, so
// the partials need gridDim.x slots whatever the lattice size.
__global__ void __launch_bounds__(EBT) __attribute__((amdgpu_waves_per_eu(3)))
energy_brick_strided(const double *y, long n3, Geo g, Phys P, long nbrick, double *part)
{
    __shared__ double tile[EBOX];
    double acc[10] = {0, 0, 0, 0, 0, 0, 0, 0, INFINITY, 0};
    for (long b = blockIdx.x; b < nbrick; b += gridDim.x) {   // the same trip count for every thread of a block
        double v[10];
        energy_brick_site<3>(y, n3, g, P, b, (int)threadIdx.x, tile, v);
#pragma unroll
        for (int q = 0; q < 10; q++) acc[q] = q == 8 ? fmin(acc[q], v[q]) : acc[q] + v[q];
    }
    block_reduce<10>(acc, 1u << 8, part);
}

bool energy_brick_fits(const ewpt_lattice &L)
{
    return L.nx % EBX == 0 && L.ny % EBY == 0 && L.nz % EBZ == 0 && L.nx >= EBX && L.ny >= EBY && L.nz >= EBZ;
}

// The reference's y-slice snapshot (slices_arr, main_rk4.jl:5700-5930,
// 6450-6720) on the plane y = j0: per (x, z), times dx^3,
//   0 PE, 1 KE, 2 GE, 3 MEW + EEW + MEY + EEY, 4 Bx + By + Bz, 5 B2x + B2y + B2z,
//   6 zero (the seventh slice type is never written).
// out: 7 planes of nx*nz, x fastest.
__global__ void __launch_bounds__(256) y_slice_kernel(const double *y, long n3, Geo g, Phys P, double sw, double cw,
                                                      int j0, double vol, double *out)
{
    const long np = (long)g.nx * g.nz;
    const long q = blockIdx.x * 256L + threadIdx.x;
    if (q >= np) return;
    const int i = (int)(q % g.nx), k = (int)(q / g.nx);
    StageBufs<double> b{};
    b.y = y;
    b.n3 = n3;
    SiteIn<double> s;
    gather_site<1>(b, g, i, j0, k, 0.0, s);
    double e[7];
    energy_terms(s, P, e);
    double Bv[3], B2[3], A[3];
    em_point(s, P, sw, cw, Bv, B2, A);
    out[0 * np + q] = e[0] * vol;
    out[1 * np + q] = e[1] * vol;
    out[2 * np + q] = e[2] * vol;
    out[3 * np + q] = e[4] * vol + e[3] * vol + e[6] * vol + e[5] * vol;
    out[4 * np + q] = Bv[0] * vol + Bv[1] * vol + Bv[2] * vol;
    out[5 * np + q] = B2[0] * vol + B2[1] * vol + B2[2] * vol;
    out[6 * np + q] = 0.0;
}

// Gauss residuals (squared), slots 0..7 as documented in ewpt.h; slots 8, 9:
// sum |E^a|^2, sum |E^Y|^2.
// Squared Gauss residuals and the E scales at one site (slots as in
// gauss_density): needs the centres, divE, divEY and the diagonal W, Y
// derivatives of s.
__device__ __forceinline__ void gauss_terms(const SiteIn<double> &s, const Phys &P, double v[10])
{
    const double p[4] = {s.c[0], s.c[1], s.c[2], s.c[3]};
    const double pi[4] = {s.c[CH_PI], s.c[CH_PI + 1], s.c[CH_PI + 2], s.c[CH_PI + 3]};
#pragma unroll
    for (int a = 0; a < 3; a++) {
        double na = 0.0;
#pragma unroll
        for (int ii = 0; ii < 3; ii++)
            na -= xprod(a, s.c[CH_W + ii], s.c[CH_W + 3 + ii], s.c[CH_W + 6 + ii], s.c[CH_E + ii],
                        s.c[CH_E + 3 + ii], s.c[CH_E + 6 + ii]);
        const double r = s.divE[a] - P.gw * (na + phiT(a, p, pi));
        v[a] = r * r;
        const double dW = s.dW[a][0][0] + s.dW[a][1][1] + s.dW[a][2][2];
        const double cc = s.c[CH_GAM + a] - dW;
        v[4 + a] = cc * cc;
    }
    const double rY = s.divEY - P.gy * phiTY(p, pi);
    v[3] = rY * rY;
    const double cY = s.c[CH_SIG] - (s.dY[0][0] + s.dY[1][1] + s.dY[2][2]);
    v[7] = cY * cY;
    // |E^a|^2 and |E^Y|^2, the scale the residuals are measured against
#pragma unroll
    for (int q = 0; q < 9; q++) v[8] += s.c[CH_E + q] * s.c[CH_E + q];
    v[9] = s.c[CH_EY] * s.c[CH_EY] + s.c[CH_EY + 1] * s.c[CH_EY + 1] + s.c[CH_EY + 2] * s.c[CH_EY + 2];
}

// slots: 0-2 G^a, 3 G^Y, 4-6 C^a, 7 C^Y (squared), 8 |E^a|^2, 9 |E^Y|^2
__global__ void __launch_bounds__(BX *BY) gauss_density(const double *y, long n3, Geo g, Phys P, double *part)
{
    const int i = blockIdx.x * BX + threadIdx.x;
    const int j = blockIdx.y * BY + threadIdx.y;
    const int k = blockIdx.z;
    double v[10] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
    if (i < g.nx && j < g.ny) {
        StageBufs<double> b{};
        b.y = y;
        b.n3 = n3;
        SiteIn<double> s;
        gather_site<1>(b, g, i, j, k, 0.0, s);
        gauss_terms(s, P, v);
    }
    block_reduce<10>(v, 0u, part);
}

// Electromagnetic fields (em_point): the channel groups selected by mask
// (bit 0: B, bit 1: B_2, bit 2: A) are written consecutively to out, three
// channels per group.
__global__ void __launch_bounds__(BX *BY) em_fields_kernel(const double *y, long n3, Geo g, Phys P, double sw,
                                                           double cw, unsigned mask, double *out)
{
    const int i = blockIdx.x * BX + threadIdx.x;
    const int j = blockIdx.y * BY + threadIdx.y;
    const int k = blockIdx.z;
    if (i >= g.nx || j >= g.ny) return;
    StageBufs<double> b{};
    b.y = y;
    b.n3 = n3;
    SiteIn<double> s;
    gather_site<1>(b, g, i, j, k, 0.0, s);
    double F[3][3];
    em_point(s, P, sw, cw, F[0], F[1], F[2]);
    const long q = i + (long)g.nx * (j + (long)g.ny * k);
    int o = 0;
#pragma unroll
    for (int grp = 0; grp < 3; grp++) {
        if (!(mask >> grp & 1u)) continue;
#pragma unroll
        for (int c = 0; c < 3; c++) out[(long)(o + c) * n3 + q] = F[grp][c];
        o += 3;
    }
}

// slots: 0 sum |B|^2, 1 sum |B_2|^2, 2 sum B_2 . A, 3 count of non-finite values
__global__ void __launch_bounds__(BX *BY) em_density(const double *y, long n3, Geo g, Phys P, double sw, double cw,
                                                     double *part)
{
    const int i = blockIdx.x * BX + threadIdx.x;
    const int j = blockIdx.y * BY + threadIdx.y;
    const int k = blockIdx.z;
    double v[4] = {0, 0, 0, 0};
    if (i < g.nx && j < g.ny) {
        StageBufs<double> b{};
        b.y = y;
        b.n3 = n3;
        SiteIn<double> s;
        gather_site<1>(b, g, i, j, k, 0.0, s);
        double Bv[3], B2[3], A[3];
        em_point(s, P, sw, cw, Bv, B2, A);
        v[0] = Bv[0] * Bv[0] + Bv[1] * Bv[1] + Bv[2] * Bv[2];
        v[1] = B2[0] * B2[0] + B2[1] * B2[1] + B2[2] * B2[2];
        v[2] = B2[0] * A[0] + B2[1] * A[1] + B2[2] * A[2];
        v[3] = (double)(!isfinite(v[0]) + !isfinite(v[1]) + !isfinite(v[2]));
    }
    block_reduce<4>(v, 0u, part);
}

// em_density on the bricks of energy_brick (the same staged derivatives,
// the same em_point: every site's values equal em_density's bit for bit).
__global__ void __launch_bounds__(EBT) __attribute__((amdgpu_waves_per_eu(3)))
em_brick(const double *y, long n3, Geo g, Phys P, double sw, double cw, double *part)
{
    __shared__ double tile[EBOX];
    const int t = threadIdx.x;
    const int nbx = g.nx / EBX, nby = g.ny / EBY;
    const int bx = blockIdx.x % nbx, by = (blockIdx.x / nbx) % nby, bz = blockIdx.x / (nbx * nby);
    const int i0 = bx * EBX, j0 = by * EBY, k0 = bz * EBZ;
    const int tx = t % EBX, ty = (t / EBX) % EBY, tz = t / (EBX * EBY);
    const long q = (i0 + tx) + (long)g.nx * ((j0 + ty) + (long)g.ny * (k0 + tz));
    SiteIn<double> s;
#pragma unroll
    for (int c = 0; c < 16; c++) s.c[c] = y[(long)c * n3 + q];
    energy_stage_all<1>(y, n3, g, i0, j0, k0, t, tile, s, std::make_integer_sequence<int, 16>{});
    double Bv[3], B2[3], A[3];
    em_point(s, P, sw, cw, Bv, B2, A);
    double v[4];
    v[0] = Bv[0] * Bv[0] + Bv[1] * Bv[1] + Bv[2] * Bv[2];
    v[1] = B2[0] * B2[0] + B2[1] * B2[1] + B2[2] * B2[2];
    v[2] = B2[0] * A[0] + B2[1] * A[1] + B2[2] * A[2];
    v[3] = (double)(!isfinite(v[0]) + !isfinite(v[1]) + !isfinite(v[2]));
    block_reduce<4>(v, 0u, part);
}

// The first derivative along AX of channel C at thread t's site, staged
// through the brick tile with the faces across AX only.
template <int Tag, int C, int AX>
__device__ __forceinline__ double brick_deriv(const double *y, long n3, const Geo &g, int i0, int j0, int k0, int t,
                                              double *tile)
{
    double d[3];
    energy_pass<Tag, AX>((const __attribute__((address_space(1))) double *)(y + (long)C * n3), g.nx, g.ny, g.nz,
                         g.inv60dx, i0, j0, k0, t, tile, d);
    return d[AX];
}

// gauss_density on the bricks of energy_brick: divergences summed in the
// order of gather_site, the same gauss_terms, so every site's values equal
// gauss_density's bit for bit.
__global__ void __launch_bounds__(EBT) __attribute__((amdgpu_waves_per_eu(3)))
gauss_brick(const double *y, long n3, Geo g, Phys P, double *part)
{
    __shared__ double tile[EBOX];
    const int t = threadIdx.x;
    const int nbx = g.nx / EBX, nby = g.ny / EBY;
    const int bx = blockIdx.x % nbx, by = (blockIdx.x / nbx) % nby, bz = blockIdx.x / (nbx * nby);
    const int i0 = bx * EBX, j0 = by * EBY, k0 = bz * EBZ;
    const int tx = t % EBX, ty = (t / EBX) % EBY, tz = t / (EBX * EBY);
    const long q = (i0 + tx) + (long)g.nx * ((j0 + ty) + (long)g.ny * (k0 + tz));
    SiteIn<double> s;
#pragma unroll
    for (int c = 0; c < NCH; c++)
        if (c < CH_Y || c >= CH_GAM) s.c[c] = y[(long)c * n3 + q];
    // divergences: channel E^a_i (CH_E + 3a + i) along i, added in axis order
    auto div3 = [&](auto c0) {
        constexpr int C0 = decltype(c0)::value;
        double dv = 0.0;
        dv += brick_deriv<2, C0, 0>(y, n3, g, i0, j0, k0, t, tile);
        dv += brick_deriv<2, C0 + 1, 1>(y, n3, g, i0, j0, k0, t, tile);
        dv += brick_deriv<2, C0 + 2, 2>(y, n3, g, i0, j0, k0, t, tile);
        return dv;
    };
    s.divE[0] = div3(std::integral_constant<int, CH_E>{});
    s.divE[1] = div3(std::integral_constant<int, CH_E + 3>{});
    s.divE[2] = div3(std::integral_constant<int, CH_E + 6>{});
    s.divEY = div3(std::integral_constant<int, CH_EY>{});
    // diagonal derivatives d_i W^a_i (channel CH_W + 3a + i) and d_i Y_i
    auto diag3 = [&](auto c0, double (&out)[3]) {
        constexpr int C0 = decltype(c0)::value;
        out[0] = brick_deriv<2, C0, 0>(y, n3, g, i0, j0, k0, t, tile);
        out[1] = brick_deriv<2, C0 + 1, 1>(y, n3, g, i0, j0, k0, t, tile);
        out[2] = brick_deriv<2, C0 + 2, 2>(y, n3, g, i0, j0, k0, t, tile);
    };
    double dd[3];
    diag3(std::integral_constant<int, CH_W>{}, dd);
    s.dW[0][0][0] = dd[0]; s.dW[0][1][1] = dd[1]; s.dW[0][2][2] = dd[2];
    diag3(std::integral_constant<int, CH_W + 3>{}, dd);
    s.dW[1][0][0] = dd[0]; s.dW[1][1][1] = dd[1]; s.dW[1][2][2] = dd[2];
    diag3(std::integral_constant<int, CH_W + 6>{}, dd);
    s.dW[2][0][0] = dd[0]; s.dW[2][1][1] = dd[1]; s.dW[2][2][2] = dd[2];
    diag3(std::integral_constant<int, CH_Y>{}, dd);
    s.dY[0][0] = dd[0]; s.dY[1][1] = dd[1]; s.dY[2][2] = dd[2];
    double v[10] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
    gauss_terms(s, P, v);
    block_reduce<10>(v, 0u, part);
}

// Fixed-order sum (or min, slots set in minmask) over nblk rows of partials,
// `stride` doubles per row, slots 0..ns-1 used.  Two passes:
// rows_partial: block g reduces the rows [g R, min((g+1) R, nblk)) of every
// slot -- thread (q, j) = (t % ns, t / ns) takes rows j, j + per, ... of slot
// q, so a sweep of the block reads per consecutive rows -- then combines the
// per threads of each slot in order of j and writes the result over the
// chunk's first row, which no other block reads; rows_final: block q reduces
// slot q of those chunk rows.  The grouping depends on nblk only, so every
// call on the same partials gives the same result.
__global__ void __launch_bounds__(256) rows_partial(double *part, long nblk, int stride, int ns, long R,
                                                    unsigned minmask)
{
    __shared__ double sh[256];
    const int t = threadIdx.x, per = 256 / ns, q = t % ns, j = t / ns;
    const long b0 = (long)blockIdx.x * R, b1 = b0 + R < nblk ? b0 + R : nblk;
    const bool mn = (minmask >> q) & 1u;
    double acc = mn ? INFINITY : 0.0;
    if (j < per)
        for (long b = b0 + j; b < b1; b += per) {
            const double v = part[b * stride + q];
            acc = mn ? fmin(acc, v) : acc + v;
        }
    sh[t] = acc;
    __syncthreads();
    if (t < ns) {
        double a = sh[t];
        for (int jj = 1; jj < per; jj++) a = mn ? fmin(a, sh[jj * ns + t]) : a + sh[jj * ns + t];
        part[b0 * stride + t] = a;
    }
}

__global__ void __launch_bounds__(256) rows_final(const double *part, int stride, long R, long G, unsigned minmask,
                                                  double *res)
{
    __shared__ double sh[256];
    const int q = blockIdx.x;
    const bool mn = (minmask >> q) & 1u;
    double acc = mn ? INFINITY : 0.0;
    for (long g = threadIdx.x; g < G; g += 256) {
        const double v = part[g * R * stride + q];
        acc = mn ? fmin(acc, v) : acc + v;
    }
    sh[threadIdx.x] = acc;
    __syncthreads();
    for (int st = 128; st > 0; st >>= 1) {
        if (threadIdx.x < st) sh[threadIdx.x] = mn ? fmin(sh[threadIdx.x], sh[threadIdx.x + st]) : sh[threadIdx.x] + sh[threadIdx.x + st];
        __syncthreads();
    }
    if (threadIdx.x == 0) res[q] = sh[0];
}

// The 13-column energy table from the reduced slots (the host-side scaling
// of ewpt_energies, done on the device): columns 0..6 times dx^3, 9 min|phi|,
// 11 mean|phi|, the rest 0.
__global__ void energy_table(const double *res, double vol, double nsites, double *out)
{
    const int q = threadIdx.x;
    if (q >= 13) return;
    out[q] = q < 7 ? res[q] * vol : q == 9 ? res[8] : q == 11 ? res[7] / nsites : 0.0;
}

Geo make_geo(const ewpt_lattice &L)
{
    Geo g;
    g.nx = L.nx; g.ny = L.ny; g.nz = L.nz;
    g.inv60dx = 1.0 / (60.0 * L.dx);
    g.invdx2 = 1.0 / (L.dx * L.dx);
    return g;
}

dim3 grid_of(const ewpt_lattice &L) { return dim3((L.nx + BX - 1) / BX, (L.ny + BY - 1) / BY, L.nz); }

}  // namespace

hipError_t launch_stage_gather(int stage, const StageBufs<double> &b, const ewpt_lattice &L, const Phys &P,
                               double dt, hipStream_t st)
{
    const Geo g = make_geo(L);
    const dim3 grid = grid_of(L), block(BX, BY, 1);
    switch (stage) {
    case 1: stage_gather<1><<<grid, block, 0, st>>>(b, g, P, dt); break;
    case 2: stage_gather<2><<<grid, block, 0, st>>>(b, g, P, dt); break;
    case 3: stage_gather<3><<<grid, block, 0, st>>>(b, g, P, dt); break;
    default: stage_gather<4><<<grid, block, 0, st>>>(b, g, P, dt); break;
    }
    return hipGetLastError();
}

hipError_t launch_rhs_gather(const double *in, double *out, const ewpt_lattice &L, const Phys &P, hipStream_t st)
{
    const Geo g = make_geo(L);
    rhs_gather<<<grid_of(L), dim3(BX, BY, 1), 0, st>>>(in, out, (long)L.nx * L.ny * L.nz, g, P);
    return hipGetLastError();
}

long reduce_blocks(const ewpt_lattice &L)
{
    const dim3 gr = grid_of(L);
    return (long)gr.x * gr.y * gr.z;
}

hipError_t launch_energies(const double *y, const ewpt_lattice &L, const Phys &P, double *part, double *res,
                           hipStream_t st)
{
    const Geo g = make_geo(L);
    const long n3 = (long)L.nx * L.ny * L.nz;
    if (energy_brick_fits(L)) {
        const long nb = n3 / EBT;   // <= reduce_blocks(L), the partials' capacity
        energy_brick<<<dim3((unsigned)nb), dim3(EBT), 0, st>>>(y, n3, g, P, part);
        return launch_fixed_reduce(part, nb, NRED, 10, 1u << 8, res, st);
    }
    energy_density<<<grid_of(L), dim3(BX, BY, 1), 0, st>>>(y, n3, g, P, part);
    return launch_fixed_reduce(part, reduce_blocks(L), NRED, 10, 1u << 8, res, st);
}

hipError_t launch_gauss(const double *y, const ewpt_lattice &L, const Phys &P, double *part, double *res,
                        hipStream_t st)
{
    const Geo g = make_geo(L);
    const long n3 = (long)L.nx * L.ny * L.nz;
    if (energy_brick_fits(L)) {
        const long nb = n3 / EBT;   // <= reduce_blocks(L), the partials' capacity
        gauss_brick<<<dim3((unsigned)nb), dim3(EBT), 0, st>>>(y, n3, g, P, part);
        return launch_fixed_reduce(part, nb, NRED, 10, 0u, res, st);
    }
    gauss_density<<<grid_of(L), dim3(BX, BY, 1), 0, st>>>(y, n3, g, P, part);
    return launch_fixed_reduce(part, reduce_blocks(L), NRED, 10, 0u, res, st);
}

hipError_t launch_fixed_reduce(double *part, long nblk, int stride, int ns, unsigned minmask, double *res,
                               hipStream_t st)
{
    if (nblk < 1 || ns < 1 || ns > 32 || stride < ns) return hipErrorInvalidValue;
    const long G0 = nblk < 1024 ? nblk : 1024;   // at most 1024 chunks of R rows
    const long R = (nblk + G0 - 1) / G0, G = (nblk + R - 1) / R;
    rows_partial<<<(unsigned)G, 256, 0, st>>>(part, nblk, stride, ns, R, minmask);
    rows_final<<<ns, 256, 0, st>>>(part, stride, R, G, minmask, res);
    return hipGetLastError();
}

hipError_t launch_em_fields(const double *y, const ewpt_lattice &L, const Phys &P, double theta_w, unsigned mask,
                            double *out, hipStream_t st)
{
    const Geo g = make_geo(L);
    const long n3 = (long)L.nx * L.ny * L.nz;
    em_fields_kernel<<<grid_of(L), dim3(BX, BY, 1), 0, st>>>(y, n3, g, P, sin(theta_w), cos(theta_w), mask, out);
    return hipGetLastError();
}

hipError_t launch_y_slice(const double *y, const ewpt_lattice &L, const Phys &P, double theta_w, int j0, double *out,
                          hipStream_t st)
{
    const Geo g = make_geo(L);
    const long n3 = (long)L.nx * L.ny * L.nz, np = (long)L.nx * L.nz;
    y_slice_kernel<<<(unsigned)((np + 255) / 256), 256, 0, st>>>(y, n3, g, P, sin(theta_w), cos(theta_w), j0,
                                                                 L.dx * L.dx * L.dx, out);
    return hipGetLastError();
}

hipError_t launch_em_energies(const double *y, const ewpt_lattice &L, const Phys &P, double theta_w, double *part,
                              double *res, hipStream_t st)
{
    const Geo g = make_geo(L);
    const long n3 = (long)L.nx * L.ny * L.nz;
    if (energy_brick_fits(L)) {
        const long nb = n3 / EBT;   // <= reduce_blocks(L), the partials' capacity
        em_brick<<<dim3((unsigned)nb), dim3(EBT), 0, st>>>(y, n3, g, P, sin(theta_w), cos(theta_w), part);
        return launch_fixed_reduce(part, nb, NRED, 4, 0u, res, st);
    }
    em_density<<<grid_of(L), dim3(BX, BY, 1), 0, st>>>(y, n3, g, P, sin(theta_w), cos(theta_w), part);
    return launch_fixed_reduce(part, reduce_blocks(L), NRED, 4, 0u, res, st);
}

// ---- the allocation-free energy table (ewpt_energies_device)
//
// Partial sums live in a file-scope device array instead of memory the
// library allocates.  Its host-side name is not a device address: the
// address is taken once per device with hipGetSymbolAddress (that device
// current) and only that address is handed to kernels.  Every call on a
// device shares the one array; calls that overlap on one device race.  Each
// block writes all of its slots, so nothing needs clearing between calls.
constexpr int NOALLOC_BLOCKS = 2048;                 // blocks of energy_brick_strided, one partial row each
constexpr int NOALLOC_DOUBLES = NOALLOC_BLOCKS * NRED + NRED;   // partials, then the reduced slots
static __device__ __align__(256) double scratch[NOALLOC_DOUBLES];

namespace {

constexpr int MAX_DEVICES = 64;
std::mutex scratch_mu;
double *scratch_ptr[MAX_DEVICES] = {};

}  // namespace

hipError_t scratch_address(double **out)
{
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return e;
    if (dev < 0 || dev >= MAX_DEVICES) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> g(scratch_mu);
    if (!scratch_ptr[dev]) {
        void *p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return e;
        scratch_ptr[dev] = static_cast<double *>(p);
    }
    *out = scratch_ptr[dev];
    return hipSuccess;
}

bool energies_noalloc_fits(const ewpt_lattice &L) { return energy_brick_fits(L); }

hipError_t launch_energies_noalloc(const double *y, const ewpt_lattice &L, const Phys &P, double *out,
                                   hipStream_t st)
{
    double *scr = nullptr;
    hipError_t e = scratch_address(&scr);
    if (e != hipSuccess) return e;
    const Geo g = make_geo(L);
    const long n3 = (long)L.nx * L.ny * L.nz, nb = n3 / EBT;
    const int grid = (int)std::min<long>(nb, NOALLOC_BLOCKS);
    double *part = scr, *res = scr + (size_t)NOALLOC_BLOCKS * NRED;
    energy_brick_strided<<<grid, EBT, 0, st>>>(y, n3, g, P, nb, part);
    if ((e = launch_fixed_reduce(part, grid, NRED, 10, 1u << 8, res, st)) != hipSuccess) return e;
    energy_table<<<1, 64, 0, st>>>(res, L.dx * L.dx * L.dx, (double)n3, out);
    return hipGetLastError();
}

}  // namespace ewpt
