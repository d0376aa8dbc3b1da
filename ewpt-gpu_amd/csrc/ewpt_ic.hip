// ewpt_ic.hip -- initial conditions on the device: bubble nucleation and the
// thermal (Bose-Einstein) state.
//
// Bubbles (main_rk4.jl:4930-4946, driver 5640-5698): each bubble adds
// rmag(r) * (p1,p2,p3,p4) to phi.  The host picks centres and orientations
// (as the reference driver does); this file evaluates the profiles.
//   profile 0 (reference): rmag = (1+a^2) e^{-mH r/sqrt2} / (1 + a^2 e^{-sqrt2 mH r}),
//     a = sqrt2 - 1, with the reference's distance
//     r^2 = sum_d sin^2(rk_d ((i-1) - (ib-0.5)) dx) / rk_d^2, rk_d = pi/((N_d-1) dx)
//     (i, ib 1-based as in the reference);
//   profile 1 (tanh wall): rmag = (1 - tanh((r - R) mH/2)) / 2, R = 3/mH, with
//     the minimum-image distance to the centre (ib, jb, kb) given 0-based.
//
// Thermal state (thermal_initializer_gauge!/_phi!, main_rk4.jl:5029-5152,
// driver 5349-5638): per k-mode, Gaussian draws N(0, sqrt2) from a
// counter-based generator (stream s, mode index), transverse polarisations
// eps1 = r x k/|r x k|, eps2 = eps1 x k/|eps1 x k|, occupation
// n = 1/(exp(omega/T) - 1), amplitude sqrt(n/(omega vol)); gauge fields get
// A(k) = amp (eps1 x1 + eps2 x2) + i amp (eps1 x3 + eps2 x4) with omega = |k|,
// phi gets amp * g with omega = sqrt(|k|^2 + meff^2).  The field is the real
// part of the unnormalised inverse FFT.  profile 0 keeps the reference's mode
// index (array index m carries the frequency of index m+1); profile 1 uses
// the signed index.  Profile 1 also sets Gamma^a = div W^a, Sigma = div Y so
// the auxiliary constraints hold; with pi = E = 0 the Gauss law holds too.
#include <hip/hip_runtime.h>
#include <hipfft/hipfft.h>

#include <algorithm>
#include <cmath>
#include <utility>
#include <vector>

#include "ewpt_internal.h"

namespace ewpt {

namespace {

__global__ void __launch_bounds__(256) add_bubbles_kernel(double *y, long n3, int nx, int ny, int nz, double dx,
                                                        const double *bub, int nb, int profile, double mH)
{
    for (long q = blockIdx.x * 256L + threadIdx.x; q < n3; q += (long)gridDim.x * 256L) {
        const int i = (int)(q % nx), j = (int)((q / nx) % ny), k = (int)(q / ((long)nx * ny));
        double add[4] = {0.0, 0.0, 0.0, 0.0};
        for (int b = 0; b < nb; b++) {
            const double *B = bub + 7 * b;
            double rmag;
            if (profile == 0) {
                const double rkx = M_PI / ((nx - 1) * dx), rky = M_PI / ((ny - 1) * dx), rkz = M_PI / ((nz - 1) * dx);
                const double sx = sin(rkx * (i - (B[0] - 0.5)) * dx), sy = sin(rky * (j - (B[1] - 0.5)) * dx),
                             sz = sin(rkz * (k - (B[2] - 0.5)) * dx);
                const double rb = sqrt(sx * sx / (rkx * rkx) + sy * sy / (rky * rky) + sz * sz / (rkz * rkz));
                const double a2 = (M_SQRT2 - 1.0) * (M_SQRT2 - 1.0);
                rmag = (1.0 + a2) * exp(-mH * rb / M_SQRT2) / (1.0 + a2 * exp(-M_SQRT2 * mH * rb));
            } else {
                double d[3] = {i - B[0], j - B[1], k - B[2]};
                const int nn[3] = {nx, ny, nz};
                for (int a = 0; a < 3; a++) {
                    d[a] -= nn[a] * rint(d[a] / nn[a]);
                    d[a] *= dx;
                }
                const double r = sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
                rmag = 0.5 * (1.0 - tanh((r - 3.0 / mH) * mH * 0.5));
            }
            for (int c = 0; c < 4; c++) add[c] += rmag * B[3 + c];
        }
        for (int c = 0; c < 4; c++) y[(long)c * n3 + q] += add[c];
    }
}

// counter-based N(0, sqrt2) draw (Box-Muller on two splitmix64 outputs)
__host__ __device__ inline unsigned long long mix64(unsigned long long z)
{
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
    return z ^ (z >> 31);
}
__device__ inline double gauss_sqrt2(unsigned long long seed, int stream, long idx)
{
    const unsigned long long base = mix64(seed + 0x9e3779b97f4a7c15ULL * (unsigned long long)(stream + 1));
    const unsigned long long a = mix64(base + 0x9e3779b97f4a7c15ULL * (2ULL * (unsigned long long)idx + 1ULL));
    const unsigned long long b = mix64(base + 0x9e3779b97f4a7c15ULL * (2ULL * (unsigned long long)idx + 2ULL));
    const double u1 = ((double)(a >> 11) + 1.0) * 0x1.0p-53;   // (0, 1]
    const double u2 = (double)(b >> 11) * 0x1.0p-53;           // [0, 1)
    return M_SQRT2 * sqrt(-2.0 * log(u1)) * cos(2.0 * M_PI * u2);
}

struct ModeGeo {
    int nx, ny, nz;
    double dx, T, meff_sq, vol;
    int profile;
    unsigned long long seed;
};

__device__ inline void mode_k(const ModeGeo &g, long q, double &kx, double &ky, double &kz, double &modk)
{
    const int i = (int)(q % g.nx), j = (int)((q / g.nx) % g.ny), k = (int)(q / ((long)g.nx * g.ny));
    int mi, mj, mk;
    if (g.profile == 0) { mi = i + 1; mj = j + 1; mk = k + 1; }
    else {
        mi = i <= g.nx / 2 ? i : i - g.nx;
        mj = j <= g.ny / 2 ? j : j - g.ny;
        mk = k <= g.nz / 2 ? k : k - g.nz;
    }
    kx = 2.0 * sin(0.5 * mi * 2.0 * M_PI / g.nx) / g.dx;
    ky = 2.0 * sin(0.5 * mj * 2.0 * M_PI / g.ny) / g.dx;
    kz = 2.0 * sin(0.5 * mk * 2.0 * M_PI / g.nz) / g.dx;
    modk = sqrt(2.0 * ((1.0 - cos(mi * 2.0 * M_PI / g.nx)) / (g.dx * g.dx) +
                       (1.0 - cos(mj * 2.0 * M_PI / g.ny)) / (g.dx * g.dx) +
                       (1.0 - cos(mk * 2.0 * M_PI / g.nz)) / (g.dx * g.dx)));
}

// gauge group grp (0..2 = W^1..3, 3 = Y): three complex component arrays.
// Draws come from streams s0 .. s0 + 9; mom = 1 multiplies the amplitude by
// omega (the conjugate momentum E of a mode: <|E_k|^2> = n omega / vol).
__global__ void __launch_bounds__(256) gauge_modes(hipfftDoubleComplex *a0, hipfftDoubleComplex *a1,
                                                   hipfftDoubleComplex *a2, long n3, ModeGeo g, int s0, int mom)
{
    for (long q = blockIdx.x * 256L + threadIdx.x; q < n3; q += (long)gridDim.x * 256L) {
        double kx, ky, kz, modk;
        mode_k(g, q, kx, ky, kz, modk);
        const double r1 = gauss_sqrt2(g.seed, s0 + 0, q), r2 = gauss_sqrt2(g.seed, s0 + 1, q),
                     r3 = gauss_sqrt2(g.seed, s0 + 2, q);
        const double x1 = gauss_sqrt2(g.seed, s0 + 6, q), x2 = gauss_sqrt2(g.seed, s0 + 7, q),
                     x3 = gauss_sqrt2(g.seed, s0 + 8, q), x4 = gauss_sqrt2(g.seed, s0 + 9, q);
        double e1x = r2 * kz - r3 * ky, e1y = r3 * kx - r1 * kz, e1z = r1 * ky - r2 * kx;
        const double l1 = sqrt(e1x * e1x + e1y * e1y + e1z * e1z);
        e1x /= l1; e1y /= l1; e1z /= l1;
        double e2x = e1y * kz - e1z * ky, e2y = e1z * kx - e1x * kz, e2z = e1x * ky - e1y * kx;
        const double l2 = sqrt(e2x * e2x + e2y * e2y + e2z * e2z);
        e2x /= l2; e2y /= l2; e2z /= l2;
        const double w = modk;
        hipfftDoubleComplex v0 = {0.0, 0.0}, v1 = {0.0, 0.0}, v2 = {0.0, 0.0};
        if (w != 0.0) {
            double amp = sqrt((1.0 / (exp(w / g.T) - 1.0)) / (w * g.vol));
            if (mom) amp *= w;
            v0 = {amp * (e1x * x1 + e2x * x2), amp * (e1x * x3 + e2x * x4)};
            v1 = {amp * (e1y * x1 + e2y * x2), amp * (e1y * x3 + e2y * x4)};
            v2 = {amp * (e1z * x1 + e2z * x2), amp * (e1z * x3 + e2z * x4)};
        }
        a0[q] = v0; a1[q] = v1; a2[q] = v2;
    }
}

// Higgs component: real amplitude times one draw of the given stream (the
// field: 40 + b; mom = 1: the momentum pi, amplitude times omega)
__global__ void __launch_bounds__(256) higgs_modes(hipfftDoubleComplex *a, long n3, ModeGeo g, int stream, int mom)
{
    for (long q = blockIdx.x * 256L + threadIdx.x; q < n3; q += (long)gridDim.x * 256L) {
        double kx, ky, kz, modk;
        mode_k(g, q, kx, ky, kz, modk);
        const double w = sqrt(modk * modk + g.meff_sq);
        // a massless zero mode has no finite occupation; it is left empty
        double amp = w > 0.0 ? sqrt((1.0 / (exp(w / g.T) - 1.0)) / (w * g.vol)) : 0.0;
        if (mom) amp *= w;
        a[q] = {amp * gauss_sqrt2(g.seed, stream, q), 0.0};
    }
}

__global__ void __launch_bounds__(256) take_real(const hipfftDoubleComplex *a, double *dst, long n3)
{
    for (long q = blockIdx.x * 256L + threadIdx.x; q < n3; q += (long)gridDim.x * 256L) dst[q] = a[q].x;
}

// Gamma^a = d_i W^a_i, Sigma = d_i Y_i with the 6th-order first derivative
__global__ void __launch_bounds__(256) aux_from_div(double *y, long n3, int nx, int ny, int nz, double inv60dx)
{
    for (long q = blockIdx.x * 256L + threadIdx.x; q < n3; q += (long)gridDim.x * 256L) {
        const int i = (int)(q % nx), j = (int)((q / nx) % ny), k = (int)(q / ((long)nx * ny));
        const long nxy = (long)nx * ny;
        for (int grp = 0; grp < 4; grp++) {
            double div = 0.0;
            for (int ax = 0; ax < 3; ax++) {
                const int c = grp < 3 ? CH_W + 3 * grp + ax : CH_Y + ax;
                const double *A = y + (long)c * n3;
                double v[7];
                for (int t = 0; t < 7; t++) {
                    long o;
                    if (ax == 0) o = wrap_idx(i + t - 3, nx) + (long)nx * j + nxy * k;
                    else if (ax == 1) o = i + (long)nx * wrap_idx(j + t - 3, ny) + nxy * k;
                    else o = i + (long)nx * j + nxy * wrap_idx(k + t - 3, nz);
                    v[t] = A[o];
                }
                div += sten_d1(v[0], v[1], v[2], v[4], v[5], v[6], inv60dx);
            }
            y[(long)(grp < 3 ? CH_GAM + grp : CH_SIG) * n3 + q] = div;
        }
    }
}

unsigned grid_for(long n)
{
    const long b = (n + 255) / 256;
    return (unsigned)(b < 65536 ? b : 65536);
}

}  // namespace

hipError_t launch_add_bubbles(double *y, const ewpt_lattice &L, const double *dbub, int nb, int profile, double mH,
                              hipStream_t st)
{
    const long n3 = (long)L.nx * L.ny * L.nz;
    add_bubbles_kernel<<<grid_for(n3), 256, 0, st>>>(y, n3, L.nx, L.ny, L.nz, L.dx, dbub, nb, profile, mH);
    return hipGetLastError();
}

// Fills channels 0..15 of y (and, profile 1, 16..19); work: 3 complex N^3 arrays.
int thermal_state(double *y, const ewpt_lattice &L, double T, double meff_sq, unsigned long long seed, int profile,
                  void *workp, hipStream_t st, const char **what)
{
    hipfftDoubleComplex *work = static_cast<hipfftDoubleComplex *>(workp);
    const long n3 = (long)L.nx * L.ny * L.nz;
    ModeGeo g{L.nx, L.ny, L.nz, L.dx, T, meff_sq, n3 * L.dx * L.dx * L.dx, profile, seed};
    hipfftHandle plan;
    // data are x fastest: the transform's slowest dimension is z
    if (hipfftPlan3d(&plan, L.nz, L.ny, L.nx, HIPFFT_Z2Z) != HIPFFT_SUCCESS) { *what = "hipfftPlan3d"; return -1; }
    int rc = 0;
    if (hipfftSetStream(plan, st) != HIPFFT_SUCCESS) { *what = "hipfftSetStream"; rc = -1; }
    hipfftDoubleComplex *a[3] = {work, work + n3, work + 2 * n3};
    for (int grp = 0; grp < 4 && rc == 0; grp++) {
        gauge_modes<<<grid_for(n3), 256, 0, st>>>(a[0], a[1], a[2], n3, g, 10 * grp, 0);
        for (int c = 0; c < 3 && rc == 0; c++) {
            if (hipfftExecZ2Z(plan, a[c], a[c], HIPFFT_BACKWARD) != HIPFFT_SUCCESS) { *what = "hipfftExecZ2Z"; rc = -1; }
            const int ch = grp < 3 ? CH_W + 3 * grp + c : CH_Y + c;
            take_real<<<grid_for(n3), 256, 0, st>>>(a[c], y + (long)ch * n3, n3);
        }
    }
    for (int b = 0; b < 4 && rc == 0; b++) {
        higgs_modes<<<grid_for(n3), 256, 0, st>>>(a[0], n3, g, 40 + b, 0);
        if (hipfftExecZ2Z(plan, a[0], a[0], HIPFFT_BACKWARD) != HIPFFT_SUCCESS) { *what = "hipfftExecZ2Z"; rc = -1; }
        take_real<<<grid_for(n3), 256, 0, st>>>(a[0], y + (long)(CH_PHI + b) * n3, n3);
    }
    if (rc == 0 && profile == 1)
        aux_from_div<<<grid_for(n3), 256, 0, st>>>(y, n3, L.nx, L.ny, L.nz, 1.0 / (60.0 * L.dx));
    if (rc == 0 && hipGetLastError() != hipSuccess) { *what = "thermal kernels"; rc = -1; }
    (void)hipStreamSynchronize(st);
    hipfftDestroy(plan);
    return rc;
}

// ---------------------------------------------------------------- Gauss-law solve
//
// Thermal momenta and electric fields that satisfy the Gauss law (the
// initial condition of the benchmark's thermal configuration; the reference
// itself starts from pi = E = 0, main_rk4.jl:5154-5162).  With the
// charges that k_Gamma / k_Sigma weight by gp2 (main_rk4.jl:603-657) the law
// reads, with the 6th-order first derivative d,
//   (G E)^a = d_i E^a_i + gw eps^{abc} W^b_i E^c_i = gw Q^a(phi, pi),
//   d_i E^Y_i = gy Q^Y(phi, pi).
// Steps: thermal E (transverse modes) and pi with amplitudes sqrt(n omega /
// vol); pi corrected on the modes d annihilates (wave numbers 0 or pi on
// every axis) so the U(1) charge has no component there; E^Y += d chi_Y by
// an exact FFT solve; E^a += (D chi)^a, (D chi)^a_i = d_i chi^a +
// gw eps^{abc} W^b_i chi^c, with D^T D chi = G E0 - gw Q (G = -D^T) solved by
// conjugate gradients preconditioned with (sum_i K_i^2 + mu)^-1 by FFT.

namespace {

struct GGeo {
    int nx, ny, nz;
    double inv60dx, gw, gy;
};

__device__ inline void site_of(const GGeo &g, long q, int &i, int &j, int &k)
{
    i = (int)(q % g.nx);
    j = (int)((q / g.nx) % g.ny);
    k = (int)(q / ((long)g.nx * g.ny));
}

// d_ax f at site (i, j, k) of the periodic lattice; f is read with stride st
// (1 for a real array, 2 for the real part of a complex one)
__device__ inline double gd1(const double *f, const GGeo &g, int i, int j, int k, int ax, int st)
{
    double v[7];
    const long nx = g.nx, nxy = (long)g.nx * g.ny;
    for (int t = 0; t < 7; t++) {
        long o;
        if (ax == 0) o = wrap_idx(i + t - 3, g.nx) + nx * j + nxy * k;
        else if (ax == 1) o = i + nx * wrap_idx(j + t - 3, g.ny) + nxy * k;
        else o = i + nx * j + nxy * wrap_idx(k + t - 3, g.nz);
        v[t] = f[o * st];
    }
    return sten_d1(v[0], v[1], v[2], v[4], v[5], v[6], g.inv60dx);
}

// Higgs charges of k_Gamma_1..3 and k_Sigma (main_rk4.jl:612-616, 631-635,
// 650-654, 660-664)
__device__ inline void higgs_charges(const double p[4], const double pi[4], double Q[3], double &QY)
{
    Q[0] = p[0] * pi[3] - p[3] * pi[0] + p[2] * pi[1] - p[1] * pi[2];
    Q[1] = p[2] * pi[0] - p[0] * pi[2] + p[3] * pi[1] - p[1] * pi[3];
    Q[2] = p[0] * pi[1] - p[1] * pi[0] + p[3] * pi[2] - p[2] * pi[3];
    QY = p[0] * pi[1] - p[1] * pi[0] + p[2] * pi[3] - p[3] * pi[2];
}

__device__ inline void load_phi_pi(const double *y, long n3, long q, double p[4], double pi[4])
{
    for (int b = 0; b < 4; b++) {
        p[b] = y[(long)(CH_PHI + b) * n3 + q];
        pi[b] = y[(long)(CH_PI + b) * n3 + q];
    }
}

// (-1)^(m . x) for the null-mode index m = mx + 2 my + 4 mz
__device__ inline double null_sign(int m, int i, int j, int k)
{
    return (((m & 1) * i + ((m >> 1) & 1) * j + ((m >> 2) & 1) * k) & 1) ? -1.0 : 1.0;
}

constexpr int GB = 256;       // threads per block of the solve's kernels
constexpr int GNB = 1024;     // blocks of its reductions (one partial each, or 16 for null_sums)

__device__ inline double block_sum(double v, double *sh)
{
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    __syncthreads();
    if (lane == 0) sh[w] = v;
    __syncthreads();
    double r = 0.0;
    if (threadIdx.x == 0)
        for (int u = 0; u < GB / 64; u++) r += sh[u];
    return r;
}

// per block: S_m = sum |phi|^2 e_m, P_m = sum gy Q^Y e_m (m = 0..7) -> part[blk * 16 + ...]
__global__ void __launch_bounds__(GB) null_sums(const double *y, long n3, GGeo g, double *part)
{
    __shared__ double sh[GB / 64];
    double acc[16];
    for (int m = 0; m < 16; m++) acc[m] = 0.0;
    for (long q = blockIdx.x * (long)GB + threadIdx.x; q < n3; q += (long)gridDim.x * GB) {
        int i, j, k;
        site_of(g, q, i, j, k);
        double p[4], pi[4], Q[3], QY;
        load_phi_pi(y, n3, q, p, pi);
        higgs_charges(p, pi, Q, QY);
        const double r2 = p[0] * p[0] + p[1] * p[1] + p[2] * p[2] + p[3] * p[3];
        for (int m = 0; m < 8; m++) {
            const double e = null_sign(m, i, j, k);
            acc[m] += r2 * e;
            acc[8 + m] += g.gy * QY * e;
        }
    }
    for (int m = 0; m < 16; m++) {
        const double r = block_sum(acc[m], sh);
        if (threadIdx.x == 0) part[blockIdx.x * 16L + m] = r;
    }
}

struct Alpha {
    double a[8];
};

// pi += c(x) J^T phi, c = sum_m alpha_m e_m, J^T phi = (-phi_2, phi_1, -phi_4, phi_3)
__global__ void __launch_bounds__(GB) pi_correct(double *y, long n3, GGeo g, Alpha al)
{
    for (long q = blockIdx.x * (long)GB + threadIdx.x; q < n3; q += (long)gridDim.x * GB) {
        int i, j, k;
        site_of(g, q, i, j, k);
        double c = 0.0;
        for (int m = 0; m < 8; m++) c += al.a[m] * null_sign(m, i, j, k);
        const double p0 = y[(CH_PHI + 0) * n3 + q], p1 = y[(CH_PHI + 1) * n3 + q];
        const double p2 = y[(CH_PHI + 2) * n3 + q], p3 = y[(CH_PHI + 3) * n3 + q];
        y[(CH_PI + 0) * n3 + q] += c * -p1;
        y[(CH_PI + 1) * n3 + q] += c * p0;
        y[(CH_PI + 2) * n3 + q] += c * -p3;
        y[(CH_PI + 3) * n3 + q] += c * p2;
    }
}

// U(1) residual d_i E^Y_i - gy Q^Y (sign = +1) or its negative, into the real
// part of a complex array (imaginary part 0)
__global__ void __launch_bounds__(GB) u1_residual(const double *y, long n3, GGeo g, double sign,
                                                  hipfftDoubleComplex *out)
{
    for (long q = blockIdx.x * (long)GB + threadIdx.x; q < n3; q += (long)gridDim.x * GB) {
        int i, j, k;
        site_of(g, q, i, j, k);
        double div = 0.0;
        for (int ax = 0; ax < 3; ax++) div += gd1(y + (long)(CH_EY + ax) * n3, g, i, j, k, ax, 1);
        double p[4], pi[4], Q[3], QY;
        load_phi_pi(y, n3, q, p, pi);
        higgs_charges(p, pi, Q, QY);
        out[q] = {sign * (div - g.gy * QY), 0.0};
    }
}

// sum_i K(theta_i)^2 of mode q, K(t) = (90 sin t - 18 sin 2t + 2 sin 3t)/(60 dx)
__device__ inline double ksym(const GGeo &g, long q)
{
    int i, j, k;
    site_of(g, q, i, j, k);
    const int m[3] = {i, j, k}, n[3] = {g.nx, g.ny, g.nz};
    double s = 0.0;
    for (int d = 0; d < 3; d++) {
        const double t = 2.0 * M_PI * m[d] / n[d];
        const double K = (90.0 * sin(t) - 18.0 * sin(2.0 * t) + 2.0 * sin(3.0 * t)) * g.inv60dx;
        s += K * K;
    }
    return s;
}

// F -> F / (sum K^2 + mu) / N^3 (mu > 0), or for mu = 0 (the exact U(1) solve)
// -F / sum K^2 / N^3 with the modes below thresh set to zero
__global__ void __launch_bounds__(GB) spectral_scale(hipfftDoubleComplex *F, long n3, GGeo g, double mu,
                                                     double thresh)
{
    for (long q = blockIdx.x * (long)GB + threadIdx.x; q < n3; q += (long)gridDim.x * GB) {
        const double sym = ksym(g, q);
        double f;
        if (mu > 0.0) f = 1.0 / ((sym + mu) * (double)n3);
        else f = sym > thresh ? -1.0 / (sym * (double)n3) : 0.0;
        F[q].x *= f;
        F[q].y *= f;
    }
}

// E^Y_i += d_i chi (chi in the real part of a complex array)
__global__ void __launch_bounds__(GB) u1_apply(double *y, long n3, GGeo g, const hipfftDoubleComplex *chi)
{
    const double *c = reinterpret_cast<const double *>(chi);
    for (long q = blockIdx.x * (long)GB + threadIdx.x; q < n3; q += (long)gridDim.x * GB) {
        int i, j, k;
        site_of(g, q, i, j, k);
        for (int ax = 0; ax < 3; ax++) y[(long)(CH_EY + ax) * n3 + q] += gd1(c, g, i, j, k, ax, 2);
    }
}

__device__ inline double Wv(const double *y, long n3, long q, int b, int i) { return y[(long)(CH_W + 3 * b + i) * n3 + q]; }

// out^a = (G E)^a - gw Q^a with E the state's electric fields
__global__ void __launch_bounds__(GB) su2_residual(const double *y, long n3, GGeo g, double *out)
{
    for (long q = blockIdx.x * (long)GB + threadIdx.x; q < n3; q += (long)gridDim.x * GB) {
        int i, j, k;
        site_of(g, q, i, j, k);
        double p[4], pi[4], Q[3], QY;
        load_phi_pi(y, n3, q, p, pi);
        higgs_charges(p, pi, Q, QY);
        for (int a = 0; a < 3; a++) {
            const int b = (a + 1) % 3, c = (a + 2) % 3;
            double v = 0.0;
            for (int ax = 0; ax < 3; ax++) v += gd1(y + (long)(CH_E + 3 * a + ax) * n3, g, i, j, k, ax, 1);
            double nl = 0.0;
            for (int ax = 0; ax < 3; ax++)
                nl += Wv(y, n3, q, b, ax) * y[(long)(CH_E + 3 * c + ax) * n3 + q] -
                      Wv(y, n3, q, c, ax) * y[(long)(CH_E + 3 * b + ax) * n3 + q];
            out[(long)a * n3 + q] = v + g.gw * nl - g.gw * Q[a];
        }
    }
}

// Dv^a_i = d_i v^a + gw eps^{abc} W^b_i v^c  (v: 3 arrays, Dv: 9 arrays a-major)
__global__ void __launch_bounds__(GB) cov_grad(const double *v, const double *y, long n3, GGeo g, double *Dv)
{
    for (long q = blockIdx.x * (long)GB + threadIdx.x; q < n3; q += (long)gridDim.x * GB) {
        int i, j, k;
        site_of(g, q, i, j, k);
        for (int a = 0; a < 3; a++) {
            const int b = (a + 1) % 3, c = (a + 2) % 3;
            for (int ax = 0; ax < 3; ax++)
                Dv[(long)(3 * a + ax) * n3 + q] =
                    gd1(v + (long)a * n3, g, i, j, k, ax, 1) +
                    g.gw * (Wv(y, n3, q, b, ax) * v[(long)c * n3 + q] - Wv(y, n3, q, c, ax) * v[(long)b * n3 + q]);
        }
    }
}

// out^a = -(G F)^a = -(d_i F^a_i + gw eps^{abc} W^b_i F^c_i)  (F: 9 arrays)
__global__ void __launch_bounds__(GB) neg_gauss(const double *F, const double *y, long n3, GGeo g, double *out)
{
    for (long q = blockIdx.x * (long)GB + threadIdx.x; q < n3; q += (long)gridDim.x * GB) {
        int i, j, k;
        site_of(g, q, i, j, k);
        for (int a = 0; a < 3; a++) {
            const int b = (a + 1) % 3, c = (a + 2) % 3;
            double v = 0.0;
            for (int ax = 0; ax < 3; ax++) v += gd1(F + (long)(3 * a + ax) * n3, g, i, j, k, ax, 1);
            double nl = 0.0;
            for (int ax = 0; ax < 3; ax++)
                nl += Wv(y, n3, q, b, ax) * F[(long)(3 * c + ax) * n3 + q] - Wv(y, n3, q, c, ax) * F[(long)(3 * b + ax) * n3 + q];
            out[(long)a * n3 + q] = -(v + g.gw * nl);
        }
    }
}

// per block: sum_q a[q] b[q] over n values
__global__ void __launch_bounds__(GB) dot_part(const double *a, const double *b, long n, double *part)
{
    __shared__ double sh[GB / 64];
    double acc = 0.0;
    for (long q = blockIdx.x * (long)GB + threadIdx.x; q < n; q += (long)gridDim.x * GB) acc += a[q] * b[q];
    const double r = block_sum(acc, sh);
    if (threadIdx.x == 0) part[blockIdx.x] = r;
}

__global__ void __launch_bounds__(GB) cg_xr(double *x, double *r, const double *p, const double *qv, double alpha, long n)
{
    for (long t = blockIdx.x * (long)GB + threadIdx.x; t < n; t += (long)gridDim.x * GB) {
        x[t] += alpha * p[t];
        r[t] -= alpha * qv[t];
    }
}

__global__ void __launch_bounds__(GB) cg_p(double *p, const double *z, double beta, long n)
{
    for (long t = blockIdx.x * (long)GB + threadIdx.x; t < n; t += (long)gridDim.x * GB) p[t] = z[t] + beta * p[t];
}

__global__ void __launch_bounds__(GB) to_complex(const double *r, hipfftDoubleComplex *c, long n)
{
    for (long t = blockIdx.x * (long)GB + threadIdx.x; t < n; t += (long)gridDim.x * GB) c[t] = {r[t], 0.0};
}

// E^a_i += F^a_i (state channels 24..32)
__global__ void __launch_bounds__(GB) add_e(double *y, long n3, const double *F)
{
    for (long t = blockIdx.x * (long)GB + threadIdx.x; t < 9 * n3; t += (long)gridDim.x * GB) y[(long)CH_E * n3 + t] += F[t];
}

unsigned gblocks(long n)
{
    const long b = (n + GB - 1) / GB;
    return (unsigned)(b < GNB ? b : GNB);
}

// deterministic dot: device partials of a fixed grid, summed on the host in order
struct Dotter {
    double *part = nullptr;
    std::vector<double> host;
    hipStream_t st;
    hipError_t dot(const double *a, const double *b, long n, double &out)
    {
        const unsigned nb = gblocks(n);
        dot_part<<<nb, GB, 0, st>>>(a, b, n, part);
        hipError_t e = hipMemcpyAsync(host.data(), part, nb * sizeof(double), hipMemcpyDeviceToHost, st);
        if (e == hipSuccess) e = hipStreamSynchronize(st);
        out = 0.0;
        for (unsigned u = 0; u < nb; u++) out += host[u];
        return e;
    }
};

// Gaussian elimination with partial pivoting on an n x n system (n <= 8)
bool solve_small(int n, double A[8][8], double b[8], double x[8])
{
    for (int c = 0; c < n; c++) {
        int piv = c;
        for (int r = c + 1; r < n; r++)
            if (fabs(A[r][c]) > fabs(A[piv][c])) piv = r;
        if (!(fabs(A[piv][c]) > 0.0)) return false;
        for (int t = 0; t < n; t++) std::swap(A[c][t], A[piv][t]);
        std::swap(b[c], b[piv]);
        for (int r = c + 1; r < n; r++) {
            const double f = A[r][c] / A[c][c];
            for (int t = c; t < n; t++) A[r][t] -= f * A[c][t];
            b[r] -= f * b[c];
        }
    }
    for (int r = n - 1; r >= 0; r--) {
        double v = b[r];
        for (int t = r + 1; t < n; t++) v -= A[r][t] * x[t];
        x[r] = v / A[r][r];
    }
    return true;
}

}  // namespace

int thermal_gauss_state(double *y, const ewpt_lattice &L, double T, double meff_sq, unsigned long long seed,
                        double gw, double gy, double tol, int maxit, void *workp, double *wa, double *wb,
                        hipStream_t st, double stats[4], const char **what)
{
    // fields as the signed-index thermal profile, Gamma = div W, Sigma = div Y
    if (thermal_state(y, L, T, meff_sq, seed, 1, workp, st, what) != 0) return -1;
    hipfftDoubleComplex *work = static_cast<hipfftDoubleComplex *>(workp);
    const long n3 = (long)L.nx * L.ny * L.nz;
    ModeGeo mg{L.nx, L.ny, L.nz, L.dx, T, meff_sq, n3 * L.dx * L.dx * L.dx, 1, seed};
    const GGeo g{L.nx, L.ny, L.nz, 1.0 / (60.0 * L.dx), gw, gy};
    hipfftHandle plan;
    if (hipfftPlan3d(&plan, L.nz, L.ny, L.nx, HIPFFT_Z2Z) != HIPFFT_SUCCESS) { *what = "hipfftPlan3d"; return -1; }
    int rc = 0;
    auto fail = [&](const char *w) { if (rc == 0) { *what = w; rc = -1; } };
    auto fft = [&](hipfftDoubleComplex *a, int dir) {
        if (rc == 0 && hipfftExecZ2Z(plan, a, a, dir) != HIPFFT_SUCCESS) fail("hipfftExecZ2Z");
    };
    if (hipfftSetStream(plan, st) != HIPFFT_SUCCESS) fail("hipfftSetStream");
    const unsigned nb = gblocks(n3);
    Dotter dt;
    dt.st = st;
    dt.host.resize(16 * GNB);
    if (hipMalloc(reinterpret_cast<void **>(&dt.part), 16 * GNB * sizeof(double)) != hipSuccess) fail("alloc partials");
    // 1. thermal momenta: E^a, E^Y (transverse modes) and pi
    hipfftDoubleComplex *a[3] = {work, work + n3, work + 2 * n3};
    for (int grp = 0; grp < 4 && rc == 0; grp++) {
        gauge_modes<<<grid_for(n3), 256, 0, st>>>(a[0], a[1], a[2], n3, mg, 100 + 10 * grp, 1);
        for (int c = 0; c < 3 && rc == 0; c++) {
            fft(a[c], HIPFFT_BACKWARD);
            const int ch = grp < 3 ? CH_E + 3 * grp + c : CH_EY + c;
            take_real<<<grid_for(n3), 256, 0, st>>>(a[c], y + (long)ch * n3, n3);
        }
    }
    for (int b = 0; b < 4 && rc == 0; b++) {
        higgs_modes<<<grid_for(n3), 256, 0, st>>>(a[0], n3, mg, 140 + b, 1);
        fft(a[0], HIPFFT_BACKWARD);
        take_real<<<grid_for(n3), 256, 0, st>>>(a[0], y + (long)(CH_PI + b) * n3, n3);
    }
    // 2. pi on the null modes of d: S_m, P_m, then M alpha = -P over the null set
    if (rc == 0) {
        null_sums<<<nb, GB, 0, st>>>(y, n3, g, dt.part);
        if (hipMemcpyAsync(dt.host.data(), dt.part, 16 * nb * sizeof(double), hipMemcpyDeviceToHost, st) != hipSuccess ||
            hipStreamSynchronize(st) != hipSuccess)
            fail("null-mode sums");
    }
    Alpha al{};
    if (rc == 0) {
        double S[8] = {}, P[8] = {};
        for (unsigned u = 0; u < nb; u++)
            for (int m = 0; m < 8; m++) {
                S[m] += dt.host[16 * u + m];
                P[m] += dt.host[16 * u + 8 + m];
            }
        int idx[8], nn = 0;
        for (int m = 0; m < 8; m++)
            if (!((m & 1) && L.nx % 2) && !(((m >> 1) & 1) && L.ny % 2) && !(((m >> 2) & 1) && L.nz % 2)) idx[nn++] = m;
        double M[8][8], rhs[8], x[8];
        for (int r = 0; r < nn; r++) {
            for (int c = 0; c < nn; c++) M[r][c] = gy * S[idx[r] ^ idx[c]];   // e_m e_n = e_(m xor n)
            rhs[r] = -P[idx[r]];
        }
        if (gy != 0.0) {
            if (!solve_small(nn, M, rhs, x)) fail("null-mode system is singular (phi = 0)");
            for (int r = 0; r < nn && rc == 0; r++) al.a[idx[r]] = x[r];
        }
        if (rc == 0) pi_correct<<<nb, GB, 0, st>>>(y, n3, g, al);
    }
    // residual norms before: SU(2) into wa[0..3n3), U(1) into the complex buffer
    double *rv = wa, *xv = wa + 3 * n3, *pv = wa + 6 * n3, *qv = wa + 9 * n3, *zv = wa + 12 * n3;
    auto residual_norm2 = [&](double &su2, double &u1) {
        su2_residual<<<nb, GB, 0, st>>>(y, n3, g, rv);
        u1_residual<<<nb, GB, 0, st>>>(y, n3, g, 1.0, a[0]);
        take_real<<<grid_for(n3), 256, 0, st>>>(a[0], zv, n3);
        if (dt.dot(rv, rv, 3 * n3, su2) != hipSuccess || dt.dot(zv, zv, n3, u1) != hipSuccess) fail("residual norm");
    };
    double su2_0 = 0.0, u1_0 = 0.0;
    if (rc == 0) residual_norm2(su2_0, u1_0);
    // 3. U(1): chi_Y = d.d^-1 (gy Q^Y - d.E^Y), E^Y += d chi_Y
    if (rc == 0) {
        u1_residual<<<nb, GB, 0, st>>>(y, n3, g, -1.0, a[0]);
        fft(a[0], HIPFFT_FORWARD);
        double kmax2 = 0.0;   // largest sum K^2 on the lattice, for the null-mode threshold
        const int n[3] = {L.nx, L.ny, L.nz};
        for (int d = 0; d < 3; d++) {
            double mx = 0.0;
            for (int m = 0; m < n[d]; m++) {
                const double t = 2.0 * M_PI * m / n[d];
                const double K = (90.0 * sin(t) - 18.0 * sin(2.0 * t) + 2.0 * sin(3.0 * t)) * g.inv60dx;
                mx = std::max(mx, K * K);
            }
            kmax2 += mx;
        }
        spectral_scale<<<nb, GB, 0, st>>>(a[0], n3, g, 0.0, 1e-12 * kmax2);
        fft(a[0], HIPFFT_BACKWARD);
        u1_apply<<<nb, GB, 0, st>>>(y, n3, g, a[0]);
    }
    // 4. SU(2): D^T D chi = G E0 - gw Q by preconditioned CG, then E += D chi
    int it = 0;
    if (rc == 0) {
        double w2 = 0.0;
        if (dt.dot(y + (long)CH_W * n3, y + (long)CH_W * n3, 9 * n3, w2) != hipSuccess) fail("W norm");
        const double mu = std::max((2.0 / 3.0) * gw * gw * w2 / (double)n3, 1e-300);
        auto precond = [&](const double *r, double *z) {
            for (int c = 0; c < 3 && rc == 0; c++) {
                to_complex<<<nb, GB, 0, st>>>(r + (long)c * n3, a[0], n3);
                fft(a[0], HIPFFT_FORWARD);
                spectral_scale<<<nb, GB, 0, st>>>(a[0], n3, g, mu, 0.0);
                fft(a[0], HIPFFT_BACKWARD);
                take_real<<<grid_for(n3), 256, 0, st>>>(a[0], z + (long)c * n3, n3);
            }
        };
        su2_residual<<<nb, GB, 0, st>>>(y, n3, g, rv);   // b = G E0 - gw Q, x = 0, r = b
        if (hipMemsetAsync(xv, 0, 3 * n3 * sizeof(double), st) != hipSuccess) fail("clear x");
        precond(rv, zv);
        if (hipMemcpyAsync(pv, zv, 3 * n3 * sizeof(double), hipMemcpyDeviceToDevice, st) != hipSuccess) fail("copy p");
        double rz = 0.0, bb = 0.0, rr = 0.0;
        if (rc == 0 && (dt.dot(rv, zv, 3 * n3, rz) != hipSuccess || dt.dot(rv, rv, 3 * n3, bb) != hipSuccess))
            fail("cg dot");
        rr = bb;
        while (rc == 0 && it < maxit && sqrt(rr) > tol * sqrt(bb)) {
            cov_grad<<<nb, GB, 0, st>>>(pv, y, n3, g, wb);
            neg_gauss<<<nb, GB, 0, st>>>(wb, y, n3, g, qv);
            double pq = 0.0;
            if (dt.dot(pv, qv, 3 * n3, pq) != hipSuccess || !(pq > 0.0)) { fail("cg breakdown"); break; }
            const double alpha = rz / pq;
            cg_xr<<<nb, GB, 0, st>>>(xv, rv, pv, qv, alpha, 3 * n3);
            precond(rv, zv);
            double rz_new = 0.0;
            if (dt.dot(rv, zv, 3 * n3, rz_new) != hipSuccess || dt.dot(rv, rv, 3 * n3, rr) != hipSuccess)
                fail("cg dot");
            cg_p<<<nb, GB, 0, st>>>(pv, zv, rz_new / rz, 3 * n3);
            rz = rz_new;
            it++;
        }
        if (rc == 0) {
            cov_grad<<<nb, GB, 0, st>>>(xv, y, n3, g, wb);
            add_e<<<nb, GB, 0, st>>>(y, n3, wb);
        }
    }
    double su2_1 = 0.0, u1_1 = 0.0;
    if (rc == 0) residual_norm2(su2_1, u1_1);
    if (rc == 0 && hipGetLastError() != hipSuccess) fail("Gauss-law solve kernels");
    (void)hipStreamSynchronize(st);
    if (dt.part) (void)hipFree(dt.part);
    hipfftDestroy(plan);
    const double vol = L.dx * L.dx * L.dx;
    stats[0] = sqrt((su2_0 + u1_0) * vol);
    stats[1] = sqrt((su2_1 + u1_1) * vol);
    stats[2] = it;
    stats[3] = sqrt(su2_1 * vol);
    return rc;
}

}  // namespace ewpt
