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

#include <cmath>

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

// gauge group grp (0..2 = W^1..3, 3 = Y): three complex component arrays
__global__ void __launch_bounds__(256) gauge_modes(hipfftDoubleComplex *a0, hipfftDoubleComplex *a1,
                                                   hipfftDoubleComplex *a2, long n3, ModeGeo g, int grp)
{
    for (long q = blockIdx.x * 256L + threadIdx.x; q < n3; q += (long)gridDim.x * 256L) {
        double kx, ky, kz, modk;
        mode_k(g, q, kx, ky, kz, modk);
        const int s0 = 10 * grp;
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
            const double amp = sqrt((1.0 / (exp(w / g.T) - 1.0)) / (w * g.vol));
            v0 = {amp * (e1x * x1 + e2x * x2), amp * (e1x * x3 + e2x * x4)};
            v1 = {amp * (e1y * x1 + e2y * x2), amp * (e1y * x3 + e2y * x4)};
            v2 = {amp * (e1z * x1 + e2z * x2), amp * (e1z * x3 + e2z * x4)};
        }
        a0[q] = v0; a1[q] = v1; a2[q] = v2;
    }
}

// Higgs component b: real amplitude times one draw (stream 40 + b)
__global__ void __launch_bounds__(256) higgs_modes(hipfftDoubleComplex *a, long n3, ModeGeo g, int b)
{
    for (long q = blockIdx.x * 256L + threadIdx.x; q < n3; q += (long)gridDim.x * 256L) {
        double kx, ky, kz, modk;
        mode_k(g, q, kx, ky, kz, modk);
        const double w = sqrt(modk * modk + g.meff_sq);
        // a massless zero mode has no finite occupation; it is left empty
        const double amp = w > 0.0 ? sqrt((1.0 / (exp(w / g.T) - 1.0)) / (w * g.vol)) : 0.0;
        a[q] = {amp * gauss_sqrt2(g.seed, 40 + b, q), 0.0};
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
        gauge_modes<<<grid_for(n3), 256, 0, st>>>(a[0], a[1], a[2], n3, g, grp);
        for (int c = 0; c < 3 && rc == 0; c++) {
            if (hipfftExecZ2Z(plan, a[c], a[c], HIPFFT_BACKWARD) != HIPFFT_SUCCESS) { *what = "hipfftExecZ2Z"; rc = -1; }
            const int ch = grp < 3 ? CH_W + 3 * grp + c : CH_Y + c;
            take_real<<<grid_for(n3), 256, 0, st>>>(a[c], y + (long)ch * n3, n3);
        }
    }
    for (int b = 0; b < 4 && rc == 0; b++) {
        higgs_modes<<<grid_for(n3), 256, 0, st>>>(a[0], n3, g, b);
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

}  // namespace ewpt
