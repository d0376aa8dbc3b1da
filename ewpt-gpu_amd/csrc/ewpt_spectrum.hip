// ewpt_spectrum.hip -- shell-averaged power spectrum of a 3-component field.
//
// Replaces the host-side spectrum of the reference: fft of B_x, B_y, B_z
// (CUFFT, gpu-kernels-rk4/main_rk4.jl:5831-5869), then convolve_1d
// (spec_routines.jl:47-77): the power sum_c |B_c(K)|^2 averaged over the
// lattice points of each distinct |K| (signed integer indices, K_c_mag
// spec_routines.jl:5-29), for the first Kc_bin_nums(cut) distinct values
// (spec_routines.jl:31-43).  Here: one batched D2Z transform (the half
// spectrum; a mode with 0 < Kx < nx/2 stands for itself and its conjugate
// partner, whose power is equal for a real field) and a histogram over
// integer K^2 with FP64 atomics, restricted to the K^2 that the requested
// shells cover.
#include <hip/hip_runtime.h>
#include <hipfft/hipfft.h>

#include <cmath>
#include <vector>

#include "ewpt_internal.h"

namespace ewpt {

namespace {

// sorted distinct K^2 = Kx^2 + Ky^2 + Kz^2 with |K_d| in 0..n_d/2 (the
// magnitudes the signed index takes on an n_d-point axis)
std::vector<long> distinct_k2(int n0, int n1, int n2)
{
    const long a0 = (long)(n0 / 2) * (n0 / 2), a1 = (long)(n1 / 2) * (n1 / 2), a2 = (long)(n2 / 2) * (n2 / 2);
    std::vector<char> pair(a0 + a1 + 1, 0), all(a0 + a1 + a2 + 1, 0);
    for (long x = 0; x <= n0 / 2; x++)
        for (long y = 0; y <= n1 / 2; y++) pair[x * x + y * y] = 1;
    for (long v = 0; v <= a0 + a1; v++)
        if (pair[v])
            for (long z = 0; z <= n2 / 2; z++) all[v + z * z] = 1;
    std::vector<long> out;
    for (long v = 0; v <= a0 + a1 + a2; v++)
        if (all[v]) out.push_back(v);
    return out;
}

__global__ void __launch_bounds__(256) shell_hist(const hipfftDoubleComplex *F, long nc, int nxh, int nx, int ny,
                                                  int nz, long k2max, double *sum, double *cnt)
{
    for (long q = blockIdx.x * 256L + threadIdx.x; q < nc; q += (long)gridDim.x * 256L) {
        const int kx = (int)(q % nxh), j = (int)((q / nxh) % ny), k = (int)(q / ((long)nxh * ny));
        const long Ky = j <= ny / 2 ? j : j - ny, Kz = k <= nz / 2 ? k : k - nz;
        const long k2 = (long)kx * kx + Ky * Ky + Kz * Kz;
        if (k2 > k2max) continue;
        const double w = (kx == 0 || 2 * kx == nx) ? 1.0 : 2.0;
        double pw = 0.0;
#pragma unroll
        for (int c = 0; c < 3; c++) {
            const hipfftDoubleComplex v = F[c * nc + q];
            pw += v.x * v.x + v.y * v.y;
        }
        atomicAdd(sum + k2, w * pw);
        atomicAdd(cnt + k2, w);
    }
}

}  // namespace

long distinct_k2_count(int n0, int n1, int n2) { return (long)distinct_k2(n0, n1, n2).size(); }

int shell_spectrum(const double *field, const ewpt_lattice &L, const int cut[3], void *cwork, hipStream_t st,
                   int maxbins, double *k_out, double *p_out, int *nbins, const char **what)
{
    const long nb = distinct_k2_count(cut[0], cut[1], cut[2]);   // Kc_bin_nums(cutx, cuty, cutz)
    *nbins = (int)nb;
    if (!k_out || !p_out) return 0;                               // size query
    if (maxbins < nb) { *what = "spectrum: maxbins is smaller than the number of shells"; return -2; }
    const std::vector<long> vals = distinct_k2(L.nx, L.ny, L.nz);
    if ((long)vals.size() < nb) { *what = "spectrum: cut has more shells than the lattice"; return -2; }
    const long k2max = vals[nb - 1];

    const long n3 = (long)L.nx * L.ny * L.nz;
    const int nxh = L.nx / 2 + 1;
    const long nc = (long)nxh * L.ny * L.nz;
    hipfftDoubleComplex *F = static_cast<hipfftDoubleComplex *>(cwork);
    double *hist = nullptr;
    if (hipMalloc(reinterpret_cast<void **>(&hist), 2 * (k2max + 1) * sizeof(double)) != hipSuccess) {
        *what = "spectrum: alloc histogram";
        return -1;
    }
    int rc = 0;
    // 64-bit planner: the batch distances exceed an int on the largest lattices
    hipfftHandle plan = 0;
    long long ldims[3] = {L.nz, L.ny, L.nx};   // x fastest
    size_t wsz = 0;
    if (hipfftCreate(&plan) != HIPFFT_SUCCESS) {
        *what = "spectrum: hipfftCreate";
        (void)hipFree(hist);
        return -1;
    }
    if (hipfftMakePlanMany64(plan, 3, ldims, nullptr, 1, n3, nullptr, 1, nc, HIPFFT_D2Z, 3, &wsz) != HIPFFT_SUCCESS) {
        *what = "spectrum: hipfftMakePlanMany64";
        rc = -1;
    }
    if (rc == 0 && hipfftSetStream(plan, st) != HIPFFT_SUCCESS) { *what = "spectrum: hipfftSetStream"; rc = -1; }
    if (rc == 0 && hipMemsetAsync(hist, 0, 2 * (k2max + 1) * sizeof(double), st) != hipSuccess) {
        *what = "spectrum: clear histogram";
        rc = -1;
    }
    if (rc == 0 && hipfftExecD2Z(plan, const_cast<double *>(field), F) != HIPFFT_SUCCESS) {
        *what = "spectrum: hipfftExecD2Z";
        rc = -1;
    }
    if (rc == 0) {
        const long blocks = (nc + 255) / 256;
        shell_hist<<<(unsigned)(blocks < 65536 ? blocks : 65536), 256, 0, st>>>(F, nc, nxh, L.nx, L.ny, L.nz, k2max,
                                                                                 hist, hist + k2max + 1);
        if (hipGetLastError() != hipSuccess) { *what = "spectrum: histogram launch"; rc = -1; }
    }
    std::vector<double> h(2 * (k2max + 1));
    if (rc == 0 && (hipMemcpyAsync(h.data(), hist, h.size() * sizeof(double), hipMemcpyDeviceToHost, st) != hipSuccess ||
                    hipStreamSynchronize(st) != hipSuccess)) {
        *what = "spectrum: copy histogram";
        rc = -1;
    }
    (void)hipStreamSynchronize(st);   // nothing may still use hist or the plan
    hipfftDestroy(plan);
    (void)hipFree(hist);
    if (rc) return rc;
    for (long b = 0; b < nb; b++) {
        const long v = vals[b];
        k_out[b] = std::sqrt((double)v);
        p_out[b] = h[v] / h[k2max + 1 + v];
    }
    return 0;
}

}  // namespace ewpt
