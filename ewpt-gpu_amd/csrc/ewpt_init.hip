// ewpt_init.hip -- seeded synthetic states generated on the device
// (ewpt_init_modes).  A 512^3 FP64 state is 38.7 GB; building it on the host
// and uploading it would dominate a benchmark run, so the field is evaluated
// in place from a small mode table.
#include <cmath>
#include <vector>

#include "ewpt_internal.h"

namespace ewpt {

namespace {

__global__ void __launch_bounds__(256) fill_modes(double *y, long n3, int nx, int ny, int nz, const double *modes,
                                                 int nmodes, double vev)
{
    // grid-stride: a lattice may hold more sites than one dispatch's 32-bit
    // work-item count
    for (long q = blockIdx.x * 256L + threadIdx.x; q < n3; q += (long)gridDim.x * 256L) {
    const int i = (int)(q % nx);
    const int j = (int)((q / nx) % ny);
    const int k = (int)(q / ((long)nx * ny));
    const double tx = 2.0 * M_PI * i / nx, ty = 2.0 * M_PI * j / ny, tz = 2.0 * M_PI * k / nz;
    for (int c = 0; c < NCH; c++) {
        double v = (c == 3) ? vev : 0.0;
        for (int m = 0; m < nmodes; m++) {
            const double *t = modes + ((long)c * nmodes + m) * 5;
            v += t[4] * cos(t[0] * tx + t[1] * ty + t[2] * tz + t[3]);
        }
        y[(long)c * n3 + q] = v;
    }
    }
}

}  // namespace

unsigned long long splitmix64(unsigned long long &s)
{
    unsigned long long z = (s += 0x9e3779b97f4a7c15ULL);
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
    return z ^ (z >> 31);
}

double u01(unsigned long long &s) { return (double)(splitmix64(s) >> 11) * 0x1.0p-53; }

void mode_table(unsigned long long seed, int nmodes, int kmax, double amp, double mom_scale, double *t)
{
    unsigned long long s = seed;
    for (int c = 0; c < NCH; c++)
        for (int m = 0; m < nmodes; m++) {
            double *e = t + ((long)c * nmodes + m) * 5;
            for (int d = 0; d < 3; d++) e[d] = std::floor(u01(s) * (2 * kmax + 1)) - kmax;
            e[3] = 2.0 * M_PI * u01(s);
            e[4] = (2.0 * u01(s) - 1.0) * amp / nmodes * (c >= CH_PI ? mom_scale : 1.0);
        }
}

hipError_t launch_fill_modes(double *y, const ewpt_lattice &L, const double *dmodes, int nmodes, double vev,
                             hipStream_t st)
{
    const long n3 = (long)L.nx * L.ny * L.nz;
    const long nb = (n3 + 255) / 256;
    fill_modes<<<(unsigned)(nb < 65536 ? nb : 65536), 256, 0, st>>>(y, n3, L.nx, L.ny, L.nz, dmodes, nmodes, vev);
    return hipGetLastError();
}

}  // namespace ewpt
