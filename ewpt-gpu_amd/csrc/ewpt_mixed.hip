// ewpt_mixed.hip -- conversions between the FP64 state and the mixed path's
// FP32 pair (high word + compensation word).  The pair represents
// y = (double)hi + (double)lo; splitting rounds y to hi and keeps the rest
// in lo, so joining a freshly split state returns y to ~2^-48 relative.
#include "ewpt_internal.h"

namespace ewpt {

namespace {

__global__ void __launch_bounds__(256) split_kernel(const double *y, float *hi, float *lo, long n)
{
    const long i = blockIdx.x * 256L + threadIdx.x;
    if (i >= n) return;
    const double v = y[i];
    const float h = (float)v;
    hi[i] = h;
    lo[i] = (float)(v - (double)h);
}

__global__ void __launch_bounds__(256) join_kernel(const float *hi, const float *lo, double *y, long n)
{
    const long i = blockIdx.x * 256L + threadIdx.x;
    if (i >= n) return;
    y[i] = (double)hi[i] + (double)lo[i];
}

}  // namespace

hipError_t launch_split(const double *y, float *hi, float *lo, long n, hipStream_t st)
{
    split_kernel<<<(unsigned)((n + 255) / 256), 256, 0, st>>>(y, hi, lo, n);
    return hipGetLastError();
}

hipError_t launch_join(const float *hi, const float *lo, double *y, long n, hipStream_t st)
{
    join_kernel<<<(unsigned)((n + 255) / 256), 256, 0, st>>>(hi, lo, y, n);
    return hipGetLastError();
}

}  // namespace ewpt
