// ewpt_mixed.hip -- conversions between the FP64 state and the mixed path's
// FP32 pair (high word + compensation word).  The pair represents
// y = (double)hi + (double)lo; splitting rounds y to hi and keeps the rest
// in lo, so joining a freshly split state returns y to ~2^-48 relative.
#include "ewpt_internal.h"

namespace ewpt {

namespace {

// A whole 36-channel state can exceed 2^32 elements (512^3: 4.8e9), more
// work-items than one dispatch holds, so these kernels stride over it.
constexpr unsigned CONV_BLOCKS = 16384;

__global__ void __launch_bounds__(256) split_kernel(const double *y, float *hi, float *lo, long n)
{
    for (long i = blockIdx.x * 256L + threadIdx.x; i < n; i += (long)gridDim.x * 256L) {
        const double v = y[i];
        const float h = (float)v;
        hi[i] = h;
        lo[i] = (float)(v - (double)h);
    }
}

__global__ void __launch_bounds__(256) join_kernel(const float *hi, const float *lo, double *y, long n)
{
    for (long i = blockIdx.x * 256L + threadIdx.x; i < n; i += (long)gridDim.x * 256L)
        y[i] = (double)hi[i] + (double)lo[i];
}

// plain FP64 copy, one 8-byte element per lane and iteration (the access
// width of the stage kernels): the reference point for memory counters
__global__ void __launch_bounds__(256) copy_kernel(const double *__restrict__ a, double *__restrict__ b, long n)
{
    for (long i = blockIdx.x * 256L + threadIdx.x; i < n; i += (long)gridDim.x * 256L) b[i] = a[i];
}

unsigned conv_grid(long n)
{
    const long b = (n + 255) / 256;
    return (unsigned)(b < (long)CONV_BLOCKS ? b : (long)CONV_BLOCKS);
}

}  // namespace

hipError_t launch_split(const double *y, float *hi, float *lo, long n, hipStream_t st)
{
    split_kernel<<<conv_grid(n), 256, 0, st>>>(y, hi, lo, n);
    return hipGetLastError();
}

hipError_t launch_join(const float *hi, const float *lo, double *y, long n, hipStream_t st)
{
    join_kernel<<<conv_grid(n), 256, 0, st>>>(hi, lo, y, n);
    return hipGetLastError();
}

hipError_t launch_copy(const double *a, double *b, long n, hipStream_t st)
{
    copy_kernel<<<conv_grid(n), 256, 0, st>>>(a, b, n);
    return hipGetLastError();
}

}  // namespace ewpt
