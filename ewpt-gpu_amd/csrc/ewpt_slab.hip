// ewpt_slab.hip -- halo planes of a z-slab context.
//
// A lattice split into z-slabs over several contexts (one per GPU in a
// multi-GPU run) keeps HALO ghost planes below and above each slab's
// interior in every array.  Before each stage the planes that stage reads
// at halo sites are exchanged: each slab packs its three lowest and three
// highest interior planes of the array into two contiguous buffers, the
// buffers travel to the z-neighbours, and each slab unpacks what it
// receives into its ghost planes.  This replaces the exact periodic wrap of
// diff_scheme.jl:161-178 at the slab faces; the reference's own multi-GPU
// driver (main_rk4_multiple.jl:69-213) exchanges boundary planes with MPI.
#include "ewpt_internal.h"

namespace ewpt {

namespace {

// One thread per element of one side: blockIdx.y 0 = low side, 1 = high.
// Channel ch's three planes are contiguous in both the array and the buffer.
template <typename T>
__global__ void __launch_bounds__(256) halo_pack(const T *a, long n3, long plane, int nz, int nch, T *lo, T *hi)
{
    const long per = HALO * plane;
    const long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (i >= nch * per) return;
    const long ch = i / per, r = i - ch * per;
    const T *base = a + ch * n3;
    if (blockIdx.y == 0) lo[i] = base[HALO * plane + r];   // interior planes 0..2
    else hi[i] = base[(long)(HALO + nz - 3) * plane + r];  // interior planes nz-3..nz-1
}

template <typename T>
__global__ void __launch_bounds__(256) halo_unpack(T *a, long n3, long plane, int nz, int nch, const T *lo,
                                                   const T *hi)
{
    const long per = HALO * plane;
    const long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (i >= nch * per) return;
    const long ch = i / per, r = i - ch * per;
    T *base = a + ch * n3;
    if (blockIdx.y == 0) base[r] = lo[i];                    // ghost planes -3..-1
    else base[(long)(HALO + nz) * plane + r] = hi[i];       // ghost planes nz..nz+2
}

template <typename T>
hipError_t launch_halo(bool pack, T *a, long n3, long plane, int nz, int nch, T *lo, T *hi, hipStream_t st)
{
    const long n = nch * HALO * plane;
    const dim3 grid((unsigned)((n + 255) / 256), 2), block(256);
    if (pack) halo_pack<T><<<grid, block, 0, st>>>(a, n3, plane, nz, nch, lo, hi);
    else halo_unpack<T><<<grid, block, 0, st>>>(a, n3, plane, nz, nch, lo, hi);
    return hipGetLastError();
}

}  // namespace

hipError_t launch_halo_pack(const double *a, long n3, long plane, int nz, int nch, double *lo, double *hi,
                            hipStream_t st)
{
    return launch_halo<double>(true, const_cast<double *>(a), n3, plane, nz, nch, lo, hi, st);
}
hipError_t launch_halo_pack(const float *a, long n3, long plane, int nz, int nch, float *lo, float *hi,
                            hipStream_t st)
{
    return launch_halo<float>(true, const_cast<float *>(a), n3, plane, nz, nch, lo, hi, st);
}
hipError_t launch_halo_unpack(double *a, long n3, long plane, int nz, int nch, const double *lo, const double *hi,
                              hipStream_t st)
{
    return launch_halo<double>(false, a, n3, plane, nz, nch, const_cast<double *>(lo), const_cast<double *>(hi), st);
}
hipError_t launch_halo_unpack(float *a, long n3, long plane, int nz, int nch, const float *lo, const float *hi,
                              hipStream_t st)
{
    return launch_halo<float>(false, a, n3, plane, nz, nch, const_cast<float *>(lo), const_cast<float *>(hi), st);
}

}  // namespace ewpt
