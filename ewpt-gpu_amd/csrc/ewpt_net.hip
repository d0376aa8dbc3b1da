// ewpt_net.hip -- learned coarse-step corrector: a periodic 3x3x3 convolution
// network 36 -> 64 -> 64 -> 36 (SiLU between layers) that maps the state
// after one RK4 step of 4 dt to a correction towards four RK4 steps of dt.
// No counterpart in the reference.
//
// Activations are channel-last BF16, [site][64] (36 state channels padded
// with zeros), weights BF16 [tap][out][in] with FP32 master copies, FP32
// accumulation on the matrix cores (v_mfma_f32_16x16x32_bf16; lane l holds
// A[row l%16][k (l/16)*8 .. +7], B[k (l/16)*8 .. +7][col l%16] and
// C[rows (l/16)*4 .. +3][col l%16]).
//
// conv_kernel: implicit GEMM, M = sites, N = 64 outputs, K = 27 taps x 64
// inputs.  A workgroup (4 waves) owns an 8x4x4 brick: the brick plus a 1-site
// halo of the input (360 sites x 64 channels) is staged once in LDS, the
// weights of one tap (64 x 64) are double-buffered in LDS, and each wave
// accumulates 32 sites x 64 outputs (8 MFMA tiles) over the 27 taps.  The
// same kernel computes the input gradient of a layer (weights transposed
// and taps mirrored) with the SiLU derivative applied in its epilogue.
// wgrad_kernel: dW[tap][out][in] = sum_x dZ[x][out] X[x + tap][in], M = out,
// N = in, K = sites: per brick the two operand tiles are written transposed
// into LDS; a workgroup owns one tap and a chunk of bricks and writes an
// FP32 partial, reduced over chunks in a fixed order (deterministic).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <mutex>
#include <new>
#include <random>
#include <type_traits>
#include <vector>

#include "ewpt_internal.h"

namespace ewpt {

namespace {

constexpr int NC = 64;        // padded channel count of every activation
constexpr int NTAP = 27;
constexpr int BRX = 8, BRY = 4, BRZ = 4, BRN = BRX * BRY * BRZ;   // conv brick
constexpr int HX = BRX + 2, HY = BRY + 2, HZ = BRZ + 2, HN = HX * HY * HZ;
constexpr int ROW = NC + 8;   // LDS row pitch (bf16) of a site / a weight row
constexpr int TROW = BRN + 8; // LDS row pitch of a transposed tile

typedef __attribute__((__vector_size__(8 * sizeof(__bf16)))) __bf16 bf16x8;
typedef __attribute__((__vector_size__(4 * sizeof(float)))) float f32x4;
typedef unsigned u32x4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ unsigned short f2bf(float f)
{
    unsigned u = __float_as_uint(f);
    if ((u & 0x7fffffffu) > 0x7f800000u) return (unsigned short)((u >> 16) | 0x40u);   // quiet NaN
    u += 0x7fffu + ((u >> 16) & 1u);                                                      // round to nearest even
    return (unsigned short)(u >> 16);
}
__device__ __forceinline__ float bf2f(unsigned short h) { return __uint_as_float((unsigned)h << 16); }

__device__ __forceinline__ float sigm(float z) { return 1.0f / (1.0f + __expf(-z)); }

__device__ __forceinline__ int wrapn(int v, int n) { return v < 0 ? v + n : (v >= n ? v - n : v); }

enum : int { EPI_HIDDEN = 0, EPI_LAST = 1, EPI_DGRAD = 2, EPI_APPLY = 3 };

// bf16 of f, round to nearest even (v_cvt_pk_bf16_f32); the values f2bf gives
// for every finite f and for the infinities
__device__ __forceinline__ unsigned short bf16_rne(float f) { return __builtin_bit_cast(unsigned short, (__bf16)f); }

// X: [n3][64] bf16; Wk: [27][64 n][64 k] bf16; out per epilogue:
//   EPI_HIDDEN: z = acc + bias -> Z (f32, optional), A = bf16(silu(z))
//   EPI_LAST:   z = acc + bias -> Z (f32)
//   EPI_DGRAD:  A = bf16(acc * silu'(zprev))
template <int EPI>
__global__ void __launch_bounds__(256) conv_kernel(const unsigned short *__restrict__ X,
                                                   const unsigned short *__restrict__ Wk,
                                                   const float *__restrict__ bias, const float *__restrict__ zprev,
                                                   int nx, int ny, int nz, float *__restrict__ Z,
                                                   unsigned short *__restrict__ A)
{
    extern __shared__ __align__(16) unsigned short lds16[];
    unsigned short *tile = lds16;                 // [HN][ROW]
    unsigned short *wbuf = lds16 + HN * ROW;      // [2][64][ROW]
    const int nbx = nx / BRX, nby = ny / BRY;
    const int bx = blockIdx.x % nbx, by = (blockIdx.x / nbx) % nby, bz = blockIdx.x / (nbx * nby);
    const int x0 = bx * BRX, y0 = by * BRY, z0 = bz * BRZ;
    const int t = threadIdx.x, lane = t & 63, w = t >> 6;

    // input brick + halo, 16-byte chunks
    for (int c = t; c < HN * 8; c += 256) {
        const int h = c >> 3, part = c & 7;
        const int hx = h % HX, hy = (h / HX) % HY, hz = h / (HX * HY);
        const long g = wrapn(x0 + hx - 1, nx) + (long)nx * (wrapn(y0 + hy - 1, ny) + (long)ny * wrapn(z0 + hz - 1, nz));
        *reinterpret_cast<uint4 *>(tile + h * ROW + part * 8) =
            *reinterpret_cast<const uint4 *>(X + g * NC + part * 8);
    }
    // weights of tap 0
    uint4 wr[2];
#pragma unroll
    for (int r = 0; r < 2; r++) wr[r] = *reinterpret_cast<const uint4 *>(Wk + (t + 256 * r) * 8);
#pragma unroll
    for (int r = 0; r < 2; r++) {
        const int c = t + 256 * r;
        *reinterpret_cast<uint4 *>(wbuf + (c >> 3) * ROW + (c & 7) * 8) = wr[r];
    }
    __syncthreads();

    f32x4 acc[2][4];
#pragma unroll
    for (int mb = 0; mb < 2; mb++)
#pragma unroll
        for (int nb = 0; nb < 4; nb++) acc[mb][nb] = f32x4{0.f, 0.f, 0.f, 0.f};

    const int r16 = lane & 15, kg = (lane >> 4) * 8;
    for (int tap = 0; tap < NTAP; tap++) {
        const unsigned short *wb = wbuf + (tap & 1) * NC * ROW;
        if (tap + 1 < NTAP) {
#pragma unroll
            for (int r = 0; r < 2; r++)
                wr[r] = *reinterpret_cast<const uint4 *>(Wk + (long)(tap + 1) * NC * NC + (t + 256 * r) * 8);
        }
        const int dx = tap % 3 - 1, dy = (tap / 3) % 3 - 1, dz = tap / 9 - 1;
#pragma unroll
        for (int ks = 0; ks < 2; ks++) {
            bf16x8 a[2], bb[4];
#pragma unroll
            for (int mb = 0; mb < 2; mb++) {
                const int s = mb * 16 + r16;      // site in this wave's z-plane of the brick
                const int h = (w + 1 + dz) * (HX * HY) + ((s >> 3) + 1 + dy) * HX + ((s & 7) + 1 + dx);
                a[mb] = __builtin_bit_cast(bf16x8, *reinterpret_cast<const uint4 *>(tile + h * ROW + ks * 32 + kg));
            }
#pragma unroll
            for (int nb = 0; nb < 4; nb++)
                bb[nb] = __builtin_bit_cast(bf16x8,
                                            *reinterpret_cast<const uint4 *>(wb + (nb * 16 + r16) * ROW + ks * 32 + kg));
#pragma unroll
            for (int mb = 0; mb < 2; mb++)
#pragma unroll
                for (int nb = 0; nb < 4; nb++)
                    acc[mb][nb] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[mb], bb[nb], acc[mb][nb], 0, 0, 0);
        }
        if (tap + 1 < NTAP) {
            unsigned short *wn = wbuf + ((tap + 1) & 1) * NC * ROW;
#pragma unroll
            for (int r = 0; r < 2; r++) {
                const int c = t + 256 * r;
                *reinterpret_cast<uint4 *>(wn + (c >> 3) * ROW + (c & 7) * 8) = wr[r];
            }
        }
        __syncthreads();
    }

    // epilogue: lane holds rows (lane/16)*4 + i of each 16-site block, column lane%16
#pragma unroll
    for (int mb = 0; mb < 2; mb++)
#pragma unroll
        for (int i = 0; i < 4; i++) {
            const int s = mb * 16 + (lane >> 4) * 4 + i;
            const long g = (x0 + (s & 7)) + (long)nx * ((y0 + (s >> 3)) + (long)ny * (z0 + w));
#pragma unroll
            for (int nb = 0; nb < 4; nb++) {
                const int co = nb * 16 + r16;
                const float v = acc[mb][nb][i];
                if constexpr (EPI == EPI_HIDDEN) {
                    const float z = v + bias[co];
                    if (Z) Z[g * NC + co] = z;
                    A[g * NC + co] = f2bf(z * sigm(z));
                } else if constexpr (EPI == EPI_LAST) {
                    Z[g * NC + co] = v + bias[co];
                } else {
                    const float zp = zprev[g * NC + co];
                    const float sg = sigm(zp);
                    A[g * NC + co] = f2bf(v * (sg * (1.0f + zp * (1.0f - sg))));
                }
            }
        }
}

// conv2_kernel: the same convolution with the weights kept out of LDS.  A
// workgroup of 2 waves (two workgroups per CU) owns an 8x8x4 brick; its
// input plus a 1-site halo (600 sites x 128 B) is loaded once, straight from
// global memory into LDS, as 128-byte rows whose 16-byte chunks are permuted
// by the low 3 bits of the site index (the 8 rows one LDS cycle serves fall
// on distinct banks).  Wave w computes the 128 sites
// of z-planes 2w, 2w+1 against NB*16 outputs: 8 x NB accumulator tiles
// (16x16, FP32) over 27 taps x 2 k-steps.  The B fragments (one tap's
// weights, 16 bytes per lane per tile, identical for every workgroup and so
// served by L1/L2) go straight into a ring of three register sets, two taps
// ahead of their use; with 128 sites per wave each weight byte feeds twice
// the matrix work of a 64-site tile, which keeps the L1 path at half its
// rate.  The A fragments are read from LDS, written once, so the tap loop
// has no barrier.  The outputs leave through LDS as whole 16-byte chunks.
// NB = 3 computes outputs 0..47 only (the last layer: its outputs 36..63
// are padding and its Z rows are written up to channel 47).
constexpr int C2X = 8, C2Y = 8, C2Z = 4, C2N = C2X * C2Y * C2Z;
constexpr int C2HX = C2X + 2, C2HY = C2Y + 2, C2HZ = C2Z + 2, C2HN = C2HX * C2HY * C2HZ;
// waves per workgroup: 2 (128 sites per wave, one wave per SIMD); 4 waves of
// 64 sites (two per SIMD) move twice the weight bytes per MFMA and measured
// slower at 256^3 (6.7 against 5.3 ms per hidden layer)
constexpr int C2W = 2;
constexpr int C2MB = 16 / C2W;         // 16-site M-blocks per wave (the brick's 256 sites over C2W waves)
constexpr int C2T = 64 * C2W;          // threads per workgroup
constexpr size_t CONV2_LDS = (size_t)C2HN * NC * sizeof(unsigned short);

// bf16 offset of 16-byte chunk c (0..7) of halo site h in the conv2 tile
__device__ __forceinline__ int c2_off(int h, int c) { return h * NC + ((c ^ (h & 7)) << 3); }

template <int EPI, int NB>
__global__ void __launch_bounds__(C2T) __attribute__((amdgpu_waves_per_eu(1, 1)))
conv2_kernel(const unsigned short *__restrict__ X, const unsigned short *__restrict__ Wk,
             const float *__restrict__ bias, const float *__restrict__ zprev, int nx, int ny, int nz,
             float *__restrict__ Z, unsigned short *__restrict__ A, double *__restrict__ Y,
             const float *__restrict__ sout)
{
    extern __shared__ __align__(16) unsigned short lds16[];
    unsigned short *tile = lds16;   // [C2HN] rows of 64 bf16, chunks permuted (c2_off)
    const int nbx = nx / C2X, nby = ny / C2Y, nbr = nbx * nby * (nz / C2Z);
    int bid = blockIdx.x;
    // consecutive block ids go to different XCDs: give XCD x a contiguous
    // range of bricks so that the halo rows neighbours share stay in its L2
    if ((nbr & 7) == 0) bid = (bid & 7) * (nbr >> 3) + (bid >> 3);
    const int bx = bid % nbx, by = (bid / nbx) % nby, bz = bid / (nbx * nby);
    const int x0 = bx * C2X, y0 = by * C2Y, z0 = bz * C2Z;
    const int t = threadIdx.x, lane = t & 63, w = t >> 6, r16 = lane & 15, q = lane >> 4;

    auto loadB = [&](int tap, bf16x8 (&b)[NB][2]) {
#pragma unroll
        for (int nb = 0; nb < NB; nb++)
#pragma unroll
            for (int ks = 0; ks < 2; ks++)
                b[nb][ks] = __builtin_bit_cast(
                    bf16x8, *reinterpret_cast<const uint4 *>(Wk + ((long)(tap * NC + nb * 16 + r16) * NC + ks * 32 + q * 8)));
    };
    bf16x8 bA[NB][2], bB[NB][2], bC[NB][2];
    loadB(0, bA);
    loadB(1, bB);

    // input brick + halo straight from global memory into LDS
    // (global_load_lds_dwordx4: lane l of a wave instruction writes the 16
    // bytes it loads to LDS slot base + l), so every chunk of the tile is in
    // flight at once without passing through registers.  Slot k = 8 h + p of
    // halo site h holds its chunk p ^ (h & 7) (the c2_off permutation).
    {
        constexpr int NINST = C2HN * 8 / 64;
        static_assert(C2HN * 8 % 64 == 0, "the tile is a whole number of 64-slot wave loads");
        for (int i = w; i < NINST; i += C2W) {
            const int k = i * 64 + lane, h = k >> 3, c = (k & 7) ^ (h & 7);
            const int hx = h % C2HX, hy = (h / C2HX) % C2HY, hz = h / (C2HX * C2HY);
            const long g =
                wrapn(x0 + hx - 1, nx) + (long)nx * (wrapn(y0 + hy - 1, ny) + (long)ny * wrapn(z0 + hz - 1, nz));
            __builtin_amdgcn_global_load_lds(reinterpret_cast<const void *>(X + g * NC + c * 8),
                                             (__attribute__((address_space(3))) void *)(tile + i * 64 * 8), 16, 0, 0);
        }
        asm volatile("s_waitcnt vmcnt(0)" ::: "memory");   // the tile has landed in LDS
    }
    __syncthreads();

    f32x4 acc[C2MB][NB];
#pragma unroll
    for (int mb = 0; mb < C2MB; mb++)
#pragma unroll
        for (int nb = 0; nb < NB; nb++) acc[mb][nb] = f32x4{0.f, 0.f, 0.f, 0.f};
    // A-fragment row r16 of M-block mb is site (r16 & 7, 2 (m & 3) + (r16 >> 3)) of z-plane m >> 2,
    // m = C2MB w + mb the M-block's index in the brick
    int hb[C2MB];
#pragma unroll
    for (int mb = 0; mb < C2MB; mb++) {
        const int m = C2MB * w + mb;
        hb[mb] = ((m >> 2) + 1) * (C2HX * C2HY) + (2 * (m & 3) + (r16 >> 3) + 1) * C2HX + (r16 & 7) + 1;
    }
    auto step = [&](int tap, const bf16x8 (&b)[NB][2]) {
        const int off = (tap / 9 - 1) * (C2HX * C2HY) + ((tap / 3) % 3 - 1) * C2HX + (tap % 3 - 1);
        bf16x8 a[C2MB][2];
#pragma unroll
        for (int ks = 0; ks < 2; ks++)
#pragma unroll
            for (int mb = 0; mb < C2MB; mb++)
                a[mb][ks] = __builtin_bit_cast(bf16x8, *reinterpret_cast<const uint4 *>(tile + c2_off(hb[mb] + off, ks * 4 + q)));
#pragma unroll
        for (int ks = 0; ks < 2; ks++)
#pragma unroll
            for (int mb = 0; mb < C2MB; mb++)
#pragma unroll
                for (int nb = 0; nb < NB; nb++)
                    acc[mb][nb] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[mb][ks], b[nb][ks], acc[mb][nb], 0, 0, 0);
    };
    for (int tap = 0; tap < NTAP; tap += 3) {
        loadB(tap + 2, bC);
        step(tap, bA);
        if (tap + 3 < NTAP) loadB(tap + 3, bA);
        step(tap + 1, bB);
        if (tap + 4 < NTAP) loadB(tap + 4, bB);
        step(tap + 2, bC);
    }

    // epilogue through LDS: lane holds rows q*4 + i of each M-block and column
    // r16 of each N-block; the brick's rows are staged in the tile region
    // (row = site s = x + 8 y + 64 z of the brick) and written as 16-byte chunks
    __syncthreads();   // every wave is done with the input tile
    auto site_of = [&](int mb, int i) {
        const int r = q * 4 + i, m = C2MB * w + mb;
        return (r & 7) + C2X * (2 * (m & 3) + (r >> 3)) + C2X * C2Y * (m >> 2);
    };
    auto gsite = [&](int s) {
        return (long)(x0 + (s & 7)) + (long)nx * ((y0 + ((s >> 3) & 7)) + (long)ny * (z0 + (s >> 6)));
    };
    constexpr int NO = NB * 16;   // output channels written per site
    if constexpr (EPI == EPI_APPLY) {
        // the correction y[c] += sout[c] * (acc + bias)[c], c < 36, into the
        // FP64 state (channel-major), in the order apply_corr adds it
        float *zs = reinterpret_cast<float *>(lds16);
        constexpr int ZP = NO + 4;
#pragma unroll
        for (int mb = 0; mb < C2MB; mb++)
#pragma unroll
            for (int i = 0; i < 4; i++)
#pragma unroll
                for (int nb = 0; nb < NB; nb++)
                    zs[site_of(mb, i) * ZP + nb * 16 + r16] = acc[mb][nb][i] + bias[nb * 16 + r16];
        __syncthreads();
        // 72 elements per thread in batches of 24 loads in flight (loads and
        // stores of y share a pointer, so the compiler keeps their order)
        const long n3 = (long)nx * ny * nz;
        constexpr int NE = NCH * C2N / C2T, YB = 24;
        static_assert(NCH * C2N % C2T == 0 && NE % YB == 0, "whole batches");
        for (int e0 = 0; e0 < NE; e0 += YB) {
            double v[YB];
#pragma unroll
            for (int j = 0; j < YB; j++) {
                const int e = t + (e0 + j) * C2T, c = e / C2N, s = e % C2N;
                v[j] = Y[c * n3 + gsite(s)] + (double)sout[c] * (double)zs[s * ZP + c];
            }
#pragma unroll
            for (int j = 0; j < YB; j++) {
                const int e = t + (e0 + j) * C2T;
                Y[(e / C2N) * n3 + gsite(e % C2N)] = v[j];
            }
        }
        return;
    }
    if constexpr (EPI != EPI_DGRAD) {
        if (Z) {
            // FP32 pre-activations, row pitch NO + 4 floats
            float *zs = reinterpret_cast<float *>(lds16);
            constexpr int ZP = NO + 4;
#pragma unroll
            for (int mb = 0; mb < C2MB; mb++)
#pragma unroll
                for (int i = 0; i < 4; i++)
#pragma unroll
                    for (int nb = 0; nb < NB; nb++)
                        zs[site_of(mb, i) * ZP + nb * 16 + r16] = acc[mb][nb][i] + bias[nb * 16 + r16];
            __syncthreads();
            constexpr int ZCH = NO / 4;   // 16-byte chunks per site
            for (int c = t; c < C2N * ZCH; c += C2T) {
                const int s = c / ZCH, part = c % ZCH;
                *reinterpret_cast<float4 *>(Z + gsite(s) * NC + part * 4) =
                    *reinterpret_cast<const float4 *>(zs + s * ZP + part * 4);
            }
            __syncthreads();
        }
    }
    if constexpr (EPI != EPI_LAST) {
        // bf16 activations (or input gradients), row pitch 72
        constexpr int AP = NC + 8;
#pragma unroll
        for (int mb = 0; mb < C2MB; mb++)
#pragma unroll
            for (int i = 0; i < 4; i++) {
                const int s = site_of(mb, i);
#pragma unroll
                for (int nb = 0; nb < NB; nb++) {
                    const int co = nb * 16 + r16;
                    float v;
                    if constexpr (EPI == EPI_HIDDEN) {
                        const float zz = acc[mb][nb][i] + bias[co];
                        v = zz * sigm(zz);
                    } else {
                        const float zp = zprev[gsite(s) * NC + co];
                        const float sg = sigm(zp);
                        v = acc[mb][nb][i] * (sg * (1.0f + zp * (1.0f - sg)));
                    }
                    lds16[s * AP + co] = bf16_rne(v);
                }
            }
        __syncthreads();
        for (int c = t; c < C2N * 8; c += C2T) {
            const int s = c >> 3, part = c & 7;
            *reinterpret_cast<uint4 *>(A + gsite(s) * NC + part * 8) = *reinterpret_cast<const uint4 *>(lds16 + s * AP + part * 8);
        }
    }
}

// per chunk of bricks and per tap: partial[chunk][tap][out][in]; the tap-0
// workgroups also sum dZ over sites into bpart[chunk][out]
__global__ void __launch_bounds__(256) wgrad_kernel(const unsigned short *__restrict__ dZ,
                                                    const unsigned short *__restrict__ X, int nx, int ny, int nz,
                                                    int bricks_per_chunk, float *__restrict__ partial,
                                                    float *__restrict__ bpart)
{
    __shared__ __align__(16) unsigned short dzT[NC * TROW];
    __shared__ __align__(16) unsigned short xT[NC * TROW];
    const int chunk = blockIdx.x, tap = blockIdx.y;
    const int dx = tap % 3 - 1, dy = (tap / 3) % 3 - 1, dz = tap / 9 - 1;
    const int nbx = nx / BRX, nby = ny / BRY, nbrick = nbx * nby * (nz / BRZ);
    const int t = threadIdx.x, lane = t & 63, w = t >> 6, r16 = lane & 15, kg = (lane >> 4) * 8;
    f32x4 acc[4];
#pragma unroll
    for (int nb = 0; nb < 4; nb++) acc[nb] = f32x4{0.f, 0.f, 0.f, 0.f};
    float bacc = 0.0f;
    const int b0 = chunk * bricks_per_chunk, b1 = min(nbrick, b0 + bricks_per_chunk);
    for (int bi = b0; bi < b1; bi++) {
        const int bx = bi % nbx, by = (bi / nbx) % nby, bz = bi / (nbx * nby);
        __syncthreads();   // previous brick's tiles consumed
        for (int c = t; c < BRN * 8; c += 256) {
            const int s = c >> 3, part = c & 7;
            const int sx = bx * BRX + (s & 7), sy = by * BRY + ((s >> 3) & 3), sz = bz * BRZ + (s >> 5);
            const long g = sx + (long)nx * (sy + (long)ny * sz);
            const long gs = wrapn(sx + dx, nx) + (long)nx * (wrapn(sy + dy, ny) + (long)ny * wrapn(sz + dz, nz));
            const uint4 vz = *reinterpret_cast<const uint4 *>(dZ + g * NC + part * 8);
            const uint4 vx = *reinterpret_cast<const uint4 *>(X + gs * NC + part * 8);
            const unsigned zz[4] = {vz.x, vz.y, vz.z, vz.w}, xx[4] = {vx.x, vx.y, vx.z, vx.w};
#pragma unroll
            for (int j = 0; j < 4; j++) {
                dzT[(part * 8 + 2 * j) * TROW + s] = (unsigned short)(zz[j] & 0xffffu);
                dzT[(part * 8 + 2 * j + 1) * TROW + s] = (unsigned short)(zz[j] >> 16);
                xT[(part * 8 + 2 * j) * TROW + s] = (unsigned short)(xx[j] & 0xffffu);
                xT[(part * 8 + 2 * j + 1) * TROW + s] = (unsigned short)(xx[j] >> 16);
            }
        }
        __syncthreads();
        if (tap == 0 && t < NC) {
            float sacc = 0.0f;
            for (int s = 0; s < BRN; s++) sacc += bf2f(dzT[t * TROW + s]);
            bacc += sacc;
        }
#pragma unroll
        for (int ks = 0; ks < BRN / 32; ks++) {
            const bf16x8 a = __builtin_bit_cast(
                bf16x8, *reinterpret_cast<const uint4 *>(dzT + (w * 16 + r16) * TROW + ks * 32 + kg));
#pragma unroll
            for (int nb = 0; nb < 4; nb++) {
                const bf16x8 b = __builtin_bit_cast(
                    bf16x8, *reinterpret_cast<const uint4 *>(xT + (nb * 16 + r16) * TROW + ks * 32 + kg));
                acc[nb] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, acc[nb], 0, 0, 0);
            }
        }
    }
    float *P = partial + ((long)chunk * NTAP + tap) * NC * NC;
#pragma unroll
    for (int nb = 0; nb < 4; nb++)
#pragma unroll
        for (int i = 0; i < 4; i++) P[(w * 16 + (lane >> 4) * 4 + i) * NC + nb * 16 + r16] = acc[nb][i];
    if (tap == 0 && t < NC) bpart[(long)chunk * NC + t] = bacc;
}

// sum the partials over chunks in chunk order
__global__ void __launch_bounds__(256) wgrad_reduce(const float *partial, const float *bpart, int nchunk, float *dW,
                                                    float *db)
{
    const int q = blockIdx.x * 256 + threadIdx.x;
    if (q < NTAP * NC * NC) {
        float s = 0.0f;
        for (int c = 0; c < nchunk; c++) s += partial[(long)c * NTAP * NC * NC + q];
        dW[q] = s;
    }
    if (q < NC) {
        float s = 0.0f;
        for (int c = 0; c < nchunk; c++) s += bpart[(long)c * NC + q];
        db[q] = s;
    }
}

// master FP32 weights [27][64 out][64 in] -> bf16 copy, and the dgrad operand
// Wt[tap][in][out] = W[26 - tap][out][in]
__global__ void __launch_bounds__(256) cast_weights(const float *W, unsigned short *Wb, unsigned short *Wt)
{
    const int q = blockIdx.x * 256 + threadIdx.x;
    if (q >= NTAP * NC * NC) return;
    const int tap = q / (NC * NC), o = (q / NC) % NC, i = q % NC;
    const unsigned short h = f2bf(W[q]);
    Wb[q] = h;
    Wt[((long)(NTAP - 1 - tap) * NC + i) * NC + o] = h;
}

// state (FP64 SoA, 36 channels) -> X [site][64] bf16, channel c scaled by
// inv[c].  mode 0: the state itself; mode 1: its 7-point lattice Laplacian
// y[x+1] + y[x-1] + ... - 6 y[x] (in lattice units), formed in FP64 before
// the rounding to bf16, which keeps the short-wavelength content that the
// coarse step gets wrong above the bf16 resolution.
__global__ void __launch_bounds__(256) pack_state(const double *y, long n3, int nx, int ny, int nz, int mode,
                                                  const float *inv, unsigned short *X)
{
    for (long q = blockIdx.x * 256L + threadIdx.x; q < n3; q += (long)gridDim.x * 256L) {
        const int i = (int)(q % nx), j = (int)((q / nx) % ny), k = (int)(q / ((long)nx * ny));
        const long nxy = (long)nx * ny;
        const long o[6] = {wrapn(i + 1, nx) - i, wrapn(i - 1, nx) - i, (long)(wrapn(j + 1, ny) - j) * nx,
                           (long)(wrapn(j - 1, ny) - j) * nx, (long)(wrapn(k + 1, nz) - k) * nxy,
                           (long)(wrapn(k - 1, nz) - k) * nxy};
        unsigned short h[NC];
#pragma unroll
        for (int c = 0; c < NC; c++) {
            if (c >= NCH) { h[c] = 0; continue; }
            const double *yc = y + c * n3 + q;
            double v = yc[0];
            if (mode == 1) v = ((yc[o[0]] + yc[o[1]]) + (yc[o[2]] + yc[o[3]]) + (yc[o[4]] + yc[o[5]])) - 6.0 * v;
            h[c] = f2bf((float)(v * (double)inv[c]));
        }
        uint4 *dst = reinterpret_cast<uint4 *>(X + q * NC);
#pragma unroll
        for (int p = 0; p < NC / 8; p++) {
            uint4 v;
            v.x = h[8 * p] | ((unsigned)h[8 * p + 1] << 16);
            v.y = h[8 * p + 2] | ((unsigned)h[8 * p + 3] << 16);
            v.z = h[8 * p + 4] | ((unsigned)h[8 * p + 5] << 16);
            v.w = h[8 * p + 6] | ((unsigned)h[8 * p + 7] << 16);
            dst[p] = v;
        }
    }
}

// block partial sums of sum_c T^2, T = (y_fine - y_coarse)/sout
__global__ void __launch_bounds__(256) target_sq(const double *yc, const double *yf, long n3, const float *sout,
                                                 double *part)
{
    __shared__ double red[256];
    double acc = 0.0;
    for (long q = blockIdx.x * 256L + threadIdx.x; q < n3; q += (long)gridDim.x * 256L)
        for (int c = 0; c < NCH; c++) {
            const double tgt = (yf[c * n3 + q] - yc[c * n3 + q]) / (double)sout[c];
            acc += tgt * tgt;
        }
    red[threadIdx.x] = acc;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}

// sum of nb partials in a fixed order into out[0] (one block)
__global__ void __launch_bounds__(256) sum_partials(const double *part, int nb, double *out)
{
    __shared__ double red[256];
    double acc = 0.0;
    for (int b = threadIdx.x; b < nb; b += 256) acc += part[b];
    red[threadIdx.x] = acc;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[0] = red[0];
}

// relative loss sum (Z - T)^2 / sum T^2 over sites and the 36 channels,
// T = (y_fine - y_coarse)/sout (each pair weighs the same whatever its
// amplitude); dZ = 2 (Z - T) / sum T^2 as bf16 (0 in the padding channels);
// block partial sums of (Z - T)^2
__global__ void __launch_bounds__(256) loss_grad(const float *Z, const double *yc, const double *yf, long n3,
                                                 const float *sout, const double *tsq, unsigned short *dZ, double *part)
{
    __shared__ double red[256];
    double acc = 0.0;
    const double scale = 2.0 / fmax(tsq[0], 1e-300);
    for (long q = blockIdx.x * 256L + threadIdx.x; q < n3; q += (long)gridDim.x * 256L) {
        unsigned short h[NC];
#pragma unroll
        for (int c = 0; c < NC; c++) {
            if (c < NCH) {
                const double tgt = (yf[c * n3 + q] - yc[c * n3 + q]) / (double)sout[c];
                const double d = (double)Z[q * NC + c] - tgt;
                acc += d * d;
                h[c] = f2bf((float)(scale * d));
            } else {
                h[c] = 0;
            }
        }
        uint4 *dst = reinterpret_cast<uint4 *>(dZ + q * NC);
#pragma unroll
        for (int p = 0; p < NC / 8; p++) {
            uint4 v;
            v.x = h[8 * p] | ((unsigned)h[8 * p + 1] << 16);
            v.y = h[8 * p + 2] | ((unsigned)h[8 * p + 3] << 16);
            v.z = h[8 * p + 4] | ((unsigned)h[8 * p + 5] << 16);
            v.w = h[8 * p + 6] | ((unsigned)h[8 * p + 7] << 16);
            dst[p] = v;
        }
    }
    red[threadIdx.x] = acc;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}

// y[c] += sout[c] * Z[site][c]
__global__ void __launch_bounds__(256) apply_corr(double *y, long n3, const float *Z, const float *sout)
{
    for (long q = blockIdx.x * 256L + threadIdx.x; q < n3; q += (long)gridDim.x * 256L)
#pragma unroll 4
        for (int c = 0; c < NCH; c++) y[c * n3 + q] += (double)sout[c] * (double)Z[q * NC + c];
}

// Adam on FP32 parameters
__global__ void __launch_bounds__(256) adam_kernel(float *p, const float *g, float *m, float *v, long n, float lr,
                                                   float b1, float b2, float eps, float c1, float c2)
{
    const long q = blockIdx.x * 256L + threadIdx.x;
    if (q >= n) return;
    const float gq = g[q];
    const float mq = b1 * m[q] + (1.0f - b1) * gq;
    const float vq = b2 * v[q] + (1.0f - b2) * gq * gq;
    m[q] = mq;
    v[q] = vq;
    p[q] -= lr * (mq / c1) / (sqrtf(vq / c2) + eps);
}

unsigned grid_of(long n, long per = 256)
{
    const long b = (n + per - 1) / per;
    return (unsigned)(b < 65536 ? b : 65536);
}

constexpr size_t CONV_LDS = (size_t)(HN * ROW + 2 * NC * ROW) * sizeof(unsigned short);

// NB: output blocks of 16 the caller needs (4: all 64 channels; 3: the last
// layer's 36 state channels, in 48); the 8x4x4-brick kernel computes all 64
bool conv2_fits(int nx, int ny, int nz) { return nx % C2X == 0 && ny % C2Y == 0 && nz % C2Z == 0; }

// EPI_APPLY (conv2 lattices only): y += sout * (last layer's output)
template <int EPI, int NB = 4>
hipError_t conv_launch(const unsigned short *X, const unsigned short *Wk, const float *bias, const float *zprev,
                       int nx, int ny, int nz, float *Z, unsigned short *A, hipStream_t st, double *Y = nullptr,
                       const float *sout = nullptr)
{
    static std::mutex attr_mu;   // several threads may launch the first convolution at once
    static bool attr = false;
    {
        std::lock_guard<std::mutex> g(attr_mu);
        if (!attr) {
            hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void *>(&conv_kernel<EPI == EPI_APPLY ? EPI_LAST : EPI>),
                                               hipFuncAttributeMaxDynamicSharedMemorySize, (int)CONV_LDS);
            if (e == hipSuccess)
                e = hipFuncSetAttribute(reinterpret_cast<const void *>(&conv2_kernel<EPI, NB>),
                                        hipFuncAttributeMaxDynamicSharedMemorySize, (int)CONV2_LDS);
            if (e != hipSuccess) return e;
            attr = true;
        }
    }
    if (conv2_fits(nx, ny, nz)) {
        const int nb = (nx / C2X) * (ny / C2Y) * (nz / C2Z);
        conv2_kernel<EPI, NB><<<nb, C2T, CONV2_LDS, st>>>(X, Wk, bias, zprev, nx, ny, nz, Z, A, Y, sout);
    } else if constexpr (EPI == EPI_APPLY) {
        return hipErrorInvalidValue;
    } else {
        const int nb = (nx / BRX) * (ny / BRY) * (nz / BRZ);
        conv_kernel<EPI><<<nb, 256, CONV_LDS, st>>>(X, Wk, bias, zprev, nx, ny, nz, Z, A);
    }
    return hipGetLastError();
}

}  // namespace

// ---------------------------------------------------------------- network object

constexpr long NW = (long)NTAP * NC * NC;        // weights per layer
constexpr long NPAR = 3 * (NW + NC);             // W1 b1 W2 b2 W3 b3

bool net_lattice_ok(const ewpt_lattice &L)
{
    return L.nx % BRX == 0 && L.ny % BRY == 0 && L.nz % BRZ == 0 && (long)L.nx * L.ny * L.nz * NC < (1L << 31);
}

struct NetBufs {
    long n3 = 0;
    unsigned short *X0 = nullptr, *A1 = nullptr, *A2 = nullptr, *dZa = nullptr, *dZb = nullptr;
    float *Z1 = nullptr, *Z2 = nullptr, *Z3 = nullptr;
    float *part = nullptr, *bpart = nullptr;
    double *lpart = nullptr, *tsq = nullptr;
    int nchunk = 0, bpc = 0;
};

struct Net {
    int device = 0;
    float *P = nullptr, *G = nullptr, *M = nullptr, *V = nullptr;   // params, grads, Adam moments (NPAR each)
    unsigned short *Wb = nullptr, *Wt = nullptr;                     // 3 layers
    float *sin_inv = nullptr, *sout = nullptr;                       // 36 each
    long adam_t = 0;
    int input_mode = 0;   // pack_state mode
    NetBufs buf;
    hipStream_t st = nullptr;
};

namespace {

float *Wl(Net *n, int l) { return n->P + l * (NW + NC); }
float *Bl(Net *n, int l) { return n->P + l * (NW + NC) + NW; }
float *dWl(Net *n, int l) { return n->G + l * (NW + NC); }
float *dBl(Net *n, int l) { return n->G + l * (NW + NC) + NW; }

void free_bufs(NetBufs &b)
{
    void *p[] = {b.X0, b.A1, b.A2, b.dZa, b.dZb, b.Z1, b.Z2, b.Z3, b.part, b.bpart, b.lpart, b.tsq};
    for (void *x : p)
        if (x) (void)hipFree(x);
    b = NetBufs{};
}

hipError_t ensure_bufs(Net *n, const ewpt_lattice &L)
{
    const long n3 = (long)L.nx * L.ny * L.nz;
    if (n->buf.n3 == n3) return hipSuccess;
    free_bufs(n->buf);
    NetBufs &b = n->buf;
    const long nbrick = n3 / BRN;
    b.bpc = (int)std::max(1L, (nbrick + 255) / 256);
    b.nchunk = (int)((nbrick + b.bpc - 1) / b.bpc);
    hipError_t e = hipSuccess;
    auto al = [&](auto **p, size_t bytes) {
        if (e == hipSuccess) e = hipMalloc(reinterpret_cast<void **>(p), bytes);
    };
    al(&b.X0, n3 * NC * 2); al(&b.A1, n3 * NC * 2); al(&b.A2, n3 * NC * 2);
    al(&b.dZa, n3 * NC * 2); al(&b.dZb, n3 * NC * 2);
    al(&b.Z1, n3 * NC * 4); al(&b.Z2, n3 * NC * 4); al(&b.Z3, n3 * NC * 4);
    al(&b.part, (size_t)b.nchunk * NW * 4); al(&b.bpart, (size_t)b.nchunk * NC * 4);
    al(&b.lpart, 65536 * sizeof(double));
    al(&b.tsq, sizeof(double));
    if (e != hipSuccess) { free_bufs(b); return e; }
    b.n3 = n3;
    return hipSuccess;
}

hipError_t refresh_weights(Net *n)
{
    for (int l = 0; l < 3; l++)
        cast_weights<<<grid_of(NW), 256, 0, n->st>>>(Wl(n, l), n->Wb + l * NW, n->Wt + l * NW);
    return hipGetLastError();
}

// forward of the three layers from X0; training keeps Z1, Z2.  With y set
// (and a lattice the brick kernel tiles) the last layer adds the correction
// to y instead of writing Z3.
hipError_t forward(Net *n, const ewpt_lattice &L, bool keep, double *y = nullptr)
{
    NetBufs &b = n->buf;
    hipError_t e = conv_launch<EPI_HIDDEN>(b.X0, n->Wb, Bl(n, 0), nullptr, L.nx, L.ny, L.nz, keep ? b.Z1 : nullptr,
                                           b.A1, n->st);
    if (e == hipSuccess)
        e = conv_launch<EPI_HIDDEN>(b.A1, n->Wb + NW, Bl(n, 1), nullptr, L.nx, L.ny, L.nz, keep ? b.Z2 : nullptr, b.A2,
                                    n->st);
    if (e == hipSuccess && y)
        return conv_launch<EPI_APPLY, 3>(b.A2, n->Wb + 2 * NW, Bl(n, 2), nullptr, L.nx, L.ny, L.nz, nullptr, nullptr,
                                         n->st, y, n->sout);
    if (e == hipSuccess)
        e = conv_launch<EPI_LAST, 3>(b.A2, n->Wb + 2 * NW, Bl(n, 2), nullptr, L.nx, L.ny, L.nz, b.Z3, nullptr, n->st);
    return e;
}

hipError_t wgrad(Net *n, const ewpt_lattice &L, const unsigned short *dZ, const unsigned short *X, int l)
{
    NetBufs &b = n->buf;
    wgrad_kernel<<<dim3(b.nchunk, NTAP), 256, 0, n->st>>>(dZ, X, L.nx, L.ny, L.nz, b.bpc, b.part, b.bpart);
    wgrad_reduce<<<grid_of(NW), 256, 0, n->st>>>(b.part, b.bpart, b.nchunk, dWl(n, l), dBl(n, l));
    return hipGetLastError();
}

}  // namespace

Net *net_create(int device, unsigned long long seed, const char **what)
{
    Net *n = new (std::nothrow) Net();
    if (!n) { *what = "net: out of host memory"; return nullptr; }
    n->device = device;
    hipError_t e = hipStreamCreateWithFlags(&n->st, hipStreamNonBlocking);
    auto al = [&](auto **p, size_t bytes) {
        if (e == hipSuccess) e = hipMalloc(reinterpret_cast<void **>(p), bytes);
    };
    al(&n->P, NPAR * 4); al(&n->G, NPAR * 4); al(&n->M, NPAR * 4); al(&n->V, NPAR * 4);
    al(&n->Wb, 3 * NW * 2); al(&n->Wt, 3 * NW * 2);
    al(&n->sin_inv, NCH * 4); al(&n->sout, NCH * 4);
    if (e != hipSuccess) { *what = "net: device allocation"; net_destroy(n); return nullptr; }
    // initial parameters: layers 1, 2 uniform(+-sqrt(6/fan_in)) on the real
    // input channels, layer 3 zero (the corrector starts as no correction)
    std::vector<float> h(NPAR, 0.0f);
    std::mt19937_64 rng(seed);
    for (int l = 0; l < 2; l++) {
        const int nin = l == 0 ? NCH : NC;
        const float a = std::sqrt(6.0f / (float)(NTAP * nin));
        std::uniform_real_distribution<float> U(-a, a);
        for (int tap = 0; tap < NTAP; tap++)
            for (int o = 0; o < NC; o++)
                for (int i = 0; i < nin; i++) h[l * (NW + NC) + ((long)tap * NC + o) * NC + i] = U(rng);
    }
    std::vector<float> ones(NCH, 1.0f);
    if (hipMemcpy(n->P, h.data(), NPAR * 4, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemset(n->M, 0, NPAR * 4) != hipSuccess || hipMemset(n->V, 0, NPAR * 4) != hipSuccess ||
        hipMemset(n->G, 0, NPAR * 4) != hipSuccess ||
        hipMemcpy(n->sin_inv, ones.data(), NCH * 4, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(n->sout, ones.data(), NCH * 4, hipMemcpyHostToDevice) != hipSuccess ||
        refresh_weights(n) != hipSuccess || hipStreamSynchronize(n->st) != hipSuccess) {
        *what = "net: initialisation";
        net_destroy(n);
        return nullptr;
    }
    return n;
}

void net_destroy(Net *n)
{
    if (!n) return;
    if (n->st) (void)hipStreamSynchronize(n->st);
    free_bufs(n->buf);
    void *p[] = {n->P, n->G, n->M, n->V, n->Wb, n->Wt, n->sin_inv, n->sout};
    for (void *x : p)
        if (x) (void)hipFree(x);
    if (n->st) (void)hipStreamDestroy(n->st);
    delete n;
}

long net_param_count() { return NPAR; }

hipError_t net_set_params(Net *n, const float *host)
{
    hipError_t e = hipMemcpyAsync(n->P, host, NPAR * 4, hipMemcpyHostToDevice, n->st);
    if (e == hipSuccess) e = refresh_weights(n);
    if (e == hipSuccess) e = hipStreamSynchronize(n->st);
    return e;
}

hipError_t net_get(Net *n, int which, float *host)
{
    hipError_t e = hipMemcpyAsync(host, which == 0 ? n->P : n->G, NPAR * 4, hipMemcpyDeviceToHost, n->st);
    return e == hipSuccess ? hipStreamSynchronize(n->st) : e;
}

void net_set_input_mode(Net *n, int mode) { n->input_mode = mode; }

hipError_t net_set_norm(Net *n, const float *in_inv, const float *out_scale)
{
    hipError_t e = hipMemcpyAsync(n->sin_inv, in_inv, NCH * 4, hipMemcpyHostToDevice, n->st);
    if (e == hipSuccess) e = hipMemcpyAsync(n->sout, out_scale, NCH * 4, hipMemcpyHostToDevice, n->st);
    return e == hipSuccess ? hipStreamSynchronize(n->st) : e;
}

// y += sout * net(y * in_inv), y an FP64 36-channel state on the device
hipError_t net_apply(Net *n, double *y, const ewpt_lattice &L)
{
    hipError_t e = ensure_bufs(n, L);
    const long n3 = n->buf.n3;
    if (e != hipSuccess) return e;
    pack_state<<<grid_of(n3), 256, 0, n->st>>>(y, n3, L.nx, L.ny, L.nz, n->input_mode, n->sin_inv, n->buf.X0);
    if (conv2_fits(L.nx, L.ny, L.nz)) {
        e = forward(n, L, false, y);   // the correction is added by the last layer
    } else {
        e = forward(n, L, false);
        if (e == hipSuccess) {
            apply_corr<<<grid_of(n3), 256, 0, n->st>>>(y, n3, n->buf.Z3, n->sout);
            e = hipGetLastError();
        }
    }
    return e == hipSuccess ? hipStreamSynchronize(n->st) : e;
}

// loss and gradients for the pair (coarse, fine); with lr > 0 also an Adam step
hipError_t net_train(Net *n, const double *yc, const double *yf, const ewpt_lattice &L, double lr, double *loss)
{
    hipError_t e = ensure_bufs(n, L);
    if (e != hipSuccess) return e;
    NetBufs &b = n->buf;
    const long n3 = b.n3;
    pack_state<<<grid_of(n3), 256, 0, n->st>>>(yc, n3, L.nx, L.ny, L.nz, n->input_mode, n->sin_inv, b.X0);
    if ((e = forward(n, L, true)) != hipSuccess) return e;
    const unsigned nlb = grid_of(n3);
    target_sq<<<nlb, 256, 0, n->st>>>(yc, yf, n3, n->sout, b.lpart);
    sum_partials<<<1, 256, 0, n->st>>>(b.lpart, (int)nlb, b.tsq);
    loss_grad<<<nlb, 256, 0, n->st>>>(b.Z3, yc, yf, n3, n->sout, b.tsq, b.dZa, b.lpart);
    // layer 3
    if ((e = wgrad(n, L, b.dZa, b.A2, 2)) != hipSuccess) return e;
    if ((e = conv_launch<EPI_DGRAD>(b.dZa, n->Wt + 2 * NW, nullptr, b.Z2, L.nx, L.ny, L.nz, nullptr, b.dZb, n->st)) !=
        hipSuccess)
        return e;
    // layer 2
    if ((e = wgrad(n, L, b.dZb, b.A1, 1)) != hipSuccess) return e;
    if ((e = conv_launch<EPI_DGRAD>(b.dZb, n->Wt + NW, nullptr, b.Z1, L.nx, L.ny, L.nz, nullptr, b.dZa, n->st)) !=
        hipSuccess)
        return e;
    // layer 1
    if ((e = wgrad(n, L, b.dZa, b.X0, 0)) != hipSuccess) return e;
    std::vector<double> lp(nlb);
    if ((e = hipMemcpyAsync(lp.data(), b.lpart, nlb * sizeof(double), hipMemcpyDeviceToHost, n->st)) != hipSuccess)
        return e;
    if (lr > 0.0) {
        n->adam_t++;
        const float b1 = 0.9f, b2 = 0.999f;
        const float c1 = 1.0f - std::pow(b1, (float)n->adam_t), c2 = 1.0f - std::pow(b2, (float)n->adam_t);
        adam_kernel<<<grid_of(NPAR, 256) , 256, 0, n->st>>>(n->P, n->G, n->M, n->V, NPAR, (float)lr, b1, b2, 1e-8f, c1, c2);
        if ((e = refresh_weights(n)) != hipSuccess) return e;
    }
    double tsq = 0.0;
    if ((e = hipMemcpyAsync(&tsq, b.tsq, sizeof(double), hipMemcpyDeviceToHost, n->st)) != hipSuccess) return e;
    if ((e = hipStreamSynchronize(n->st)) != hipSuccess) return e;
    double s = 0.0;
    for (double v : lp) s += v;
    *loss = s / std::max(tsq, 1e-300);
    return hipSuccess;
}

// one conv layer (or its input gradient) on host buffers, for tests
hipError_t net_conv_host(int nx, int ny, int nz, int epi, const unsigned short *X, const unsigned short *W,
                         const float *bias, const float *zprev, float *Zout, unsigned short *Aout)
{
    const ewpt_lattice L{nx, ny, nz, 1.0};
    if (!net_lattice_ok(L)) return hipErrorInvalidValue;
    const long n3 = (long)nx * ny * nz;
    unsigned short *dX = nullptr, *dW = nullptr, *dA = nullptr;
    float *db = nullptr, *dzp = nullptr, *dZ = nullptr;
    hipError_t e = hipMalloc(reinterpret_cast<void **>(&dX), n3 * NC * 2);
    if (e == hipSuccess) e = hipMalloc(reinterpret_cast<void **>(&dW), NW * 2);
    if (e == hipSuccess) e = hipMalloc(reinterpret_cast<void **>(&dA), n3 * NC * 2);
    if (e == hipSuccess) e = hipMalloc(reinterpret_cast<void **>(&db), NC * 4);
    if (e == hipSuccess) e = hipMalloc(reinterpret_cast<void **>(&dzp), n3 * NC * 4);
    if (e == hipSuccess) e = hipMalloc(reinterpret_cast<void **>(&dZ), n3 * NC * 4);
    if (e == hipSuccess) e = hipMemcpy(dX, X, n3 * NC * 2, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(dW, W, NW * 2, hipMemcpyHostToDevice);
    if (e == hipSuccess && bias) e = hipMemcpy(db, bias, NC * 4, hipMemcpyHostToDevice);
    if (e == hipSuccess && zprev) e = hipMemcpy(dzp, zprev, n3 * NC * 4, hipMemcpyHostToDevice);
    if (e == hipSuccess) {
        if (epi == EPI_HIDDEN) e = conv_launch<EPI_HIDDEN>(dX, dW, db, nullptr, nx, ny, nz, dZ, dA, nullptr);
        else if (epi == EPI_LAST) e = conv_launch<EPI_LAST>(dX, dW, db, nullptr, nx, ny, nz, dZ, nullptr, nullptr);
        else e = conv_launch<EPI_DGRAD>(dX, dW, nullptr, dzp, nx, ny, nz, nullptr, dA, nullptr);
    }
    if (e == hipSuccess) e = hipDeviceSynchronize();
    if (e == hipSuccess && Zout && epi != EPI_DGRAD) e = hipMemcpy(Zout, dZ, n3 * NC * 4, hipMemcpyDeviceToHost);
    if (e == hipSuccess && Aout && epi != EPI_LAST) e = hipMemcpy(Aout, dA, n3 * NC * 2, hipMemcpyDeviceToHost);
    void *p[] = {dX, dW, dA, db, dzp, dZ};
    for (void *x : p)
        if (x) (void)hipFree(x);
    return e;
}

// weight gradient of one layer on host buffers, for tests
hipError_t net_wgrad_host(int nx, int ny, int nz, const unsigned short *dZ, const unsigned short *X, float *dW,
                          float *db)
{
    const ewpt_lattice L{nx, ny, nz, 1.0};
    if (!net_lattice_ok(L)) return hipErrorInvalidValue;
    Net n;
    NetBufs &b = n.buf;
    hipError_t e = hipStreamCreateWithFlags(&n.st, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipMalloc(reinterpret_cast<void **>(&n.G), NPAR * 4);
    if (e == hipSuccess) e = ensure_bufs(&n, L);
    const long n3 = (long)nx * ny * nz;
    if (e == hipSuccess) e = hipMemcpy(b.dZa, dZ, n3 * NC * 2, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(b.X0, X, n3 * NC * 2, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = wgrad(&n, L, b.dZa, b.X0, 0);
    if (e == hipSuccess) e = hipStreamSynchronize(n.st);
    if (e == hipSuccess) e = hipMemcpy(dW, n.G, NW * 4, hipMemcpyDeviceToHost);
    if (e == hipSuccess) e = hipMemcpy(db, n.G + NW, NC * 4, hipMemcpyDeviceToHost);
    free_bufs(b);
    if (n.G) (void)hipFree(n.G);
    if (n.st) (void)hipStreamDestroy(n.st);
    n.G = nullptr;
    n.st = nullptr;
    return e;
}

}  // namespace ewpt
