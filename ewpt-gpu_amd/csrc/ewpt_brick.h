// ewpt_brick.h -- fused RK4 stage kernel, LDS-tiled (template; instantiated
// by the ewpt_brick_v*.hip files, one variant per file).
//
// One workgroup owns a TX*TY*TZ brick of sites, one site per thread.  The 32
// stencilled channels of the stage input are streamed one per pass through a
// double-buffered LDS tile holding the brick plus the star-shaped halo the
// radius-3 stencils of diff_scheme.jl:157-311 touch (the six 3-deep face
// slabs; edges and corners are never read).  The halo of pass p+PF is in
// flight in registers while pass p is computed.  The stage input y + c*k is
// formed when a value is staged (ewpt_stage.h), so the reference's broadcast
// temporaries (main_rk4.jl:6226-6242 ...) never exist.  Periodic wrap is done
// on the global index; there is no halo exchange.
//
// The right-hand side of rk4_step! (main_rk4.jl:337-657) is linear in the
// spatial derivatives once the 36 centre values are known.  Each channel's
// derivatives are therefore folded into 20 output accumulators (plus the
// magnetic-field accumulators B^a, B^Y and u_j = phi.d_j phi) as soon as they
// are computed, instead of holding all 80 derivative values.  The formulas
// are those of ewpt_point.h, distributed over the derivative terms.  The
// passes run E, Gamma/Sigma, phi, Y, W so that centre values die early.
// BETA = false drops the beta_W/beta_Y terms (the reference's default runs
// have beta = 0), which removes u, B^Y and the E centre values from the
// registers after their own passes.
//
// Stage S=4 also finishes the step: y' = k4*(dt/6) + n with
// n = ((k1*(dt/6) + y) + k2*(dt/3)) + k3*(dt/3) (main_rk4.jl:6400-6437).  The
// field channels of y' and the partial sums n of channels 16..35 are computed
// first (n is parked in LDS until k4 is known).
#pragma once

#include <mutex>
#include <type_traits>
#include <utility>

#include "ewpt_internal.h"

namespace ewpt {

namespace {

template <int TX, int TY, int TZ>
struct BrickShape {
    static constexpr int NT = TX * TY * TZ;
    static constexpr int HX = TX + 6, HY = TY + 6, HZ = TZ + 6;
    static constexpr int BOX = HX * HY * HZ;   // LDS values of the halo'd brick
    static constexpr int BUF = BOX + 64;        // + one trash slot per lane for empty halo slots
    static constexpr int NXF = 6 * TY * TZ;    // x-face halo elements
    static constexpr int NYF = 6 * TX * TZ;
    static constexpr int NZF = 6 * TX * TY;
};

struct BrickGeo {
    int nx, ny, nz;
    int nbx, nby, nbz;      // bricks per axis
    int xcd_map;            // 1: remap block ids so each XCD walks a compact region
    int order;              // brick enumeration, see brick_coords
    double inv60dx, invdx2;
    int zg = 0;             // ghost planes below/above in z (StageBufs::zghost); 0 = periodic in z
};

constexpr int NPASS = 32;

// Pass schedule.  kind: 0 = E^a_i / E^Y_i (one direction), 1 = Gamma^a /
// Sigma (gradient), 2 = phi_b, 3 = Y_i, 4 = W^a_i (gradient + Laplacian).
__host__ __device__ constexpr int pass_kind(int p)
{
    return p < 12 ? 0 : p < 16 ? 1 : p < 20 ? 2 : p < 23 ? 3 : 4;
}
__host__ __device__ constexpr int pass_channel(int p)
{
    return p < 12 ? ((p % 4) < 3 ? CH_E + 3 * (p % 4) + p / 4 : CH_EY + p / 4)   // E^a_i, i = p/4
         : p < 16 ? CH_GAM + (p - 12)                                             // Gamma^1..3, Sigma
         : p < 20 ? CH_PHI + (p - 16)                                             // phi_1..4
         : p < 23 ? CH_Y + (p - 20)                                               // Y_x, Y_y, Y_z
         : CH_W + 3 * ((p - 23) % 3) + (p - 23) / 3;                              // W^a_i, component-major
}
// halo needed: 0 = full star, 1/2/3 = x/y/z faces only
__host__ __device__ constexpr int pass_dir(int p) { return p < 12 ? 1 + p / 4 : 0; }

__device__ __forceinline__ int wrapc(int v, int n) { return v < 0 ? v + n : (v >= n ? v - n : v); }

// Brick coordinates of enumeration index i.  Blocks with nearby indices run
// at nearby times, so the order decides which halo re-reads find their data
// still in L2 / the memory-side cache:
//   0: x fastest, then y, then z
//   1: x fastest, then z, then y
//   2: x fastest, then (y,z) in 4x4 tiles of brick rows (z fastest inside)
//   3: cubes of 2x4x4 bricks (x,y,z), 32 consecutive indices per cube: the
//      workgroups one XCD runs at the same time share most of their halos
//   4: cubes of 4x4x4 bricks (64 per cube), for two workgroups per CU
//   5: 2x8x4, 6: 2x4x8, 7: 1x8x8, 8: 2x8x8, 9: 1x4x4 bricks per cube
// Orders 3-9 fall back to 0 where the brick counts do not divide.
__host__ __device__ constexpr int cube_bricks(int order, int axis)
{
    // x, y, z extents packed one hex digit each, orders 3..9
    constexpr int kCube[7] = {0x244, 0x444, 0x284, 0x248, 0x188, 0x288, 0x144};
    return (kCube[order - 3] >> (4 * (2 - axis))) & 15;
}

__device__ __forceinline__ void brick_coords(const BrickGeo &g, int i, int &bx, int &by, int &bz)
{
    if (g.order >= 3) {
        const int cx = cube_bricks(g.order, 0), cy = cube_bricks(g.order, 1), cz = cube_bricks(g.order, 2);
        const int cn = cx * cy * cz;
        if (g.nbx % cx == 0 && g.nby % cy == 0 && g.nbz % cz == 0) {
            const int cube = i / cn, in = i % cn;
            const int ncx = g.nbx / cx, ncy = g.nby / cy;
            bx = (cube % ncx) * cx + in % cx;
            by = ((cube / ncx) % ncy) * cy + (in / cx) % cy;
            bz = (cube / (ncx * ncy)) * cz + in / (cx * cy);
            return;
        }
        bx = i % g.nbx;
        by = (i / g.nbx) % g.nby;
        bz = i / (g.nbx * g.nby);
        return;
    }
    bx = i % g.nbx;
    const int r = i / g.nbx;   // brick-row index over (y, z)
    if (g.order == 1) {
        bz = r % g.nbz;
        by = r / g.nbz;
    } else if (g.order == 2 && g.nby % 4 == 0 && g.nbz % 4 == 0) {
        const int tile = r / 16, in = r % 16;
        const int tz = tile % (g.nbz / 4), ty = tile / (g.nbz / 4);
        bz = tz * 4 + in % 4;
        by = ty * 4 + in / 4;
    } else {
        by = r % g.nby;
        bz = r / g.nby;
    }
}

// box coordinates of halo element f of the faces selected by dir
template <int TX, int TY, int TZ>
__device__ __forceinline__ void face_elem(int dir, int f, int &lx, int &ly, int &lz)
{
    using B = BrickShape<TX, TY, TZ>;
    int d = dir;
    if (dir == 0) {
        if (f < B::NXF) d = 1;
        else if (f < B::NXF + B::NYF) { d = 2; f -= B::NXF; }
        else { d = 3; f -= B::NXF + B::NYF; }
    }
    if (d == 1) {
        const int h = f % 6, r = f / 6;
        lx = h < 3 ? h : TX + h;
        ly = 3 + r % TY;
        lz = 3 + r / TY;
    } else if (d == 2) {
        const int h = (f / TX) % 6;
        lx = 3 + f % TX;
        ly = h < 3 ? h : TY + h;
        lz = 3 + f / (TX * 6);
    } else {
        const int h = f / (TX * TY);
        lx = 3 + f % TX;
        ly = 3 + (f / TX) % TY;
        lz = h < 3 ? h : TZ + h;
    }
}

template <int TX, int TY, int TZ>
__host__ __device__ constexpr int face_count(int dir)
{
    using B = BrickShape<TX, TY, TZ>;
    return dir == 0 ? B::NXF + B::NYF + B::NZF : dir == 1 ? B::NXF : dir == 2 ? B::NYF : B::NZF;
}

// column b of A_i = h(W1 T1 + W2 T2 + W3 T3) + hy Y T_Y applied to val, added
// to acc (rows read off apply_gen in ewpt_point.h)
template <int b, typename R>
__device__ __forceinline__ void add_gen_col(R h, R W1, R W2, R W3, R hy, R Yv, R val, R acc[4])
{
    if constexpr (b == 0) {
        acc[1] -= (h * W3 + hy * Yv) * val; acc[2] += h * W2 * val; acc[3] -= h * W1 * val;
    } else if constexpr (b == 1) {
        acc[0] += (h * W3 + hy * Yv) * val; acc[2] += h * W1 * val; acc[3] += h * W2 * val;
    } else if constexpr (b == 2) {
        acc[0] -= h * W2 * val; acc[1] -= h * W1 * val; acc[3] += (h * W3 - hy * Yv) * val;
    } else {
        acc[0] += h * W1 * val; acc[1] -= h * W2 * val; acc[2] += (hy * Yv - h * W3) * val;
    }
}

// coefficient of v_b in phi.T_a v (a = 0,1,2) and phi.T_Y v
template <int a, int b, typename R>
__device__ __forceinline__ R phiT_coef(const R p[4])
{
    constexpr int tab[3][4][2] = {{{3, -1}, {2, 1}, {1, -1}, {0, 1}},
                                  {{2, 1}, {3, 1}, {0, -1}, {1, -1}},
                                  {{1, -1}, {0, 1}, {3, 1}, {2, -1}}};
    return tab[a][b][1] > 0 ? p[tab[a][b][0]] : -p[tab[a][b][0]];
}
template <int b, typename R>
__device__ __forceinline__ R phiTY_coef(const R p[4])
{
    constexpr int tab[4][2] = {{1, -1}, {0, 1}, {3, -1}, {2, 1}};
    return tab[b][1] > 0 ? p[tab[b][0]] : -p[tab[b][0]];
}

// Load/store at a 32-bit byte offset from a (uniform) channel base pointer:
// lets the compiler use the SGPR-base + VGPR-offset addressing form instead
// of materialising one 64-bit address per channel.
// The accesses are made through global-address-space pointers: the batched
// kernels take their buffer pointers from a descriptor in memory, which the
// compiler would otherwise address as generic (FLAT) memory, and FLAT loads
// also count against the LDS wait of lds_barrier, so every barrier would wait
// for the prefetched halo loads.
template <typename T>
using global_ptr = __attribute__((address_space(1))) T *;
template <typename T>
__device__ __forceinline__ T ldo(const T *base, unsigned boff)
{
    return *(global_ptr<const T>)(reinterpret_cast<const char *>(base) + boff);
}
template <typename T>
__device__ __forceinline__ void sto(T *base, unsigned boff, T v)
{
    *(global_ptr<T>)(reinterpret_cast<char *>(base) + boff) = v;
}
// Streaming (non-temporal) store: no launch re-reads what it writes, so
// its lines are evicted first and leave L2 to the halo lines that
// neighbouring bricks re-read (streaming loads of the centres measured
// slower).
template <typename T>
__device__ __forceinline__ void stn(T *base, unsigned boff, T v)
{
    __builtin_nontemporal_store(v, (global_ptr<T>)(reinterpret_cast<char *>(base) + boff));
}
// Streaming load of a value no other brick reads, so that it does not push
// the halo lines neighbours re-read out of L2: stage 4's update inputs
// (512^3: FP64 stage 4 33.4 -> 32.7 ms together with a prefetch depth of 3,
// mixed 25.8 -> 25.6 ms) and the y / k1 values of the next stage input that
// stages 2-3 load during their last passes (FP64 stages 2 + 3 -0.5 ms, mixed
// -0.7 ms).
template <typename T>
__device__ __forceinline__ T ldn(const T *base, unsigned boff)
{
    return __builtin_nontemporal_load((global_ptr<const T>)(reinterpret_cast<const char *>(base) + boff));
}

// Workgroup barrier for LDS only.  __syncthreads() also fences global
// memory, which makes every in-flight global load (the prefetched halo of
// the next passes) complete before the barrier; here only the LDS accesses
// are waited for.  The "memory" clobber keeps the compiler from moving
// memory accesses across it.
__device__ __forceinline__ void lds_barrier() { asm volatile("s_waitcnt lgkmcnt(0)\n\ts_barrier" ::: "memory"); }

// Opaque use/def of a register value: keeps the compiler from sinking the
// arithmetic that produced it past the next barrier (which would keep every
// staged LDS value alive across passes).
__device__ __forceinline__ void pin(double &x) { asm volatile("" : "+v"(x)); }
__device__ __forceinline__ void pin(float &x) { asm volatile("" : "+v"(x)); }

// Knuth's TwoSum: s + e == a + b exactly (round-to-nearest, no reassociation)
template <typename R>
__device__ __forceinline__ void two_sum(R a, R b, R &s, R &e)
{
    s = a + b;
    const R bb = s - a;
    e = (a - (s - bb)) + (b - bb);
}

// sign of eps^{a b c} for the third index c
__host__ __device__ constexpr int eps3(int a, int b) { return ((b - a + 3) % 3 == 1) ? 1 : -1; }

// In-flight loads of one pass: the halo values (up to 3 arrays per element).
template <typename T, int MAXF>
struct PassRegs {
    T v[MAXF][3];
    T vc[3];    // the site's own value, for passes that load their centre
};

// Prefetch depth of stage S from the profile PF: a single digit applies to
// all stages, four digits d1 d2 d3 d4 give stages 1..4 (deeper where a
// halo element is one array, shallower where it is three).
//
// A fifth digit 1 (PF >= 10000) pairs the passes of stages 1-3: two channels
// are staged per LDS barrier (four staging buffers), halving the barriers;
// stage 4 stays single (its LDS also parks the partial sums).
__host__ __device__ constexpr int pf_of_stage(int PF, int S)
{
    return PF >= 100000 ? pf_of_stage(PF % 100000, S)
         : PF >= 10000 ? pf_of_stage(PF % 10000, S)
         : PF < 10 ? PF : (PF / (S == 1 ? 1000 : S == 2 ? 100 : S == 3 ? 10 : 1)) % 10;
}
__host__ __device__ constexpr bool paired_of(int PF, int S) { return PF % 100000 >= 10000 && S < 4; }
// PF / 100000 = m: stored stage inputs.  Bit s-2 of m set: stage s (2..4)
// reads its input u_s as one array (every staged value, centre or halo, is a
// single load instead of two or three), and stage s-1 writes it (36 channels:
// y + c k_{s-1}, the fields' k being the stage momenta).
__host__ __device__ constexpr bool stored_in(int PF, int S) { return S >= 2 && (((PF / 100000) % 10) >> (S - 2)) & 1; }
// PF / 1000000 = 1: late centres.  The E and Gamma/Sigma centres (state
// channels 16..19, 24..35) are not loaded up front but with their own pass,
// so a brick reads its interior of such a channel while its neighbours read
// the same lines as their halo; the terms that need these centres earlier
// (W x E in k_Gamma, W x Gamma in kt_W, the Gamma/Sigma generator in kt_phi)
// are added once their passes are done.
__host__ __device__ constexpr bool late_centres(int PF) { return (PF / 1000000) % 10 == 1; }
__host__ __device__ constexpr bool is_late_channel(int ch) { return (ch >= 16 && ch < 20) || ch >= 24; }
// PF / 10000000 = 1: stages 2-3 that write the next stage input load the y
// values it needs during their last passes instead of after them.  = 2: the
// same, and the 16 field channels of that input (which need no derivative
// of this stage) are stored as soon as their y arrives, so only the other 20
// y values are held in registers to the end (fewer live registers: the FP32
// kernels at four waves per SIMD stop spilling; the FP64 kernels, which do
// not spill, measured slower with the stores during the passes).
__host__ __device__ constexpr bool y_early(int PF) { return (PF / 10000000) % 10 == 1 || (PF / 10000000) % 10 == 2; }
__host__ __device__ constexpr bool early_fields(int PF) { return (PF / 10000000) % 10 == 2; }
// Stage 1 of an early_fields variant also parks the 20 centre values of
// channels 16..35 (E, Gamma/Sigma, pi) in LDS once their passes have staged
// them: after that only the epilogue (the next stage input y + c k_1) reads
// them, so their registers are free for the phi, Y and W passes.
__host__ __device__ constexpr bool park_centres(int PF, int S)
{
    return early_fields(PF) && S == 1 && stored_in(PF, 2) && !late_centres(PF);
}
// PF / 100000000 = 1 or 2: stage 4 loads y, k1..k3 for the partial sums of
// channels 16..35 (mixed: the k's and the compensation word for the
// increments) one channel per pass during its first 20 passes instead of all
// before the first pass (2: fused_fields as well).
__host__ __device__ constexpr bool spread_update(int PF)
{
    return (PF / 100000000) % 10 >= 1 && (PF / 100000000) % 10 <= 3;
}
__host__ __device__ constexpr bool stores_next(int PF, int S) { return S < 4 && stored_in(PF, S + 1); }
// PF / 1000000000 = 1 (mixed): stage 4 parks only the increment in LDS and
// reads the state's high word again for the final TwoSum, which halves its
// LDS and lets two workgroups share a CU (four waves per SIMD).
__host__ __device__ constexpr bool lean_stash(int PF) { return (PF / 1000000000) % 10 == 1; }
// PF / 100000000 = 2: spread_update as 1, and stage 4 updates the
// field channels of y' in the same passes (field channel ch with the partial
// sum of channel ch + 20, its momentum) from the same loads of the momentum
// y and k's: a separate field loop at the start loaded those 64 channels a
// second time, after L2 had evicted them.
__host__ __device__ constexpr bool fused_fields(int PF) { return (PF / 100000000) % 10 == 2 || (PF / 100000000) % 10 == 3; }
// PF / 100000000 = 3 (FP64, stored stage inputs): as 2, and y is never read
// after stage 1: it is reconstructed from the stage input and the stored
// derivatives.  Stage 4: y = u4 - k3 dt (momenta, Gamma, Sigma) and
// y_f = u4_f - k3_f dt with k3_f = y_pi + kt2 dt/2 (fields), from the u4 centres
// and the k loads of the update; its partial sums run in the order the
// passes stage the centres (E, Gamma/Sigma, then pi).  Stage 2 forms u3 from
// u2 and k1 (20 channels) instead of y (36).  Each reconstruction rounds once
// more than the reference's y + c k (about one unit in the last place).
__host__ __device__ constexpr bool recon_y(int PF) { return (PF / 100000000) % 10 == 3; }
// Channel 16+m whose stage-4 update the spread passes form in pass q: with
// recon_y the order in which the passes stage the centres y is rebuilt from
// (E, Gamma/Sigma, then pi), else m = q.
__host__ __device__ constexpr int upd_channel(int PF, int q)
{
    return !recon_y(PF) ? q : q < 16 ? pass_channel(q) - 16 : 4 + (q - 16);
}
__host__ __device__ constexpr int waves_per_simd(bool f32, int PF, int S) { return f32 && (S < 4 || lean_stash(PF)) ? 4 : 2; }

// T: storage and arithmetic type (double, or float for the mixed path).
// PF: prefetch profile (pf_of_stage).  Two waves per SIMD are requested for
// FP64 (256 VGPRs, also for 256-thread bricks at two workgroups per CU) and
// four for FP32 stages 1-3 (two 512-thread workgroups per CU); the FP32
// stage 4 parks more in LDS and runs one workgroup per CU.
//
// brick_body computes brick `bid` (enumeration index, brick_coords) of the
// lattice described by b and g, using the dynamic LDS at lds_raw; t is the
// thread's index in the workgroup.  It is called by the single-lattice
// kernel stage_brick and by the batched kernels stage_brick_batch.
// Slots of the diagnostics a stage-1 brick can reduce on the way (DIAG):
// the energy densities of the reference's compute_* kernels
// (main_rk4.jl:1334-2301), the Gauss-law residuals of the k_Gamma /
// k_Sigma charges (main_rk4.jl:603-657) and the electromagnetic densities
// of the magnetic field B, its B_2 form and B_2.A (the EM columns of the
// energy table, main_rk4.jl:2303-2822, 3718-4250; em_point in ewpt_point.h),
// in the layout of the standalone reductions: 0 PE, 1 KE, 2 GE, 3 EEW,
// 4 MEW, 5 EEY, 6 MEY, 7 sum|phi|, 8 min|phi|, 9 non-finite count,
// 10-12 (G^a)^2, 13 (G^Y)^2, 14-16 (C^a)^2, 17 (C^Y)^2, 18 |B|^2,
// 19 |B_2|^2, 20 B_2.A.
constexpr int NDIAG = 21;

__device__ __forceinline__ double wave_reduce(double v, bool mn)
{
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        const double o = __shfl_xor(v, off, 64);
        v = mn ? fmin(v, o) : v + o;
    }
    return v;
}

// DIAG: 0 = a stage, 1 = stage 1 that also reduces the diagnostics of its
// input, 2 = those diagnostics alone (no stage outputs written).
template <int TX, int TY, int TZ, int PF, bool BETA, int S, typename T, int DIAG = 0>
__device__ __forceinline__ void brick_body(const StageBufs<T> &b, const BrickGeo &g, const Phys &P, T dt, int bid,
                                           unsigned char *lds_raw, int t, double *diag = nullptr, double sw = 0.0,
                                           double cw = 0.0)
{
    // T = float is the mixed-precision path: y is stored as an FP32 pair
    // (b.y = high word, b.ylo = compensation word), stages read the high
    // word only, and stage 4 adds the step increment with a compensated sum:
    // t = lo + dy, then (hi', lo') = TwoSum(hi, t) (Knuth's error-free sum).
    constexpr bool MIXED = std::is_same_v<T, float>;
    using B = BrickShape<TX, TY, TZ>;
    using R = T;
    // dynamic LDS: two staging buffers, then (stage 4) the 20 partial sums n
    R *lds_base = reinterpret_cast<R *>(lds_raw);
    R *lds[4] = {lds_base, lds_base + B::BUF, lds_base + 2 * B::BUF, lds_base + 3 * B::BUF};   // [2..3]: paired
    R *nstash = lds_base + 2 * B::BUF;
    R *park = lds_base + 2 * B::BUF;   // stage 1 with park_centres: [NK][NT]
    constexpr bool PARK = park_centres(PF, S);

    // ---- brick of this workgroup
    int bx, by, bz;
    brick_coords(g, bid, bx, by, bz);
    const int x0 = bx * TX, y0 = by * TY, z0 = bz * TZ;
    const int tx = t % TX, ty = (t / TX) % TY, tz = t / (TX * TY);
    const long nxl = g.nx, nxy = (long)g.nx * g.ny;
    const unsigned qb = (unsigned)((x0 + tx) + nxl * (y0 + ty) + nxy * (z0 + tz + g.zg)) * (unsigned)sizeof(T);
    auto Y = [&](int ch) { return b.y + (long)ch * b.n3; };
    auto K = [&](const T *k, int m) { return k + (long)m * b.n3; };
    const R half_dt = dt * R(0.5), d6 = dt / R(6), d3 = dt / R(3);

    // ---- stage 4: the field channels of y' and the partial sums n of
    // channels 16..35 (parked in LDS until k4 is known) need only centre
    // values; they are done first, in two batches of loads (SPREAD: the
    // partial sums one per pass during the first NK passes instead).
    constexpr bool SPREAD = S == 4 && spread_update(PF);
    static_assert(S < 4 || !fused_fields(PF) || SPREAD, "the field update is fused into the spread passes");
    if constexpr (S == 4) {
        if constexpr (!SPREAD) {
#pragma unroll
            for (int m = 0; m < NK; m++) {
                if constexpr (MIXED) {
                    // increment without k4, plus the compensation word
                    R dy = ldo(K(b.k1, m), qb) * d6;
                    dy = ldo(K(b.k2, m), qb) * d3 + dy;
                    dy = ldo(K(b.k3, m), qb) * d3 + dy;
                    nstash[m * B::NT + t] = ldo(b.ylo + (long)(16 + m) * b.n3, qb) + dy;
                    if constexpr (!lean_stash(PF)) nstash[(NK + m) * B::NT + t] = ldo(Y(16 + m), qb);
                } else {
                    R n = ldo(K(b.k1, m), qb) * d6 + ldo(Y(16 + m), qb);
                    n = ldo(K(b.k2, m), qb) * d3 + n;
                    nstash[m * B::NT + t] = ldo(K(b.k3, m), qb) * d3 + n;
                }
            }
            __builtin_amdgcn_sched_barrier(0);
        }
#pragma unroll
        for (int ch = 0; ch < (fused_fields(PF) ? 0 : 16); ch++) {
            const R piv = ldo(Y(ch + 20), qb);
            const R k2 = piv + ldo(K(b.k1, ch + 4), qb) * half_dt;
            const R k3 = piv + ldo(K(b.k2, ch + 4), qb) * half_dt;
            const R k4 = piv + ldo(K(b.k3, ch + 4), qb) * dt;
            if constexpr (MIXED) {
                R dy = piv * d6;
                dy = k2 * d3 + dy;
                dy = k3 * d3 + dy;
                dy = k4 * d6 + dy;
                R nh, nl;
                two_sum(ldo(Y(ch), qb), ldo(b.ylo + (long)ch * b.n3, qb) + dy, nh, nl);
                stn(b.ynew + (long)ch * b.n3, qb, nh);
                stn(b.ylo_new + (long)ch * b.n3, qb, nl);
            } else {
                R n = piv * d6 + ldo(Y(ch), qb);
                n = k2 * d3 + n;
                n = k3 * d3 + n;
                stn(b.ynew + (long)ch * b.n3, qb, k4 * d6 + n);
            }
        }
        __builtin_amdgcn_sched_barrier(0);
    }

    // ---- centre values of the stage input (ewpt_stage.h)
    constexpr bool US = stored_in(PF, S);
    auto U = [&](int ch) { return b.u + (long)ch * b.n3; };
    R c[NCH];
#pragma unroll
    for (int ch = 0; ch < NCH; ch++) {
        if (late_centres(PF) && is_late_channel(ch)) {
            c[ch] = R(0);   // set when the channel's pass is staged
            continue;
        }
        const R yv = ldo(US ? U(ch) : Y(ch), qb);
        if constexpr (S == 1 || US) {
            c[ch] = yv;
        } else if (ch < 16) {
            const R piv = ldo(Y(ch + 20), qb);
            if constexpr (S == 2) c[ch] = yv + piv * half_dt;
            if constexpr (S == 3) c[ch] = yv + (piv + ldo(K(b.k1, ch + 4), qb) * half_dt) * half_dt;
            if constexpr (S == 4) c[ch] = yv + (piv + ldo(K(b.k2, ch + 4), qb) * half_dt) * dt;
        } else {
            const T *kp = (S == 2) ? b.k1 : (S == 3) ? b.k2 : b.k3;
            c[ch] = yv + ldo(K(kp, ch - 16), qb) * (S == 4 ? dt : half_dt);
        }
    }

    const R gw = R(P.gw), gy = R(P.gy), gp2 = R(P.gp2), h = R(0.5) * gw, hy = R(0.5) * gy;
    const R bw = R(P.bw), byc = R(P.by);
    const R p[4] = {c[0], c[1], c[2], c[3]};
    auto W = [&](int a, int i) -> R { return c[CH_W + 3 * a + i]; };
    auto E = [&](int a, int i) -> R { return c[CH_E + 3 * a + i]; };
    const R r2 = p[0] * p[0] + p[1] * p[1] + p[2] * p[2] + p[3] * p[3];
    const R phipi = c[CH_PI] * p[0] + c[CH_PI + 1] * p[1] + c[CH_PI + 2] * p[2] + c[CH_PI + 3] * p[3];

    // ---- accumulators
    R ktphi[4], ktW[3][3], ktY[3], kG[3], kS, u[3] = {R(0), R(0), R(0)};
    R Bw[3][3], BYv[3] = {R(0), R(0), R(0)};
    {
        const R pot = R(2) * R(P.lam) * (r2 - R(P.vev) * R(P.vev)) + R(0.5) * R(P.jex);
        R Gp[4] = {R(0), R(0), R(0), R(0)};
        if constexpr (!late_centres(PF)) apply_gen(h, c[CH_GAM], c[CH_GAM + 1], c[CH_GAM + 2], hy, c[CH_SIG], p, Gp);
        const R inv_r2 = R(1) / r2;   // keeps the reference's 0/0 at |phi| = 0
#pragma unroll
        for (int bb = 0; bb < 4; bb++) {
            ktphi[bb] = -pot * p[bb] - R(P.gam) * p[bb] * (p[bb] * c[CH_PI + bb]) * inv_r2;
            if constexpr (!late_centres(PF)) ktphi[bb] = Gp[bb] + ktphi[bb];
        }
        const R pi[4] = {c[CH_PI], c[CH_PI + 1], c[CH_PI + 2], c[CH_PI + 3]};
#pragma unroll
        for (int a = 0; a < 3; a++) {
            if constexpr (late_centres(PF)) {
                kG[a] = gp2 * gw * phiT(a, p, pi);
#pragma unroll
                for (int i = 0; i < 3; i++) ktW[a][i] = R(0);
            } else {
                R na = R(0);
#pragma unroll
                for (int i = 0; i < 3; i++) na -= xprod(a, W(0, i), W(1, i), W(2, i), E(0, i), E(1, i), E(2, i));
                kG[a] = gp2 * gw * (na + phiT(a, p, pi));
#pragma unroll
                for (int i = 0; i < 3; i++)
                    ktW[a][i] = -gw * xprod(a, W(0, i), W(1, i), W(2, i), c[CH_GAM], c[CH_GAM + 1], c[CH_GAM + 2]);
            }
        }
        kS = gp2 * gy * phiTY(p, pi);
#pragma unroll
        for (int i = 0; i < 3; i++) ktY[i] = R(0);
    }

    if constexpr (PARK) {
#pragma unroll
        for (int ch = CH_PI; ch < CH_PI + 4; ch++) park[(ch - 16) * B::NT + t] = c[ch];   // never staged
    }

    // ---- diagnostics (DIAG, stage 1).  Every per-site value that needs the
    // derivatives is parked in the thread's own LDS slots dgl[k][t] the moment
    // it is known (the derivative sums div E^a, div E^Y, |D phi|^2, div Y,
    // div W^a accumulate there beside the right-hand side's); the densities of
    // centre values alone (PE, KE, |phi|, EEW, EEY) are formed after the
    // passes from the centres the next stage input needs anyway.  After the
    // passes each slot is wave-reduced (dgw[wave][slot]) and the waves
    // combined in a fixed order.  Per-thread slots k:
    //   0-3 div E^a, div E^Y, then (G^a)^2, (G^Y)^2   4 |D phi|^2
    //   5 div Y, then (C^Y)^2   6-8 div W^a, then (C^a)^2   9 MEY  10 MEW
    //   11-13 D_i phi_1 (pass 16), then the cross products X_i of the B field
    //         (passes 17, 19), then |B|^2, |B_2|^2, B_2.A (pass 31)
    //   14-16 D_i phi_3 (pass 18), then B^Y_i (pass 22)
    // sw, cw: sin and cos of the Weinberg angle of the EM densities.
    constexpr bool DG = DIAG && S == 1;
    static_assert(!DIAG || (S == 1 && !PARK), "diagnostics: stage 1 without parked centres");
    // the FP32 (mixed) form computes each site's densities in FP64 from the
    // FP32 centre values and pass derivatives (|phi|^2 - v^2 without FP32
    // cancellation); the FP64 form keeps its own arithmetic
    auto dd = [](R v) { return (double)v; };
    constexpr int NKD = 17;
    double *dgw = reinterpret_cast<double *>(lds_raw + 2 * B::BUF * sizeof(T));
    double *dgl = dgw + (B::NT / 64) * NDIAG;
    auto dput = [&](int k, double v) { dgl[k * B::NT + t] = v; };
    auto dacc = [&](int k, double v) { dgl[k * B::NT + t] += v; };
    auto dget = [&](int k) { return dgl[k * B::NT + t]; };
    if constexpr (DG) {
#pragma unroll
        for (int k = 0; k < 9; k++) dput(k, 0.0);
    }

    // ---- staging of one channel into LDS: centre from registers, faces from
    // global.  Per face class (0 = star, 1/2/3 = x/y/z faces) each thread owns
    // up to MAXF halo elements; their in-channel byte offsets and LDS offsets
    // are computed once.  An empty slot loads element 0 and writes it to the
    // lane's trash slot past B::BOX, so the pipeline has no branches.
    constexpr int MAXF = (face_count<TX, TY, TZ>(0) + B::NT - 1) / B::NT;
    const R inv60dx = R(g.inv60dx), invdx2 = R(g.invdx2);
    int goff[4][MAXF], loff[4][MAXF];
#pragma unroll
    for (int dir = 0; dir < 4; dir++) {
#pragma unroll
        for (int s = 0; s < MAXF; s++) {
            const int f = t + s * B::NT;
            goff[dir][s] = 0;
            loff[dir][s] = B::BOX + (t & 63);
            if (f < face_count<TX, TY, TZ>(dir)) {
                int lx, ly, lz;
                face_elem<TX, TY, TZ>(dir, f, lx, ly, lz);
                // with ghost planes (g.zg >= 3) the z halo is read from them
                const int zz = g.zg ? z0 + lz - 3 + g.zg : wrapc(z0 + lz - 3, g.nz);
                goff[dir][s] = (int)((wrapc(x0 + lx - 3, g.nx) + nxl * wrapc(y0 + ly - 3, g.ny) + nxy * zz) *
                                     (long)sizeof(T));
                loff[dir][s] = lz * B::HX * B::HY + ly * B::HX + lx;
            }
        }
    }
    const int lcen = (tz + 3) * B::HX * B::HY + (ty + 3) * B::HX + (tx + 3);

    using Regs = PassRegs<T, MAXF>;
    // passes whose centre value is loaded with the pass (late_centres)
    constexpr auto pass_loads_centre = [](int pp) constexpr {
        return late_centres(PF) && is_late_channel(pass_channel(pp));
    };
    auto fetch = [&](auto pc, Regs &rg) {
        constexpr int pp = decltype(pc)::value;
        constexpr int ch = pass_channel(pp);
        constexpr int dir = pass_dir(pp);
        constexpr int nf = face_count<TX, TY, TZ>(dir);
        const T *y0p = US ? U(ch) : Y(ch);
        const T *y1p = nullptr, *k2p = nullptr;
        if constexpr (S > 1 && !US) {
            if constexpr (ch < 16) {
                y1p = Y(ch + 20);
                if constexpr (S == 3) k2p = K(b.k1, ch + 4);
                if constexpr (S == 4) k2p = K(b.k2, ch + 4);
            } else {
                const T *kp = (S == 2) ? b.k1 : (S == 3) ? b.k2 : b.k3;
                y1p = K(kp, ch - 16);
            }
        }
#pragma unroll
        for (int s = 0; s < MAXF; s++) {
            if (s * B::NT < nf) {
                const unsigned o = (unsigned)goff[dir][s];
                rg.v[s][0] = ldo(y0p, o);
                if constexpr (S > 1 && !US) rg.v[s][1] = ldo(y1p, o);
                if constexpr (S > 2 && !US && ch < 16) rg.v[s][2] = ldo(k2p, o);
            }
        }
        if constexpr (pass_loads_centre(pp)) {
            rg.vc[0] = ldo(y0p, qb);
            if constexpr (S > 1 && !US) rg.vc[1] = ldo(y1p, qb);
            if constexpr (S > 2 && !US && ch < 16) rg.vc[2] = ldo(k2p, qb);
        }
    };
    auto stage_value = [&](auto pc, const T (&v)[3]) -> R {
        constexpr int ch = pass_channel(decltype(pc)::value);
        if constexpr (S == 1 || US) return v[0];
        else if constexpr (ch < 16) {
            if constexpr (S == 2) return v[0] + v[1] * half_dt;
            else if constexpr (S == 3) return v[0] + (v[1] + v[2] * half_dt) * half_dt;
            else return v[0] + (v[1] + v[2] * half_dt) * dt;
        } else {
            return v[0] + v[1] * (S == 4 ? dt : half_dt);
        }
    };
    auto store = [&](auto pc, const Regs &rg, R *buf) {
        constexpr int pp = decltype(pc)::value;
        constexpr int dir = pass_dir(pp);
        constexpr int nf = face_count<TX, TY, TZ>(dir);
        if constexpr (pass_loads_centre(pp)) {
            const R cv = stage_value(pc, rg.vc);
            c[pass_channel(pp)] = cv;
            buf[lcen] = cv;
        } else {
            buf[lcen] = c[pass_channel(pp)];
            if constexpr (PARK && pass_channel(pp) >= 16) park[(pass_channel(pp) - 16) * B::NT + t] = c[pass_channel(pp)];
        }
#pragma unroll
        for (int s = 0; s < MAXF; s++)
            if (s * B::NT < nf) buf[loff[dir][s]] = stage_value(pc, rg.v[s]);
    };

    // derivatives of the staged channel at this thread's site
    auto deriv = [&](const R *buf, int ax, R cen, bool second, R &d1, R &d2) {
        const int stride = ax == 0 ? 1 : ax == 1 ? B::HX : B::HX * B::HY;
        const R *o = buf + lcen;
        const R m3 = o[-3 * stride], m2 = o[-2 * stride], m1 = o[-stride];
        const R a1 = o[stride], a2 = o[2 * stride], a3 = o[3 * stride];
        d1 = sten_d1(m3, m2, m1, a1, a2, a3, inv60dx);
        if (second) d2 = sten_d2(m3, m2, m1, cen, a1, a2, a3, invdx2);
    };

    // fold the derivatives of pass pp into the accumulators
    auto consume = [&](auto pc, const R *buf) {
        constexpr int pp = decltype(pc)::value;
        constexpr int kind = pass_kind(pp);
        constexpr int ch = pass_channel(pp);
        const R cen = c[ch];
        if constexpr (kind == 0) {
            // E^a_i, E^Y_i: d_i only, divergences into k_Gamma, k_Sigma (603-657)
            constexpr int i = pp / 4, a = pp % 4;
            R d1, d2;
            deriv(buf, i, cen, false, d1, d2);
            if constexpr (a < 3) kG[a] += (R(1) - gp2) * d1;
            else kS += (R(1) - gp2) * d1;
            if constexpr (DG) dacc(a, d1);
            if constexpr (late_centres(PF) && pp == 11) {
                // every E centre is staged: the W x E term of k_Gamma
#pragma unroll
                for (int aa = 0; aa < 3; aa++) {
                    R na = R(0);
#pragma unroll
                    for (int ii = 0; ii < 3; ii++)
                        na -= xprod(aa, W(0, ii), W(1, ii), W(2, ii), E(0, ii), E(1, ii), E(2, ii));
                    kG[aa] += gp2 * gw * na;
                }
            }
            if constexpr (DG && pp == 11 && !MIXED) {
                // every E centre and div E is known: the Gauss residuals
                const R piv[4] = {c[CH_PI], c[CH_PI + 1], c[CH_PI + 2], c[CH_PI + 3]};
#pragma unroll
                for (int aa = 0; aa < 3; aa++) {
                    R na = R(0);
#pragma unroll
                    for (int ii = 0; ii < 3; ii++)
                        na -= xprod(aa, W(0, ii), W(1, ii), W(2, ii), E(0, ii), E(1, ii), E(2, ii));
                    const R r = dget(aa) - gw * (na + phiT(aa, p, piv));
                    dput(aa, r * r);
                }
                const R rY = dget(3) - gy * phiTY(p, piv);
                dput(3, rY * rY);
            } else if constexpr (DG && pp == 11) {
                // the same in FP64 from the FP32 values
                const double pd[4] = {dd(p[0]), dd(p[1]), dd(p[2]), dd(p[3])};
                const double piv[4] = {dd(c[CH_PI]), dd(c[CH_PI + 1]), dd(c[CH_PI + 2]), dd(c[CH_PI + 3])};
#pragma unroll
                for (int aa = 0; aa < 3; aa++) {
                    double na = 0.0;
#pragma unroll
                    for (int ii = 0; ii < 3; ii++)
                        na -= xprod(aa, dd(W(0, ii)), dd(W(1, ii)), dd(W(2, ii)), dd(E(0, ii)), dd(E(1, ii)), dd(E(2, ii)));
                    const double r = dget(aa) - P.gw * (na + phiT(aa, pd, piv));
                    dput(aa, r * r);
                }
                const double rY = dget(3) - P.gy * phiTY(pd, piv);
                dput(3, rY * rY);
            }
        } else if constexpr (kind == 1) {
            // Gamma^a, Sigma: -d_i Gamma^a into kt_W^a_i, -d_i Sigma into kt_Y_i
#pragma unroll
            for (int i = 0; i < 3; i++) {
                R d1, d2;
                deriv(buf, i, cen, false, d1, d2);
                if constexpr (pp < 15) ktW[pp - 12][i] -= d1;
                else ktY[i] -= d1;
            }
            if constexpr (late_centres(PF) && pp == 15) {
                // every Gamma/Sigma centre is staged: W x Gamma in kt_W, the generator in kt_phi
#pragma unroll
                for (int aa = 0; aa < 3; aa++)
#pragma unroll
                    for (int ii = 0; ii < 3; ii++)
                        ktW[aa][ii] -= gw * xprod(aa, W(0, ii), W(1, ii), W(2, ii), c[CH_GAM], c[CH_GAM + 1],
                                                  c[CH_GAM + 2]);
                R Gp[4];
                apply_gen(h, c[CH_GAM], c[CH_GAM + 1], c[CH_GAM + 2], hy, c[CH_SIG], p, Gp);
#pragma unroll
                for (int bb = 0; bb < 4; bb++) ktphi[bb] += Gp[bb];
            }
        } else if constexpr (kind == 2) {
            // phi_b: kt_phi (Laplacian, A_i(d_i phi + D_i phi)), phi.T D into kt_W, kt_Y, u_j
            constexpr int bb = pp - 16;
            R lap = R(0);
            [[maybe_unused]] double Dk[3];   // D_i phi_bb, for the B field's cross products
#pragma unroll
            for (int i = 0; i < 3; i++) {
                R d1, d2;
                deriv(buf, i, cen, true, d1, d2);
                lap += d2;
                R Ap[4];
                apply_gen(h, W(0, i), W(1, i), W(2, i), hy, c[CH_Y + i], p, Ap);
                const R D = d1 + Ap[bb];
                if constexpr (DG) {
                    dacc(4, dd(D) * dd(D));
                    Dk[i] = dd(D);
                }
                add_gen_col<bb>(h, W(0, i), W(1, i), W(2, i), hy, c[CH_Y + i], d1 + D, ktphi);
                ktW[0][i] += gw * phiT_coef<0, bb>(p) * D;
                ktW[1][i] += gw * phiT_coef<1, bb>(p) * D;
                ktW[2][i] += gw * phiT_coef<2, bb>(p) * D;
                ktY[i] += gy * phiTY_coef<bb>(p) * D;
                if constexpr (BETA) u[i] += p[bb] * d1;
            }
            ktphi[bb] += lap;
            if constexpr (DG) {
                // X_i = D_j phi_1 D_k phi_2 - D_j phi_2 D_k phi_1 + D_j phi_3 D_k phi_4
                //       - D_j phi_4 D_k phi_3, (i, j, k) cyclic (em_point)
                if constexpr (bb == 0 || bb == 2) {
#pragma unroll
                    for (int i = 0; i < 3; i++) dput((bb == 0 ? 11 : 14) + i, Dk[i]);
                } else {
                    const double Dp[3] = {dget(bb == 1 ? 11 : 14), dget(bb == 1 ? 12 : 15), dget(bb == 1 ? 13 : 16)};
                    double Xn[3];
#pragma unroll
                    for (int i = 0; i < 3; i++) {
                        const int j = (i + 1) % 3, k = (i + 2) % 3;
                        Xn[i] = bb == 1 ? Dp[j] * Dk[k] - Dk[j] * Dp[k] : (dget(11 + i) + Dp[j] * Dk[k]) - Dk[j] * Dp[k];
                    }
#pragma unroll
                    for (int i = 0; i < 3; i++) dput(11 + i, Xn[i]);
                }
            }
        } else if constexpr (kind == 3) {
            // Y_i: Laplacian into kt_Y, d_j Y_i into B^Y (beta terms only)
            constexpr int i = pp - 20;
            R lap = R(0);
#pragma unroll
            for (int j = 0; j < 3; j++) {
                R d1, d2;
                deriv(buf, j, cen, true, d1, d2);
                lap += d2;
                if constexpr (BETA || DG) {
                    if (j != i) {
                        if (i == (j + 1) % 3) BYv[(j + 2) % 3] += d1;
                        else BYv[(j + 1) % 3] -= d1;
                    }
                }
                if constexpr (DG)
                    if (j == i) dacc(5, d1);
            }
            ktY[i] += lap;
            if constexpr (DG && pp == 22) {
                dput(9, 0.5 * (dd(BYv[0]) * dd(BYv[0]) + dd(BYv[1]) * dd(BYv[1]) + dd(BYv[2]) * dd(BYv[2])));
                const double cY = dd(c[CH_SIG]) - dget(5);
                dput(5, cY * cY);
#pragma unroll
                for (int i = 0; i < 3; i++) dput(14 + i, dd(BYv[i]));
            }
            if constexpr (BETA && pp == 22) {
                // B^Y complete: beta_Y terms (main_rk4.jl:354, 553-581)
                const R QY = c[CH_EY] * BYv[0] + c[CH_EY + 1] * BYv[1] + c[CH_EY + 2] * BYv[2];
#pragma unroll
                for (int bq = 0; bq < 4; bq++) ktphi[bq] += byc * QY * p[bq];
#pragma unroll
                for (int ii = 0; ii < 3; ii++) {
                    const int j = (ii + 1) % 3, k = (ii + 2) % 3;
                    ktY[ii] += R(2) * byc * (u[j] * c[CH_EY + k] - u[k] * c[CH_EY + j] - phipi * BYv[ii]);
                }
            }
        } else {
            // W^a_i: Laplacian, -gw eps W^c_j d_j W^a_i into kt_W^{a'}_i, d_j W^a_i into B^a
            constexpr int i = (pp - 23) / 3, a = (pp - 23) % 3;
            if constexpr (pp == 23) {
#pragma unroll
                for (int aa = 0; aa < 3; aa++)
#pragma unroll
                    for (int ii = 0; ii < 3; ii++) {
                        const int j = (ii + 1) % 3, k = (ii + 2) % 3;
                        Bw[aa][ii] = gw * xprod(aa, W(0, j), W(1, j), W(2, j), W(0, k), W(1, k), W(2, k));
                    }
            }
            R lap = R(0);
#pragma unroll
            for (int j = 0; j < 3; j++) {
                R d1, d2;
                deriv(buf, j, cen, true, d1, d2);
                lap += d2;
                // (d_j W_i x W_j)^{a'} = eps^{a' a c} d_j W^a_i W^c_j
                constexpr int a1 = (a + 1) % 3, a2 = (a + 2) % 3;   // a' values
                ktW[a1][i] -= gw * R(eps3(a1, a)) * W(3 - a1 - a, j) * d1;
                ktW[a2][i] -= gw * R(eps3(a2, a)) * W(3 - a2 - a, j) * d1;
                if constexpr (DG)
                    if (j == i) dacc(6 + a, d1);
                if (j != i) {
                    if (i == (j + 1) % 3) Bw[a][(j + 2) % 3] += d1;
                    else Bw[a][(j + 1) % 3] -= d1;
                }
            }
            ktW[a][i] += lap;
            if constexpr (DG && pp == 31) {
                double me = 0.0;
#pragma unroll
                for (int aa = 0; aa < 3; aa++)
                    me += dd(Bw[aa][0]) * dd(Bw[aa][0]) + dd(Bw[aa][1]) * dd(Bw[aa][1]) + dd(Bw[aa][2]) * dd(Bw[aa][2]);
                dput(10, 0.5 * me);
#pragma unroll
                for (int aa = 0; aa < 3; aa++) {
                    const double cc = dd(c[CH_GAM + aa]) - dget(6 + aa);
                    dput(6 + aa, cc * cc);
                }
                // the electromagnetic densities of em_point (main_rk4.jl:2303-2822,
                // 3718-4250): B = -(F + c1 X), B_2 = -(F + c2 X), A
                const double p1 = dd(p[0]), p2 = dd(p[1]), p3 = dd(p[2]), p4 = dd(p[3]);
                const double mag = p1 * p1 + p2 * p2 + p3 * p3 + p4 * p4;
                double nv[3] = {0.0, 0.0, 0.0};
                if (mag != 0.0) {
                    nv[0] = -2.0 * (p1 * p3 + p2 * p4) / mag;
                    nv[1] = 2.0 * (p2 * p3 - p1 * p4) / mag;
                    nv[2] = (-p1 * p1 - p2 * p2 + p3 * p3 + p4 * p4) / mag;
                }
                const double c1 = 4.0 * sw / (P.gw * P.vev * P.vev);
                const double c2 = mag != 0.0 ? 4.0 * sw / (P.gw * mag * mag) : 0.0;
                double bb2 = 0.0, b22 = 0.0, b2a = 0.0;
#pragma unroll
                for (int i = 0; i < 3; i++) {
                    const double Xi = dget(11 + i);
                    const double F = sw * (nv[0] * dd(Bw[0][i]) + nv[1] * dd(Bw[1][i]) + nv[2] * dd(Bw[2][i])) +
                                     cw * dget(14 + i);
                    const double Be = -(F + c1 * Xi);
                    const double B2 = mag != 0.0 ? -(F + c2 * Xi) : 0.0;
                    const double Ai = sw * (nv[0] * dd(W(0, i)) + nv[1] * dd(W(1, i)) + nv[2] * dd(W(2, i))) +
                                      cw * dd(c[CH_Y + i]);
                    bb2 += Be * Be;
                    b22 += B2 * B2;
                    b2a += B2 * Ai;
                }
                dput(11, bb2);
                dput(12, b22);
                dput(13, b2a);
            }
            if constexpr (pp == 31) {
                // B^a complete: -gw sum_j (W_j x F_ij)^a, beta_W terms
                if constexpr (BETA) {
                    R QW = R(0);
#pragma unroll
                    for (int aa = 0; aa < 3; aa++)
                        QW += E(aa, 0) * Bw[aa][0] + E(0, 1) * Bw[aa][1] + E(0, 2) * Bw[aa][2];
#pragma unroll
                    for (int bq = 0; bq < 4; bq++) ktphi[bq] += bw * QW * p[bq];
                }
#pragma unroll
                for (int aa = 0; aa < 3; aa++) {
#pragma unroll
                    for (int ii = 0; ii < 3; ii++) {
                        const int j1 = (ii + 1) % 3, j2 = (ii + 2) % 3;
                        // F_{i,i+1} = B_{i+2}, F_{i,i+2} = -B_{i+1}
                        const R t2 = xprod(aa, W(0, j1), W(1, j1), W(2, j1), Bw[0][j2], Bw[1][j2], Bw[2][j2]) -
                                     xprod(aa, W(0, j2), W(1, j2), W(2, j2), Bw[0][j1], Bw[1][j1], Bw[2][j1]);
                        ktW[aa][ii] += -gw * t2;
                        if constexpr (BETA)
                            ktW[aa][ii] += R(2) * bw * (u[j1] * E(aa, j2) - u[j2] * E(aa, j1) - phipi * Bw[aa][ii]);
                    }
                }
            }
        }
    };

    // pin the accumulators still live after pass pp
    auto pin_all = [&](auto pc) {
        constexpr int pp = decltype(pc)::value;
#pragma unroll
        for (int i = 0; i < 4; i++) pin(ktphi[i]);
#pragma unroll
        for (int i = 0; i < 3; i++) {
            pin(ktY[i]); pin(kG[i]);
            if constexpr (BETA && pp >= 16) pin(u[i]);
            if constexpr ((BETA || DG) && pp >= 20 && pp < 22) pin(BYv[i]);
#pragma unroll
            for (int j = 0; j < 3; j++) {
                pin(ktW[i][j]);
                if constexpr (pp >= 23 && pp < 31) pin(Bw[i][j]);
            }
        }
        pin(kS);
    };

    // ---- the channel pipeline: the halos of passes p+1 .. p+PFS are in
    // flight (in a ring of PFS register sets) while pass p is computed from
    // LDS.  Set k % PFS holds pass k; the set of pass p is refilled with pass
    // p + PFS once pass p has been staged (during pass p - 1).
    constexpr int PFS = pf_of_stage(PF, S);
    Regs rr[PFS];
    [&]<int... I>(std::integer_sequence<int, I...>) {
        (fetch(std::integral_constant<int, I>{}, rr[I]), ...);
    }(std::make_integer_sequence<int, PFS>{});
    // y for the next stage input, loaded during the last YP passes (y_early).
    // With early_fields the 16 field channels of that input, y + c * (this
    // stage's momentum input), are stored as soon as their y arrives (passes
    // NPASS-YP .. NPASS-YP/2-1); the other 20 need this stage's derivatives,
    // so their y is held until the end (loaded in the last YP/2 passes).
    constexpr bool YE = y_early(PF) && S > 1 && stores_next(PF, S) && !paired_of(PF, S);
    constexpr bool YF = YE && early_fields(PF);
    // recon_y, stage 2: k1 (20 channels) instead of y (36) in the last passes;
    // u3 = y + k2 dt/2 with y = u2 - k1 dt/2 (momenta, Gamma, Sigma) and, for a
    // field, y_f = u2_f - y_pi dt/2, y_pi = u2_pi - kt1 dt/2
    constexpr bool YK = YE && recon_y(PF) && S == 2;
    // passes that load those values: the last 8, the k1 values of YK in the
    // last 4 (held over 8 passes they spilled the FP64 stage 2: 20 B per lane;
    // 512^3: stage 2 31.0 -> 30.6 ms)
    constexpr int YP = YK ? 4 : 8;
    R ypre[YE && !YK ? NCH : 1];
    R kpre[YK ? NK : 1];
    if constexpr (!paired_of(PF, S) && !SPREAD) {
        store(std::integral_constant<int, 0>{}, rr[0], lds[0]);
        lds_barrier();
        auto body = [&](auto pc) {
            constexpr int pp = decltype(pc)::value;
            if constexpr (YK && pp >= NPASS - YP) {
                constexpr int q = pp - (NPASS - YP);
#pragma unroll
                for (int m = q * NK / YP; m < (q + 1) * NK / YP; m++) kpre[m] = ldn(K(b.k1, m), qb);
            } else if constexpr (YE && !YF && pp >= NPASS - YP) {
                constexpr int q = pp - (NPASS - YP);
#pragma unroll
                for (int ch = q * NCH / YP; ch < (q + 1) * NCH / YP; ch++) ypre[ch] = ldn(Y(ch), qb);
            }
            if constexpr (YF && pp >= NPASS - YP) {
                constexpr int q = pp - (NPASS - YP);
                if constexpr (q < YP / 2) {
#pragma unroll
                    for (int ch = q * 16 / (YP / 2); ch < (q + 1) * 16 / (YP / 2); ch++)
                        stn(b.unext + (long)ch * b.n3, qb, R(ldn(Y(ch), qb)) + c[ch + 20] * (S == 3 ? dt : half_dt));
                } else {
                    constexpr int q2 = q - YP / 2;
#pragma unroll
                    for (int ch = 16 + q2 * 20 / (YP / 2); ch < 16 + (q2 + 1) * 20 / (YP / 2); ch++)
                        ypre[ch] = ldn(Y(ch), qb);
                }
            }
            if constexpr (pp + PFS < NPASS) fetch(std::integral_constant<int, pp + PFS>{}, rr[pp % PFS]);
            consume(pc, lds[pp & 1]);
            pin_all(pc);
            if constexpr (pp + 1 < NPASS)
                store(std::integral_constant<int, pp + 1>{}, rr[(pp + 1) % PFS], lds[(pp + 1) & 1]);
            lds_barrier();
        };
        [&]<int... I>(std::integer_sequence<int, I...>) {
            (body(std::integral_constant<int, I>{}), ...);
        }(std::make_integer_sequence<int, NPASS>{});
    } else if constexpr (!paired_of(PF, S)) {
        // stage 4 with the partial sums spread over the first NK passes.  The
        // inputs of pass q's update (channel 16+m with its k's, and with
        // fused_fields the field channel whose momentum it is) are loaded one
        // pass earlier, before that pass's halo fetch: a use of a loaded value
        // waits for every load issued before it, so inputs loaded after the
        // fetch of pass q + PFS and used in pass q made the pass wait for that
        // halo too (512^3: FP64 334 stage 4 37.0 -> 33.2 ms, mixed 333 28.9 ->
        // 26.1 ms; loading them in their own pass before the fetch gave 37.0 /
        // 26.3 ms, two passes ahead 32.9 / 28.5 ms, the mixed kernel spilling).
        constexpr int UA = 1;
        struct Upd {
            R k1, k2, k3;   // k1..k3 of channel 16+m
            R y, lo;        // y (mixed: high and compensation word) of channel 16+m
            R f, lf;        // y of field channel m-4 (and its compensation word)
        };
        auto upd_load = [&](auto qc, Upd &u) {
            constexpr int m = upd_channel(PF, decltype(qc)::value);
            u.k1 = ldn(K(b.k1, m), qb);
            u.k2 = ldn(K(b.k2, m), qb);
            u.k3 = ldn(K(b.k3, m), qb);
            if constexpr (MIXED) {
                u.lo = ldn(b.ylo + (long)(16 + m) * b.n3, qb);
                if constexpr (!lean_stash(PF) || (fused_fields(PF) && m >= 4)) u.y = ldn(Y(16 + m), qb);
                if constexpr (fused_fields(PF) && m >= 4) {
                    u.f = ldn(Y(m - 4), qb);
                    u.lf = ldn(b.ylo + (long)(m - 4) * b.n3, qb);
                }
            } else if constexpr (!recon_y(PF)) {
                u.y = ldn(Y(16 + m), qb);
                if constexpr (fused_fields(PF) && m >= 4) u.f = ldn(Y(m - 4), qb);
            }
        };
        auto upd_apply = [&](auto qc, const Upd &u) {
            constexpr int m = upd_channel(PF, decltype(qc)::value);
            if constexpr (MIXED) {
                // the increment of channel 16+m (and with m >= 4 the field
                // channel m - 4, whose momentum it is), as the prologue forms them
                R dy = u.k1 * d6;
                dy = u.k2 * d3 + dy;
                dy = u.k3 * d3 + dy;
                nstash[m * B::NT + t] = u.lo + dy;
                if constexpr (!lean_stash(PF)) nstash[(NK + m) * B::NT + t] = u.y;
                if constexpr (fused_fields(PF) && m >= 4) {
                    constexpr int ch = m - 4;
                    const R piv = u.y;
                    const R k2 = piv + u.k1 * half_dt;
                    const R k3 = piv + u.k2 * half_dt;
                    const R k4 = piv + u.k3 * dt;
                    R dyf = piv * d6;
                    dyf = k2 * d3 + dyf;
                    dyf = k3 * d3 + dyf;
                    dyf = k4 * d6 + dyf;
                    R nh, nl;
                    two_sum(u.f, u.lf + dyf, nh, nl);
                    stn(b.ynew + (long)ch * b.n3, qb, nh);
                    stn(b.ylo_new + (long)ch * b.n3, qb, nl);
                }
            } else if constexpr (recon_y(PF)) {
                // the partial sum of channel 16+m with y reconstructed from the
                // centre of the stage input this pass has set: u4 = y + k3 dt
                static_assert(stored_in(PF, 4), "y is reconstructed from the stored stage input");
                const R yv = c[16 + m] - u.k3 * dt;
                R n = u.k1 * d6 + yv;
                n = u.k2 * d3 + n;
                nstash[m * B::NT + t] = u.k3 * d3 + n;
                if constexpr (m >= 4) {
                    // field channel m - 4, whose momentum is channel 16 + m:
                    // u4_f = y_f + k3_f dt
                    constexpr int ch = m - 4;
                    const R piv = yv;
                    const R k2 = piv + u.k1 * half_dt;
                    const R k3 = piv + u.k2 * half_dt;
                    const R k4 = piv + u.k3 * dt;
                    R nf = piv * d6 + (c[ch] - k3 * dt);
                    nf = k2 * d3 + nf;
                    nf = k3 * d3 + nf;
                    stn(b.ynew + (long)ch * b.n3, qb, k4 * d6 + nf);
                }
            } else {
                R n = u.k1 * d6 + u.y;
                n = u.k2 * d3 + n;
                nstash[m * B::NT + t] = u.k3 * d3 + n;
                if constexpr (fused_fields(PF) && m >= 4) {
                    // field channel m - 4: its momentum is channel 16 + m,
                    // the expressions of the field loop
                    constexpr int ch = m - 4;
                    const R piv = u.y;
                    const R k2 = piv + u.k1 * half_dt;
                    const R k3 = piv + u.k2 * half_dt;
                    const R k4 = piv + u.k3 * dt;
                    R nf = piv * d6 + u.f;
                    nf = k2 * d3 + nf;
                    nf = k3 * d3 + nf;
                    stn(b.ynew + (long)ch * b.n3, qb, k4 * d6 + nf);
                }
            }
        };
        Upd up[UA + 1];
        [&]<int... I>(std::integer_sequence<int, I...>) {
            (upd_load(std::integral_constant<int, I>{}, up[I]), ...);
        }(std::make_integer_sequence<int, UA>{});
        store(std::integral_constant<int, 0>{}, rr[0], lds[0]);
        lds_barrier();
        auto body = [&](auto pc) {
            constexpr int pp = decltype(pc)::value;
            if constexpr (pp + UA < NK) {
                upd_load(std::integral_constant<int, pp + UA>{}, up[(pp + UA) % (UA + 1)]);
                __builtin_amdgcn_sched_barrier(0);
            }
            if constexpr (pp + PFS < NPASS) fetch(std::integral_constant<int, pp + PFS>{}, rr[pp % PFS]);
            consume(pc, lds[pp & 1]);
            pin_all(pc);
            if constexpr (pp < NK) upd_apply(pc, up[pp % (UA + 1)]);
            if constexpr (pp + 1 < NPASS)
                store(std::integral_constant<int, pp + 1>{}, rr[(pp + 1) % PFS], lds[(pp + 1) & 1]);
            lds_barrier();
        };
        [&]<int... I>(std::integer_sequence<int, I...>) {
            (body(std::integral_constant<int, I>{}), ...);
        }(std::make_integer_sequence<int, NPASS>{});
    } else {
        // passes 2q, 2q+1 are computed between two barriers from buffers
        // (2q)%4, (2q+1)%4 while passes 2q+2, 2q+3 are staged into the other two
        static_assert(PFS >= 3, "paired passes need a prefetch depth of at least 3");
        store(std::integral_constant<int, 0>{}, rr[0], lds[0]);
        store(std::integral_constant<int, 1>{}, rr[1], lds[1]);
        lds_barrier();
        auto body2 = [&](auto qc) {
            constexpr int p0 = 2 * decltype(qc)::value, p1 = p0 + 1;
            if constexpr (p0 + PFS < NPASS) fetch(std::integral_constant<int, p0 + PFS>{}, rr[p0 % PFS]);
            if constexpr (p1 + PFS < NPASS) fetch(std::integral_constant<int, p1 + PFS>{}, rr[p1 % PFS]);
            // one pin per pair: the scheduler may interleave the two passes
            consume(std::integral_constant<int, p0>{}, lds[p0 & 3]);
            consume(std::integral_constant<int, p1>{}, lds[p1 & 3]);
            pin_all(std::integral_constant<int, p1>{});
            if constexpr (p0 + 2 < NPASS)
                store(std::integral_constant<int, p0 + 2>{}, rr[(p0 + 2) % PFS], lds[(p0 + 2) & 3]);
            if constexpr (p1 + 2 < NPASS)
                store(std::integral_constant<int, p1 + 2>{}, rr[(p1 + 2) % PFS], lds[(p1 + 2) & 3]);
            lds_barrier();
        };
        [&]<int... I>(std::integer_sequence<int, I...>) {
            (body2(std::integral_constant<int, I>{}), ...);
        }(std::make_integer_sequence<int, NPASS / 2>{});
    }

    // ---- outputs: k_Gamma(3), k_Sigma, kt_phi(4), kt_W(9), kt_Y(3)
    R o[NK];
#pragma unroll
    for (int a = 0; a < 3; a++) o[a] = kG[a];
    o[3] = kS;
#pragma unroll
    for (int bb = 0; bb < 4; bb++) o[4 + bb] = ktphi[bb];
#pragma unroll
    for (int a = 0; a < 3; a++)
#pragma unroll
        for (int i = 0; i < 3; i++) o[8 + 3 * a + i] = ktW[a][i];
#pragma unroll
    for (int i = 0; i < 3; i++) o[17 + i] = ktY[i];

    if constexpr (S < 4 && DIAG == 2) {
        // diagnostics alone: the stage outputs (and the arithmetic only they
        // need) are dropped
    } else if constexpr (S < 4) {
#pragma unroll
        for (int m = 0; m < NK; m++) stn(b.kout + (long)m * b.n3, qb, o[m]);
        if constexpr (stores_next(PF, S)) {
            // next stage input: fields y + c * (this stage's momentum input),
            // the other channels y + c * k_s (the expressions of stage_in)
            const R cc = S == 3 ? dt : half_dt;
#pragma unroll
            for (int ch = YF ? 16 : 0; ch < NCH; ch++) {   // early_fields: the field channels are stored
                R yv;
                if constexpr (PARK) yv = ch >= 16 ? park[(ch - 16) * B::NT + t] : c[ch];
                else if constexpr (S == 1) yv = c[ch];
                else if constexpr (YK) yv = ch < 16 ? c[ch] - (c[ch + 20] - kpre[ch + 4] * half_dt) * half_dt
                                                    : c[ch] - kpre[ch - 16] * half_dt;
                else if constexpr (YE) yv = ypre[ch];
                else yv = R(ldo(Y(ch), qb));
                const R kv = ch < 16 ? c[ch + 20] : o[ch - 16];
                stn(b.unext + (long)ch * b.n3, qb, yv + kv * cc);
            }
        }
    } else if constexpr (MIXED) {
#pragma unroll
        for (int m = 0; m < NK; m++) {
            R nh, nl;
            const R hi = lean_stash(PF) ? R(ldn(Y(16 + m), qb)) : nstash[(NK + m) * B::NT + t];
            two_sum(hi, o[m] * d6 + nstash[m * B::NT + t], nh, nl);
            stn(b.ynew + (long)(16 + m) * b.n3, qb, nh);
            stn(b.ylo_new + (long)(16 + m) * b.n3, qb, nl);
        }
    } else {
#pragma unroll
        for (int m = 0; m < NK; m++) stn(b.ynew + (long)(16 + m) * b.n3, qb, o[m] * d6 + nstash[m * B::NT + t]);
    }
    if constexpr (DG) {
        // the brick's partial sums: per-thread values wave-reduced, then the
        // waves combined in a fixed order.  The densities of centre values
        // alone, from the centres (the same expressions as the slots had).
        double cen[5];   // PE, KE, |phi|, EEW, EEY
        if constexpr (MIXED) {
            const double r2d = dd(p[0]) * dd(p[0]) + dd(p[1]) * dd(p[1]) + dd(p[2]) * dd(p[2]) + dd(p[3]) * dd(p[3]);
            const double X = r2d - P.vev * P.vev;
            cen[0] = P.lam * (X * X);
            cen[1] = dd(c[CH_PI]) * dd(c[CH_PI]) + dd(c[CH_PI + 1]) * dd(c[CH_PI + 1]) +
                     dd(c[CH_PI + 2]) * dd(c[CH_PI + 2]) + dd(c[CH_PI + 3]) * dd(c[CH_PI + 3]);
            cen[2] = sqrt(r2d);
            double ee = 0.0;
#pragma unroll
            for (int qq = 0; qq < 9; qq++) ee += dd(c[CH_E + qq]) * dd(c[CH_E + qq]);
            cen[3] = 0.5 * ee;
            cen[4] = 0.5 * (dd(c[CH_EY]) * dd(c[CH_EY]) + dd(c[CH_EY + 1]) * dd(c[CH_EY + 1]) +
                            dd(c[CH_EY + 2]) * dd(c[CH_EY + 2]));
        } else {
            const R X = r2 - R(P.vev) * R(P.vev);
            cen[0] = R(P.lam) * (X * X);
            cen[1] = c[CH_PI] * c[CH_PI] + c[CH_PI + 1] * c[CH_PI + 1] + c[CH_PI + 2] * c[CH_PI + 2] +
                     c[CH_PI + 3] * c[CH_PI + 3];
            cen[2] = sqrt(r2);
            R ee = R(0);
#pragma unroll
            for (int qq = 0; qq < 9; qq++) ee += c[CH_E + qq] * c[CH_E + qq];
            cen[3] = R(0.5) * ee;
            cen[4] = R(0.5) * (c[CH_EY] * c[CH_EY] + c[CH_EY + 1] * c[CH_EY + 1] + c[CH_EY + 2] * c[CH_EY + 2]);
        }
        constexpr int cen_slot[5] = {0, 1, 7, 3, 5};
        // slots 0..13 (14-16 are scratch)
        constexpr int slot_of[14] = {10, 11, 12, 13, 2, 17, 14, 15, 16, 6, 4, 18, 19, 20};
        int bad = 0;
#pragma unroll
        for (int k = 0; k < 5; k++) {
            const double v = cen[k];
            bad += !isfinite(v);
            const double r = wave_reduce(v, false);
            if ((t & 63) == 0) dgw[(t >> 6) * NDIAG + cen_slot[k]] = r;
            if (k == 2) {
                const double rm = wave_reduce(v, true);
                if ((t & 63) == 0) dgw[(t >> 6) * NDIAG + 8] = rm;
            }
        }
#pragma unroll
        for (int k = 0; k < 14; k++) {
            const double v = dget(k);
            if (slot_of[k] < 8 || slot_of[k] >= 18) bad += !isfinite(v);
            const double r = wave_reduce(v, false);
            if ((t & 63) == 0) dgw[(t >> 6) * NDIAG + slot_of[k]] = r;
        }
        const double rb = wave_reduce(double(bad), false);
        if ((t & 63) == 0) dgw[(t >> 6) * NDIAG + 9] = rb;
        lds_barrier();
        if (t < NDIAG) {
            double acc = t == 8 ? INFINITY : 0.0;
#pragma unroll
            for (int w = 0; w < B::NT / 64; w++) acc = t == 8 ? fmin(acc, dgw[w * NDIAG + t]) : acc + dgw[w * NDIAG + t];
            diag[(long)bid * NDIAG + t] = acc;
        }
    }
}

template <int TX, int TY, int TZ, int PF, bool BETA, int S, typename T = double>
__global__ void __launch_bounds__(TX *TY *TZ)
__attribute__((amdgpu_waves_per_eu(waves_per_simd(std::is_same_v<T, float>, PF, S))))
stage_brick(StageBufs<T> b, BrickGeo g, Phys P, T dt)
{
    extern __shared__ __align__(16) unsigned char lds_raw[];
    int bid = blockIdx.x;
    if (g.xcd_map) {
        // consecutive ids go to different XCDs; give XCD x the contiguous range
        // [x*nbrick/8, (x+1)*nbrick/8) so its resident workgroups are neighbours
        const int per = (g.nbx * g.nby * g.nbz) >> 3;
        bid = (bid & 7) * per + (bid >> 3);
    }
    brick_body<TX, TY, TZ, PF, BETA, S, T>(b, g, P, dt, bid, lds_raw, (int)threadIdx.x);
}

// Stage 1 with the diagnostics of its input (the state) reduced on the way:
// per brick NDIAG partial sums to diag[brick * NDIAG + slot].  Two waves per
// SIMD (the per-site slots in LDS leave room for two 256-thread or one
// 512-thread workgroup per CU).
template <int TX, int TY, int TZ, int PF, bool BETA, typename T = double, int MODE = 1>
__global__ void __launch_bounds__(TX *TY *TZ) __attribute__((amdgpu_waves_per_eu(waves_per_simd(false, PF, 1))))
stage1_diag(StageBufs<T> b, BrickGeo g, Phys P, T dt, double *diag, double sw, double cw)
{
    extern __shared__ __align__(16) unsigned char lds_raw[];
    int bid = blockIdx.x;
    if (g.xcd_map) {
        const int per = (g.nbx * g.nby * g.nbz) >> 3;
        bid = (bid & 7) * per + (bid >> 3);
    }
    brick_body<TX, TY, TZ, PF, BETA, 1, T, MODE>(b, g, P, dt, bid, lds_raw, (int)threadIdx.x, diag, sw, cw);
}

template <int TX, int TY, int TZ>
bool brick_fits(const ewpt_lattice &L)
{
    // byte offsets within one channel are 32-bit (FP64 bound, also used for FP32)
    return L.nx % TX == 0 && L.ny % TY == 0 && L.nz % TZ == 0 && L.nx >= TX && L.ny >= TY && L.nz >= TZ &&
           (long)L.nx * L.ny * L.nz * 8 < (1L << 31);
}

template <int TX, int TY, int TZ, int PF, bool BETA, typename T>
hipError_t launch_tile_beta(int stage, const StageBufs<T> &b, const BrickGeo &g, const Phys &P, T dt, hipStream_t st)
{
    using B = BrickShape<TX, TY, TZ>;
    const dim3 grid(g.nbx * g.nby * g.nbz), block(B::NT);
    const size_t lds13 = (paired_of(PF, 1) ? 4 : 2) * B::BUF * sizeof(T) + (park_centres(PF, 1) ? NK * B::NT * sizeof(T) : 0);
    // stage 4 (never paired) parks n (FP64) or the increment and the high word (mixed)
    const size_t lds4 =
        2 * B::BUF * sizeof(T) + (std::is_same_v<T, float> && !lean_stash(PF) ? 2 : 1) * NK * B::NT * sizeof(T);
    // the kernels may need more than 64 KB of LDS: raise their limit once per
    // process (a failed attempt is retried by the next launch); the mutex
    // keeps threads that launch the same kernels from racing on the flag
    static std::mutex attr_mu;
    static bool attr_set = false;
    {
        std::lock_guard<std::mutex> g(attr_mu);
        if (!attr_set) {
            hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void *>(&stage_brick<TX, TY, TZ, PF, BETA, 4, T>),
                                               hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds4);
            const void *k13[3] = {reinterpret_cast<const void *>(&stage_brick<TX, TY, TZ, PF, BETA, 1, T>),
                                  reinterpret_cast<const void *>(&stage_brick<TX, TY, TZ, PF, BETA, 2, T>),
                                  reinterpret_cast<const void *>(&stage_brick<TX, TY, TZ, PF, BETA, 3, T>)};
            for (int i = 0; i < 3 && e == hipSuccess; i++)
                e = hipFuncSetAttribute(k13[i], hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds13);
            if (e != hipSuccess) return e;
            attr_set = true;
        }
    }
    switch (stage) {
    case 1: stage_brick<TX, TY, TZ, PF, BETA, 1, T><<<grid, block, lds13, st>>>(b, g, P, dt); break;
    case 2: stage_brick<TX, TY, TZ, PF, BETA, 2, T><<<grid, block, lds13, st>>>(b, g, P, dt); break;
    case 3: stage_brick<TX, TY, TZ, PF, BETA, 3, T><<<grid, block, lds13, st>>>(b, g, P, dt); break;
    default: stage_brick<TX, TY, TZ, PF, BETA, 4, T><<<grid, block, lds4, st>>>(b, g, P, dt); break;
    }
    return hipGetLastError();
}

template <int TX, int TY, int TZ, int PF, bool BETA, typename T, int MODE = 1>
hipError_t launch_diag_beta(const StageBufs<T> &b, const BrickGeo &g, const Phys &P, T dt, double *diag, double sw,
                            double cw, hipStream_t st)
{
    using B = BrickShape<TX, TY, TZ>;
    const size_t lds = 2 * B::BUF * sizeof(T) + ((B::NT / 64) * NDIAG + 17 * B::NT) * sizeof(double);
    static std::mutex attr_mu;
    static bool attr_set = false;
    {
        std::lock_guard<std::mutex> gd(attr_mu);
        if (!attr_set) {
            const hipError_t e =
                hipFuncSetAttribute(reinterpret_cast<const void *>(&stage1_diag<TX, TY, TZ, PF, BETA, T, MODE>),
                                    hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
            if (e != hipSuccess) return e;
            attr_set = true;
        }
    }
    stage1_diag<TX, TY, TZ, PF, BETA, T, MODE><<<dim3(g.nbx * g.nby * g.nbz), dim3(B::NT), lds, st>>>(b, g, P, dt, diag,
                                                                                                      sw, cw);
    return hipGetLastError();
}

BrickGeo make_brick_geo(int TX, int TY, int TZ, const ewpt_lattice &L, int xcd_map, int order, int zg)
{
    BrickGeo g;
    g.order = order;
    g.nx = L.nx; g.ny = L.ny; g.nz = L.nz;
    g.nbx = L.nx / TX; g.nby = L.ny / TY; g.nbz = L.nz / TZ;
    const int nb = g.nbx * g.nby * g.nbz;
    g.xcd_map = (xcd_map && nb % 8 == 0) ? 1 : 0;
    g.inv60dx = 1.0 / (60.0 * L.dx);
    g.invdx2 = 1.0 / (L.dx * L.dx);
    g.zg = zg;
    return g;
}

// stage 1 with the diagnostics (diag: one NDIAG row per brick; sw, cw: the
// Weinberg angle of the EM densities)
template <int TX, int TY, int TZ, int PF, typename T = double>
hipError_t launch_tile_diag(const StageBufs<T> &b, const ewpt_lattice &L, const Phys &P, double dt, int xcd_map,
                            int order, double *diag, double sw, double cw, hipStream_t st)
{
    const BrickGeo g = make_brick_geo(TX, TY, TZ, L, xcd_map, order, b.zghost);
    if (P.bw != 0.0 || P.by != 0.0) return launch_diag_beta<TX, TY, TZ, PF, true, T>(b, g, P, T(dt), diag, sw, cw, st);
    return launch_diag_beta<TX, TY, TZ, PF, false, T>(b, g, P, T(dt), diag, sw, cw, st);
}

// the diagnostics of b.y alone (stage 1's passes without its outputs; the
// beta terms enter only the outputs)
template <int TX, int TY, int TZ, int PF>
hipError_t launch_tile_diag_only(const StageBufs<double> &b, const ewpt_lattice &L, const Phys &P, int xcd_map,
                                 int order, double *diag, double sw, double cw, hipStream_t st)
{
    const BrickGeo g = make_brick_geo(TX, TY, TZ, L, xcd_map, order, b.zghost);
    return launch_diag_beta<TX, TY, TZ, PF, false, double, 2>(b, g, P, 0.0, diag, sw, cw, st);
}

template <int TX, int TY, int TZ, int PF, typename T = double>
hipError_t launch_tile(int stage, const StageBufs<T> &b, const ewpt_lattice &L, const Phys &P, double dt,
                       int xcd_map, int order, hipStream_t st)
{
    BrickGeo g;
    g.order = order;
    g.nx = L.nx; g.ny = L.ny; g.nz = L.nz;
    g.nbx = L.nx / TX; g.nby = L.ny / TY; g.nbz = L.nz / TZ;
    const int nb = g.nbx * g.nby * g.nbz;
    g.xcd_map = (xcd_map && nb % 8 == 0) ? 1 : 0;
    g.inv60dx = 1.0 / (60.0 * L.dx);
    g.invdx2 = 1.0 / (L.dx * L.dx);
    g.zg = b.zghost;
    if (P.bw != 0.0 || P.by != 0.0) return launch_tile_beta<TX, TY, TZ, PF, true, T>(stage, b, g, P, T(dt), st);
    return launch_tile_beta<TX, TY, TZ, PF, false, T>(stage, b, g, P, T(dt), st);
}

}  // namespace

}  // namespace ewpt
