// ewpt_batch.hip -- RK4 stages of many independent lattices (of different
// sizes) in one launch per stage.
//
// A plan lists every brick of every lattice in a tile table (lattice id,
// brick index), lattice by lattice, each lattice's bricks in the brick
// enumeration order of the chosen variant.  XCD x (workgroups with
// blockIdx % 8 == x) owns the x-th eighth of the table, so the bricks an XCD
// works on at one time are neighbours and share its L2.  Two launch forms:
// one workgroup per brick, or a persistent grid of (resident workgroups per
// CU) x (CUs) whose workgroups stride through their XCD's eighth; the
// autotuner picks.  Every brick runs the same brick_body as the
// single-lattice kernel, so a batched step gives bitwise the same result as
// stepping each lattice alone with the same variant.
//
// The per-lattice stage buffers and scalars live in a device descriptor
// table [parity][stage][lattice]; parity 1 is parity 0 with the state and
// the updated state swapped, so two consecutive steps need no host work and
// can be captured once in a hipGraph and replayed.
#include <algorithm>
#include <cstring>
#include <new>
#include <vector>

#include "ewpt_brick.h"

namespace ewpt {

namespace {

template <typename T>
struct BatchDesc {
    StageBufs<T> b;
    BrickGeo g;
    Phys P;
    T dt;
};

// PERSIST = false: one workgroup per brick, grid = 8 * ceil(ntiles / 8);
// workgroup b runs on XCD b % 8 and takes tile (b % 8) * chunk + b / 8.
// PERSIST = true: a fixed grid (resident workgroups x CUs); the workgroups of
// XCD x stride through the x-th eighth of the table.  The thread index is
// passed through an opaque move each iteration, which keeps the compiler
// from hoisting the per-brick address set-up out of the loop (that costs
// registers for the whole loop body).
template <int TX, int TY, int TZ, int PF, bool BETA, int S, typename T, bool PERSIST>
__global__ void __launch_bounds__(TX *TY *TZ)
__attribute__((amdgpu_waves_per_eu(waves_per_simd(std::is_same_v<T, float>, PF, S))))
stage_brick_batch(const BatchDesc<T> *desc, const int2 *tiles, int ntiles)
{
    extern __shared__ __align__(16) unsigned char lds_raw[];
    const int xcd = blockIdx.x & 7;
    const int chunk = (ntiles + 7) >> 3;
    if constexpr (!PERSIST) {
        const int w = xcd * chunk + (int)(blockIdx.x >> 3);
        if (w >= ntiles) return;
        const int2 tl = tiles[w];
        const BatchDesc<T> d = desc[tl.x];
        brick_body<TX, TY, TZ, PF, BETA, S, T>(d.b, d.g, d.P, d.dt, tl.y, lds_raw, (int)threadIdx.x);
    } else {
        const int nper = gridDim.x >> 3;
        for (int j = (int)(blockIdx.x >> 3); j < chunk; j += nper) {
            const int w = xcd * chunk + j;
            if (w >= ntiles) break;
            const int2 tl = tiles[w];
            const BatchDesc<T> d = desc[tl.x];
            int t;
            asm volatile("v_mov_b32 %0, %1" : "=v"(t) : "v"((int)threadIdx.x));
            brick_body<TX, TY, TZ, PF, BETA, S, T>(d.b, d.g, d.P, d.dt, tl.y, lds_raw, t);
            lds_barrier();   // the next brick reuses the LDS
        }
    }
}

// tile shape of a variant (variant % 100), as in ewpt_brick*.hip
bool tile_of(int v, bool f32, int &tx, int &ty, int &tz)
{
    switch (v) {
    case 10: tx = 16; ty = 8; tz = 4; return true;
    case 11: tx = 32; ty = 4; tz = 4; return true;
    default: break;
    }
    if (f32) {
        // stored stage inputs for stages 2-4 (the mixed path's single-lattice forms)
        switch (v) {
        case 15: case 31: case 33: tx = 32; ty = 4; tz = 4; return true;
        default: return false;
        }
    }
    switch (v) {
    case 2: tx = 8; ty = 8; tz = 8; return true;
    case 32: case 34: tx = 16; ty = 4; tz = 4; return true;
    default: return false;
    }
}

template <int TX, int TY, int TZ, int PF, typename T>
struct Kern {
    using B = BrickShape<TX, TY, TZ>;
    static size_t lds(int stage)
    {
        const size_t l13 = 2 * B::BUF * sizeof(T);
        if (stage == 1 && park_centres(PF, 1)) return l13 + NK * B::NT * sizeof(T);
        return stage < 4 ? l13 : l13 + (std::is_same_v<T, float> && !lean_stash(PF) ? 2 : 1) * NK * B::NT * sizeof(T);
    }
    template <bool BETA, int S, bool PS>
    static const void *fn()
    {
        return reinterpret_cast<const void *>(&stage_brick_batch<TX, TY, TZ, PF, BETA, S, T, PS>);
    }
    template <bool PS>
    static const void *fn(bool beta, int s)
    {
        if (beta)
            return s == 1 ? fn<true, 1, PS>() : s == 2 ? fn<true, 2, PS>() : s == 3 ? fn<true, 3, PS>() : fn<true, 4, PS>();
        return s == 1 ? fn<false, 1, PS>() : s == 2 ? fn<false, 2, PS>() : s == 3 ? fn<false, 3, PS>() : fn<false, 4, PS>();
    }
    static hipError_t launch(bool beta, int s, bool persist, int grid, const void *desc, const int2 *tiles, int ntiles,
                             hipStream_t st)
    {
        const BatchDesc<T> *d = static_cast<const BatchDesc<T> *>(desc);
        const dim3 g(grid), blk(B::NT);
        const size_t l = lds(s);
#define EWPT_BATCH_LAUNCH(BT, SS, PS) stage_brick_batch<TX, TY, TZ, PF, BT, SS, T, PS><<<g, blk, l, st>>>(d, tiles, ntiles)
#define EWPT_BATCH_STAGES(BT, PS)                                   \
    do {                                                            \
        if (s == 1) EWPT_BATCH_LAUNCH(BT, 1, PS);                   \
        else if (s == 2) EWPT_BATCH_LAUNCH(BT, 2, PS);              \
        else if (s == 3) EWPT_BATCH_LAUNCH(BT, 3, PS);              \
        else EWPT_BATCH_LAUNCH(BT, 4, PS);                          \
    } while (0)
        if (beta) {
            if (persist) EWPT_BATCH_STAGES(true, true); else EWPT_BATCH_STAGES(true, false);
        } else {
            if (persist) EWPT_BATCH_STAGES(false, true); else EWPT_BATCH_STAGES(false, false);
        }
#undef EWPT_BATCH_STAGES
#undef EWPT_BATCH_LAUNCH
        return hipGetLastError();
    }
    // set the dynamic-LDS limit of every kernel; for the persistent form return
    // (resident workgroups per CU of the most demanding stage) x CUs, a multiple of 8
    template <bool PS>
    static int prepare(bool beta, int device)
    {
        int ncu = 0, best = 1 << 30;
        if (hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, device) != hipSuccess) return 0;
        for (int s = 1; s <= 4; s++) {
            int nb = 0;
            if (hipFuncSetAttribute(fn<PS>(beta, s), hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds(s)) !=
                    hipSuccess ||
                hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, fn<PS>(beta, s), B::NT, lds(s)) != hipSuccess)
                return 0;
            best = std::min(best, std::max(nb, 1));
        }
        return ((best * ncu + 7) / 8) * 8;
    }
};

}  // namespace

struct BatchPlan {
    int variant = 0, order = 0;
    bool f32 = false, persist = false;
    int tx = 0, ty = 0, tz = 0;
    int n = 0;
    int ntiles = 0;
    int grid[2] = {0, 0};   // [beta]
    int2 *dtiles = nullptr;
    void *ddesc = nullptr;   // [2][4][n] BatchDesc<T>
    void *hdesc = nullptr;   // pinned staging copy
    size_t desc_bytes = 0;
    int device = 0;
};

// batch variant = brick variant (shape % 100, order / 100 % 10) + 1000 for the
// persistent launch
bool batch_variant_fits(int variant, bool f32, const ewpt_lattice &L)
{
    int tx, ty, tz;
    if (variant < 0 || variant >= 2000 || (variant % 1000) / 100 > 9 || !tile_of(variant % 100, f32, tx, ty, tz))
        return false;
    return L.nx % tx == 0 && L.ny % ty == 0 && L.nz % tz == 0 && (long)L.nx * L.ny * L.nz * 8 < (1L << 31);
}

void batch_plan_destroy(BatchPlan *p)
{
    if (!p) return;
    if (p->dtiles) (void)hipFree(p->dtiles);
    if (p->ddesc) (void)hipFree(p->ddesc);
    if (p->hdesc) (void)hipHostFree(p->hdesc);
    delete p;
}

// the kernels are instantiated for these (shape, prefetch) pairs only
template <typename T, typename F>
auto with_kern(int v, F &&f)
{
    if (v == 11) return f(Kern<32, 4, 4, 4432, T>{});
    if constexpr (std::is_same_v<T, float>) {
        if (v == 15) return f(Kern<32, 4, 4, 704432, T>{});
        if (v == 31) return f(Kern<32, 4, 4, 1010704443, T>{});
        if (v == 33) return f(Kern<32, 4, 4, 1220704443, T>{});
    }
    if constexpr (std::is_same_v<T, double>) {
        if (v == 2) return f(Kern<8, 8, 8, 1, T>{});
        if (v == 32) return f(Kern<16, 4, 4, 211704432, T>{});
        if (v == 34) return f(Kern<16, 4, 4, 311704433, T>{});
    }
    return f(Kern<16, 8, 4, 4432, T>{});
}

BatchPlan *batch_plan_create(int variant, bool f32, const std::vector<ewpt_lattice> &lats, int device,
                             const char **what)
{
    BatchPlan *p = new (std::nothrow) BatchPlan();
    if (!p) { *what = "batch plan: out of host memory"; return nullptr; }
    p->variant = variant % 100;
    p->order = (variant % 1000) / 100;
    p->persist = variant >= 1000;
    p->f32 = f32;
    p->n = (int)lats.size();
    p->device = device;
    if (!tile_of(p->variant, f32, p->tx, p->ty, p->tz)) { *what = "batch plan: no such variant"; delete p; return nullptr; }
    std::vector<int2> tiles;
    for (int l = 0; l < p->n; l++) {
        if (!batch_variant_fits(variant, f32, lats[l])) {
            *what = "batch plan: a lattice does not tile with this variant";
            delete p;
            return nullptr;
        }
        const int nb = (lats[l].nx / p->tx) * (lats[l].ny / p->ty) * (lats[l].nz / p->tz);
        for (int b = 0; b < nb; b++) tiles.push_back(make_int2(l, b));
    }
    p->ntiles = (int)tiles.size();
    const size_t one = f32 ? sizeof(BatchDesc<float>) : sizeof(BatchDesc<double>);
    p->desc_bytes = 8 * (size_t)p->n * one;
    if (hipMalloc(reinterpret_cast<void **>(&p->dtiles), tiles.size() * sizeof(int2)) != hipSuccess ||
        hipMalloc(&p->ddesc, p->desc_bytes) != hipSuccess || hipHostMalloc(&p->hdesc, p->desc_bytes) != hipSuccess ||
        hipMemcpy(p->dtiles, tiles.data(), tiles.size() * sizeof(int2), hipMemcpyHostToDevice) != hipSuccess) {
        *what = "batch plan: device allocation";
        batch_plan_destroy(p);
        return nullptr;
    }
    for (int beta = 0; beta < 2; beta++) {
        auto prep = [&](auto k) {
            return p->persist ? decltype(k)::template prepare<true>(beta, device)
                              : decltype(k)::template prepare<false>(beta, device);
        };
        const int g = f32 ? with_kern<float>(p->variant, prep) : with_kern<double>(p->variant, prep);
        if (g <= 0) { *what = "batch plan: kernel attributes / occupancy query"; batch_plan_destroy(p); return nullptr; }
        p->grid[beta] = p->persist ? std::min(g, ((p->ntiles + 7) / 8) * 8) : ((p->ntiles + 7) / 8) * 8;
    }
    return p;
}

template <typename T>
static void fill_desc(const BatchPlan *p, const BatchLat<T> *lats, BatchDesc<T> *d)
{
    for (int par = 0; par < 2; par++)
        for (int s = 1; s <= 4; s++)
            for (int l = 0; l < p->n; l++) {
                BatchDesc<T> &e = d[((size_t)par * 4 + (s - 1)) * p->n + l];
                std::memset(&e, 0, sizeof e);
                const BatchLat<T> &L = lats[l];
                e.b = L.b[par][s - 1];
                const ewpt_lattice &G = L.L;
                e.g.nx = G.nx; e.g.ny = G.ny; e.g.nz = G.nz;
                e.g.nbx = G.nx / p->tx; e.g.nby = G.ny / p->ty; e.g.nbz = G.nz / p->tz;
                e.g.xcd_map = 0;
                e.g.order = p->order;
                e.g.inv60dx = 1.0 / (60.0 * G.dx);
                e.g.invdx2 = 1.0 / (G.dx * G.dx);
                e.P = L.P;
                e.dt = T(L.dt);
            }
}

int batch_plan_set(BatchPlan *p, const void *lats, hipStream_t st)
{
    if (p->f32) fill_desc<float>(p, static_cast<const BatchLat<float> *>(lats), static_cast<BatchDesc<float> *>(p->hdesc));
    else fill_desc<double>(p, static_cast<const BatchLat<double> *>(lats), static_cast<BatchDesc<double> *>(p->hdesc));
    if (hipMemcpyAsync(p->ddesc, p->hdesc, p->desc_bytes, hipMemcpyHostToDevice, st) != hipSuccess) return -1;
    // the staging copy is rewritten by the next call: wait for this one
    return hipStreamSynchronize(st) == hipSuccess ? 0 : -1;
}

hipError_t batch_plan_launch(const BatchPlan *p, int parity, int stage, bool beta, hipStream_t st)
{
    const size_t one = p->f32 ? sizeof(BatchDesc<float>) : sizeof(BatchDesc<double>);
    const void *d = static_cast<const char *>(p->ddesc) + ((size_t)parity * 4 + (stage - 1)) * p->n * one;
    const int grid = p->grid[beta ? 1 : 0];
    if (p->f32)
        return with_kern<float>(p->variant, [&](auto k) {
            return decltype(k)::launch(beta, stage, p->persist, grid, d, p->dtiles, p->ntiles, st);
        });
    return with_kern<double>(p->variant, [&](auto k) {
        return decltype(k)::launch(beta, stage, p->persist, grid, d, p->dtiles, p->ntiles, st);
    });
}

int batch_plan_tiles(const BatchPlan *p) { return p->ntiles; }

}  // namespace ewpt
