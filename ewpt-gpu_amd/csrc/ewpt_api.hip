// ewpt_api.hip -- the C ABI of libewpt (declared in include/ewpt.h).
//
// Replaces the device-side surface of run() in gpu-kernels-rk4/main_rk4.jl:
// allocation (5176-5299), the RK4 evolve loop (6146-6441) and the energy
// reductions (5700-5815, 6449-6560).  The context owns every device buffer
// and one HIP stream; all entry points are synchronous except
// ewpt_rhs_device.
#include <cmath>
#include <cstdio>
#include <cstring>
#include <algorithm>
#include <map>
#include <mutex>
#include <new>
#include <string>
#include <vector>

#include "ewpt_internal.h"

using namespace ewpt;

struct ewpt_ctx {
    int device = 0;
    int precision = EWPT_FP64;
    int variant = 0;
    ewpt_lattice lat{};
    long n3 = 0;            // channel stride in elements (sites, plus ghost planes for a slab)
    hipStream_t stream = nullptr;
    // z-slab contexts (ewpt_create_slab): HALO ghost planes below and above
    // the interior in every array, the send buffers of the halo exchange and
    // the events that order it between contexts
    int zg = 0;
    long plane = 0;         // nx * ny
    void *halo_lo = nullptr, *halo_hi = nullptr;
    hipEvent_t ev_packed = nullptr, ev_unpacked = nullptr;
    // FP64 path: state ping-pong and the three stored stage derivatives.
    // Mixed path: y is an FP64 view of the state (for state I/O and the
    // FP64 reductions), rebuilt from the FP32 pair when view_valid is false.
    double *y = nullptr, *ynext = nullptr, *k[3] = {nullptr, nullptr, nullptr};
    float *fhi = nullptr, *flo = nullptr, *fhi_n = nullptr, *flo_n = nullptr;
    float *fk[3] = {nullptr, nullptr, nullptr};
    // stage-input buffer of the stored-stage-input variants
    // (brick_stored_inputs), allocated when such a variant is first used: u2
    // and u4 live here, u3 in the next-state buffer, which stage 4 only
    // writes after stage 3 has read it
    double *u = nullptr;
    float *fu = nullptr;
    bool stored_ok = true;   // false: no room for it, AUTO skips those variants
    bool view_valid = true;
    // host-buffer RHS path (allocated on first use)
    double *rhs_in = nullptr, *rhs_out = nullptr;
    // reductions
    double *part = nullptr, *res = nullptr;
    long nred = 0;
    // partial sums of the diagnostic stage 1 (rows x DIAG_SLOTS, then the
    // DIAG_SLOTS reduced values), allocated on first use
    double *diag = nullptr;
    bool has_state = false;
    unsigned long long bytes = 0;
    std::string err;
};

namespace {

constexpr int NRED_SLOTS = 16;

int fail(ewpt_ctx *c, int code, const char *what, hipError_t e = hipSuccess)
{
    if (c) {
        char buf[512];
        if (e != hipSuccess)
            std::snprintf(buf, sizeof buf, "%s: %s (%s)", what, hipGetErrorString(e), hipGetErrorName(e));
        else
            std::snprintf(buf, sizeof buf, "%s", what);
        c->err = buf;
    }
    return code;
}

#define HIPCHK(ctx, call, what)                                 \
    do {                                                        \
        hipError_t e_ = (call);                                 \
        if (e_ != hipSuccess) return fail(ctx, EWPT_ERR_HIP, what, e_); \
    } while (0)

template <typename T>
int dev_alloc(ewpt_ctx *c, T **p, size_t n, const char *what)
{
    hipError_t e = hipMalloc(reinterpret_cast<void **>(p), n * sizeof(T));
    if (e != hipSuccess) {
        *p = nullptr;
        (void)hipGetLastError();
        return fail(c, e == hipErrorOutOfMemory ? EWPT_ERR_OOM : EWPT_ERR_HIP, what, e);
    }
    c->bytes += n * sizeof(T);
    return EWPT_OK;
}

void release(ewpt_ctx *c)
{
    double **bufs[] = {&c->y, &c->ynext, &c->k[0], &c->k[1], &c->k[2], &c->rhs_in, &c->rhs_out, &c->part, &c->res,
                       &c->u, &c->diag};
    for (double **b : bufs)
        if (*b) { (void)hipFree(*b); *b = nullptr; }
    float **fbufs[] = {&c->fhi, &c->flo, &c->fhi_n, &c->flo_n, &c->fk[0], &c->fk[1], &c->fk[2], &c->fu};
    for (float **b : fbufs)
        if (*b) { (void)hipFree(*b); *b = nullptr; }
    void **hbufs[] = {&c->halo_lo, &c->halo_hi};
    for (void **b : hbufs)
        if (*b) { (void)hipFree(*b); *b = nullptr; }
    if (c->ev_packed) { (void)hipEventDestroy(c->ev_packed); c->ev_packed = nullptr; }
    if (c->ev_unpacked) { (void)hipEventDestroy(c->ev_unpacked); c->ev_unpacked = nullptr; }
    if (c->stream) { (void)hipStreamDestroy(c->stream); c->stream = nullptr; }
}

bool params_ok(const ewpt_params *p)
{
    if (!p) return false;
    const double v[] = {p->gw, p->gy, p->gp2, p->beta_w, p->beta_y, p->vev, p->lambda, p->dt, p->jex, p->gamma};
    for (double x : v)
        if (!std::isfinite(x)) return false;
    return true;
}

bool lattice_ok(const ewpt_lattice *L)
{
    return L && L->nx >= 8 && L->ny >= 8 && L->nz >= 8 && std::isfinite(L->dx) && L->dx > 0.0 &&
           (long)L->nx * L->ny <= (1L << 30) && (long)L->nx * L->ny * (long)L->nz <= (1L << 34);
}

// stage inputs of the stored-stage-input variants: u2 -> U, u3 -> the
// next-state buffer, u4 -> U (stage 4 then writes the next state)
template <typename T>
void set_stage_inputs(StageBufs<T> &b, int stage, T *U, T *next)
{
    b.u = stage == 3 ? next : stage == 1 ? nullptr : U;
    b.unext = stage == 2 ? next : stage == 4 ? nullptr : U;
}

StageBufs<double> bufs_for(ewpt_ctx *c, int stage, bool ustaged = false)
{
    StageBufs<double> b{};
    b.y = c->y;
    b.k1 = c->k[0];
    b.k2 = c->k[1];
    b.k3 = c->k[2];
    b.n3 = c->n3;
    if (stage < 4) b.kout = c->k[stage - 1];
    else b.ynew = c->ynext;
    if (ustaged) set_stage_inputs(b, stage, c->u, c->ynext);
    b.zghost = c->zg;
    return b;
}

int ensure_stage_input_buffer(ewpt_ctx *c, int variant)
{
    if (!brick_stored_inputs(variant)) return EWPT_OK;
    const size_t n = NCH * (size_t)c->n3;
    if (c->precision == EWPT_MIXED) return c->fu ? EWPT_OK : dev_alloc(c, &c->fu, n, "alloc stage input (fp32)");
    return c->u ? EWPT_OK : dev_alloc(c, &c->u, n, "alloc stage input");
}

int set_device(ewpt_ctx *c) { HIPCHK(c, hipSetDevice(c->device), "hipSetDevice"); return EWPT_OK; }

// sites of the lattice (a slab's interior); equals n3 without ghost planes
long sites(const ewpt_ctx *c) { return c->plane * c->lat.nz; }

// copy nch channels of `count` elements between arrays of different channel
// strides (a slab's interior and a dense array)
hipError_t copy_channels(void *dst, long dstride, long doff, const void *src, long sstride, long soff, int nch,
                         long count, size_t elem, hipMemcpyKind kind, hipStream_t st)
{
    if (dstride == count && sstride == count && doff == 0 && soff == 0)
        return hipMemcpyAsync(dst, src, nch * count * elem, kind, st);
    for (int ch = 0; ch < nch; ch++) {
        hipError_t e = hipMemcpyAsync(static_cast<char *>(dst) + (ch * dstride + doff) * elem,
                                      static_cast<const char *>(src) + (ch * sstride + soff) * elem, count * elem,
                                      kind, st);
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
}

// entry points that need the whole periodic lattice on one context
int no_slab(ewpt_ctx *c, const char *what)
{
    return fail(c, EWPT_ERR_UNSUPPORTED, (std::string(what) + ": not available on a slab context").c_str());
}

bool mixed(const ewpt_ctx *c) { return c->precision == EWPT_MIXED; }

// mixed path: refresh the FP64 view from the FP32 pair
int ensure_view(ewpt_ctx *c)
{
    if (!mixed(c) || c->view_valid) return EWPT_OK;
    HIPCHK(c, launch_join(c->fhi, c->flo, c->y, NCH * c->n3, c->stream), "join state");
    c->view_valid = true;
    return EWPT_OK;
}

// mixed path: the FP64 view was written; split it into the FP32 pair
int push_view(ewpt_ctx *c)
{
    if (!mixed(c)) return EWPT_OK;
    HIPCHK(c, launch_split(c->y, c->fhi, c->flo, NCH * c->n3, c->stream), "split state");
    c->view_valid = true;
    return EWPT_OK;
}

// variant actually launched: AUTO takes the first brick variant that tiles
// the lattice from the measured ranking (512^3 on MI355X, FP64: 334 124.3,
// 332 140.6, 324 170.3, 320 181.2, 311 184.4, 310 187.0 ms/step; mixed:
// 333 75.8, 331 82.7, 315 89.4, 311 96.3, 211 103.2 ms/step), else the
// gather kernel.  The stored-stage-input variants (334, 333, ...) hold one
// more state-sized buffer; without room for it AUTO skips them.
int resolved_variant(const ewpt_ctx *c)
{
    if (c->variant != VARIANT_AUTO) return c->variant;
    if (mixed(c)) {
        for (int v : {333, 331, 315, 324, 311, 211, 310, 210, 325})
            if ((c->stored_ok || !brick_stored_inputs(v)) && brick_supported_f32(v, c->lat)) return v;
        return -1;
    }
    for (int v : {334, 332, 324, 325, 320, 311, 310, 210, 211, 2})
        if ((c->stored_ok || !brick_stored_inputs(v)) && brick_supported(v, c->lat)) return v;
    return VARIANT_GATHER;
}

// the mixed path's stage buffers (FP32 pair, k's, stored stage input)
StageBufs<float> fbufs_for(ewpt_ctx *c, int stage, bool us)
{
    StageBufs<float> b{};
    b.y = c->fhi;
    b.ylo = c->flo;
    b.k1 = c->fk[0];
    b.k2 = c->fk[1];
    b.k3 = c->fk[2];
    b.n3 = c->n3;
    if (stage < 4) b.kout = c->fk[stage - 1];
    else { b.ynew = c->fhi_n; b.ylo_new = c->flo_n; }
    if (us) set_stage_inputs(b, stage, c->fu, c->fhi_n);
    b.zghost = c->zg;
    return b;
}

int launch_stage_kernel(ewpt_ctx *c, int stage, const ewpt_params *p)
{
    const Phys P = phys_of(*p);
    const int v = resolved_variant(c);
    const bool us = brick_stored_inputs(v);
    if (us && stage == 1) {
        const int rc = ensure_stage_input_buffer(c, v);
        if (rc) return rc;
    }
    if (mixed(c)) {
        HIPCHK(c, launch_stage_brick_f32(v, stage, fbufs_for(c, stage, us), c->lat, P, p->dt, c->stream),
               "stage launch");
        if (stage == 4) {
            std::swap(c->fhi, c->fhi_n);
            std::swap(c->flo, c->flo_n);
            c->view_valid = false;
        }
        return EWPT_OK;
    }
    if (v == VARIANT_GATHER)
        HIPCHK(c, launch_stage_gather(stage, bufs_for(c, stage), c->lat, P, p->dt, c->stream), "stage launch");
    else
        HIPCHK(c, launch_stage_brick(v, stage, bufs_for(c, stage, us), c->lat, P, p->dt, c->stream), "stage launch");
    if (stage == 4) std::swap(c->y, c->ynext);
    return EWPT_OK;
}

// ---- halo exchange of z-slab contexts
// what = 0: the state (mixed: its FP32 high word), which stage 1 reads at
// halo sites.  what = s in 1..3, before stage s+1: what that stage reads at
// halo sites -- k_s for the variants that form their input on load from the
// state and the k's (ewpt_stage.h; the state's and k_1..k_{s-1}'s ghost
// planes are still current), or the stored stage input u_{s+1} (36
// channels: u2 and u4 in the stage-input buffer, u3 in the next-state
// buffer) for the variants that store it; their other reads are interior.
bool halo_is_stage_input(const ewpt_ctx *c, int what)
{
    return what > 0 && brick_stored_inputs(resolved_variant(c));
}

void *halo_array(ewpt_ctx *c, int what, int &nch)
{
    if (halo_is_stage_input(c, what)) {
        nch = NCH;
        if (mixed(c)) return what == 2 ? static_cast<void *>(c->fhi_n) : static_cast<void *>(c->fu);
        return what == 2 ? static_cast<void *>(c->ynext) : static_cast<void *>(c->u);
    }
    nch = what == 0 ? NCH : NK;
    if (mixed(c)) return what == 0 ? static_cast<void *>(c->fhi) : static_cast<void *>(c->fk[what - 1]);
    return what == 0 ? static_cast<void *>(c->y) : static_cast<void *>(c->k[what - 1]);
}

size_t halo_bytes(const ewpt_ctx *c, int what)
{
    const int nch = what == 0 || halo_is_stage_input(c, what) ? NCH : NK;
    return (size_t)nch * HALO * c->plane * (mixed(c) ? sizeof(float) : sizeof(double));
}

int halo_pack_into(ewpt_ctx *c, int what, void *lo, void *hi)
{
    int nch;
    void *a = halo_array(c, what, nch);
    if (!a) return fail(c, EWPT_ERR_STATE, "halo pack of a stage input before stage 1 wrote it");
    hipError_t e = mixed(c) ? launch_halo_pack(static_cast<const float *>(a), c->n3, c->plane, c->lat.nz, nch,
                                               static_cast<float *>(lo), static_cast<float *>(hi), c->stream)
                            : launch_halo_pack(static_cast<const double *>(a), c->n3, c->plane, c->lat.nz, nch,
                                               static_cast<double *>(lo), static_cast<double *>(hi), c->stream);
    HIPCHK(c, e, "halo pack");
    return EWPT_OK;
}

int halo_unpack_from(ewpt_ctx *c, int what, const void *lo, const void *hi)
{
    int nch;
    void *a = halo_array(c, what, nch);
    if (!a) return fail(c, EWPT_ERR_STATE, "halo unpack of a stage input before stage 1 wrote it");
    hipError_t e = mixed(c) ? launch_halo_unpack(static_cast<float *>(a), c->n3, c->plane, c->lat.nz, nch,
                                                 static_cast<const float *>(lo), static_cast<const float *>(hi),
                                                 c->stream)
                            : launch_halo_unpack(static_cast<double *>(a), c->n3, c->plane, c->lat.nz, nch,
                                                 static_cast<const double *>(lo), static_cast<const double *>(hi),
                                                 c->stream);
    HIPCHK(c, e, "halo unpack");
    return EWPT_OK;
}

// One exchange over a ring of slabs on one device (slab i's z-neighbours
// are i-1 and i+1 mod n; n = 1 is periodic in z): every slab packs into its
// own send buffers, then unpacks its neighbours'.  Events order the streams:
// an unpack waits for the neighbours' packs, and a pack waits for the
// neighbours' unpacks of the previous exchange, which read its buffers.
int ring_exchange(ewpt_ctx *const *sl, int n, int what)
{
    for (int i = 0; i < n; i++) {
        ewpt_ctx *c = sl[i], *lo = sl[(i + n - 1) % n], *hi = sl[(i + 1) % n];
        if (n > 1) {
            HIPCHK(c, hipStreamWaitEvent(c->stream, lo->ev_unpacked, 0), "wait unpack");
            HIPCHK(c, hipStreamWaitEvent(c->stream, hi->ev_unpacked, 0), "wait unpack");
        }
        if (int rc = halo_pack_into(c, what, c->halo_lo, c->halo_hi)) return rc;
        HIPCHK(c, hipEventRecord(c->ev_packed, c->stream), "record pack");
    }
    for (int i = 0; i < n; i++) {
        ewpt_ctx *c = sl[i], *lo = sl[(i + n - 1) % n], *hi = sl[(i + 1) % n];
        if (n > 1) {
            HIPCHK(c, hipStreamWaitEvent(c->stream, lo->ev_packed, 0), "wait pack");
            HIPCHK(c, hipStreamWaitEvent(c->stream, hi->ev_packed, 0), "wait pack");
        }
        // the ghost planes below come from the lower neighbour's top planes
        if (int rc = halo_unpack_from(c, what, lo->halo_hi, hi->halo_lo)) return rc;
        HIPCHK(c, hipEventRecord(c->ev_unpacked, c->stream), "record unpack");
    }
    return EWPT_OK;
}

// One stage of a step; a slab context on its own is periodic in z and
// exchanges with itself first.
int launch_step_stage(ewpt_ctx *c, int stage, const ewpt_params *p)
{
    if (c->zg) {
        ewpt_ctx *self[1] = {c};
        if (int rc = ring_exchange(self, 1, stage - 1)) return rc;
    }
    return launch_stage_kernel(c, stage, p);
}

}  // namespace

extern "C" {

const char *ewpt_version(void) { return "libewpt 0.1 (gfx950)"; }

const char *ewpt_strerror(int code)
{
    switch (code) {
    case EWPT_OK: return "ok";
    case EWPT_ERR_INVAL: return "invalid argument";
    case EWPT_ERR_HIP: return "HIP runtime error";
    case EWPT_ERR_OOM: return "device out of memory";
    case EWPT_ERR_NONFINITE: return "non-finite value in state";
    case EWPT_ERR_STATE: return "no state set";
    case EWPT_ERR_UNSUPPORTED: return "unsupported request";
    default: return "unknown error";
    }
}

const char *ewpt_last_error(ewpt_ctx *ctx) { return ctx ? ctx->err.c_str() : ""; }

namespace {

// zg = 0: periodic lattice; zg = HALO: z-slab with ghost planes
int create_ctx(ewpt_ctx **out, const ewpt_lattice *lat, int precision, int device, int zg)
{
    if (!out) return EWPT_ERR_INVAL;
    *out = nullptr;
    if (!lattice_ok(lat) || (precision != EWPT_FP64 && precision != EWPT_MIXED) || device < 0)
        return EWPT_ERR_INVAL;
    // stage kernels address a channel with 32-bit byte offsets
    if (zg && (long)lat->nx * lat->ny * (lat->nz + 2L * zg) * 8 >= (1L << 31)) return EWPT_ERR_UNSUPPORTED;
    if (precision == EWPT_MIXED && !brick_supported_f32(210, *lat) && !brick_supported_f32(211, *lat) &&
        !brick_supported_f32(325, *lat))
        return EWPT_ERR_UNSUPPORTED;   // the mixed path has LDS-brick kernels only
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || device >= ndev) {
        (void)hipGetLastError();
        return EWPT_ERR_HIP;
    }
    ewpt_ctx *c = new (std::nothrow) ewpt_ctx();
    if (!c) return EWPT_ERR_OOM;
    c->device = device;
    c->precision = precision;
    c->lat = *lat;
    c->zg = zg;
    c->plane = (long)lat->nx * lat->ny;
    c->n3 = c->plane * (lat->nz + 2L * zg);
    int rc = set_device(c);
    if (rc == EWPT_OK && hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess)
        rc = fail(c, EWPT_ERR_HIP, "hipStreamCreate");
    const size_t n3 = (size_t)c->n3;
    if (rc == EWPT_OK) rc = dev_alloc(c, &c->y, NCH * n3, "alloc y");
    if (precision == EWPT_FP64) {
        if (rc == EWPT_OK) rc = dev_alloc(c, &c->ynext, NCH * n3, "alloc y'");
        for (int s = 0; s < 3 && rc == EWPT_OK; s++) rc = dev_alloc(c, &c->k[s], NK * n3, "alloc k");
    } else {
        float **st[] = {&c->fhi, &c->flo, &c->fhi_n, &c->flo_n};
        for (float **b : st)
            if (rc == EWPT_OK) rc = dev_alloc(c, b, NCH * n3, "alloc y (fp32)");
        for (int s = 0; s < 3 && rc == EWPT_OK; s++) rc = dev_alloc(c, &c->fk[s], NK * n3, "alloc k (fp32)");
    }
    c->nred = reduce_blocks(c->lat);
    if (rc == EWPT_OK) rc = dev_alloc(c, &c->part, (size_t)c->nred * NRED_SLOTS, "alloc partials");
    if (rc == EWPT_OK) rc = dev_alloc(c, &c->res, NRED_SLOTS, "alloc results");
    if (rc == EWPT_OK && zg) {
        const int v = resolved_variant(c);
        if (v < 0 || v == VARIANT_GATHER) rc = fail(c, EWPT_ERR_UNSUPPORTED, "no brick variant tiles this slab");
        const size_t hb = (size_t)NCH * HALO * c->plane * (precision == EWPT_MIXED ? sizeof(float) : sizeof(double));
        unsigned char **hp[] = {reinterpret_cast<unsigned char **>(&c->halo_lo),
                                reinterpret_cast<unsigned char **>(&c->halo_hi)};
        for (unsigned char **b : hp)
            if (rc == EWPT_OK) rc = dev_alloc(c, b, hb, "alloc halo buffers");
        if (rc == EWPT_OK && (hipEventCreateWithFlags(&c->ev_packed, hipEventDisableTiming) != hipSuccess ||
                              hipEventCreateWithFlags(&c->ev_unpacked, hipEventDisableTiming) != hipSuccess))
            rc = fail(c, EWPT_ERR_HIP, "hipEventCreate");
    }
    if (rc == EWPT_OK && brick_stored_inputs(resolved_variant(c))) {
        // AUTO's first choice wants the stage-input buffer: take it now, or
        // fall back to the variants that do without
        rc = ensure_stage_input_buffer(c, resolved_variant(c));
        if (rc == EWPT_ERR_OOM) {
            c->stored_ok = false;
            c->err.clear();
            rc = EWPT_OK;
        }
    }
    if (rc != EWPT_OK) {
        release(c);
        delete c;
        return rc;
    }
    *out = c;
    return EWPT_OK;
}

}  // namespace

int ewpt_create(ewpt_ctx **out, const ewpt_lattice *lat, int precision, int device)
{
    return create_ctx(out, lat, precision, device, 0);
}

int ewpt_create_slab(ewpt_ctx **out, const ewpt_lattice *local, int precision, int device)
{
    return create_ctx(out, local, precision, device, HALO);
}

int ewpt_destroy(ewpt_ctx *ctx)
{
    if (!ctx) return EWPT_ERR_INVAL;
    (void)hipSetDevice(ctx->device);
    if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
    release(ctx);
    delete ctx;
    return EWPT_OK;
}

int ewpt_set_state(ewpt_ctx *ctx, const double *host)
{
    if (!ctx || !host) return EWPT_ERR_INVAL;
    int rc = set_device(ctx);
    if (rc) return rc;
    HIPCHK(ctx, copy_channels(ctx->y, ctx->n3, ctx->zg * ctx->plane, host, sites(ctx), 0, NCH, sites(ctx),
                              sizeof(double), hipMemcpyHostToDevice, ctx->stream),
           "set_state copy");
    if ((rc = push_view(ctx))) return rc;
    HIPCHK(ctx, hipStreamSynchronize(ctx->stream), "set_state sync");
    ctx->has_state = true;
    return EWPT_OK;
}

int ewpt_get_state(ewpt_ctx *ctx, double *host)
{
    if (!ctx || !host) return EWPT_ERR_INVAL;
    if (!ctx->has_state) return fail(ctx, EWPT_ERR_STATE, "get_state before set_state");
    int rc = set_device(ctx);
    if (rc) return rc;
    if ((rc = ensure_view(ctx))) return rc;
    HIPCHK(ctx, copy_channels(host, sites(ctx), 0, ctx->y, ctx->n3, ctx->zg * ctx->plane, NCH, sites(ctx),
                              sizeof(double), hipMemcpyDeviceToHost, ctx->stream),
           "get_state copy");
    HIPCHK(ctx, hipStreamSynchronize(ctx->stream), "get_state sync");
    return EWPT_OK;
}

int ewpt_set_state_device(ewpt_ctx *ctx, const double *dev_state)
{
    if (!ctx || !dev_state) return EWPT_ERR_INVAL;
    int rc = set_device(ctx);
    if (rc) return rc;
    HIPCHK(ctx, copy_channels(ctx->y, ctx->n3, ctx->zg * ctx->plane, dev_state, sites(ctx), 0, NCH, sites(ctx),
                              sizeof(double), hipMemcpyDeviceToDevice, ctx->stream),
           "set_state_device copy");
    if ((rc = push_view(ctx))) return rc;
    HIPCHK(ctx, hipStreamSynchronize(ctx->stream), "set_state_device sync");
    ctx->has_state = true;
    return EWPT_OK;
}

int ewpt_get_state_device(ewpt_ctx *ctx, double *dev_state)
{
    if (!ctx || !dev_state) return EWPT_ERR_INVAL;
    if (!ctx->has_state) return fail(ctx, EWPT_ERR_STATE, "get_state_device before set_state");
    int rc = set_device(ctx);
    if (rc) return rc;
    if ((rc = ensure_view(ctx))) return rc;
    HIPCHK(ctx, copy_channels(dev_state, sites(ctx), 0, ctx->y, ctx->n3, ctx->zg * ctx->plane, NCH, sites(ctx),
                              sizeof(double), hipMemcpyDeviceToDevice, ctx->stream),
           "get_state_device copy");
    HIPCHK(ctx, hipStreamSynchronize(ctx->stream), "get_state_device sync");
    return EWPT_OK;
}

int ewpt_get_state_pair(ewpt_ctx *ctx, float *hi, float *lo)
{
    if (!ctx || !hi || !lo) return EWPT_ERR_INVAL;
    if (!mixed(ctx)) return fail(ctx, EWPT_ERR_INVAL, "state pair of an FP64 context");
    if (!ctx->has_state) return fail(ctx, EWPT_ERR_STATE, "get_state_pair before set_state");
    int rc = set_device(ctx);
    if (rc) return rc;
    const long off = ctx->zg * ctx->plane;
    HIPCHK(ctx, copy_channels(hi, sites(ctx), 0, ctx->fhi, ctx->n3, off, NCH, sites(ctx), sizeof(float),
                              hipMemcpyDeviceToHost, ctx->stream),
           "get_state_pair copy");
    HIPCHK(ctx, copy_channels(lo, sites(ctx), 0, ctx->flo, ctx->n3, off, NCH, sites(ctx), sizeof(float),
                              hipMemcpyDeviceToHost, ctx->stream),
           "get_state_pair copy");
    HIPCHK(ctx, hipStreamSynchronize(ctx->stream), "get_state_pair sync");
    return EWPT_OK;
}

int ewpt_set_state_pair(ewpt_ctx *ctx, const float *hi, const float *lo)
{
    if (!ctx || !hi || !lo) return EWPT_ERR_INVAL;
    if (!mixed(ctx)) return fail(ctx, EWPT_ERR_INVAL, "state pair of an FP64 context");
    int rc = set_device(ctx);
    if (rc) return rc;
    const long off = ctx->zg * ctx->plane;
    HIPCHK(ctx, copy_channels(ctx->fhi, ctx->n3, off, hi, sites(ctx), 0, NCH, sites(ctx), sizeof(float),
                              hipMemcpyHostToDevice, ctx->stream),
           "set_state_pair copy");
    HIPCHK(ctx, copy_channels(ctx->flo, ctx->n3, off, lo, sites(ctx), 0, NCH, sites(ctx), sizeof(float),
                              hipMemcpyHostToDevice, ctx->stream),
           "set_state_pair copy");
    HIPCHK(ctx, hipStreamSynchronize(ctx->stream), "set_state_pair sync");
    ctx->view_valid = false;   // the FP64 view is joined from the pair when next read
    ctx->has_state = true;
    return EWPT_OK;
}

int ewpt_copy_probe(const double *dev_src, double *dev_dst, long n, int reps, void *stream, double *ms)
{
    if ((!dev_src) != (!dev_dst) || n < 1 || reps < 1 || !ms) return EWPT_ERR_INVAL;
    if (!dev_src) {
        // own buffers
        double *a = nullptr, *b = nullptr;
        if (hipMalloc(reinterpret_cast<void **>(&a), n * sizeof(double)) != hipSuccess) {
            (void)hipGetLastError();
            return EWPT_ERR_OOM;
        }
        if (hipMalloc(reinterpret_cast<void **>(&b), n * sizeof(double)) != hipSuccess) {
            (void)hipGetLastError();
            (void)hipFree(a);
            return EWPT_ERR_OOM;
        }
        int rc = hipMemsetAsync(a, 0, n * sizeof(double), static_cast<hipStream_t>(stream)) == hipSuccess
                     ? ewpt_copy_probe(a, b, n, reps, stream, ms)
                     : EWPT_ERR_HIP;
        (void)hipStreamSynchronize(static_cast<hipStream_t>(stream));
        (void)hipFree(a);
        (void)hipFree(b);
        return rc;
    }
    hipStream_t st = static_cast<hipStream_t>(stream);
    hipEvent_t e0, e1;
    if (hipEventCreate(&e0) != hipSuccess) return EWPT_ERR_HIP;
    if (hipEventCreate(&e1) != hipSuccess) { (void)hipEventDestroy(e0); return EWPT_ERR_HIP; }
    hipError_t e = launch_copy(dev_src, dev_dst, n, st);   // untimed first launch
    if (e == hipSuccess) e = hipEventRecord(e0, st);
    for (int r = 0; r < reps && e == hipSuccess; r++) e = launch_copy(dev_src, dev_dst, n, st);
    if (e == hipSuccess) e = hipEventRecord(e1, st);
    if (e == hipSuccess) e = hipEventSynchronize(e1);
    float f = 0.f;
    if (e == hipSuccess) e = hipEventElapsedTime(&f, e0, e1);
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    *ms = f / reps;
    return e == hipSuccess ? EWPT_OK : EWPT_ERR_HIP;
}

int ewpt_copy_state(ewpt_ctx *dst, ewpt_ctx *src)
{
    if (!dst || !src || dst == src) return EWPT_ERR_INVAL;
    if (dst->device != src->device || dst->precision != src->precision || dst->lat.nx != src->lat.nx ||
        dst->lat.ny != src->lat.ny || dst->lat.nz != src->lat.nz)
        return fail(dst, EWPT_ERR_INVAL, "copy_state: contexts differ in device, precision or lattice");
    if (!src->has_state) return fail(dst, EWPT_ERR_STATE, "copy_state: source has no state");
    int rc = set_device(src);
    if (rc) return rc;
    // interior only, so slab and periodic contexts of one shape copy both ways
    const long n = sites(src), so = src->zg * src->plane, d0 = dst->zg * dst->plane;
    const hipMemcpyKind dd = hipMemcpyDeviceToDevice;
    if (mixed(src)) {
        HIPCHK(dst, copy_channels(dst->fhi, dst->n3, d0, src->fhi, src->n3, so, NCH, n, sizeof(float), dd, src->stream),
               "copy hi");
        HIPCHK(dst, copy_channels(dst->flo, dst->n3, d0, src->flo, src->n3, so, NCH, n, sizeof(float), dd, src->stream),
               "copy lo");
        dst->view_valid = false;
    } else {
        HIPCHK(dst, copy_channels(dst->y, dst->n3, d0, src->y, src->n3, so, NCH, n, sizeof(double), dd, src->stream),
               "copy y");
    }
    HIPCHK(dst, hipStreamSynchronize(src->stream), "copy sync");
    dst->has_state = true;
    return EWPT_OK;
}

int ewpt_init_modes(ewpt_ctx *ctx, unsigned long long seed, int nmodes, int kmax, double amp, double vev,
                    double mom_scale, double *modes_out)
{
    if (!ctx || nmodes < 1 || nmodes > 64 || kmax < 0 || !std::isfinite(amp) || !std::isfinite(vev) ||
        !std::isfinite(mom_scale))
        return EWPT_ERR_INVAL;
    if (ctx->zg) return no_slab(ctx, "init_modes");
    int rc = set_device(ctx);
    if (rc) return rc;
    const size_t nt = (size_t)NCH * nmodes * 5;
    std::vector<double> t(nt);
    mode_table(seed, nmodes, kmax, amp, mom_scale, t.data());
    double *dt = nullptr;
    HIPCHK(ctx, hipMalloc(reinterpret_cast<void **>(&dt), nt * sizeof(double)), "alloc mode table");
    hipError_t e = hipMemcpyAsync(dt, t.data(), nt * sizeof(double), hipMemcpyHostToDevice, ctx->stream);
    if (e == hipSuccess) e = launch_fill_modes(ctx->y, ctx->lat, dt, nmodes, vev, ctx->stream);
    if (e == hipSuccess && mixed(ctx)) e = launch_split(ctx->y, ctx->fhi, ctx->flo, NCH * ctx->n3, ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    (void)hipFree(dt);
    if (e != hipSuccess) return fail(ctx, EWPT_ERR_HIP, "init_modes", e);
    if (modes_out) std::memcpy(modes_out, t.data(), nt * sizeof(double));
    ctx->has_state = true;
    return EWPT_OK;
}

int ewpt_init_bubbles(ewpt_ctx *ctx, const double *bubbles, int nb, int profile, double mH)
{
    if (!ctx || nb < 0 || nb > (1 << 20) || (nb > 0 && !bubbles) || (profile != 0 && profile != 1) ||
        !std::isfinite(mH) || !(mH > 0.0))
        return EWPT_ERR_INVAL;
    if (ctx->zg) return no_slab(ctx, "init_bubbles");
    for (long q = 0; q < 7L * nb; q++)
        if (!std::isfinite(bubbles[q])) return fail(ctx, EWPT_ERR_INVAL, "init_bubbles: non-finite bubble table");
    int rc = set_device(ctx);
    if (rc) return rc;
    double *db = nullptr;
    if (nb > 0) HIPCHK(ctx, hipMalloc(reinterpret_cast<void **>(&db), 7L * nb * sizeof(double)), "alloc bubbles");
    hipError_t e = hipMemsetAsync(ctx->y, 0, NCH * ctx->n3 * sizeof(double), ctx->stream);
    if (e == hipSuccess && nb > 0)
        e = hipMemcpyAsync(db, bubbles, 7L * nb * sizeof(double), hipMemcpyHostToDevice, ctx->stream);
    if (e == hipSuccess && nb > 0) e = launch_add_bubbles(ctx->y, ctx->lat, db, nb, profile, mH, ctx->stream);
    if (e == hipSuccess && mixed(ctx)) e = launch_split(ctx->y, ctx->fhi, ctx->flo, NCH * ctx->n3, ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    if (db) (void)hipFree(db);
    if (e != hipSuccess) return fail(ctx, EWPT_ERR_HIP, "init_bubbles", e);
    ctx->view_valid = true;
    ctx->has_state = true;
    return EWPT_OK;
}

int ewpt_init_thermal(ewpt_ctx *ctx, double T, double meff_sq, unsigned long long seed, int profile)
{
    if (!ctx || !std::isfinite(T) || !(T > 0.0) || !std::isfinite(meff_sq) || (profile != 0 && profile != 1))
        return EWPT_ERR_INVAL;
    if (ctx->zg) return no_slab(ctx, "init_thermal");
    int rc = set_device(ctx);
    if (rc) return rc;
    // the stored stage derivatives are free between steps: they hold the
    // 3 N^3 complex transform buffers (FP64: 20 N^3 doubles, mixed: 20 N^3
    // floats, both >= 6 N^3 doubles)
    void *work = mixed(ctx) ? static_cast<void *>(ctx->fk[0]) : static_cast<void *>(ctx->k[0]);
    HIPCHK(ctx, hipMemsetAsync(ctx->y, 0, NCH * ctx->n3 * sizeof(double), ctx->stream), "init_thermal clear");
    const char *what = "";
    if (thermal_state(ctx->y, ctx->lat, T, meff_sq, seed, profile, work, ctx->stream, &what) != 0) {
        (void)hipGetLastError();
        return fail(ctx, EWPT_ERR_HIP, what);
    }
    if ((rc = push_view(ctx))) return rc;
    HIPCHK(ctx, hipStreamSynchronize(ctx->stream), "init_thermal sync");
    ctx->has_state = true;
    return EWPT_OK;
}

int ewpt_init_thermal_gauss(ewpt_ctx *ctx, double T, double meff_sq, unsigned long long seed, const ewpt_params *p,
                            double stats[4])
{
    if (!ctx || !std::isfinite(T) || !(T > 0.0) || !std::isfinite(meff_sq) || !params_ok(p)) return EWPT_ERR_INVAL;
    if (ctx->zg) return no_slab(ctx, "init_thermal_gauss");
    int rc = set_device(ctx);
    if (rc) return rc;
    // free between steps: the stored derivatives and the next-state arrays
    void *work;
    double *wa, *wb;
    if (mixed(ctx)) {
        work = ctx->fk[0];                                  // 20 N^3 floats >= 3 N^3 complex doubles
        wa = reinterpret_cast<double *>(ctx->fhi_n);       // 36 N^3 floats >= 15 N^3 doubles
        wb = reinterpret_cast<double *>(ctx->flo_n);       // 36 N^3 floats >= 9 N^3 doubles
    } else {
        work = ctx->k[0];
        wa = ctx->ynext;
        wb = ctx->k[1];
    }
    HIPCHK(ctx, hipMemsetAsync(ctx->y, 0, NCH * ctx->n3 * sizeof(double), ctx->stream), "init_thermal_gauss clear");
    double st[4] = {0.0, 0.0, 0.0, 0.0};
    const char *what = "";
    if (thermal_gauss_state(ctx->y, ctx->lat, T, meff_sq, seed, p->gw, p->gy, 1e-13, 4000, work, wa, wb, ctx->stream,
                            st, &what) != 0) {
        (void)hipGetLastError();
        return fail(ctx, EWPT_ERR_HIP, what);
    }
    if ((rc = push_view(ctx))) return rc;
    HIPCHK(ctx, hipStreamSynchronize(ctx->stream), "init_thermal_gauss sync");
    ctx->has_state = true;
    if (stats)
        for (int q = 0; q < 4; q++) stats[q] = st[q];
    if (!std::isfinite(st[1])) return fail(ctx, EWPT_ERR_NONFINITE, "Gauss-law solve did not converge");
    return EWPT_OK;
}

namespace {

// One rk4_step! evaluation on device buffers: the stage-1 brick kernel of a
// variant that forms its input on load (it reads only `in`) writes k_Gamma,
// k_Sigma and kt into channels 16..35 of out; channels 0..15 are the momenta
// (k_f = pi), copied.  Lattices no brick tiles take the gather kernel.
hipError_t launch_rhs(const double *in, double *out, const ewpt_lattice &L, const Phys &P, hipStream_t st)
{
    for (int v : {325, 311, 310, 2}) {
        if (!brick_supported(v, L) || brick_stored_inputs(v)) continue;
        const long n3 = (long)L.nx * L.ny * L.nz;
        StageBufs<double> b{};
        b.y = in;
        b.kout = out + 16 * n3;
        b.n3 = n3;
        hipError_t e = launch_stage_brick(v, 1, b, L, P, 0.0, st);
        if (e == hipSuccess) e = hipMemcpyAsync(out, in + 20 * n3, 16 * n3 * sizeof(double), hipMemcpyDeviceToDevice, st);
        return e;
    }
    return launch_rhs_gather(in, out, L, P, st);
}

}  // namespace

int ewpt_rhs_device(const ewpt_lattice *lat, const ewpt_params *p, const double *din, double *dout, void *stream)
{
    if (!lattice_ok(lat) || !params_ok(p) || !din || !dout || din == dout) return EWPT_ERR_INVAL;
    const hipError_t e = launch_rhs(din, dout, *lat, phys_of(*p), static_cast<hipStream_t>(stream));
    return e == hipSuccess ? EWPT_OK : EWPT_ERR_HIP;
}

int ewpt_rhs(ewpt_ctx *ctx, const ewpt_params *p, const double *hin, double *hout)
{
    if (!ctx || !hin || !hout || !params_ok(p)) return EWPT_ERR_INVAL;
    if (ctx->zg) return no_slab(ctx, "rhs");
    int rc = set_device(ctx);
    if (rc) return rc;
    const size_t n = NCH * (size_t)ctx->n3;
    if (!ctx->rhs_in && (rc = dev_alloc(ctx, &ctx->rhs_in, n, "alloc rhs in"))) return rc;
    if (!ctx->rhs_out && (rc = dev_alloc(ctx, &ctx->rhs_out, n, "alloc rhs out"))) return rc;
    HIPCHK(ctx, hipMemcpyAsync(ctx->rhs_in, hin, n * sizeof(double), hipMemcpyHostToDevice, ctx->stream), "rhs in copy");
    HIPCHK(ctx, launch_rhs(ctx->rhs_in, ctx->rhs_out, ctx->lat, phys_of(*p), ctx->stream), "rhs launch");
    HIPCHK(ctx, hipMemcpyAsync(hout, ctx->rhs_out, n * sizeof(double), hipMemcpyDeviceToHost, ctx->stream), "rhs out copy");
    HIPCHK(ctx, hipStreamSynchronize(ctx->stream), "rhs sync");
    return EWPT_OK;
}

int ewpt_step(ewpt_ctx *ctx, const ewpt_params *p, int nsteps)
{
    if (!ctx || !params_ok(p) || nsteps < 0) return EWPT_ERR_INVAL;
    if (!ctx->has_state) return fail(ctx, EWPT_ERR_STATE, "step before set_state");
    int rc = set_device(ctx);
    if (rc) return rc;
    for (int it = 0; it < nsteps; it++)
        for (int s = 1; s <= 4; s++)
            if ((rc = launch_step_stage(ctx, s, p))) return rc;
    HIPCHK(ctx, hipStreamSynchronize(ctx->stream), "step sync");
    return EWPT_OK;
}

namespace {

// the energy table (ewpt_energies layout; with em the EM columns 7, 8, 12
// of ewpt_em_energies) and Gauss norms (ewpt_gauss layout) from the
// DIAG_SLOTS reduced sums of the diagnostic stage 1
void diag_outputs(const ewpt_ctx *c, const double *r, bool em, double e[13], double gs[8])
{
    const double vol = c->lat.dx * c->lat.dx * c->lat.dx;
    for (int q = 0; q < 13; q++) e[q] = 0.0;
    for (int q = 0; q < 7; q++) e[q] = r[q] * vol;
    e[9] = r[8];
    e[11] = r[7] / (double)sites(c);
    if (em) {
        e[7] = 0.5 * r[18] * vol;
        e[8] = 0.5 * r[19] * vol;
        e[12] = r[20] * vol;
    }
    for (int q = 0; q < 8; q++) gs[q] = std::sqrt(r[10 + q] * vol);
}

// nsteps steps with the diagnostics of the input state reduced by the first
// stage 1 (theta_w: the Weinberg angle of the EM columns when em)
int step_diag_impl(ewpt_ctx *ctx, const ewpt_params *p, int nsteps, bool em, double theta_w, double energies[13],
                   double gauss[8])
{
    if (!ctx || !params_ok(p) || nsteps < 1 || !energies || !gauss || (em && !std::isfinite(theta_w)))
        return EWPT_ERR_INVAL;
    if (ctx->zg) return no_slab(ctx, "step_diag");
    if (!ctx->has_state) return fail(ctx, EWPT_ERR_STATE, "step before set_state");
    int rc = set_device(ctx);
    if (rc) return rc;
    const int v = resolved_variant(ctx);
    const bool fdiag = mixed(ctx) && brick_has_diag_f32(v) && brick_supported_f32(v, ctx->lat);
    if (!fdiag && (mixed(ctx) || !brick_has_diag(v))) {
        // no diagnostic stage 1 for this precision / variant: the standalone
        // reductions of the state, then the steps
        const int re = ewpt_energies(ctx, p, energies);
        if (re && re != EWPT_ERR_NONFINITE) return re;
        const int rg = ewpt_gauss(ctx, p, gauss);
        if (rg && rg != EWPT_ERR_NONFINITE) return rg;
        int rm = EWPT_OK;
        if (em) {
            double m[3];
            rm = ewpt_em_energies(ctx, p, theta_w, m);
            if (rm && rm != EWPT_ERR_NONFINITE) return rm;
            energies[7] = m[0];
            energies[8] = m[1];
            energies[12] = m[2];
        }
        if ((rc = ewpt_step(ctx, p, nsteps))) return rc;
        return re ? re : rg ? rg : rm;
    }
    // rows: one per brick of the diagnostic stage 1 (the FP64 kernels' 16x4x4
    // bricks outnumber the mixed kernels' 32x4x4 ones, so one allocation
    // serves both precisions)
    const long rows = fdiag ? diag_rows_f32(ctx->lat) : diag_rows(ctx->lat);
    if (!ctx->diag &&
        (rc = dev_alloc(ctx, &ctx->diag, (size_t)(diag_rows(ctx->lat) + 1) * DIAG_SLOTS, "alloc diagnostics")))
        return rc;
    if ((rc = ensure_stage_input_buffer(ctx, v))) return rc;
    double *res = ctx->diag + rows * DIAG_SLOTS;
    const double sw = em ? std::sin(theta_w) : 0.0, cw = em ? std::cos(theta_w) : 0.0;
    if (fdiag)
        HIPCHK(ctx, launch_stage1_diag_f32(v, fbufs_for(ctx, 1, brick_stored_inputs(v)), ctx->lat, phys_of(*p), p->dt,
                                           ctx->diag, sw, cw, ctx->stream),
               "diagnostic stage launch");
    else
        HIPCHK(ctx, launch_stage1_diag(v, bufs_for(ctx, 1, brick_stored_inputs(v)), ctx->lat, phys_of(*p), p->dt,
                                       ctx->diag, sw, cw, ctx->stream),
               "diagnostic stage launch");
    HIPCHK(ctx, launch_fixed_reduce(ctx->diag, rows, DIAG_SLOTS, DIAG_SLOTS, 1u << 8, res, ctx->stream), "diagnostic reduction");
    double r[DIAG_SLOTS];
    HIPCHK(ctx, hipMemcpyAsync(r, res, sizeof r, hipMemcpyDeviceToHost, ctx->stream), "diagnostic copy");
    for (int s = 2; s <= 4; s++)
        if ((rc = launch_stage_kernel(ctx, s, p))) return rc;
    for (int it = 1; it < nsteps; it++)
        for (int s = 1; s <= 4; s++)
            if ((rc = launch_stage_kernel(ctx, s, p))) return rc;
    HIPCHK(ctx, hipStreamSynchronize(ctx->stream), "step sync");
    diag_outputs(ctx, r, em, energies, gauss);
    bool bad = r[9] != 0.0;
    for (int q = 0; q < 8; q++) bad |= !std::isfinite(gauss[q]);
    if (bad) return fail(ctx, EWPT_ERR_NONFINITE, "energy density, EM density or Gauss residual is NaN/Inf");
    return EWPT_OK;
}

}  // namespace

int ewpt_step_diag(ewpt_ctx *ctx, const ewpt_params *p, int nsteps, double energies[13], double gauss[8])
{
    return step_diag_impl(ctx, p, nsteps, false, 0.0, energies, gauss);
}

int ewpt_step_diag_em(ewpt_ctx *ctx, const ewpt_params *p, double theta_w, int nsteps, double energies[13],
                      double gauss[8])
{
    return step_diag_impl(ctx, p, nsteps, true, theta_w, energies, gauss);
}

int ewpt_step_timed(ewpt_ctx *ctx, const ewpt_params *p, int nsteps, double stage_ms[4], double *total_ms)
{
    if (!ctx || !params_ok(p) || nsteps <= 0 || !stage_ms || !total_ms) return EWPT_ERR_INVAL;
    if (!ctx->has_state) return fail(ctx, EWPT_ERR_STATE, "step before set_state");
    int rc = set_device(ctx);
    if (rc) return rc;
    const int nev = 4 * nsteps + 1;
    hipEvent_t *ev = new (std::nothrow) hipEvent_t[nev];
    if (!ev) return EWPT_ERR_OOM;
    int made = 0;
    for (; made < nev; made++)
        if (hipEventCreate(&ev[made]) != hipSuccess) { rc = fail(ctx, EWPT_ERR_HIP, "hipEventCreate"); break; }
    if (rc == EWPT_OK && hipEventRecord(ev[0], ctx->stream) != hipSuccess) rc = fail(ctx, EWPT_ERR_HIP, "event record");
    for (int it = 0; it < nsteps && rc == EWPT_OK; it++)
        for (int s = 1; s <= 4 && rc == EWPT_OK; s++) {
            rc = launch_step_stage(ctx, s, p);
            if (rc == EWPT_OK && hipEventRecord(ev[4 * it + s], ctx->stream) != hipSuccess)
                rc = fail(ctx, EWPT_ERR_HIP, "event record");
        }
    if (rc == EWPT_OK && hipStreamSynchronize(ctx->stream) != hipSuccess) rc = fail(ctx, EWPT_ERR_HIP, "sync");
    if (rc == EWPT_OK) {
        for (int s = 0; s < 4; s++) stage_ms[s] = 0.0;
        for (int it = 0; it < nsteps; it++)
            for (int s = 1; s <= 4; s++) {
                float ms = 0.f;
                (void)hipEventElapsedTime(&ms, ev[4 * it + s - 1], ev[4 * it + s]);
                stage_ms[s - 1] += ms;
            }
        for (int s = 0; s < 4; s++) stage_ms[s] /= nsteps;
        float tot = 0.f;
        (void)hipEventElapsedTime(&tot, ev[0], ev[nev - 1]);
        *total_ms = tot;
    }
    for (int e = 0; e < made; e++) (void)hipEventDestroy(ev[e]);
    delete[] ev;
    return rc;
}

// ---------------------------------------------------------------- z-slabs

namespace {

int slab_ok(ewpt_ctx *c, int what)
{
    if (!c->zg) return fail(c, EWPT_ERR_INVAL, "not a slab context");
    if (what < 0 || what > 3) return fail(c, EWPT_ERR_INVAL, "halo array must be 0 (state) or 1..3 (before stage 2..4)");
    if (!c->has_state) return fail(c, EWPT_ERR_STATE, "halo exchange before set_state");
    return EWPT_OK;
}

int ring_ok(ewpt_ctx *const *sl, int n)
{
    if (!sl || n < 1) return EWPT_ERR_INVAL;
    for (int i = 0; i < n; i++) {
        if (!sl[i]) return EWPT_ERR_INVAL;
        for (int j = 0; j < i; j++)
            if (sl[j] == sl[i]) return fail(sl[0], EWPT_ERR_INVAL, "slab_step: a context appears twice");
        if (!sl[i]->zg) return fail(sl[0], EWPT_ERR_INVAL, "slab_step: not a slab context");
        if (sl[i]->device != sl[0]->device || sl[i]->precision != sl[0]->precision ||
            sl[i]->lat.nx != sl[0]->lat.nx || sl[i]->lat.ny != sl[0]->lat.ny || sl[i]->lat.dx != sl[0]->lat.dx)
            return fail(sl[0], EWPT_ERR_INVAL, "slab_step: slabs differ in device, precision, nx, ny or dx");
        // the exchange carries k_s or the stored stage input, by variant
        if (brick_stored_inputs(resolved_variant(sl[i])) != brick_stored_inputs(resolved_variant(sl[0])))
            return fail(sl[0], EWPT_ERR_INVAL, "slab_step: slabs mix variants with and without stored stage inputs");
        if (!sl[i]->has_state) return fail(sl[0], EWPT_ERR_STATE, "slab_step before set_state");
    }
    return EWPT_OK;
}

// nsteps steps of a ring of slabs; with timing, events on slab 0's stream
// split each stage into its exchange and its kernel
int ring_steps(ewpt_ctx *const *sl, int n, const ewpt_params *p, int nsteps, double *stage_ms, double *exch_ms,
               double *total_ms)
{
    ewpt_ctx *c0 = sl[0];
    std::vector<hipEvent_t> ev;
    int rc = EWPT_OK;
    auto mark = [&]() {
        hipEvent_t e = nullptr;
        if (hipEventCreate(&e) != hipSuccess) return fail(c0, EWPT_ERR_HIP, "hipEventCreate");
        ev.push_back(e);
        if (hipEventRecord(e, c0->stream) != hipSuccess) return fail(c0, EWPT_ERR_HIP, "event record");
        return EWPT_OK;
    };
    if (stage_ms) rc = mark();
    for (int it = 0; it < nsteps && rc == EWPT_OK; it++)
        for (int s = 1; s <= 4 && rc == EWPT_OK; s++) {
            rc = ring_exchange(sl, n, s - 1);
            if (rc == EWPT_OK && stage_ms) rc = mark();
            for (int i = 0; i < n && rc == EWPT_OK; i++) rc = launch_stage_kernel(sl[i], s, p);
            if (rc == EWPT_OK && stage_ms) rc = mark();
        }
    for (int i = 0; i < n; i++)
        if (hipStreamSynchronize(sl[i]->stream) != hipSuccess && rc == EWPT_OK) rc = fail(sl[i], EWPT_ERR_HIP, "sync");
    if (rc == EWPT_OK && stage_ms) {
        for (int s = 0; s < 4; s++) stage_ms[s] = 0.0;
        *exch_ms = 0.0;
        for (int it = 0; it < nsteps; it++)
            for (int s = 0; s < 4; s++) {
                const size_t k = 1 + 2 * (4 * (size_t)it + s);
                float a = 0.f, b = 0.f;
                (void)hipEventElapsedTime(&a, ev[k - 1], ev[k]);
                (void)hipEventElapsedTime(&b, ev[k], ev[k + 1]);
                *exch_ms += a;
                stage_ms[s] += b;
            }
        for (int s = 0; s < 4; s++) stage_ms[s] /= nsteps;
        *exch_ms /= nsteps;
        float tot = 0.f;
        (void)hipEventElapsedTime(&tot, ev.front(), ev.back());
        *total_ms = tot;
    }
    for (hipEvent_t e : ev) (void)hipEventDestroy(e);
    return rc;
}

}  // namespace

int ewpt_slab_halo_bytes(ewpt_ctx *ctx, int what, unsigned long long *bytes)
{
    if (!ctx || !bytes) return EWPT_ERR_INVAL;
    if (!ctx->zg) return fail(ctx, EWPT_ERR_INVAL, "not a slab context");
    if (what < 0 || what > 3) return EWPT_ERR_INVAL;
    *bytes = halo_bytes(ctx, what);
    return EWPT_OK;
}

int ewpt_slab_buffers(ewpt_ctx *ctx, void **send_lo, void **send_hi)
{
    if (!ctx || !send_lo || !send_hi) return EWPT_ERR_INVAL;
    if (!ctx->zg) return fail(ctx, EWPT_ERR_INVAL, "not a slab context");
    *send_lo = ctx->halo_lo;
    *send_hi = ctx->halo_hi;
    return EWPT_OK;
}

int ewpt_slab_pack(ewpt_ctx *ctx, int what, void *send_lo, void *send_hi)
{
    if (!ctx || !send_lo || !send_hi) return EWPT_ERR_INVAL;
    int rc = slab_ok(ctx, what);
    if (rc || (rc = set_device(ctx))) return rc;
    return halo_pack_into(ctx, what, send_lo, send_hi);
}

int ewpt_slab_unpack(ewpt_ctx *ctx, int what, const void *recv_lo, const void *recv_hi)
{
    if (!ctx || !recv_lo || !recv_hi) return EWPT_ERR_INVAL;
    int rc = slab_ok(ctx, what);
    if (rc || (rc = set_device(ctx))) return rc;
    return halo_unpack_from(ctx, what, recv_lo, recv_hi);
}

int ewpt_stage(ewpt_ctx *ctx, const ewpt_params *p, int stage)
{
    if (!ctx || !params_ok(p) || stage < 1 || stage > 4) return EWPT_ERR_INVAL;
    if (!ctx->has_state) return fail(ctx, EWPT_ERR_STATE, "stage before set_state");
    int rc = set_device(ctx);
    if (rc) return rc;
    return launch_stage_kernel(ctx, stage, p);
}

int ewpt_get_stream(ewpt_ctx *ctx, void **stream)
{
    if (!ctx || !stream) return EWPT_ERR_INVAL;
    *stream = static_cast<void *>(ctx->stream);
    return EWPT_OK;
}

int ewpt_sync(ewpt_ctx *ctx)
{
    if (!ctx) return EWPT_ERR_INVAL;
    int rc = set_device(ctx);
    if (rc) return rc;
    HIPCHK(ctx, hipStreamSynchronize(ctx->stream), "sync");
    return EWPT_OK;
}

int ewpt_slab_step(ewpt_ctx *const *slabs, int n, const ewpt_params *p, int nsteps)
{
    if (!params_ok(p) || nsteps < 0) return EWPT_ERR_INVAL;
    int rc = ring_ok(slabs, n);
    if (rc || (rc = set_device(slabs[0]))) return rc;
    return ring_steps(slabs, n, p, nsteps, nullptr, nullptr, nullptr);
}

int ewpt_slab_step_timed(ewpt_ctx *const *slabs, int n, const ewpt_params *p, int nsteps, double stage_ms[4],
                         double *exchange_ms, double *total_ms)
{
    if (!params_ok(p) || nsteps <= 0 || !stage_ms || !exchange_ms || !total_ms) return EWPT_ERR_INVAL;
    int rc = ring_ok(slabs, n);
    if (rc || (rc = set_device(slabs[0]))) return rc;
    return ring_steps(slabs, n, p, nsteps, stage_ms, exchange_ms, total_ms);
}

namespace {

// The 21 diagnostic sums of the state (energy densities, Gauss residuals, EM
// densities with the Weinberg angle theta_w) from one pass of the diagnostic
// stage-1 kernel without its stage outputs, where its 16x4x4 bricks tile the
// lattice (*done = false otherwise); mixed contexts pass their FP64 view.
int state_diag_query(ewpt_ctx *ctx, const ewpt_params *p, double theta_w, double r[DIAG_SLOTS], bool *done)
{
    *done = false;
    if (!state_diag_fits(ctx->lat)) return EWPT_OK;
    int rc = ensure_view(ctx);
    if (rc) return rc;
    const long rows = diag_rows(ctx->lat);
    if (!ctx->diag && (rc = dev_alloc(ctx, &ctx->diag, (size_t)(rows + 1) * DIAG_SLOTS, "alloc diagnostics")))
        return rc;
    double *res = ctx->diag + rows * DIAG_SLOTS;
    HIPCHK(ctx, launch_state_diag(ctx->y, ctx->n3, ctx->lat, phys_of(*p), ctx->diag, std::sin(theta_w),
                                  std::cos(theta_w), ctx->stream),
           "diagnostic launch");
    HIPCHK(ctx, launch_fixed_reduce(ctx->diag, rows, DIAG_SLOTS, DIAG_SLOTS, 1u << 8, res, ctx->stream),
           "diagnostic reduction");
    HIPCHK(ctx, hipMemcpyAsync(r, res, DIAG_SLOTS * sizeof(double), hipMemcpyDeviceToHost, ctx->stream),
           "diagnostic copy");
    HIPCHK(ctx, hipStreamSynchronize(ctx->stream), "diagnostic sync");
    *done = true;
    return EWPT_OK;
}

}  // namespace

int ewpt_energies(ewpt_ctx *ctx, const ewpt_params *p, double out[13])
{
    if (!ctx || !out || !params_ok(p)) return EWPT_ERR_INVAL;
    if (ctx->zg) return no_slab(ctx, "energies");
    if (!ctx->has_state) return fail(ctx, EWPT_ERR_STATE, "energies before set_state");
    int rc = set_device(ctx);
    if (rc) return rc;
    double rd[DIAG_SLOTS];
    bool done = false;
    if ((rc = state_diag_query(ctx, p, 0.0, rd, &done))) return rc;
    if (done) {
        double gs[8];
        diag_outputs(ctx, rd, false, out, gs);
        if (rd[9] != 0.0) return fail(ctx, EWPT_ERR_NONFINITE, "energy density is NaN/Inf");
        return EWPT_OK;
    }
    if ((rc = ensure_view(ctx))) return rc;
    HIPCHK(ctx, launch_energies(ctx->y, ctx->lat, phys_of(*p), ctx->part, ctx->res, ctx->stream), "energy launch");
    double r[16];
    HIPCHK(ctx, hipMemcpyAsync(r, ctx->res, 10 * sizeof(double), hipMemcpyDeviceToHost, ctx->stream), "energy copy");
    HIPCHK(ctx, hipStreamSynchronize(ctx->stream), "energy sync");
    const double vol = ctx->lat.dx * ctx->lat.dx * ctx->lat.dx;
    for (int q = 0; q < 7; q++) out[q] = r[q] * vol;
    out[7] = 0.0;
    out[8] = 0.0;
    out[9] = r[8];
    out[10] = 0.0;
    out[11] = r[7] / (double)ctx->n3;
    out[12] = 0.0;
    if (r[9] != 0.0) return fail(ctx, EWPT_ERR_NONFINITE, "energy density is NaN/Inf");
    return EWPT_OK;
}

int ewpt_energies_device(const ewpt_lattice *lat, const ewpt_params *p, const double *dev_state, double *dev_out,
                         void *stream)
{
    if (!lattice_ok(lat) || !params_ok(p) || !dev_state || !dev_out) return EWPT_ERR_INVAL;
    if (!energies_noalloc_fits(*lat)) return EWPT_ERR_UNSUPPORTED;   // the 16x4x4-brick reduction only
    const hipError_t e = launch_energies_noalloc(dev_state, *lat, phys_of(*p), dev_out, static_cast<hipStream_t>(stream));
    return e == hipSuccess ? EWPT_OK : EWPT_ERR_HIP;
}

namespace {

// scratch for the diagnostics: the stored stage derivatives are free between
// steps (FP64: 20 N^3 doubles each, mixed: 20 N^3 floats = 10 N^3 doubles)
void *diag_work(ewpt_ctx *c, int which)
{
    return mixed(c) ? static_cast<void *>(c->fk[which]) : static_cast<void *>(c->k[which]);
}

bool theta_ok(double th) { return std::isfinite(th); }

}  // namespace

int ewpt_em_energies(ewpt_ctx *ctx, const ewpt_params *p, double theta_w, double out[3])
{
    if (!ctx || !out || !params_ok(p) || !theta_ok(theta_w)) return EWPT_ERR_INVAL;
    if (ctx->zg) return no_slab(ctx, "em_energies");
    if (!ctx->has_state) return fail(ctx, EWPT_ERR_STATE, "em_energies before set_state");
    int rc = set_device(ctx);
    if (rc) return rc;
    double rd[DIAG_SLOTS];
    bool done = false;
    if ((rc = state_diag_query(ctx, p, theta_w, rd, &done))) return rc;
    if (done) {
        double e[13], gs[8];
        diag_outputs(ctx, rd, true, e, gs);
        out[0] = e[7];
        out[1] = e[8];
        out[2] = e[12];
        if (!std::isfinite(out[0]) || !std::isfinite(out[1]) || !std::isfinite(out[2]))
            return fail(ctx, EWPT_ERR_NONFINITE, "EM field is NaN/Inf");
        return EWPT_OK;
    }
    if ((rc = ensure_view(ctx))) return rc;
    HIPCHK(ctx, launch_em_energies(ctx->y, ctx->lat, phys_of(*p), theta_w, ctx->part, ctx->res, ctx->stream),
           "em energy launch");
    double r[4];
    HIPCHK(ctx, hipMemcpyAsync(r, ctx->res, 4 * sizeof(double), hipMemcpyDeviceToHost, ctx->stream), "em energy copy");
    HIPCHK(ctx, hipStreamSynchronize(ctx->stream), "em energy sync");
    const double vol = ctx->lat.dx * ctx->lat.dx * ctx->lat.dx;
    out[0] = 0.5 * r[0] * vol;
    out[1] = 0.5 * r[1] * vol;
    out[2] = r[2] * vol;
    if (r[3] != 0.0) return fail(ctx, EWPT_ERR_NONFINITE, "EM field is NaN/Inf");
    return EWPT_OK;
}

int ewpt_em_fields(ewpt_ctx *ctx, const ewpt_params *p, double theta_w, double *host_out)
{
    if (!ctx || !host_out || !params_ok(p) || !theta_ok(theta_w)) return EWPT_ERR_INVAL;
    if (ctx->zg) return no_slab(ctx, "em_fields");
    if (!ctx->has_state) return fail(ctx, EWPT_ERR_STATE, "em_fields before set_state");
    int rc = set_device(ctx);
    if (rc) return rc;
    if ((rc = ensure_view(ctx))) return rc;
    double *w = static_cast<double *>(diag_work(ctx, 0));   // 9 N^3 doubles
    HIPCHK(ctx, launch_em_fields(ctx->y, ctx->lat, phys_of(*p), theta_w, 7u, w, ctx->stream), "em fields launch");
    HIPCHK(ctx, hipMemcpyAsync(host_out, w, 9 * ctx->n3 * sizeof(double), hipMemcpyDeviceToHost, ctx->stream),
           "em fields copy");
    HIPCHK(ctx, hipStreamSynchronize(ctx->stream), "em fields sync");
    return EWPT_OK;
}

int ewpt_y_slice(ewpt_ctx *ctx, const ewpt_params *p, double theta_w, int j, double *host_out)
{
    if (!ctx || !host_out || !params_ok(p) || !theta_ok(theta_w) || j < 0 || j >= ctx->lat.ny)
        return EWPT_ERR_INVAL;
    if (ctx->zg) return no_slab(ctx, "y_slice");
    if (!ctx->has_state) return fail(ctx, EWPT_ERR_STATE, "y_slice before set_state");
    int rc = set_device(ctx);
    if (rc) return rc;
    if ((rc = ensure_view(ctx))) return rc;
    double *w = static_cast<double *>(diag_work(ctx, 0));   // 7 nx nz doubles
    const size_t n = 7 * (size_t)ctx->lat.nx * ctx->lat.nz;
    HIPCHK(ctx, launch_y_slice(ctx->y, ctx->lat, phys_of(*p), theta_w, j, w, ctx->stream), "y_slice launch");
    HIPCHK(ctx, hipMemcpyAsync(host_out, w, n * sizeof(double), hipMemcpyDeviceToHost, ctx->stream), "y_slice copy");
    HIPCHK(ctx, hipStreamSynchronize(ctx->stream), "y_slice sync");
    return EWPT_OK;
}

int ewpt_spectrum(ewpt_ctx *ctx, const ewpt_params *p, double theta_w, int field, const int cut[3], int maxbins,
                  double *k_out, double *p_out, int *nbins)
{
    if (!ctx || !params_ok(p) || !theta_ok(theta_w) || (field != 0 && field != 1) || !cut || !nbins ||
        (!k_out) != (!p_out))
        return EWPT_ERR_INVAL;
    if (ctx->zg) return no_slab(ctx, "spectrum");
    for (int d = 0; d < 3; d++)
        if (cut[d] < 1 || cut[d] > 4096) return fail(ctx, EWPT_ERR_INVAL, "spectrum: cut out of range");
    const char *what = "";
    if (!k_out) {   // size query only
        int nb = 0;
        shell_spectrum(nullptr, ctx->lat, cut, nullptr, nullptr, 0, nullptr, nullptr, &nb, &what);
        *nbins = nb;
        return EWPT_OK;
    }
    if (!ctx->has_state) return fail(ctx, EWPT_ERR_STATE, "spectrum before set_state");
    int rc = set_device(ctx);
    if (rc) return rc;
    if ((rc = ensure_view(ctx))) return rc;
    double *f = static_cast<double *>(diag_work(ctx, 1));   // 3 N^3 doubles
    HIPCHK(ctx, launch_em_fields(ctx->y, ctx->lat, phys_of(*p), theta_w, field == 0 ? 1u : 2u, f, ctx->stream),
           "spectrum field launch");
    const int r = shell_spectrum(f, ctx->lat, cut, diag_work(ctx, 0), ctx->stream, maxbins, k_out, p_out, nbins, &what);
    if (r == -2) return fail(ctx, EWPT_ERR_INVAL, what);
    if (r != 0) {
        (void)hipGetLastError();
        return fail(ctx, EWPT_ERR_HIP, what);
    }
    for (int b = 0; b < *nbins; b++)
        if (!std::isfinite(p_out[b])) return fail(ctx, EWPT_ERR_NONFINITE, "spectrum is NaN/Inf");
    return EWPT_OK;
}

int ewpt_gauss(ewpt_ctx *ctx, const ewpt_params *p, double out[8])
{
    if (!ctx || !out || !params_ok(p)) return EWPT_ERR_INVAL;
    if (ctx->zg) return no_slab(ctx, "gauss");
    if (!ctx->has_state) return fail(ctx, EWPT_ERR_STATE, "gauss before set_state");
    int rc = set_device(ctx);
    if (rc) return rc;
    double rd[DIAG_SLOTS];
    bool done = false;
    if ((rc = state_diag_query(ctx, p, 0.0, rd, &done))) return rc;
    if (done) {
        double e[13];
        diag_outputs(ctx, rd, false, e, out);
        for (int q = 0; q < 8; q++)
            if (!std::isfinite(out[q])) return fail(ctx, EWPT_ERR_NONFINITE, "Gauss residual is NaN/Inf");
        return EWPT_OK;
    }
    if ((rc = ensure_view(ctx))) return rc;
    HIPCHK(ctx, launch_gauss(ctx->y, ctx->lat, phys_of(*p), ctx->part, ctx->res, ctx->stream), "gauss launch");
    double r[8];
    HIPCHK(ctx, hipMemcpyAsync(r, ctx->res, 8 * sizeof(double), hipMemcpyDeviceToHost, ctx->stream), "gauss copy");
    HIPCHK(ctx, hipStreamSynchronize(ctx->stream), "gauss sync");
    const double vol = ctx->lat.dx * ctx->lat.dx * ctx->lat.dx;
    bool bad = false;
    for (int q = 0; q < 8; q++) {
        out[q] = std::sqrt(r[q] * vol);
        bad |= !std::isfinite(out[q]);
    }
    if (bad) return fail(ctx, EWPT_ERR_NONFINITE, "Gauss residual is NaN/Inf");
    return EWPT_OK;
}

int ewpt_set_variant(ewpt_ctx *ctx, int variant)
{
    if (!ctx) return EWPT_ERR_INVAL;
    if (variant == VARIANT_AUTO) {
        ctx->variant = variant;
        return EWPT_OK;
    }
    if (ctx->zg && variant == VARIANT_GATHER)
        return fail(ctx, EWPT_ERR_UNSUPPORTED, "slab contexts take the brick variants only");
    if (mixed(ctx)) {
        if (!brick_exists_f32(variant)) return EWPT_ERR_INVAL;
        if (!brick_supported_f32(variant, ctx->lat))
            return fail(ctx, EWPT_ERR_UNSUPPORTED, "stage variant does not tile this lattice");
        if (int rc = ensure_stage_input_buffer(ctx, variant)) return rc;
        ctx->variant = variant;
        return EWPT_OK;
    }
    if (variant == VARIANT_GATHER) {
        ctx->variant = variant;
        return EWPT_OK;
    }
    if (!brick_exists(variant)) return EWPT_ERR_INVAL;
    if (!brick_supported(variant, ctx->lat))
        return fail(ctx, EWPT_ERR_UNSUPPORTED, "stage variant does not tile this lattice");
    if (int rc = ensure_stage_input_buffer(ctx, variant)) return rc;
    ctx->variant = variant;
    return EWPT_OK;
}

int ewpt_get_variant(ewpt_ctx *ctx, int *variant)
{
    if (!ctx || !variant) return EWPT_ERR_INVAL;
    *variant = ctx->variant;
    return EWPT_OK;
}

int ewpt_active_variant(ewpt_ctx *ctx, int *variant)
{
    if (!ctx || !variant) return EWPT_ERR_INVAL;
    *variant = resolved_variant(ctx);
    return EWPT_OK;
}

int ewpt_device_bytes(ewpt_ctx *ctx, unsigned long long *bytes)
{
    if (!ctx || !bytes) return EWPT_ERR_INVAL;
    *bytes = ctx->bytes;
    return EWPT_OK;
}

// ---------------------------------------------------------------- batches

}  // extern "C"

struct ewpt_batch {
    int n = 0, device = 0, precision = EWPT_FP64;
    std::vector<ewpt_ctx *> ctx;
    int variant = 0;        // 0: first candidate that tiles every lattice
    BatchPlan *plan = nullptr;
    hipStream_t stream = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    int graph_beta = -1;    // beta flavour the captured graph was built for
    bool no_stored = false; // automatic choice: a context had no room for a stage-input buffer
    std::string err;
};

namespace {

int bfail(ewpt_batch *b, int code, const char *what, hipError_t e = hipSuccess)
{
    char buf[512];
    if (e != hipSuccess) std::snprintf(buf, sizeof buf, "%s: %s", what, hipGetErrorString(e));
    else std::snprintf(buf, sizeof buf, "%s", what);
    b->err = buf;
    return code;
}

void drop_graph(ewpt_batch *b)
{
    if (b->exec) (void)hipGraphExecDestroy(b->exec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    b->exec = nullptr;
    b->graph = nullptr;
    b->graph_beta = -1;
}

void drop_plan(ewpt_batch *b)
{
    drop_graph(b);
    batch_plan_destroy(b->plan);
    b->plan = nullptr;
}

// candidate variants in preference order (measured single-lattice ranking)
std::vector<int> batch_candidates(bool f32)
{
    std::vector<int> v = f32 ? std::vector<int>{333, 331, 315, 311, 211, 310, 210, 11, 10}
                             : std::vector<int>{334, 332, 311, 310, 211, 210, 11, 10, 2};
    const size_t n = v.size();
    for (size_t i = 0; i < n; i++) v.push_back(1000 + v[i]);   // persistent forms
    return v;
}

bool batch_fits_all(const ewpt_batch *b, int v)
{
    for (const ewpt_ctx *c : b->ctx)
        if (!batch_variant_fits(v, b->precision == EWPT_MIXED, c->lat)) return false;
    return true;
}

int batch_resolved(const ewpt_batch *b)
{
    if (b->variant) return b->variant;
    for (int v : batch_candidates(b->precision == EWPT_MIXED))
        if (batch_fits_all(b, v) && (!b->no_stored || !brick_stored_inputs(v % 1000))) return v;
    return -1;
}

int ensure_plan(ewpt_batch *b)
{
    if (b->plan) return EWPT_OK;
    int v = batch_resolved(b);
    if (v < 0) return bfail(b, EWPT_ERR_UNSUPPORTED, "no batch variant tiles every lattice");
    std::vector<ewpt_lattice> lats;
    for (ewpt_ctx *c : b->ctx) {
        lats.push_back(c->lat);
        // stored-stage-input variants need each context's stage-input buffer
        int rc = ensure_stage_input_buffer(c, v % 1000);
        if (rc == EWPT_ERR_OOM && !b->variant) {
            // the automatic choice falls back to a candidate without that buffer
            c->err.clear();
            b->no_stored = true;
            v = batch_resolved(b);
            if (v < 0) return bfail(b, EWPT_ERR_OOM, "batch: no room for stage inputs and no variant without them");
            rc = EWPT_OK;
        }
        if (rc) return bfail(b, rc, ("batch: " + c->err).c_str());
    }
    const char *what = "";
    b->plan = batch_plan_create(v, b->precision == EWPT_MIXED, lats, b->device, &what);
    if (!b->plan) return bfail(b, EWPT_ERR_HIP, what);
    return EWPT_OK;
}

template <typename T>
void batch_lat(const ewpt_ctx *c, const ewpt_params &p, bool stored, BatchLat<T> &L)
{
    std::memset(&L, 0, sizeof L);
    L.L = c->lat;
    L.P = phys_of(p);
    L.dt = p.dt;
    for (int par = 0; par < 2; par++)
        for (int s = 1; s <= 4; s++) {
            StageBufs<T> &sb = L.b[par][s - 1];
            sb.n3 = c->n3;
            if constexpr (std::is_same_v<T, float>) {
                sb.y = par ? c->fhi_n : c->fhi;
                sb.ylo = par ? c->flo_n : c->flo;
                sb.ynew = par ? c->fhi : c->fhi_n;
                sb.ylo_new = par ? c->flo : c->flo_n;
                sb.k1 = c->fk[0]; sb.k2 = c->fk[1]; sb.k3 = c->fk[2];
                sb.kout = s < 4 ? c->fk[s - 1] : nullptr;
                if (stored) set_stage_inputs(sb, s, c->fu, par ? c->fhi : c->fhi_n);
            } else {
                sb.y = par ? c->ynext : c->y;
                sb.ynew = par ? c->y : c->ynext;
                sb.k1 = c->k[0]; sb.k2 = c->k[1]; sb.k3 = c->k[2];
                sb.kout = s < 4 ? c->k[s - 1] : nullptr;
                if (stored) set_stage_inputs(sb, s, c->u, par ? c->y : c->ynext);
            }
        }
}

// upload the descriptors for params p (one per lattice)
int batch_upload(ewpt_batch *b, const ewpt_params *p)
{
    int r;
    const bool stored = brick_stored_inputs(batch_resolved(b) % 1000);
    if (b->precision == EWPT_MIXED) {
        std::vector<BatchLat<float>> v(b->n);
        for (int i = 0; i < b->n; i++) batch_lat(b->ctx[i], p[i], stored, v[i]);
        r = batch_plan_set(b->plan, v.data(), b->stream);
    } else {
        std::vector<BatchLat<double>> v(b->n);
        for (int i = 0; i < b->n; i++) batch_lat(b->ctx[i], p[i], stored, v[i]);
        r = batch_plan_set(b->plan, v.data(), b->stream);
    }
    return r ? bfail(b, EWPT_ERR_HIP, "batch descriptor upload") : EWPT_OK;
}

int batch_launch_step(ewpt_batch *b, int parity, bool beta)
{
    for (int s = 1; s <= 4; s++) {
        const hipError_t e = batch_plan_launch(b->plan, parity, s, beta, b->stream);
        if (e != hipSuccess) return bfail(b, EWPT_ERR_HIP, "batch stage launch", e);
    }
    return EWPT_OK;
}

int ensure_graph(ewpt_batch *b, bool beta)
{
    if (b->exec && b->graph_beta == (int)beta) return EWPT_OK;
    drop_graph(b);
    hipError_t e = hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal);
    if (e != hipSuccess) return bfail(b, EWPT_ERR_HIP, "graph capture begin", e);
    int rc = batch_launch_step(b, 0, beta);
    if (rc == EWPT_OK) rc = batch_launch_step(b, 1, beta);
    hipGraph_t g = nullptr;
    e = hipStreamEndCapture(b->stream, &g);
    if (rc) {
        if (g) (void)hipGraphDestroy(g);
        return rc;
    }
    if (e != hipSuccess) return bfail(b, EWPT_ERR_HIP, "graph capture end", e);
    b->graph = g;
    e = hipGraphInstantiate(&b->exec, g, nullptr, nullptr, 0);
    if (e != hipSuccess) {
        drop_graph(b);
        return bfail(b, EWPT_ERR_HIP, "graph instantiate", e);
    }
    b->graph_beta = beta;
    return EWPT_OK;
}

// after nsteps batched steps: an odd count leaves the update in the other buffers
void batch_commit(ewpt_batch *b, int nsteps)
{
    for (ewpt_ctx *c : b->ctx) {
        if (nsteps % 2) {
            if (mixed(c)) { std::swap(c->fhi, c->fhi_n); std::swap(c->flo, c->flo_n); }
            else std::swap(c->y, c->ynext);
        }
        if (nsteps > 0 && mixed(c)) c->view_valid = false;
    }
}

int batch_check(ewpt_batch *b, const ewpt_params *p, int nsteps)
{
    if (!p || nsteps < 0) return EWPT_ERR_INVAL;
    for (int i = 0; i < b->n; i++) {
        if (!params_ok(&p[i])) return bfail(b, EWPT_ERR_INVAL, "batch: invalid parameters");
        if (!b->ctx[i]->has_state) return bfail(b, EWPT_ERR_STATE, "batch: a lattice has no state");
    }
    if (hipSetDevice(b->device) != hipSuccess) return bfail(b, EWPT_ERR_HIP, "hipSetDevice");
    return ensure_plan(b);
}

bool any_beta(const ewpt_batch *b, const ewpt_params *p)
{
    for (int i = 0; i < b->n; i++)
        if (p[i].beta_w != 0.0 || p[i].beta_y != 0.0) return true;
    return false;
}

// nsteps steps: pairs through the captured two-step graph, an odd last step eagerly
int batch_run(ewpt_batch *b, const ewpt_params *p, int nsteps, bool use_graph)
{
    const bool beta = any_beta(b, p);
    int rc = batch_upload(b, p);
    if (rc) return rc;
    if (use_graph && nsteps >= 2 && (rc = ensure_graph(b, beta))) return rc;
    for (int it = 0; it + 1 < nsteps; it += 2) {
        if (use_graph) {
            const hipError_t e = hipGraphLaunch(b->exec, b->stream);
            if (e != hipSuccess) return bfail(b, EWPT_ERR_HIP, "graph launch", e);
        } else if ((rc = batch_launch_step(b, 0, beta)) || (rc = batch_launch_step(b, 1, beta))) {
            return rc;
        }
    }
    if (nsteps % 2 && (rc = batch_launch_step(b, 0, beta))) return rc;
    return EWPT_OK;
}

std::string batch_key(const ewpt_batch *b)
{
    std::vector<std::string> sh;
    for (const ewpt_ctx *c : b->ctx)
        sh.push_back(std::to_string(c->lat.nx) + "x" + std::to_string(c->lat.ny) + "x" + std::to_string(c->lat.nz));
    std::sort(sh.begin(), sh.end());
    std::string k = std::to_string(b->precision) + ":" + std::to_string(b->device);
    for (const std::string &x : sh) k += "," + x;
    return k;
}

// tuned variant per set of lattice shapes, shared by every batch of the
// process; guarded because batches may be tuned from several threads
struct TuneCache {
    std::mutex mu;
    std::map<std::string, int> m;
    bool get(const std::string &k, int &v)
    {
        std::lock_guard<std::mutex> g(mu);
        auto it = m.find(k);
        if (it == m.end()) return false;
        v = it->second;
        return true;
    }
    void put(const std::string &k, int v)
    {
        std::lock_guard<std::mutex> g(mu);
        m[k] = v;
    }
};

TuneCache &tune_cache()
{
    static TuneCache c;
    return c;
}

}  // namespace

extern "C" {

// ---------------------------------------------------------------- learned corrector

}  // extern "C"

struct ewpt_net {
    Net *net = nullptr;
    int device = 0;
    double *save[2] = {nullptr, nullptr};   // states before the last two coarse steps (fallback)
    long save_n = 0;
    std::string err;
};

namespace {

int nfail(ewpt_net *n, int code, const char *what, hipError_t e = hipSuccess)
{
    char buf[512];
    if (e != hipSuccess) std::snprintf(buf, sizeof buf, "%s: %s", what, hipGetErrorString(e));
    else std::snprintf(buf, sizeof buf, "%s", what);
    n->err = buf;
    return code;
}

// one reduction pass: G = sqrt(sum of the squared Gauss-law norms G^1..3,
// G^Y) and ||E|| = sqrt(sum (E^a)^2 + (E^Y)^2) dx^3 (= sqrt(2 (EEW + EEY)))
int gauss_and_enorm(ewpt_ctx *c, const ewpt_params *p, double *G, double *En)
{
    int rc = set_device(c);
    if (rc) return rc;
    double rd[DIAG_SLOTS];
    bool done = false;
    if ((rc = state_diag_query(c, p, 0.0, rd, &done))) return rc;
    if (done) {
        double e[13], gs[8];
        diag_outputs(c, rd, false, e, gs);
        *G = std::sqrt(gs[0] * gs[0] + gs[1] * gs[1] + gs[2] * gs[2] + gs[3] * gs[3]);
        *En = std::sqrt(2.0 * (e[3] + e[5]));
        if (!std::isfinite(*G) || !std::isfinite(*En)) return EWPT_ERR_NONFINITE;
        return EWPT_OK;
    }
    if ((rc = ensure_view(c))) return rc;
    HIPCHK(c, launch_gauss(c->y, c->lat, phys_of(*p), c->part, c->res, c->stream), "gauss launch");
    double r[10];
    HIPCHK(c, hipMemcpyAsync(r, c->res, 10 * sizeof(double), hipMemcpyDeviceToHost, c->stream), "gauss copy");
    HIPCHK(c, hipStreamSynchronize(c->stream), "gauss sync");
    const double vol = c->lat.dx * c->lat.dx * c->lat.dx;
    *G = std::sqrt((r[0] + r[1] + r[2] + r[3]) * vol);
    *En = std::sqrt((r[8] + r[9]) * vol);
    if (!std::isfinite(*G) || !std::isfinite(*En)) return EWPT_ERR_NONFINITE;
    return EWPT_OK;
}

}  // namespace

extern "C" {

int ewpt_net_create(ewpt_net **out, int device, unsigned long long seed)
{
    if (!out || device < 0) return EWPT_ERR_INVAL;
    *out = nullptr;
    if (hipSetDevice(device) != hipSuccess) return EWPT_ERR_HIP;
    ewpt_net *n = new (std::nothrow) ewpt_net();
    if (!n) return EWPT_ERR_OOM;
    const char *what = "";
    n->device = device;
    n->net = net_create(device, seed, &what);
    if (!n->net) {
        delete n;
        return EWPT_ERR_HIP;
    }
    *out = n;
    return EWPT_OK;
}

int ewpt_net_destroy(ewpt_net *n)
{
    if (!n) return EWPT_ERR_INVAL;
    (void)hipSetDevice(n->device);
    net_destroy(n->net);
    for (double *b : n->save)
        if (b) (void)hipFree(b);
    delete n;
    return EWPT_OK;
}

const char *ewpt_net_last_error(ewpt_net *n) { return n ? n->err.c_str() : ""; }

long ewpt_net_param_count(void) { return net_param_count(); }

int ewpt_net_set_params(ewpt_net *n, const float *params)
{
    if (!n || !params) return EWPT_ERR_INVAL;
    for (long q = 0; q < net_param_count(); q++)
        if (!std::isfinite(params[q])) return nfail(n, EWPT_ERR_INVAL, "net: non-finite parameter");
    if (hipSetDevice(n->device) != hipSuccess) return EWPT_ERR_HIP;
    const hipError_t e = net_set_params(n->net, params);
    return e == hipSuccess ? EWPT_OK : nfail(n, EWPT_ERR_HIP, "net: set parameters", e);
}

int ewpt_net_get_params(ewpt_net *n, float *params)
{
    if (!n || !params) return EWPT_ERR_INVAL;
    if (hipSetDevice(n->device) != hipSuccess) return EWPT_ERR_HIP;
    const hipError_t e = net_get(n->net, 0, params);
    return e == hipSuccess ? EWPT_OK : nfail(n, EWPT_ERR_HIP, "net: get parameters", e);
}

int ewpt_net_get_grads(ewpt_net *n, float *grads)
{
    if (!n || !grads) return EWPT_ERR_INVAL;
    if (hipSetDevice(n->device) != hipSuccess) return EWPT_ERR_HIP;
    const hipError_t e = net_get(n->net, 1, grads);
    return e == hipSuccess ? EWPT_OK : nfail(n, EWPT_ERR_HIP, "net: get gradients", e);
}

int ewpt_net_set_norm(ewpt_net *n, const float in_scale[36], const float out_scale[36])
{
    if (!n || !in_scale || !out_scale) return EWPT_ERR_INVAL;
    float inv[NCH];
    for (int c = 0; c < NCH; c++) {
        if (!std::isfinite(in_scale[c]) || !(in_scale[c] > 0.f) || !std::isfinite(out_scale[c]))
            return nfail(n, EWPT_ERR_INVAL, "net: scales must be finite, input scales positive");
        inv[c] = 1.0f / in_scale[c];
    }
    if (hipSetDevice(n->device) != hipSuccess) return EWPT_ERR_HIP;
    const hipError_t e = net_set_norm(n->net, inv, out_scale);
    return e == hipSuccess ? EWPT_OK : nfail(n, EWPT_ERR_HIP, "net: set scales", e);
}

static int net_ctx_ok(ewpt_net *n, ewpt_ctx *c)
{
    if (c->device != n->device) return nfail(n, EWPT_ERR_INVAL, "net: context on another device");
    if (c->precision != EWPT_FP64) return nfail(n, EWPT_ERR_UNSUPPORTED, "net: FP64 contexts only");
    if (!net_lattice_ok(c->lat)) return nfail(n, EWPT_ERR_UNSUPPORTED, "net: lattice must tile 8x4x4");
    if (c->zg) return nfail(n, EWPT_ERR_UNSUPPORTED, "net: not available on a slab context");
    if (!c->has_state) return nfail(n, EWPT_ERR_STATE, "net: context has no state");
    return EWPT_OK;
}

int ewpt_net_set_input(ewpt_net *n, int mode)
{
    if (!n || (mode != 0 && mode != 1)) return EWPT_ERR_INVAL;
    net_set_input_mode(n->net, mode);
    return EWPT_OK;
}

int ewpt_net_apply(ewpt_net *n, ewpt_ctx *ctx)
{
    if (!n || !ctx) return EWPT_ERR_INVAL;
    int rc = net_ctx_ok(n, ctx);
    if (rc) return rc;
    if (hipSetDevice(n->device) != hipSuccess) return EWPT_ERR_HIP;
    const hipError_t e = net_apply(n->net, ctx->y, ctx->lat);
    return e == hipSuccess ? EWPT_OK : nfail(n, EWPT_ERR_HIP, "net: apply", e);
}

int ewpt_net_train_step(ewpt_net *n, ewpt_ctx *coarse, ewpt_ctx *fine, double lr, double *loss)
{
    if (!n || !coarse || !fine || !loss || !std::isfinite(lr) || lr < 0.0) return EWPT_ERR_INVAL;
    int rc = net_ctx_ok(n, coarse);
    if (rc || (rc = net_ctx_ok(n, fine))) return rc;
    if (coarse->lat.nx != fine->lat.nx || coarse->lat.ny != fine->lat.ny || coarse->lat.nz != fine->lat.nz)
        return nfail(n, EWPT_ERR_INVAL, "net: pair of different lattices");
    if (hipSetDevice(n->device) != hipSuccess) return EWPT_ERR_HIP;
    const hipError_t e = net_train(n->net, coarse->y, fine->y, coarse->lat, lr, loss);
    return e == hipSuccess ? EWPT_OK : nfail(n, EWPT_ERR_HIP, "net: train step", e);
}

int ewpt_step_corrected(ewpt_ctx *ctx, ewpt_net *n, const ewpt_params *p, int factor, int ncoarse,
                        double gauss_tol, int *nfallback)
{
    if (!ctx || !n || !params_ok(p) || factor < 1 || ncoarse < 0 || std::isnan(gauss_tol)) return EWPT_ERR_INVAL;
    int rc = net_ctx_ok(n, ctx);
    if (rc) return rc;
    if (hipSetDevice(n->device) != hipSuccess) return EWPT_ERR_HIP;
    ewpt_params pc = *p;
    pc.dt = p->dt * factor;
    const bool check = !std::isinf(gauss_tol) || gauss_tol < 0;   // +inf: never fall back, no checks
    if (!check) {
        for (int it = 0; it < ncoarse; it++) {
            if ((rc = ewpt_step(ctx, &pc, 1))) return rc;
            if ((rc = ewpt_net_apply(n, ctx))) return rc;
        }
        if (nfallback) *nfallback = 0;
        return EWPT_OK;
    }
    const long need = NCH * ctx->n3;
    if (n->save_n != need) {
        for (double *&b : n->save) {
            if (b) (void)hipFree(b);
            b = nullptr;
        }
        n->save_n = 0;
        for (double *&b : n->save)
            if (hipMalloc(reinterpret_cast<void **>(&b), need * sizeof(double)) != hipSuccess)
                return nfail(n, EWPT_ERR_OOM, "net: alloc fallback copies");
        n->save_n = need;
    }
    // The Gauss residual norm G of a corrected state is reduced inside stage 1
    // of the next coarse step (ewpt_step_diag), so the check costs no extra
    // pass over the lattice; only the last corrected state takes the
    // standalone reduction.  A state whose drift (G - G_prev) / |E| exceeds
    // the tolerance is replaced by `factor` fine steps from the accepted state
    // before its coarse step (save[base]), and the coarse step that had
    // started from it is redone.  The decisions and states are those of
    // checking every corrected state right after its correction.
    auto copy = [&](double *dst, const double *src) {
        return hipMemcpyAsync(dst, src, need * sizeof(double), hipMemcpyDeviceToDevice, ctx->stream);
    };
    // one coarse step with the Gauss norm G and |E| of its input state
    auto coarse = [&](double *G, double *En) {
        double e[13], g[8];
        const int r = ewpt_step_diag(ctx, &pc, 1, e, g);
        *G = std::sqrt(g[0] * g[0] + g[1] * g[1] + g[2] * g[2] + g[3] * g[3]);
        *En = std::sqrt(2.0 * (e[3] + e[5]));
        return r;
    };
    // replace the state by `factor` fine steps from save[b]
    auto fine_from = [&](int b) {
        HIPCHK(ctx, copy(ctx->y, n->save[b]), "restore state");
        HIPCHK(ctx, hipStreamSynchronize(ctx->stream), "restore sync");
        return ewpt_step(ctx, p, factor);
    };
    int fb = 0, base = 0;   // save[base]: the accepted state before the pending coarse step
    double g0 = 0.0;        // G of the last accepted state
    for (int it = 0; it < ncoarse; it++) {
        double G = 0.0, En = 0.0;
        HIPCHK(ctx, copy(n->save[base ^ 1], ctx->y), "save state");
        rc = coarse(&G, &En);
        if (rc && rc != EWPT_ERR_NONFINITE) return rc;
        if (it == 0) {
            if (rc) {   // the input state itself is not finite: leave it as it was
                HIPCHK(ctx, copy(ctx->y, n->save[base ^ 1]), "restore state");
                HIPCHK(ctx, hipStreamSynchronize(ctx->stream), "restore sync");
                return rc;
            }
        } else if (rc == EWPT_ERR_NONFINITE || !((G - g0) / std::max(En, 1e-12) <= gauss_tol)) {
            // the corrected state of the previous coarse step failed: redo that
            // step with fine steps, then this coarse step from the result
            if ((rc = fine_from(base))) return rc;
            fb++;
            HIPCHK(ctx, copy(n->save[base ^ 1], ctx->y), "save state");
            if ((rc = coarse(&G, &En))) return rc;
        }
        g0 = G;
        base ^= 1;
        if ((rc = ewpt_net_apply(n, ctx))) return rc;
    }
    if (ncoarse > 0) {
        double G = 0.0, En = 0.0;
        rc = gauss_and_enorm(ctx, p, &G, &En);
        if (rc && rc != EWPT_ERR_NONFINITE) return rc;
        if (rc == EWPT_ERR_NONFINITE || !((G - g0) / std::max(En, 1e-12) <= gauss_tol)) {
            if ((rc = fine_from(base))) return rc;
            fb++;
        }
    }
    if (nfallback) *nfallback = fb;
    return EWPT_OK;
}

int ewpt_net_conv_test(int nx, int ny, int nz, int mode, const unsigned short *X, const unsigned short *W,
                       const float *bias, const float *zprev, float *Z, unsigned short *A)
{
    if (!X || !W || mode < 0 || mode > 2 || (mode < 2 && !bias) || (mode == 2 && (!zprev || !A)) ||
        (mode < 2 && !Z) || (mode == 0 && !A))
        return EWPT_ERR_INVAL;
    if (!net_lattice_ok(ewpt_lattice{nx, ny, nz, 1.0}) || nx < 8 || ny < 4 || nz < 4) return EWPT_ERR_INVAL;
    return net_conv_host(nx, ny, nz, mode, X, W, bias, zprev, Z, A) == hipSuccess ? EWPT_OK : EWPT_ERR_HIP;
}

int ewpt_net_wgrad_test(int nx, int ny, int nz, const unsigned short *dZ, const unsigned short *X, float *dW,
                        float *db)
{
    if (!dZ || !X || !dW || !db) return EWPT_ERR_INVAL;
    if (!net_lattice_ok(ewpt_lattice{nx, ny, nz, 1.0}) || nx < 8 || ny < 4 || nz < 4) return EWPT_ERR_INVAL;
    return net_wgrad_host(nx, ny, nz, dZ, X, dW, db) == hipSuccess ? EWPT_OK : EWPT_ERR_HIP;
}

int ewpt_batch_create(ewpt_batch **out, ewpt_ctx *const *ctxs, int n)
{
    if (!out || !ctxs || n < 1) return EWPT_ERR_INVAL;
    for (int i = 0; i < n; i++) {
        if (!ctxs[i]) return EWPT_ERR_INVAL;
        if (ctxs[i]->device != ctxs[0]->device || ctxs[i]->precision != ctxs[0]->precision) return EWPT_ERR_INVAL;
        if (ctxs[i]->zg) return EWPT_ERR_UNSUPPORTED;   // batches hold periodic lattices
        for (int j = 0; j < i; j++)
            if (ctxs[j] == ctxs[i]) return EWPT_ERR_INVAL;
    }
    ewpt_batch *b = new (std::nothrow) ewpt_batch();
    if (!b) return EWPT_ERR_OOM;
    b->n = n;
    b->device = ctxs[0]->device;
    b->precision = ctxs[0]->precision;
    b->ctx.assign(ctxs, ctxs + n);
    if (hipSetDevice(b->device) != hipSuccess || hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking) != hipSuccess) {
        delete b;
        return EWPT_ERR_HIP;
    }
    if (batch_resolved(b) < 0) {
        (void)hipStreamDestroy(b->stream);
        delete b;
        return EWPT_ERR_UNSUPPORTED;
    }
    *out = b;
    return EWPT_OK;
}

int ewpt_batch_destroy(ewpt_batch *b)
{
    if (!b) return EWPT_ERR_INVAL;
    (void)hipSetDevice(b->device);
    if (b->stream) (void)hipStreamSynchronize(b->stream);
    drop_plan(b);
    if (b->stream) (void)hipStreamDestroy(b->stream);
    delete b;
    return EWPT_OK;
}

const char *ewpt_batch_last_error(ewpt_batch *b) { return b ? b->err.c_str() : ""; }

int ewpt_batch_set_variant(ewpt_batch *b, int variant)
{
    if (!b) return EWPT_ERR_INVAL;
    if (variant != 0 && !batch_fits_all(b, variant))
        return bfail(b, EWPT_ERR_UNSUPPORTED, "batch variant does not tile every lattice");
    if (variant != b->variant) {
        if (hipSetDevice(b->device) == hipSuccess && b->stream) (void)hipStreamSynchronize(b->stream);
        drop_plan(b);
        b->variant = variant;
    }
    return EWPT_OK;
}

int ewpt_batch_get_variant(ewpt_batch *b, int *variant)
{
    if (!b || !variant) return EWPT_ERR_INVAL;
    *variant = batch_resolved(b);
    return EWPT_OK;
}

int ewpt_batch_step(ewpt_batch *b, const ewpt_params *p, int nsteps)
{
    if (!b) return EWPT_ERR_INVAL;
    int rc = batch_check(b, p, nsteps);
    if (rc) return rc;
    rc = batch_run(b, p, nsteps, true);
    const hipError_t e = hipStreamSynchronize(b->stream);
    if (rc) return rc;
    if (e != hipSuccess) return bfail(b, EWPT_ERR_HIP, "batch sync", e);
    batch_commit(b, nsteps);
    return EWPT_OK;
}

int ewpt_batch_step_timed(ewpt_batch *b, const ewpt_params *p, int nsteps, double *ms)
{
    if (!b || !ms || nsteps < 1) return EWPT_ERR_INVAL;
    int rc = batch_check(b, p, nsteps);
    if (rc) return rc;
    // capture and upload outside the timed region
    if (nsteps >= 2 && (rc = ensure_graph(b, any_beta(b, p)))) return rc;
    hipEvent_t e0, e1;
    if (hipEventCreate(&e0) != hipSuccess) return bfail(b, EWPT_ERR_HIP, "event");
    if (hipEventCreate(&e1) != hipSuccess) { (void)hipEventDestroy(e0); return bfail(b, EWPT_ERR_HIP, "event"); }
    if ((rc = batch_upload(b, p)) == EWPT_OK) {
        (void)hipEventRecord(e0, b->stream);
        rc = batch_run(b, p, nsteps, true);
        (void)hipEventRecord(e1, b->stream);
    }
    const hipError_t e = hipStreamSynchronize(b->stream);
    float f = 0.f;
    if (rc == EWPT_OK && e == hipSuccess) (void)hipEventElapsedTime(&f, e0, e1);
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (rc) return rc;
    if (e != hipSuccess) return bfail(b, EWPT_ERR_HIP, "batch sync", e);
    batch_commit(b, nsteps);
    *ms = f;
    return EWPT_OK;
}

int ewpt_batch_tune(ewpt_batch *b, const ewpt_params *p, int nsteps, int *best_variant, double *best_ms)
{
    if (!b || !p || nsteps < 1) return EWPT_ERR_INVAL;
    const std::string key = batch_key(b);
    int cached = 0;
    if (tune_cache().get(key, cached)) {
        int rc = ewpt_batch_set_variant(b, cached);
        if (rc) return rc;
        if (best_variant) *best_variant = cached;
        if (best_ms) *best_ms = -1.0;   // served from the cache
        return EWPT_OK;
    }
    // dt = 0 leaves every state bitwise unchanged (y' = y + 0), so timing does not
    // disturb the lattices
    std::vector<ewpt_params> q(p, p + b->n);
    for (ewpt_params &x : q) x.dt = 0.0;
    int best = -1;
    double tbest = 1e300;
    for (int v : batch_candidates(b->precision == EWPT_MIXED)) {
        if (!batch_fits_all(b, v)) continue;
        int rc = ewpt_batch_set_variant(b, v);
        if (rc == EWPT_OK) rc = ewpt_batch_step(b, q.data(), 2);   // warm-up: capture, first launches
        double ms = 0.0;
        if (rc == EWPT_OK) rc = ewpt_batch_step_timed(b, q.data(), nsteps, &ms);
        if (rc == EWPT_ERR_OOM && brick_stored_inputs(v % 1000)) continue;   // no room for the stage inputs
        if (rc) return rc;
        if (ms < tbest) { tbest = ms; best = v; }
    }
    if (best < 0) return bfail(b, EWPT_ERR_UNSUPPORTED, "no batch variant tiles every lattice");
    tune_cache().put(key, best);
    int rc = ewpt_batch_set_variant(b, best);
    if (rc) return rc;
    if (best_variant) *best_variant = best;
    if (best_ms) *best_ms = tbest / nsteps;
    return EWPT_OK;
}

}  // extern "C"
