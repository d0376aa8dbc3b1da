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
    long n3 = 0;
    hipStream_t stream = nullptr;
    // FP64 path: state ping-pong and the three stored stage derivatives.
    // Mixed path: y is an FP64 view of the state (for state I/O and the
    // FP64 reductions), rebuilt from the FP32 pair when view_valid is false.
    double *y = nullptr, *ynext = nullptr, *k[3] = {nullptr, nullptr, nullptr};
    float *fhi = nullptr, *flo = nullptr, *fhi_n = nullptr, *flo_n = nullptr;
    float *fk[3] = {nullptr, nullptr, nullptr};
    bool view_valid = true;
    // host-buffer RHS path (allocated on first use)
    double *rhs_in = nullptr, *rhs_out = nullptr;
    // reductions
    double *part = nullptr, *res = nullptr;
    long nred = 0;
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
    double **bufs[] = {&c->y, &c->ynext, &c->k[0], &c->k[1], &c->k[2], &c->rhs_in, &c->rhs_out, &c->part, &c->res};
    for (double **b : bufs)
        if (*b) { (void)hipFree(*b); *b = nullptr; }
    float **fbufs[] = {&c->fhi, &c->flo, &c->fhi_n, &c->flo_n, &c->fk[0], &c->fk[1], &c->fk[2]};
    for (float **b : fbufs)
        if (*b) { (void)hipFree(*b); *b = nullptr; }
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

StageBufs<double> bufs_for(ewpt_ctx *c, int stage)
{
    StageBufs<double> b{};
    b.y = c->y;
    b.k1 = c->k[0];
    b.k2 = c->k[1];
    b.k3 = c->k[2];
    b.n3 = c->n3;
    if (stage < 4) b.kout = c->k[stage - 1];
    else b.ynew = c->ynext;
    return b;
}

int set_device(ewpt_ctx *c) { HIPCHK(c, hipSetDevice(c->device), "hipSetDevice"); return EWPT_OK; }

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
// the lattice from the measured ranking (512^3 FP64 on MI355X: 211, 210, 2),
// else the gather kernel
int resolved_variant(const ewpt_ctx *c)
{
    if (c->variant != VARIANT_AUTO) return c->variant;
    if (mixed(c)) {
        for (int v : {211, 210})
            if (brick_supported_f32(v, c->lat)) return v;
        return -1;
    }
    for (int v : {211, 210, 2})
        if (brick_supported(v, c->lat)) return v;
    return VARIANT_GATHER;
}

int launch_step_stage(ewpt_ctx *c, int stage, const ewpt_params *p)
{
    const Phys P = phys_of(*p);
    const int v = resolved_variant(c);
    if (mixed(c)) {
        StageBufs<float> b{};
        b.y = c->fhi;
        b.ylo = c->flo;
        b.k1 = c->fk[0];
        b.k2 = c->fk[1];
        b.k3 = c->fk[2];
        b.n3 = c->n3;
        if (stage < 4) b.kout = c->fk[stage - 1];
        else { b.ynew = c->fhi_n; b.ylo_new = c->flo_n; }
        HIPCHK(c, launch_stage_brick_f32(v, stage, b, c->lat, P, p->dt, c->stream), "stage launch");
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
        HIPCHK(c, launch_stage_brick(v, stage, bufs_for(c, stage), c->lat, P, p->dt, c->stream), "stage launch");
    if (stage == 4) std::swap(c->y, c->ynext);
    return EWPT_OK;
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

int ewpt_create(ewpt_ctx **out, const ewpt_lattice *lat, int precision, int device)
{
    if (!out) return EWPT_ERR_INVAL;
    *out = nullptr;
    if (!lattice_ok(lat) || (precision != EWPT_FP64 && precision != EWPT_MIXED) || device < 0)
        return EWPT_ERR_INVAL;
    if (precision == EWPT_MIXED && !brick_supported_f32(210, *lat) && !brick_supported_f32(211, *lat))
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
    c->n3 = (long)lat->nx * lat->ny * lat->nz;
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
    if (rc != EWPT_OK) {
        release(c);
        delete c;
        return rc;
    }
    *out = c;
    return EWPT_OK;
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
    HIPCHK(ctx, hipMemcpyAsync(ctx->y, host, NCH * ctx->n3 * sizeof(double), hipMemcpyHostToDevice, ctx->stream),
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
    HIPCHK(ctx, hipMemcpyAsync(host, ctx->y, NCH * ctx->n3 * sizeof(double), hipMemcpyDeviceToHost, ctx->stream),
           "get_state copy");
    HIPCHK(ctx, hipStreamSynchronize(ctx->stream), "get_state sync");
    return EWPT_OK;
}

int ewpt_init_modes(ewpt_ctx *ctx, unsigned long long seed, int nmodes, int kmax, double amp, double vev,
                    double mom_scale, double *modes_out)
{
    if (!ctx || nmodes < 1 || nmodes > 64 || kmax < 0 || !std::isfinite(amp) || !std::isfinite(vev) ||
        !std::isfinite(mom_scale))
        return EWPT_ERR_INVAL;
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

int ewpt_rhs_device(const ewpt_lattice *lat, const ewpt_params *p, const double *din, double *dout, void *stream)
{
    if (!lattice_ok(lat) || !params_ok(p) || !din || !dout || din == dout) return EWPT_ERR_INVAL;
    const hipError_t e = launch_rhs_gather(din, dout, *lat, phys_of(*p), static_cast<hipStream_t>(stream));
    return e == hipSuccess ? EWPT_OK : EWPT_ERR_HIP;
}

int ewpt_rhs(ewpt_ctx *ctx, const ewpt_params *p, const double *hin, double *hout)
{
    if (!ctx || !hin || !hout || !params_ok(p)) return EWPT_ERR_INVAL;
    int rc = set_device(ctx);
    if (rc) return rc;
    const size_t n = NCH * (size_t)ctx->n3;
    if (!ctx->rhs_in && (rc = dev_alloc(ctx, &ctx->rhs_in, n, "alloc rhs in"))) return rc;
    if (!ctx->rhs_out && (rc = dev_alloc(ctx, &ctx->rhs_out, n, "alloc rhs out"))) return rc;
    HIPCHK(ctx, hipMemcpyAsync(ctx->rhs_in, hin, n * sizeof(double), hipMemcpyHostToDevice, ctx->stream), "rhs in copy");
    HIPCHK(ctx, launch_rhs_gather(ctx->rhs_in, ctx->rhs_out, ctx->lat, phys_of(*p), ctx->stream), "rhs launch");
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

int ewpt_energies(ewpt_ctx *ctx, const ewpt_params *p, double out[13])
{
    if (!ctx || !out || !params_ok(p)) return EWPT_ERR_INVAL;
    if (!ctx->has_state) return fail(ctx, EWPT_ERR_STATE, "energies before set_state");
    int rc = set_device(ctx);
    if (rc) return rc;
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
    if (!ctx->has_state) return fail(ctx, EWPT_ERR_STATE, "em_energies before set_state");
    int rc = set_device(ctx);
    if (rc) return rc;
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
    if (!ctx->has_state) return fail(ctx, EWPT_ERR_STATE, "gauss before set_state");
    int rc = set_device(ctx);
    if (rc) return rc;
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
    if (mixed(ctx)) {
        if (variant % 100 != 10 && variant % 100 != 11) return EWPT_ERR_INVAL;
        if (!brick_exists(variant)) return EWPT_ERR_INVAL;
        if (!brick_supported_f32(variant, ctx->lat))
            return fail(ctx, EWPT_ERR_UNSUPPORTED, "stage variant does not tile this lattice");
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
    ctx->variant = variant;
    return EWPT_OK;
}

int ewpt_get_variant(ewpt_ctx *ctx, int *variant)
{
    if (!ctx || !variant) return EWPT_ERR_INVAL;
    *variant = ctx->variant;
    return EWPT_OK;
}

int ewpt_device_bytes(ewpt_ctx *ctx, unsigned long long *bytes)
{
    if (!ctx || !bytes) return EWPT_ERR_INVAL;
    *bytes = ctx->bytes;
    return EWPT_OK;
}

}  // extern "C"
