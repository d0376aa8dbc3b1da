// ewpt_internal.h -- declarations shared between the kernel files and the
// C-ABI implementation of libewpt.
#pragma once

#include <hip/hip_runtime.h>

#include <vector>

#include "../../include/ewpt.h"
#include "ewpt_point.h"
#include "ewpt_stage.h"
#include "ewpt_stencil.h"

namespace ewpt {

// ewpt_gather.hip
hipError_t launch_stage_gather(int stage, const StageBufs<double> &b, const ewpt_lattice &L, const Phys &P,
                               double dt, hipStream_t st);
hipError_t launch_rhs_gather(const double *in, double *out, const ewpt_lattice &L, const Phys &P, hipStream_t st);
long reduce_blocks(const ewpt_lattice &L);
hipError_t launch_energies(const double *y, const ewpt_lattice &L, const Phys &P, double *part, double *res,
                           hipStream_t st);
hipError_t launch_gauss(const double *y, const ewpt_lattice &L, const Phys &P, double *part, double *res,
                        hipStream_t st);
// The energy table of launch_energies with no memory of the library's own:
// partials in a fixed file-scope device array; the 13 columns (as
// ewpt_energies returns them) written to the device buffer out.  Lattices
// that tile 16x4x4 only.  Asynchronous on st.
bool energies_noalloc_fits(const ewpt_lattice &L);
hipError_t launch_energies_noalloc(const double *y, const ewpt_lattice &L, const Phys &P, double *out,
                                   hipStream_t st);

// EM diagnostics (em_point): fields of the groups in mask (bit 0 B, 1 B_2,
// 2 A) to out; energy sums 0.5|B|^2, 0.5|B_2|^2, B_2.A (unscaled) to res[0..2],
// non-finite count to res[3]
hipError_t launch_em_fields(const double *y, const ewpt_lattice &L, const Phys &P, double theta_w, unsigned mask,
                            double *out, hipStream_t st);
hipError_t launch_em_energies(const double *y, const ewpt_lattice &L, const Phys &P, double theta_w, double *part,
                              double *res, hipStream_t st);
// the seven y-slice snapshot planes at y = j0 (7 * nx * nz doubles to out)
hipError_t launch_y_slice(const double *y, const ewpt_lattice &L, const Phys &P, double theta_w, int j0, double *out,
                          hipStream_t st);

// ewpt_spectrum.hip -- shell-averaged power spectrum of a 3-component field
// (3 contiguous real N^3 channels in `field`); cwork holds 3 (nx/2+1) ny nz
// complex doubles.  Returns 0 or -1 with *what naming the failed call.
int shell_spectrum(const double *field, const ewpt_lattice &L, const int cut[3], void *cwork, hipStream_t st,
                   int maxbins, double *k_out, double *p_out, int *nbins, const char **what);
// number of distinct integer |K|^2 of an n0*n1*n2 lattice (Kc_bin_nums)
long distinct_k2_count(int n0, int n1, int n2);

// ewpt_brick.hip -- LDS-tiled fused stage kernels (FP64 variants)
bool brick_exists(int variant);
// variants in which some of stages 2-4 read their input as one array written
// by the stage before (StageBufs::u / unext)
bool brick_stored_inputs(int variant);
bool brick_supported(int variant, const ewpt_lattice &L);
hipError_t launch_stage_brick(int variant, int stage, const StageBufs<double> &b, const ewpt_lattice &L,
                              const Phys &P, double dt, hipStream_t st);
// stage 1 that also reduces the energy densities, Gauss residuals and EM
// densities of its input (the state): one row of DIAG_SLOTS partial sums per
// brick in diag (slots: PE KE GE EEW MEW EEY MEY sum|phi| min|phi|
// non-finite, then (G^a)^2 (G^Y)^2 (C^a)^2 (C^Y)^2, then |B|^2 |B_2|^2
// B_2.A with sw, cw the sine and cosine of the Weinberg angle).  FP64
// variants 32, 34 (+100*order).
constexpr int DIAG_SLOTS = 21;
bool brick_has_diag(int variant);
long diag_rows(const ewpt_lattice &L);   // rows of diag (bricks of the diagnostic stage 1)
hipError_t launch_stage1_diag(int variant, const StageBufs<double> &b, const ewpt_lattice &L, const Phys &P, double dt,
                              double *diag, double sw, double cw, hipStream_t st);
// The same rows for an FP64 state y (36 * n3, periodic) without a stage: the
// passes of variant 32's stage 1 with its outputs dropped, diag_rows(L) rows.
bool state_diag_fits(const ewpt_lattice &L);
hipError_t launch_state_diag(const double *y, long n3, const ewpt_lattice &L, const Phys &P, double *diag, double sw,
                             double cw, hipStream_t st);
// Fixed-order sum / min (slots in minmask) of nblk rows of partials, stride
// doubles per row, slots 0..ns-1 (ns <= 32) into res[0..ns-1]; overwrites
// some rows of part.
hipError_t launch_fixed_reduce(double *part, long nblk, int stride, int ns, unsigned minmask, double *res,
                               hipStream_t st);

// ewpt_brick_f32.hip -- mixed-precision stage kernels
bool brick_exists_f32(int variant);
bool brick_supported_f32(int variant, const ewpt_lattice &L);
hipError_t launch_stage_brick_f32(int variant, int stage, const StageBufs<float> &b, const ewpt_lattice &L,
                                  const Phys &P, double dt, hipStream_t st);
// mixed variants 31, 33 (+100*order): stage 1 with the diagnostics as above,
// densities in FP64 from the FP32 high words and pass derivatives
bool brick_has_diag_f32(int variant);
long diag_rows_f32(const ewpt_lattice &L);
hipError_t launch_stage1_diag_f32(int variant, const StageBufs<float> &b, const ewpt_lattice &L, const Phys &P,
                                  double dt, double *diag, double sw, double cw, hipStream_t st);

// ewpt_mixed.hip
hipError_t launch_split(const double *y, float *hi, float *lo, long n, hipStream_t st);
hipError_t launch_join(const float *hi, const float *lo, double *y, long n, hipStream_t st);
hipError_t launch_copy(const double *a, double *b, long n, hipStream_t st);

// ewpt_slab.hip -- ghost planes of z-slab contexts: HALO planes below and
// above the nz interior planes of every array (the stencil radius)
constexpr int HALO = 3;
hipError_t launch_halo_pack(const double *a, long n3, long plane, int nz, int nch, double *lo, double *hi,
                            hipStream_t st);
hipError_t launch_halo_pack(const float *a, long n3, long plane, int nz, int nch, float *lo, float *hi,
                            hipStream_t st);
hipError_t launch_halo_unpack(double *a, long n3, long plane, int nz, int nch, const double *lo, const double *hi,
                              hipStream_t st);
hipError_t launch_halo_unpack(float *a, long n3, long plane, int nz, int nch, const float *lo, const float *hi,
                              hipStream_t st);

// stage kernel variants selectable through ewpt_set_variant
constexpr int VARIANT_AUTO = 0;
constexpr int VARIANT_GATHER = 9;

// ewpt_init.hip
void mode_table(unsigned long long seed, int nmodes, int kmax, double amp, double mom_scale, double *t);
hipError_t launch_fill_modes(double *y, const ewpt_lattice &L, const double *dmodes, int nmodes, double vev,
                             hipStream_t st);

// ewpt_ic.hip -- bubble and thermal initial conditions
hipError_t launch_add_bubbles(double *y, const ewpt_lattice &L, const double *dbub, int nb, int profile, double mH,
                              hipStream_t st);
// returns 0, or -1 with *what naming the failed call; work holds 3*N^3 complex doubles
int thermal_state(double *y, const ewpt_lattice &L, double T, double meff_sq, unsigned long long seed, int profile,
                  void *work, hipStream_t st, const char **what);
// thermal_state (signed modes, Gamma = div W, Sigma = div Y) plus thermal E
// and pi, with E projected onto the Gauss law (FP64 CG + FFT).  work: 3 N^3
// complex doubles; wa: 15 N^3 doubles; wb: 9 N^3 doubles.  stats: residual
// norm before, after, CG iterations, SU(2) part after.  0 or -1 (*what).
int thermal_gauss_state(double *y, const ewpt_lattice &L, double T, double meff_sq, unsigned long long seed,
                        double gw, double gy, double tol, int maxit, void *work, double *wa, double *wb,
                        hipStream_t st, double stats[4], const char **what);

// ewpt_batch.hip -- one persistent launch per stage over the bricks of many
// lattices.  BatchLat describes one lattice: its stage buffers for parity 0
// (the context's current state) and parity 1 (state and update swapped).
template <typename T>
struct BatchLat {
    ewpt_lattice L;
    Phys P;
    double dt;
    StageBufs<T> b[2][4];
};
struct BatchPlan;
bool batch_variant_fits(int variant, bool f32, const ewpt_lattice &L);
BatchPlan *batch_plan_create(int variant, bool f32, const std::vector<ewpt_lattice> &lats, int device,
                             const char **what);
void batch_plan_destroy(BatchPlan *p);
// upload the descriptors of lats (BatchLat<float> or BatchLat<double>, one per lattice); 0 or -1
int batch_plan_set(BatchPlan *p, const void *lats, hipStream_t st);
hipError_t batch_plan_launch(const BatchPlan *p, int parity, int stage, bool beta, hipStream_t st);
int batch_plan_tiles(const BatchPlan *p);

// ewpt_net.hip -- learned corrector (3x3x3 conv net on the matrix cores)
struct Net;
bool net_lattice_ok(const ewpt_lattice &L);
Net *net_create(int device, unsigned long long seed, const char **what);
void net_destroy(Net *n);
long net_param_count();
hipError_t net_set_params(Net *n, const float *host);
hipError_t net_get(Net *n, int which, float *host);   // 0 = parameters, 1 = last gradients
hipError_t net_set_norm(Net *n, const float *in_inv, const float *out_scale);
void net_set_input_mode(Net *n, int mode);   // 0 state, 1 lattice Laplacian of the state
hipError_t net_apply(Net *n, double *y, const ewpt_lattice &L);
hipError_t net_train(Net *n, const double *yc, const double *yf, const ewpt_lattice &L, double lr, double *loss);
hipError_t net_conv_host(int nx, int ny, int nz, int epi, const unsigned short *X, const unsigned short *W,
                         const float *bias, const float *zprev, float *Zout, unsigned short *Aout);
hipError_t net_wgrad_host(int nx, int ny, int nz, const unsigned short *dZ, const unsigned short *X, float *dW,
                          float *db);

inline Phys phys_of(const ewpt_params &p)
{
    return Phys{p.gw, p.gy, p.gp2, p.beta_w, p.beta_y, p.vev, p.lambda, p.jex, p.gamma};
}

}  // namespace ewpt
