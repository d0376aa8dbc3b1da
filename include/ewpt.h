/*
 * ewpt.h -- C ABI of libewpt, the MI355X implementation of the RK4 time step
 * of the SU(2)xU(1) Higgs-gauge lattice evolution of Teerthal/EWPT-GPU.
 *
 * Plain C types only; callable through Julia `ccall`, Python `ctypes`, or C.
 * Every entry point returns 0 on success or a negative EWPT_ERR_* code; none
 * throws or aborts.  A context owns its device memory and one HIP stream; the
 * caller owns every host buffer.  Calls are synchronous (they return after the
 * device work finished) unless the name says otherwise.  A context is not
 * re-entrant; distinct contexts may be used from distinct threads.
 *
 * State layout ("36-channel state"): 36 channels, each an nx*ny*nz FP64 array
 * with x fastest (element (i,j,k) at i + nx*(j + ny*k)), channels in the
 * argument order of rk4_step! (gpu-kernels-rk4/main_rk4.jl:52-74):
 *    0-3   phi_1..phi_4                 20-23 dphi_a/dt   (pi)
 *    4-12  W^a_i, a=1..3 outer, i=x,y,z  24-32 dW^a_i/dt  (E^a_i)
 *    13-15 Y_x, Y_y, Y_z                 33-35 dY_i/dt    (E^Y_i)
 *    16-18 Gamma^1..3, 19 Sigma
 * This is the column-major Julia Array{Float64,3} of the reference, seen
 * through a C pointer.
 */
#ifndef EWPT_H
#define EWPT_H

#ifdef __cplusplus
extern "C" {
#endif

#define EWPT_NCH 36

/* error codes */
#define EWPT_OK                0
#define EWPT_ERR_INVAL        -1  /* bad argument (null pointer, size, enum) */
#define EWPT_ERR_HIP          -2  /* a HIP runtime call failed */
#define EWPT_ERR_OOM          -3  /* device allocation failed */
#define EWPT_ERR_NONFINITE    -4  /* a reduction met NaN/Inf in the state */
#define EWPT_ERR_STATE        -5  /* call made in the wrong order (no state) */
#define EWPT_ERR_UNSUPPORTED  -6  /* valid request the library cannot serve */

/* precision of a context */
#define EWPT_FP64   0   /* FP64 storage and arithmetic */
#define EWPT_MIXED  1   /* FP32 storage/stencils, compensated RK4 sum, FP64 reductions */

typedef struct ewpt_ctx ewpt_ctx;

/* Lattice: periodic nx*ny*nz, spacing dx (parameters.jl:14, 45). */
typedef struct {
    int nx, ny, nz;
    double dx;
} ewpt_lattice;

/* Per-step physics scalars: the ten trailing scalars of rk4_step!
 * (main_rk4.jl:74: gw,gy,gp2,beta_W,beta_Y,vev,lambda,dx,Jex,gamma) with dx
 * moved into ewpt_lattice and dt added (parameters.jl:46). */
typedef struct {
    double gw, gy, gp2, beta_w, beta_y, vev, lambda, dt, jex, gamma;
} ewpt_params;

/* Replaces the CUDA.zeros allocation block of run() (main_rk4.jl:5176-5299).
 * nx,ny,nz >= 8.  device = HIP ordinal. */
int ewpt_create(ewpt_ctx **out, const ewpt_lattice *lat, int precision, int device);
int ewpt_destroy(ewpt_ctx *ctx);

/* Copy the 36-channel state in / out (host buffers of 36*nx*ny*nz doubles).
 * Replaces the IC upload and the Array(...) downloads of run()
 * (main_rk4.jl:5640-5698 write into, 6094-6103 read from the state). */
int ewpt_set_state(ewpt_ctx *ctx, const double *host_state);
int ewpt_get_state(ewpt_ctx *ctx, double *host_state);

/* The same on a device buffer of 36*nx*ny*nz doubles on the context's
 * device (a CuArray the caller owns, as the reference's state arrays are):
 * device-to-device copies, synchronous.  Replaces the copyto! of
 * CuArrays into and out of the evolved fields (main_rk4.jl:5640-5698,
 * 6094-6103) for callers that keep their state on the GPU. */
int ewpt_set_state_device(ewpt_ctx *ctx, const double *dev_state);
int ewpt_get_state_device(ewpt_ctx *ctx, double *dev_state);

/* The mixed-precision state as the FP32 pair it is evolved in (host buffers
 * of 36*nx*ny*nz floats each; the state is hi + lo), so that a checkpoint
 * resumes bit for bit; the FP64 view of ewpt_get_state rounds the pair when
 * lo is below FP64 resolution of hi.  No counterpart in the reference, which
 * keeps FP64 state and writes no checkpoint; the state get/set it would bind
 * is that above.  EWPT_ERR_INVAL on an FP64 context. */
int ewpt_get_state_pair(ewpt_ctx *ctx, float *host_hi, float *host_lo);
int ewpt_set_state_pair(ewpt_ctx *ctx, const float *host_hi, const float *host_lo);

/* Measurement aid, no counterpart in the reference: a plain FP64 device copy
 * of n doubles with one 8-byte access per lane (the stage kernels' access
 * width), launched reps + 1 times on stream; *ms = mean time of the last
 * reps.  The reference point for memory-traffic counters.  With dev_src and
 * dev_dst both NULL the library allocates (and frees) the two buffers on
 * the current device. */
int ewpt_copy_probe(const double *dev_src, double *dev_dst, long n, int reps, void *stream, double *ms);

/* Device-to-device copy of the state of src into dst (same device,
 * precision and lattice). */
int ewpt_copy_state(ewpt_ctx *dst, ewpt_ctx *src);

/* Fill the context state on the device with a seeded smooth field: channel c
 * is the sum of nmodes cosines with wave numbers |k_x|,|k_y|,|k_z| <= kmax,
 * phases uniform in [0,2pi) and amplitudes uniform in [-amp,amp]/nmodes,
 * plus vev on phi_4 (channel 3); channels 20..35 (momenta) are scaled by
 * mom_scale.  The mode table comes from splitmix64(seed); the same table is
 * returned in modes_out (if not NULL, 36*nmodes*5 doubles: kx,ky,kz,phase,amp)
 * so that a CPU can rebuild the identical field.  Synthetic input for tests
 * and benchmarks at sizes where a host upload is impractical. */
int ewpt_init_modes(ewpt_ctx *ctx, unsigned long long seed, int nmodes, int kmax, double amp,
                    double vev, double mom_scale, double *modes_out);

/* Bubble-nucleation initial condition, replacing the initializer! launches of
 * the bubble branch of run() (kernel main_rk4.jl:4930-4946, driver
 * 5640-5698).  Zeroes the whole state, then adds rmag_b(x)*(p1,p2,p3,p4) to
 * phi for every bubble b.  bubbles holds nb rows of 7 doubles
 * (ib, jb, kb, p1, p2, p3, p4); the caller draws centres and orientations
 * (gen_phi, convenients.jl:8-14), as the reference driver does on the host.
 *   profile 0 (reference): ib,jb,kb are 1-based sub-lattice coordinates and
 *     rmag = (1+a^2) e^{-mH r/sqrt2} / (1 + a^2 e^{-sqrt2 mH r}), a = sqrt2-1,
 *     with r^2 = sum_d sin^2(rk_d ((i_d-1) - (ib_d-0.5)) dx)/rk_d^2,
 *     rk_d = pi/((N_d-1) dx), i_d 1-based (main_rk4.jl:4935-4936);
 *   profile 1 (tanh wall): ib,jb,kb are real 0-based positions and
 *     rmag = (1 - tanh((r - 3/mH) mH/2))/2 with r the periodic
 *     minimum-image distance.
 * 0 <= nb <= 2^20; mH > 0. */
int ewpt_init_bubbles(ewpt_ctx *ctx, const double *bubbles, int nb, int profile, double mH);

/* Thermal initial condition, replacing the thermal branch of run()
 * (thermal_initializer_gauge! / thermal_initializer_phi!,
 * main_rk4.jl:5029-5152, driver 5349-5638).  Zeroes the state, then for each
 * gauge field (W^1, W^2, W^3, Y) and each Higgs component builds Gaussian
 * modes weighted by sqrt(n/(omega vol)), n = 1/(exp(omega/T)-1),
 * vol = N^3 dx^3 (gauge: transverse polarisations, omega = |k|; Higgs:
 * omega = sqrt(|k|^2 + meff_sq)), and stores the real part of their
 * unnormalised inverse FFT.  The normal draws come from a counter-based
 * generator keyed by (seed, stream, mode index), so a CPU can rebuild the
 * same state.  profile 0 keeps the reference's mode numbering (array index m
 * carries the lattice momentum of index m+1, main_rk4.jl:5040-5046) and leaves
 * Gamma = Sigma = 0; profile 1 uses the signed index and sets Gamma^a = div W^a,
 * Sigma = div Y.  Momenta stay zero, so the Gauss law holds with E = 0.
 * Needs T > 0 and finite meff_sq. */
int ewpt_init_thermal(ewpt_ctx *ctx, double T, double meff_sq, unsigned long long seed, int profile);

/* Thermal initial condition with thermal momenta that satisfy the Gauss
 * law: the configuration's "Gauss law is solved for the initial electric
 * fields".  No counterpart in the reference, whose thermal runs start from
 * pi = E = 0 and drive the Higgs with J_ext instead (main_rk4.jl:5154-5162).
 * Builds the fields of ewpt_init_thermal profile 1 (signed mode index,
 * Gamma = div W, Sigma = div Y), then E^a, E^Y (transverse modes) and pi
 * with Bose-Einstein amplitudes sqrt(n omega / vol), and projects them, in
 * FP64, onto the Gauss law with the charges k_Gamma / k_Sigma weight by gp2
 * (main_rk4.jl:603-657):
 *   d_i E^a_i + gw eps^{abc} W^b_i E^c_i = gw Q^a(phi, pi),  d_i E^Y_i = gy Q^Y(phi, pi)
 * (d the 6th-order first derivative): pi is corrected on the lattice modes
 * d annihilates so the U(1) charge has no component there, E^Y gains a
 * gradient from an exact FFT solve, E^a a covariant gradient from a
 * conjugate-gradient solve of D^T D chi = G E0 - gw Q (FFT-preconditioned,
 * relative residual 1e-13).  p supplies gw and gy.  stats (if not NULL):
 * [0] residual norm sqrt(sum r^2 dx^3) over G^1..3, G^Y before the
 * projection, [1] after it, [2] CG iterations, [3] the SU(2) part of [1].
 * Needs T > 0; returns EWPT_ERR_NONFINITE if the solve produced NaN/Inf. */
int ewpt_init_thermal_gauss(ewpt_ctx *ctx, double T, double meff_sq, unsigned long long seed, const ewpt_params *p,
                            double stats[4]);

/* One RHS evaluation, host buffers: the @cuda rk4_step!(...) launch
 * (main_rk4.jl:52-661; launch 6155).  in = 36-channel stage state, out =
 * the 36 k/kt arrays in the same channel order (0-15 k = momenta copies,
 * 16-19 k_Gamma/k_Sigma, 20-35 kt). */
int ewpt_rhs(ewpt_ctx *ctx, const ewpt_params *p, const double *host_in, double *host_out);

/* Same on device buffers (FP64, 36*nx*ny*nz each, distinct), enqueued on
 * `stream` (a hipStream_t, NULL = default stream); asynchronous.  Both run
 * the LDS-brick stage-1 kernel (16x4x4, 32x4x4, 16x8x4 or 8x8x8 bricks,
 * whichever tiles the lattice; else the one-site-per-thread gather kernel). */
int ewpt_rhs_device(const ewpt_lattice *lat, const ewpt_params *p,
                    const double *dev_in, double *dev_out, void *stream);

/* nsteps full RK4 steps of the context state with fixed scalars: the
 * `for it in 1:nte` body of main_rk4.jl:6146-6441 (4 stage launches + the
 * broadcast combinations), with the stage combinations fused into the
 * stage kernels. */
int ewpt_step(ewpt_ctx *ctx, const ewpt_params *p, int nsteps);

/* nsteps RK4 steps as ewpt_step, and the energy table (ewpt_energies
 * layout: columns 0-6, 9, 11; the EM columns 7, 8, 12 are 0 here and filled
 * by ewpt_step_diag_em below) and the Gauss
 * norms (ewpt_gauss layout) of the state before the first step.  On the
 * FP64 path with stage variant 32 or 34 (34 is the automatic one where it
 * tiles) the first step's stage 1 reduces them from the values it stages anyway (the
 * reductions of main_rk4.jl:6449-6560 and the Gauss residuals of the
 * k_Gamma / k_Sigma charges, wave-reduced in the stage kernel): no extra
 * pass over the state; equal to ewpt_energies / ewpt_gauss to rounding (the
 * sums are grouped differently).  On the mixed path with variant 31 or 33
 * (33 is the automatic one) the FP32 stage 1 does the same, the densities
 * computed in FP64 from the FP32 high words and the FP32 pass derivatives:
 * equal to the FP64 reductions of the state (high + compensation word) to
 * the FP32 rounding of the fields (~1e-7 relative).  Otherwise (other
 * variants) the standalone reductions run first.  A snapshot loop calls it
 * for the step after the snapshot.  Returns EWPT_ERR_NONFINITE (after
 * stepping and filling the outputs) if a density is NaN/Inf. */
int ewpt_step_diag(ewpt_ctx *ctx, const ewpt_params *p, int nsteps, double energies[13], double gauss[8]);
/* ewpt_step_diag with the EM columns 7, 8, 12 of the energy table as well
 * (ewpt_em_energies of the same state with Weinberg angle theta_w: the
 * magnetic energies of compute_B_* / compute_B_2_*, main_rk4.jl:2303-2822,
 * 3718-4250, as the snapshot block takes them, 6861-6901), reduced by the
 * same stage 1: every column of a snapshot's energy row from one pass. */
int ewpt_step_diag_em(ewpt_ctx *ctx, const ewpt_params *p, double theta_w, int nsteps, double energies[13],
                      double gauss[8]);

/* Like ewpt_step, and reports mean device time per launch (ms) of each
 * stage kernel, measured with HIP events on the context stream:
 * stage_ms[0..3] = stages 1..4.  *total_ms = event time of the whole run.
 * On a slab context each stage's time includes the halo exchange before it. */
int ewpt_step_timed(ewpt_ctx *ctx, const ewpt_params *p, int nsteps,
                    double stage_ms[4], double *total_ms);

/* ---- z-slabs: one lattice split over several contexts (one per GPU).
 * No reference counterpart on the single-GPU path; the reference's own
 * multi-GPU driver (main_rk4_multiple.jl:69-213, which its README says no
 * longer works) exchanges boundary planes with MPI.  A slab context holds the
 * local lattice nx x ny x nz (nz = its share of the z axis) with 3 ghost
 * planes below and above in every array; its stage kernels read the z halo
 * from the ghost planes instead of wrapping.  Before stage s the planes that
 * stage reads at halo sites are exchanged: the state before stage 1 (what =
 * 0; mixed: its FP32 high word) and, before stage s > 1 (what = s-1), k_{s-1}
 * (20 channels) for the variants that form their input on load, or the
 * stored stage input u_s (36 channels) for the variants that store it (the
 * automatic choice where it fits).  A multi-GPU driver runs, per stage and rank:
 *   ewpt_slab_pack(ctx, what, send_lo, send_hi)   lowest / highest 3 interior planes
 *   send send_lo to the lower and send_hi to the upper z-neighbour; receive
 *   recv_lo from the lower neighbour (its send_hi), recv_hi from the upper
 *   ewpt_slab_unpack(ctx, what, recv_lo, recv_hi)
 *   ewpt_stage(ctx, p, s)
 * pack, unpack and stage are asynchronous on the context's stream
 * (ewpt_get_stream); buffers are device memory of ewpt_slab_halo_bytes each.
 * ewpt_step on a slab context alone is periodic in z (it exchanges with
 * itself).  Slab contexts take set/get/copy_state, steps, the brick stage
 * variants and the calls below; the
 * whole-lattice calls (ICs, RHS, energies, Gauss, EM fields, spectrum,
 * corrector, batches) return EWPT_ERR_UNSUPPORTED on them. */
int ewpt_create_slab(ewpt_ctx **out, const ewpt_lattice *local, int precision, int device);
/* Bytes of one side's halo buffer for array `what` (0 state, 1..3 before
 * stage 2..4: k_s or the stored stage input, by the active variant). */
int ewpt_slab_halo_bytes(ewpt_ctx *ctx, int what, unsigned long long *bytes);
/* The context's own send buffers (device, large enough for what = 0). */
int ewpt_slab_buffers(ewpt_ctx *ctx, void **send_lo, void **send_hi);
int ewpt_slab_pack(ewpt_ctx *ctx, int what, void *send_lo, void *send_hi);
int ewpt_slab_unpack(ewpt_ctx *ctx, int what, const void *recv_lo, const void *recv_hi);
/* One stage launch (1..4) without exchange, asynchronous; stage 4 makes
 * the new state current.  Any context.  Call the stages of a step in order
 * 1, 2, 3, 4: stage s reads the derivatives (and, with stored stage inputs,
 * the input array) that stage s-1 wrote. */
int ewpt_stage(ewpt_ctx *ctx, const ewpt_params *p, int stage);
/* The context's HIP stream (hipStream_t) and a wait for it. */
int ewpt_get_stream(ewpt_ctx *ctx, void **stream);
int ewpt_sync(ewpt_ctx *ctx);
/* nsteps steps of n slab contexts on one device forming a ring along z
 * (slab i's neighbours are i-1 and i+1 mod n), the exchange done with
 * device copies between the contexts: the single-GPU stand-in for the
 * exchange a multi-GPU driver does over xGMI.  All slabs share device,
 * precision, nx, ny and dx; the result equals stepping the stacked lattice
 * (slab 0 lowest) in one context with the same variant. */
int ewpt_slab_step(ewpt_ctx *const *slabs, int n, const ewpt_params *p, int nsteps);
/* Like ewpt_slab_step; on slab 0's stream: mean kernel time per stage,
 * mean exchange time per step (the four exchanges) and the total. */
int ewpt_slab_step_timed(ewpt_ctx *const *slabs, int n, const ewpt_params *p, int nsteps,
                         double stage_ms[4], double *exchange_ms, double *total_ms);

/* Energy table in the reference's column order (main_rk4.jl:6104-6116):
 * out[0..6] = PE, KE, GE, EEW, MEW, EEY, MEY (sum * dx^3, FP64),
 * out[9] = min|phi|, out[10] = 0, out[11] = mean|phi|.  The EM columns
 * out[7], out[8], out[12] are left 0: ewpt_em_energies computes them (they
 * need the weak mixing angle).
 * Returns EWPT_ERR_NONFINITE (after filling out) if any density is NaN/Inf. */
int ewpt_energies(ewpt_ctx *ctx, const ewpt_params *p, double out[13]);

/* The energy table of ewpt_energies for a state in a caller's device buffer,
 * with no allocation: no counterpart in the reference; for callers that
 * cannot allocate device memory (code under HIP graph capture, serving
 * loops, harnesses that forbid hipMalloc).  dev_state = 36*nx*ny*nz doubles
 * and dev_out = 13 doubles, both on the current device; the 13 columns are
 * those ewpt_energies returns (sums grouped differently: equal to rounding).
 * Enqueued on `stream` (a hipStream_t of the current device, NULL = default
 * stream), asynchronous, nothing allocated or freed.  The partial sums live
 * in a fixed library array of 2048 x 16 doubles per device, whatever the
 * lattice size; every call on a device shares it, so calls that overlap on
 * one device race.  A non-finite density shows as a NaN/Inf column (no
 * EWPT_ERR_NONFINITE: the call does not wait for the result).  Lattices that
 * tile 16x4x4 (else EWPT_ERR_UNSUPPORTED); sizes as ewpt_create (else
 * EWPT_ERR_INVAL). */
int ewpt_energies_device(const ewpt_lattice *lat, const ewpt_params *p, const double *dev_state, double *dev_out,
                         void *stream);

/* Electromagnetic columns of the energy table, FP64:
 *   out[0] = column 8:  0.5 sum |B|^2 dx^3   (main_rk4.jl:5830-5868, 6113)
 *   out[1] = column 9:  0.5 sum |B_2|^2 dx^3 (5887-5926, 6114)
 *   out[2] = column 13: sum B_2 . A dx^3     (6901; the reference's first
 *            snapshot, 6118, omits the dx^3)
 * with B from compute_Bx!/By!/Bz! (2303-2822), B_2 from compute_B_2_x/y/z!
 * (3718-4250: |phi|^4 in place of vev^2, 0 where phi = 0) and A from
 * compute_Ax!/Ay!/Az! (3012-3528).  theta_w is the weak mixing angle the
 * reference passes to these kernels (asin(sqrt(0.22)), parameters.jl:17,26).
 * Returns EWPT_ERR_NONFINITE if any field value is NaN/Inf. */
int ewpt_em_energies(ewpt_ctx *ctx, const ewpt_params *p, double theta_w, double out[3]);

/* The EM fields themselves, for snapshots (phys_stack, main_rk4.jl:5832-6056):
 * host_out = 9 N^3 doubles, channels Bx, By, Bz, B2x, B2y, B2z, Ax, Ay, Az,
 * each x fastest. */
int ewpt_em_fields(ewpt_ctx *ctx, const ewpt_params *p, double theta_w, double *host_out);

/* The y-slice snapshot of the reference (slices_arr[snp, 1:7, :, :],
 * main_rk4.jl:5700-5930, 6450-6720) on the plane y = j (0-based; the
 * reference takes its 1-based Ny/2, i.e. j = ny/2 - 1): host_out = 7 planes of
 * nx*nz doubles, x fastest, each value times dx^3:
 *   0 PE, 1 KE, 2 GE densities, 3 MEW + EEW + MEY + EEY, 4 Bx + By + Bz,
 *   5 B2x + B2y + B2z, 6 zero (the reference never fills the seventh type). */
int ewpt_y_slice(ewpt_ctx *ctx, const ewpt_params *p, double theta_w, int j, double *host_out);

/* Shell-averaged magnetic power spectrum: the reference's fft of B_x, B_y,
 * B_z (main_rk4.jl:5831-5869) followed by convolve_1d(|B_x|^2 + |B_y|^2 +
 * |B_z|^2, Nx, Ny, Nz, cut...) (spec_routines.jl:47-77, called at
 * main_rk4.jl:6066-6069 and 7069-7072).  field 0 = B, 1 = B_2.  The shells are
 * the first Kc_bin_nums(cut[0], cut[1], cut[2]) distinct |K| of the lattice
 * (K the signed integer mode index, spec_routines.jl:5-43); the reference's
 * cut is spec_cut = (N/4, N/4, N/4) (parameters.jl:78).  On return
 * k_out[b] = |K| and p_out[b] = the mean of sum_c |B_c(K)|^2 (unnormalised
 * forward DFT) over the lattice points of shell b, b < *nbins.  With
 * k_out = p_out = NULL only *nbins is set (no state needed).  Fails with
 * EWPT_ERR_INVAL when maxbins < *nbins. */
int ewpt_spectrum(ewpt_ctx *ctx, const ewpt_params *p, double theta_w, int field, const int cut[3],
                  int maxbins, double *k_out, double *p_out, int *nbins);

/* Gauss-law residual norms, sqrt(sum r^2 dx^3):
 * out[0..2] G^a = div E^a - rho^a, out[3] G^Y = div E^Y - rho^Y,
 * out[4..6] C^a = Gamma^a - div W^a, out[7] C^Y = Sigma - div Y,
 * where rho is the charge that k_Gamma/k_Sigma weight by gp2
 * (main_rk4.jl:603-657).  Returns EWPT_ERR_NONFINITE if any is NaN/Inf. */
int ewpt_gauss(ewpt_ctx *ctx, const ewpt_params *p, double out[8]);

/* Kernel variant of the stage launches.  All variants compute the same
 * result (to the rounding of the arithmetic); they exist for the autotuner
 * and for measurements.
 *   0 = automatic: FP64 the first of 334, 332, 324, 325, 320, 311, 310, 210,
 *       211, 2 that tiles the lattice, else 9; mixed the first of 333, 331,
 *       315, 324, 311, 211, 310, 210, 325.  The stored-stage-input variants
 *       (15, 24, 31, 32, 33, 34) only when their buffer fits in device memory
 *       and not on slab contexts.
 *   LDS-brick stage kernels (brick shape, threads, form):
 *       2 (FP64) = 8x8x8, 512 threads, prefetch depth 1;
 *      10 = 16x8x4, 512 threads, every stage input formed on load;
 *      11 = 32x4x4, 512 threads, formed on load;
 *      15 (mixed) = 11 with stored stage inputs (stages 1-3 also write the
 *           next stage's input, 36 channels, which stages 2-4 read as one array);
 *      20 (FP64) = 11 with the E and Gamma/Sigma centre values loaded with
 *           their own passes (late centres);
 *      24 = 16x4x4 bricks of 256 threads (two workgroups per CU in FP64),
 *           late centres, stored stage inputs;
 *      25 = 24 with every stage input formed on load;
 *      31 (mixed) = 15 with the next input's y loaded during the last passes,
 *           a lean stage-4 stash (two workgroups per CU), prefetch 4/4/4/3;
 *      32 (FP64) = 24 with the next input's y loaded during the last passes
 *           and stage 4's update loads spread over its passes (bitwise 24);
 *      33 (mixed) = 31 with the same spread stage 4, stages 2-3 storing the
 *           next input's field channels during their last passes and stage 1
 *           parking its non-field centres in LDS (bitwise 31);
 *      34 (FP64) = 32 with y rebuilt instead of read: stage 2 forms the next
 *           input from u2 - k1 dt/2, stage 4 its update from u4 - k3 dt
 *           (one more rounding per rebuilt value; 416 B/site less per step).
 *       15, 24, 31, 32, 33, 34 hold one more state-sized buffer (allocated
 *       when first selected).  Prefetch depth 4/4/3/2 by stage unless stated.
 *   9 = gather kernel (FP64; one site per thread, any lattice size)
 * A brick variant plus 100*o selects brick enumeration order o (0 = x,y,z;
 * 1 = x,z,y; 2 = x then 4x4 tiles of (y,z) brick rows; 3-9 = cubes of
 * 2x4x4, 4x4x4, 2x8x4, 2x4x8, 1x8x8, 2x8x8, 1x4x4 bricks), e.g. 311.
 * Setting a variant the precision does not have returns EWPT_ERR_INVAL; one
 * that does not tile the lattice returns EWPT_ERR_UNSUPPORTED. */
int ewpt_set_variant(ewpt_ctx *ctx, int variant);
int ewpt_get_variant(ewpt_ctx *ctx, int *variant);
/* The variant the next step launches (what 0 = automatic resolves to). */
int ewpt_active_variant(ewpt_ctx *ctx, int *variant);

/* Device bytes the context holds. */
int ewpt_device_bytes(ewpt_ctx *ctx, unsigned long long *bytes);

const char *ewpt_strerror(int code);
/* Last error message of this context ("" if none). */
const char *ewpt_last_error(ewpt_ctx *ctx);
/* Library version string. */
const char *ewpt_version(void);

/* ---- batches: many independent lattices stepped together.
 * No counterpart in the reference, which runs one lattice per process
 * (the configuration with many seeded realisations of uneven sizes).
 * A batch refers to existing contexts (same device and precision; each
 * keeps its own state, ICs, energies, spectra).  One step of the batch is
 * one persistent launch per RK4 stage over the bricks of every lattice; two
 * steps are captured once in a hipGraph and replayed.  The result for each
 * lattice is bitwise that of ewpt_step on it alone with the same stage
 * variant.  The contexts must not be used by other calls while a batch call
 * runs; destroying a context invalidates the batch. */
typedef struct ewpt_batch ewpt_batch;
int ewpt_batch_create(ewpt_batch **out, ewpt_ctx *const *ctxs, int n);
int ewpt_batch_destroy(ewpt_batch *b);
const char *ewpt_batch_last_error(ewpt_batch *b);
/* Brick variant for every lattice (as ewpt_set_variant: 10/11 (+100*order)
 * for both precisions, also 2 and the stored-stage-input form 32 for FP64 and
 * the stored-stage-input forms 15/31/33 for mixed, which allocate each
 * context's stage-input buffer), plus 1000 for the persistent launch form (a
 * resident grid striding through the bricks instead of one workgroup per
 * brick); 0 = the first candidate that tiles all (when a context has no room
 * for a stage-input buffer, the first that needs none). */
int ewpt_batch_set_variant(ewpt_batch *b, int variant);
int ewpt_batch_get_variant(ewpt_batch *b, int *variant);
/* nsteps RK4 steps of every lattice; p holds one ewpt_params per lattice. */
int ewpt_batch_step(ewpt_batch *b, const ewpt_params *p, int nsteps);
/* Same, and *ms = device time of the nsteps steps (HIP events). */
int ewpt_batch_step_timed(ewpt_batch *b, const ewpt_params *p, int nsteps, double *ms);
/* Autotuner: times every candidate variant that tiles all lattices over
 * nsteps steps with dt set to 0 (which leaves every state bitwise
 * unchanged), selects the fastest, and caches the choice for this set of
 * lattice shapes for the life of the process.  *best_ms = ms per step of the
 * winner (-1 when the choice came from the cache). */
int ewpt_batch_tune(ewpt_batch *b, const ewpt_params *p, int nsteps, int *best_variant, double *best_ms);

/* ---- learned coarse-step corrector.
 * No counterpart in the reference; for callers that trade the fine RK4
 * step for one step of factor*dt plus a learned correction.  A periodic
 * 3x3x3 convolution network 36 -> 64 -> 64 -> 36 with SiLU, BF16 weights and
 * activations, FP32 accumulation on the matrix cores, maps the coarse state
 * y_c (channel c scaled by 1/in_scale[c]) to out; the corrected state is
 * y_c + out_scale[c] * out[c].  Parameters: FP32, ewpt_net_param_count()
 * values, layers 1..3 each [27 taps][64 out][64 in] weights then 64 biases
 * (tap = (dz+1)*9 + (dy+1)*3 + (dx+1); in/out channels >= 36 of the first
 * and last layer are padding).  Lattices must tile 8x4x4; FP64 contexts. */
typedef struct ewpt_net ewpt_net;
int ewpt_net_create(ewpt_net **out, int device, unsigned long long seed);
int ewpt_net_destroy(ewpt_net *n);
const char *ewpt_net_last_error(ewpt_net *n);
long ewpt_net_param_count(void);
int ewpt_net_set_params(ewpt_net *n, const float *params);
int ewpt_net_get_params(ewpt_net *n, float *params);
/* gradients of the last ewpt_net_train_step, same layout as the parameters */
int ewpt_net_get_grads(ewpt_net *n, float *grads);
int ewpt_net_set_norm(ewpt_net *n, const float in_scale[36], const float out_scale[36]);
/* Network input: mode 0 = the coarse state, mode 1 = its 7-point lattice
 * Laplacian (sum of the 6 neighbours - 6 x centre, FP64 before the bf16
 * rounding); in_scale then scales the Laplacian. */
int ewpt_net_set_input(ewpt_net *n, int mode);
/* y <- y + out_scale * net(y / in_scale) on the context state */
int ewpt_net_apply(ewpt_net *n, ewpt_ctx *ctx);
/* One training step on the pair (coarse state, fine state): relative loss
 * sum (out - T)^2 / sum T^2 over sites and the 36 channels,
 * T = (y_fine - y_coarse)/out_scale (every pair weighs the same),
 * gradients by the backward kernels, then Adam (beta 0.9, 0.999) with
 * learning rate lr; lr = 0 computes loss and gradients only. */
int ewpt_net_train_step(ewpt_net *n, ewpt_ctx *coarse, ewpt_ctx *fine, double lr, double *loss);
/* ncoarse steps of factor*dt (p->dt is the fine step), each followed by the
 * correction.  When the Gauss-law residual norm sqrt(sum G^2) grows over a
 * coarse step by more than gauss_tol * ||E|| (||E||^2 = 2 (EEW + EEY)), the
 * step is redone as factor fine steps; *nfallback counts those
 * (gauss_tol = +inf never falls back and skips the residual evaluations).
 * The residual of each corrected state is reduced inside stage 1 of the next
 * coarse step where the stage variant has the diagnostic stage 1 (FP64
 * contexts); the context then keeps two saved states (2 x 36 N^3 doubles on
 * the corrector) to redo a failed step. */
int ewpt_step_corrected(ewpt_ctx *ctx, ewpt_net *n, const ewpt_params *p, int factor, int ncoarse,
                        double gauss_tol, int *nfallback);
/* Test hooks on host buffers: one convolution layer (mode 0: hidden layer,
 * Z = conv + bias, A = bf16(silu(Z)); mode 1: last layer, Z only; mode 2:
 * input gradient, A = bf16(conv(X, W) * silu'(zprev)) with W the mirrored,
 * transposed weights), and one weight gradient (dW[tap][out][in],
 * db[out]).  X, A, dZ: [site][64] bf16 bit patterns. */
int ewpt_net_conv_test(int nx, int ny, int nz, int mode, const unsigned short *X, const unsigned short *W,
                       const float *bias, const float *zprev, float *Z, unsigned short *A);
int ewpt_net_wgrad_test(int nx, int ny, int nz, const unsigned short *dZ, const unsigned short *X,
                        float *dW, float *db);

#ifdef __cplusplus
}
#endif
#endif /* EWPT_H */
