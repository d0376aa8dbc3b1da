/*
 * ewpt_oracle.h -- CPU restatement of the reference RK4 hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This code is the checker that the HIP library is
 * compared against.  Only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg may load it.  The product library never links it.
 *
 * State layout: 36 channels, each an nx*ny*nz FP64 array, x fastest
 * (idx = i + nx*(j + ny*k)), channel order as in the argument list of
 * rk4_step! (gpu-kernels-rk4/main_rk4.jl:52-74):
 *   0-3  phi_1..phi_4        4-12 W^a_i (a outer, i=x,y,z inner)
 *   13-15 Y_x,Y_y,Y_z        16-18 Gamma^1..3     19 Sigma
 *   20-23 dphi_a/dt          24-32 dW^a_i/dt      33-35 dY_i/dt
 */
#ifndef EWPT_ORACLE_H
#define EWPT_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
    double gw, gy, gp2, beta_w, beta_y, vev, lambda, dt, jex, gamma;
} oracle_params;

/* One RHS evaluation (main_rk4.jl:52-661): in = 36-channel stage state,
 * out = 36-channel k (0-19) / kt (20-35) arrays. */
void oracle_rhs(int nx, int ny, int nz, double dx, const oracle_params *p,
                const double *in, double *out);

/* nsteps full RK4 steps in place on y (main_rk4.jl:6146-6441). */
void oracle_rk4(int nx, int ny, int nz, double dx, const oracle_params *p,
                double *y, int nsteps);

/* Energy table, reference column order (main_rk4.jl:6104-6116):
 * out[0..6] = PE KE GE EEW MEW EEY MEY (each sum*dx^3), out[9] = min|phi|,
 * out[11] = mean|phi|.  Columns 7,8 (EM magnetic) and 12 (helicity) are
 * written by oracle_bfield_energies; column 10 is 0 as in the reference. */
void oracle_energies(int nx, int ny, int nz, double dx, const oracle_params *p,
                     const double *y, double out[13]);

/* Gauss-law residuals: out[0..3] = L2 of G^1,G^2,G^3,G^Y; out[4..7] = L2 of
 * C^1,C^2,C^3,C^Y (see ewpt_oracle.c for the definitions). */
void oracle_gauss(int nx, int ny, int nz, double dx, const oracle_params *p,
                  const double *y, double out[8]);

/* Electromagnetic fields (compute_Bx!/By!/Bz!, compute_B_2_*!, compute_A*!,
 * main_rk4.jl:2303-2822, 3012-3530, 3718-4250): out = 9 channels
 * Bx, By, Bz, B2x, B2y, B2z, Ax, Ay, Az. */
void oracle_em_fields(int nx, int ny, int nz, double dx, const oracle_params *p, double theta_w,
                      const double *y, double *out);

#ifdef __cplusplus
}
#endif
#endif
