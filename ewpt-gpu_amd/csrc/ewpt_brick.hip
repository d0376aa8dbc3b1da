// ewpt_brick.hip -- dispatch over the LDS-brick stage-kernel variants
// (kernel in ewpt_brick.h, one instantiation per ewpt_brick_v*.hip file so
// they compile in parallel).
//
//   1: 16x8x4 brick, 512 threads, prefetch distance 1, XCD-aware block order
//   2: 8x8x8 brick, 512 threads, prefetch distance 1, XCD-aware block order
//   3: 16x8x4 brick, 512 threads, prefetch distance 2, XCD-aware block order
//   4: 8x8x4 brick, 256 threads, prefetch distance 1, XCD-aware block order
//   5: 16x8x4 brick, 512 threads, prefetch distance 1, linear block order
//   6: 16x4x4 brick, 256 threads, prefetch distance 1, XCD-aware block order
//   7: 8x8x4 brick, 256 threads, prefetch distance 2, XCD-aware block order
//   8: 16x4x4 brick, 256 threads, prefetch distance 2, XCD-aware block order
//   10: 16x8x4 brick, 512 threads, prefetch depth 4/4/3/2 by stage, XCD-aware block order
//   11: 32x4x4 brick, 512 threads, prefetch depth 4/4/3/2 by stage, XCD-aware block order
//   12, 13: as 10, 11 with two channels per LDS barrier in stages 1-3
//   14, 15: as 10, 11 with stored stage inputs for stages 2-4 (brick_stored_inputs)
//   16, 17: as 10, 11 with a stored stage input for stage 4
//   18, 19: as 10, 11 with stored stage inputs for stages 3-4
//   20, 22: as 11, 10 with the E and Gamma/Sigma centres loaded with their
//       passes (late_centres); 21, 23: the same with stored stage inputs for
//       stages 2-4
//   24, 25: 16x4x4 bricks of 256 threads (two workgroups per CU in FP64) as
//       21 and 20; 26: 8x8x4 bricks of 256 threads as 21
//   27: 24 with stages 2-3 loading the y values of the next stage input
//       during their last passes (y_early); 28 exists for the mixed path only
//   29 (FP64): 27 with stage 4's partial-sum loads spread over its passes
//   32 (FP64): 29 with the field channels of y' updated in those passes
//   30 exists for the mixed path only
#include "ewpt_internal.h"

namespace ewpt {

bool brick_fits_v1(const ewpt_lattice &L);
hipError_t launch_brick_v1(int stage, const StageBufs<double> &b, const ewpt_lattice &L, const Phys &P, double dt,
                            int order, hipStream_t st);
bool brick_fits_v2(const ewpt_lattice &L);
hipError_t launch_brick_v2(int stage, const StageBufs<double> &b, const ewpt_lattice &L, const Phys &P, double dt,
                            int order, hipStream_t st);
bool brick_fits_v3(const ewpt_lattice &L);
hipError_t launch_brick_v3(int stage, const StageBufs<double> &b, const ewpt_lattice &L, const Phys &P, double dt,
                            int order, hipStream_t st);
bool brick_fits_v4(const ewpt_lattice &L);
hipError_t launch_brick_v4(int stage, const StageBufs<double> &b, const ewpt_lattice &L, const Phys &P, double dt,
                            int order, hipStream_t st);
bool brick_fits_v5(const ewpt_lattice &L);
hipError_t launch_brick_v5(int stage, const StageBufs<double> &b, const ewpt_lattice &L, const Phys &P, double dt,
                            int order, hipStream_t st);
bool brick_fits_v6(const ewpt_lattice &L);
hipError_t launch_brick_v6(int stage, const StageBufs<double> &b, const ewpt_lattice &L, const Phys &P, double dt,
                            int order, hipStream_t st);
bool brick_fits_v7(const ewpt_lattice &L);
hipError_t launch_brick_v7(int stage, const StageBufs<double> &b, const ewpt_lattice &L, const Phys &P, double dt,
                            int order, hipStream_t st);
bool brick_fits_v8(const ewpt_lattice &L);
hipError_t launch_brick_v8(int stage, const StageBufs<double> &b, const ewpt_lattice &L, const Phys &P, double dt,
                            int order, hipStream_t st);

bool brick_fits_v10(const ewpt_lattice &L);
hipError_t launch_brick_v10(int stage, const StageBufs<double> &b, const ewpt_lattice &L, const Phys &P, double dt,
                             int order, hipStream_t st);

bool brick_fits_v11(const ewpt_lattice &L);
hipError_t launch_brick_v11(int stage, const StageBufs<double> &b, const ewpt_lattice &L, const Phys &P, double dt,
                             int order, hipStream_t st);

bool brick_fits_v12(const ewpt_lattice &L);
hipError_t launch_brick_v12(int stage, const StageBufs<double> &b, const ewpt_lattice &L, const Phys &P, double dt,
                             int order, hipStream_t st);

bool brick_fits_v13(const ewpt_lattice &L);
hipError_t launch_brick_v13(int stage, const StageBufs<double> &b, const ewpt_lattice &L, const Phys &P, double dt,
                             int order, hipStream_t st);

bool brick_fits_v14(const ewpt_lattice &L);
hipError_t launch_brick_v14(int stage, const StageBufs<double> &b, const ewpt_lattice &L, const Phys &P, double dt,
                             int order, hipStream_t st);

bool brick_fits_v15(const ewpt_lattice &L);
hipError_t launch_brick_v15(int stage, const StageBufs<double> &b, const ewpt_lattice &L, const Phys &P, double dt,
                             int order, hipStream_t st);

bool brick_fits_v16(const ewpt_lattice &L);
hipError_t launch_brick_v16(int stage, const StageBufs<double> &b, const ewpt_lattice &L, const Phys &P, double dt,
                             int order, hipStream_t st);

bool brick_fits_v17(const ewpt_lattice &L);
hipError_t launch_brick_v17(int stage, const StageBufs<double> &b, const ewpt_lattice &L, const Phys &P, double dt,
                             int order, hipStream_t st);

bool brick_fits_v18(const ewpt_lattice &L);
hipError_t launch_brick_v18(int stage, const StageBufs<double> &b, const ewpt_lattice &L, const Phys &P, double dt,
                             int order, hipStream_t st);

bool brick_fits_v19(const ewpt_lattice &L);
hipError_t launch_brick_v19(int stage, const StageBufs<double> &b, const ewpt_lattice &L, const Phys &P, double dt,
                             int order, hipStream_t st);

bool brick_fits_v20(const ewpt_lattice &L);
hipError_t launch_brick_v20(int stage, const StageBufs<double> &b, const ewpt_lattice &L, const Phys &P, double dt,
                             int order, hipStream_t st);

bool brick_fits_v21(const ewpt_lattice &L);
hipError_t launch_brick_v21(int stage, const StageBufs<double> &b, const ewpt_lattice &L, const Phys &P, double dt,
                             int order, hipStream_t st);

bool brick_fits_v22(const ewpt_lattice &L);
hipError_t launch_brick_v22(int stage, const StageBufs<double> &b, const ewpt_lattice &L, const Phys &P, double dt,
                             int order, hipStream_t st);

bool brick_fits_v23(const ewpt_lattice &L);
hipError_t launch_brick_v23(int stage, const StageBufs<double> &b, const ewpt_lattice &L, const Phys &P, double dt,
                             int order, hipStream_t st);

bool brick_fits_v24(const ewpt_lattice &L);
hipError_t launch_brick_v24(int stage, const StageBufs<double> &b, const ewpt_lattice &L, const Phys &P, double dt,
                             int order, hipStream_t st);

bool brick_fits_v25(const ewpt_lattice &L);
hipError_t launch_brick_v25(int stage, const StageBufs<double> &b, const ewpt_lattice &L, const Phys &P, double dt,
                             int order, hipStream_t st);

bool brick_fits_v26(const ewpt_lattice &L);
hipError_t launch_brick_v26(int stage, const StageBufs<double> &b, const ewpt_lattice &L, const Phys &P, double dt,
                             int order, hipStream_t st);

bool brick_fits_v27(const ewpt_lattice &L);
hipError_t launch_brick_v27(int stage, const StageBufs<double> &b, const ewpt_lattice &L, const Phys &P, double dt,
                             int order, hipStream_t st);

bool brick_fits_v29(const ewpt_lattice &L);
hipError_t launch_brick_v29(int stage, const StageBufs<double> &b, const ewpt_lattice &L, const Phys &P, double dt,
                             int order, hipStream_t st);

bool brick_fits_v32(const ewpt_lattice &L);
hipError_t launch_brick_v32(int stage, const StageBufs<double> &b, const ewpt_lattice &L, const Phys &P, double dt,
                             int order, hipStream_t st);
int brick_variant_count() { return 29; }

bool brick_exists(int variant)
{
    const int v = variant % 100, order = variant / 100;
    return v >= 1 && v <= 33 && v != 9 && order <= 4;
}

bool brick_stored_inputs(int variant)
{
    const int v = variant % 100;
    return variant >= 0 && ((v >= 14 && v <= 19) || v == 21 || v == 23 || v == 24 || v >= 26);
}

bool brick_supported(int variant, const ewpt_lattice &L)
{
    switch (variant % 100) {
    case 1: return brick_fits_v1(L);
    case 2: return brick_fits_v2(L);
    case 3: return brick_fits_v3(L);
    case 4: return brick_fits_v4(L);
    case 5: return brick_fits_v5(L);
    case 6: return brick_fits_v6(L);
    case 7: return brick_fits_v7(L);
    case 8: return brick_fits_v8(L);
    case 10: return brick_fits_v10(L);
    case 11: return brick_fits_v11(L);
    case 12: return brick_fits_v12(L);
    case 13: return brick_fits_v13(L);
    case 14: return brick_fits_v14(L);
    case 15: return brick_fits_v15(L);
    case 16: return brick_fits_v16(L);
    case 17: return brick_fits_v17(L);
    case 18: return brick_fits_v18(L);
    case 19: return brick_fits_v19(L);
    case 20: return brick_fits_v20(L);
    case 21: return brick_fits_v21(L);
    case 22: return brick_fits_v22(L);
    case 23: return brick_fits_v23(L);
    case 24: return brick_fits_v24(L);
    case 25: return brick_fits_v25(L);
    case 26: return brick_fits_v26(L);
    case 27: return brick_fits_v27(L);
    case 29: return brick_fits_v29(L);
    case 32: return brick_fits_v32(L);
    default: return false;
    }
}

hipError_t launch_stage_brick(int variant, int stage, const StageBufs<double> &b, const ewpt_lattice &L,
                              const Phys &P, double dt, hipStream_t st)
{
    const int order = variant / 100;   // brick enumeration order (brick_coords)
    variant %= 100;
    switch (variant) {
    case 1: return launch_brick_v1(stage, b, L, P, dt, order, st);
    case 2: return launch_brick_v2(stage, b, L, P, dt, order, st);
    case 3: return launch_brick_v3(stage, b, L, P, dt, order, st);
    case 4: return launch_brick_v4(stage, b, L, P, dt, order, st);
    case 5: return launch_brick_v5(stage, b, L, P, dt, order, st);
    case 6: return launch_brick_v6(stage, b, L, P, dt, order, st);
    case 7: return launch_brick_v7(stage, b, L, P, dt, order, st);
    case 8: return launch_brick_v8(stage, b, L, P, dt, order, st);
    case 10: return launch_brick_v10(stage, b, L, P, dt, order, st);
    case 11: return launch_brick_v11(stage, b, L, P, dt, order, st);
    case 12: return launch_brick_v12(stage, b, L, P, dt, order, st);
    case 13: return launch_brick_v13(stage, b, L, P, dt, order, st);
    case 14: return launch_brick_v14(stage, b, L, P, dt, order, st);
    case 15: return launch_brick_v15(stage, b, L, P, dt, order, st);
    case 16: return launch_brick_v16(stage, b, L, P, dt, order, st);
    case 17: return launch_brick_v17(stage, b, L, P, dt, order, st);
    case 18: return launch_brick_v18(stage, b, L, P, dt, order, st);
    case 19: return launch_brick_v19(stage, b, L, P, dt, order, st);
    case 20: return launch_brick_v20(stage, b, L, P, dt, order, st);
    case 21: return launch_brick_v21(stage, b, L, P, dt, order, st);
    case 22: return launch_brick_v22(stage, b, L, P, dt, order, st);
    case 23: return launch_brick_v23(stage, b, L, P, dt, order, st);
    case 24: return launch_brick_v24(stage, b, L, P, dt, order, st);
    case 25: return launch_brick_v25(stage, b, L, P, dt, order, st);
    case 26: return launch_brick_v26(stage, b, L, P, dt, order, st);
    case 27: return launch_brick_v27(stage, b, L, P, dt, order, st);
    case 29: return launch_brick_v29(stage, b, L, P, dt, order, st);
    case 32: return launch_brick_v32(stage, b, L, P, dt, order, st);
    default: return hipErrorInvalidValue;
    }
}

}  // namespace ewpt
