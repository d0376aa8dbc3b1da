// ewpt_brick_f32.hip -- mixed-precision stage kernels: FP32 storage and
// arithmetic with the compensated stage-4 update (ewpt_brick.h, T = float).
// Variant numbers follow the FP64 ones where the form is the same:
//   10 (+100*order): 16x8x4 brick, prefetch depth 4/4/3/2, inputs formed on load
//   11 (+100*order): 32x4x4 brick, prefetch depth 4/4/3/2, inputs formed on load
//   15 (+100*order): 11 with stored stage inputs for stages 2-4
//   24 (+100*order): 16x4x4 brick of 256 threads, late E and Gamma/Sigma
//                    centres, stored stage inputs for stages 2-4
//   25 (+100*order): 24 with every stage input formed on load
//   31 (+100*order): 15 with the next input's y loaded during the last
//                    passes of stages 2-3, stage 4 keeping only the increment
//                    in LDS (two workgroups per CU) and prefetch depth 4/4/4/3;
//                    the same arithmetic in the same order as 15
//   33 (+100*order): 31 with stage 4's increments formed one channel per pass
//                    and the field channels updated from the same k loads, and
//                    stages 2-3 storing the field channels of the next input
//                    during their last passes (no register spills at four
//                    waves per SIMD); bitwise equal to 31
#include "ewpt_brick.h"

namespace ewpt {

bool brick_exists_f32(int variant)
{
    if (variant < 0 || variant / 100 > 9) return false;
    switch (variant % 100) {
    case 10: case 11: case 15: case 24: case 25: case 31: case 33: return true;
    default: return false;
    }
}

bool brick_supported_f32(int variant, const ewpt_lattice &L)
{
    if (!brick_exists_f32(variant)) return false;
    switch (variant % 100) {
    case 10: return brick_fits<16, 8, 4>(L);
    case 11: case 15: case 31: case 33: return brick_fits<32, 4, 4>(L);
    case 24: case 25: return brick_fits<16, 4, 4>(L);
    default: return false;
    }
}

hipError_t launch_stage_brick_f32(int variant, int stage, const StageBufs<float> &b, const ewpt_lattice &L,
                                  const Phys &P, double dt, hipStream_t st)
{
    const int order = variant / 100;
    switch (variant % 100) {
    case 10: return launch_tile<16, 8, 4, 4432, float>(stage, b, L, P, dt, 1, order, st);
    case 11: return launch_tile<32, 4, 4, 4432, float>(stage, b, L, P, dt, 1, order, st);
    case 15: return launch_tile<32, 4, 4, 704432, float>(stage, b, L, P, dt, 1, order, st);
    case 24: return launch_tile<16, 4, 4, 1704432, float>(stage, b, L, P, dt, 1, order, st);
    case 25: return launch_tile<16, 4, 4, 1004432, float>(stage, b, L, P, dt, 1, order, st);
    case 31: return launch_tile<32, 4, 4, 1010704443, float>(stage, b, L, P, dt, 1, order, st);
    case 33: return launch_tile<32, 4, 4, 1220704443, float>(stage, b, L, P, dt, 1, order, st);
    default: return hipErrorInvalidValue;
    }
}

// Stage 1 of variants 31 and 33 (the same kernel up to where the non-field
// centres wait: in registers or parked in LDS; bitwise equal outputs) with
// the energy table and Gauss norms of its input reduced on the way, one row
// per 32x4x4 brick.
bool brick_has_diag_f32(int variant)
{
    return brick_exists_f32(variant) && (variant % 100 == 31 || variant % 100 == 33);
}

long diag_rows_f32(const ewpt_lattice &L) { return (long)(L.nx / 32) * (L.ny / 4) * (L.nz / 4); }

hipError_t launch_stage1_diag_f32(int variant, const StageBufs<float> &b, const ewpt_lattice &L, const Phys &P,
                                  double dt, double *diag, double sw, double cw, hipStream_t st)
{
    if (!brick_has_diag_f32(variant) || !brick_supported_f32(variant, L)) return hipErrorInvalidValue;
    return launch_tile_diag<32, 4, 4, 1010704443, float>(b, L, P, dt, 1, variant / 100, diag, sw, cw, st);
}

}  // namespace ewpt
