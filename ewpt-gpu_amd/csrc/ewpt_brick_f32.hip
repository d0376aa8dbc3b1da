// ewpt_brick_f32.hip -- mixed-precision stage kernels: FP32 storage and
// arithmetic with the compensated stage-4 update (ewpt_brick.h, T = float).
//   10 (+100*order): 16x8x4 brick, prefetch depth 4/4/3/2
//   11 (+100*order): 32x4x4 brick, prefetch depth 4/4/3/2
//   12, 13 (+100*order): as 10, 11 with two channels per LDS barrier in stages 1-3
//   14, 15 (+100*order): as 10, 11 with stored stage inputs for stages 2-4
//   16, 17 (+100*order): as 10, 11 with a stored stage input for stage 4
//   18, 19 (+100*order): as 10, 11 with stored stage inputs for stages 3-4
//   20, 22 (+100*order): as 11, 10 with late E and Gamma/Sigma centres
//   21, 23 (+100*order): 20, 22 with stored stage inputs for stages 2-4
//   24, 25 (+100*order): 16x4x4 bricks of 256 threads as 21, 20
//   26 (+100*order): 8x8x4 bricks of 256 threads as 21
//   27 (+100*order): 16x4x4 as 24 with the next input's y loaded early
//   28 (+100*order): 32x4x4 as 15 with the next input's y loaded early
//   30 (+100*order): 28 with the lean stage-4 stash (two workgroups per CU)
//   31 (+100*order): 30 with prefetch depth 4/4/4/3 (stages 3-4 read one input array)
//   33 (+100*order): 31 with stage 4's increments formed one channel per pass
//                    and the field channels updated from the same k loads
#include "ewpt_brick.h"

namespace ewpt {

bool brick_supported_f32(int variant, const ewpt_lattice &L)
{
    switch (variant % 100) {
    case 10: return brick_fits<16, 8, 4>(L);
    case 11: return brick_fits<32, 4, 4>(L);
    case 12: return brick_fits<16, 8, 4>(L);
    case 13: return brick_fits<32, 4, 4>(L);
    case 14: return brick_fits<16, 8, 4>(L);
    case 15: return brick_fits<32, 4, 4>(L);
    case 16: return brick_fits<16, 8, 4>(L);
    case 17: return brick_fits<32, 4, 4>(L);
    case 18: return brick_fits<16, 8, 4>(L);
    case 19: return brick_fits<32, 4, 4>(L);
    case 20: return brick_fits<32, 4, 4>(L);
    case 21: return brick_fits<32, 4, 4>(L);
    case 22: return brick_fits<16, 8, 4>(L);
    case 23: return brick_fits<16, 8, 4>(L);
    case 24: return brick_fits<16, 4, 4>(L);
    case 25: return brick_fits<16, 4, 4>(L);
    case 26: return brick_fits<8, 8, 4>(L);
    case 27: return brick_fits<16, 4, 4>(L);
    case 28: return brick_fits<32, 4, 4>(L);
    case 30: return brick_fits<32, 4, 4>(L);
    case 31: return brick_fits<32, 4, 4>(L);
    case 33: return brick_fits<32, 4, 4>(L);
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
    case 12: return launch_tile<16, 8, 4, 14432, float>(stage, b, L, P, dt, 1, order, st);
    case 13: return launch_tile<32, 4, 4, 14432, float>(stage, b, L, P, dt, 1, order, st);
    case 14: return launch_tile<16, 8, 4, 704432, float>(stage, b, L, P, dt, 1, order, st);
    case 15: return launch_tile<32, 4, 4, 704432, float>(stage, b, L, P, dt, 1, order, st);
    case 16: return launch_tile<16, 8, 4, 404432, float>(stage, b, L, P, dt, 1, order, st);
    case 17: return launch_tile<32, 4, 4, 404432, float>(stage, b, L, P, dt, 1, order, st);
    case 18: return launch_tile<16, 8, 4, 604432, float>(stage, b, L, P, dt, 1, order, st);
    case 19: return launch_tile<32, 4, 4, 604432, float>(stage, b, L, P, dt, 1, order, st);
    case 20: return launch_tile<32, 4, 4, 1004432, float>(stage, b, L, P, dt, 1, order, st);
    case 21: return launch_tile<32, 4, 4, 1704432, float>(stage, b, L, P, dt, 1, order, st);
    case 22: return launch_tile<16, 8, 4, 1004432, float>(stage, b, L, P, dt, 1, order, st);
    case 23: return launch_tile<16, 8, 4, 1704432, float>(stage, b, L, P, dt, 1, order, st);
    case 24: return launch_tile<16, 4, 4, 1704432, float>(stage, b, L, P, dt, 1, order, st);
    case 25: return launch_tile<16, 4, 4, 1004432, float>(stage, b, L, P, dt, 1, order, st);
    case 26: return launch_tile<8, 8, 4, 1704432, float>(stage, b, L, P, dt, 1, order, st);
    case 27: return launch_tile<16, 4, 4, 11704432, float>(stage, b, L, P, dt, 1, order, st);
    case 28: return launch_tile<32, 4, 4, 10704432, float>(stage, b, L, P, dt, 1, order, st);
    case 30: return launch_tile<32, 4, 4, 1010704432, float>(stage, b, L, P, dt, 1, order, st);
    case 31: return launch_tile<32, 4, 4, 1010704443, float>(stage, b, L, P, dt, 1, order, st);
    case 33: return launch_tile<32, 4, 4, 1210704443, float>(stage, b, L, P, dt, 1, order, st);
    default: return hipErrorInvalidValue;
    }
}

}  // namespace ewpt
