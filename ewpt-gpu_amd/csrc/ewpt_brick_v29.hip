// ewpt_brick_v29.hip -- stage-kernel variant 29: as 27 (16x4x4 brick, 256
// threads, late centres, stored stage inputs, early y) with stage 4 loading
// the inputs of its partial sums one channel per pass (spread_update).
#include "ewpt_brick.h"

namespace ewpt {

bool brick_fits_v29(const ewpt_lattice &L) { return brick_fits<16, 4, 4>(L); }

hipError_t launch_brick_v29(int stage, const StageBufs<double> &b, const ewpt_lattice &L, const Phys &P, double dt,
                             int order, hipStream_t st)
{
    return launch_tile<16, 4, 4, 111704432>(stage, b, L, P, dt, 1, order, st);
}

}  // namespace ewpt
