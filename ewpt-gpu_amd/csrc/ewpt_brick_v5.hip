// ewpt_brick_v5.hip -- stage-kernel variant 5: 16x8x4 brick, prefetch
// distance 1, linear block order.
#include "ewpt_brick.h"

namespace ewpt {

bool brick_fits_v5(const ewpt_lattice &L) { return brick_fits<16, 8, 4>(L); }

hipError_t launch_brick_v5(int stage, const StageBufs<double> &b, const ewpt_lattice &L, const Phys &P, double dt,
                            int order, hipStream_t st)
{
    return launch_tile<16, 8, 4, 1>(stage, b, L, P, dt, 0, order, st);
}

}  // namespace ewpt
