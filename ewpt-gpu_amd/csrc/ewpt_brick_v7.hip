// ewpt_brick_v7.hip -- stage-kernel variant 7: 8x8x4 brick, prefetch
// distance 2, XCD-aware block order.
#include "ewpt_brick.h"

namespace ewpt {

bool brick_fits_v7(const ewpt_lattice &L) { return brick_fits<8, 8, 4>(L); }

hipError_t launch_brick_v7(int stage, const StageBufs<double> &b, const ewpt_lattice &L, const Phys &P, double dt,
                            int order, hipStream_t st)
{
    return launch_tile<8, 8, 4, 2>(stage, b, L, P, dt, 1, order, st);
}

}  // namespace ewpt
