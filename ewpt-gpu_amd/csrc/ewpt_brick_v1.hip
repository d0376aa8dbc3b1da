// ewpt_brick_v1.hip -- stage-kernel variant 1: 16x8x4 brick, prefetch
// distance 1, XCD-aware block order.
#include "ewpt_brick.h"

namespace ewpt {

bool brick_fits_v1(const ewpt_lattice &L) { return brick_fits<16, 8, 4>(L); }

hipError_t launch_brick_v1(int stage, const StageBufs<double> &b, const ewpt_lattice &L, const Phys &P, double dt,
                            int order, hipStream_t st)
{
    return launch_tile<16, 8, 4, 1>(stage, b, L, P, dt, 1, order, st);
}

}  // namespace ewpt
