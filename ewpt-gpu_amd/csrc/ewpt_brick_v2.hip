// ewpt_brick_v2.hip -- stage-kernel variant 2: 8x8x8 brick, prefetch
// distance 1, XCD-aware block order.
#include "ewpt_brick.h"

namespace ewpt {

bool brick_fits_v2(const ewpt_lattice &L) { return brick_fits<8, 8, 8>(L); }

hipError_t launch_brick_v2(int stage, const StageBufs<double> &b, const ewpt_lattice &L, const Phys &P, double dt,
                            int order, hipStream_t st)
{
    return launch_tile<8, 8, 8, 1>(stage, b, L, P, dt, 1, order, st);
}

}  // namespace ewpt
