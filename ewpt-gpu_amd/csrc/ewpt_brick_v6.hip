// ewpt_brick_v6.hip -- stage-kernel variant 6: 16x4x4 brick, prefetch
// distance 1, XCD-aware block order.
#include "ewpt_brick.h"

namespace ewpt {

bool brick_fits_v6(const ewpt_lattice &L) { return brick_fits<16, 4, 4>(L); }

hipError_t launch_brick_v6(int stage, const StageBufs<double> &b, const ewpt_lattice &L, const Phys &P, double dt,
                            int order, hipStream_t st)
{
    return launch_tile<16, 4, 4, 1>(stage, b, L, P, dt, 1, order, st);
}

}  // namespace ewpt
