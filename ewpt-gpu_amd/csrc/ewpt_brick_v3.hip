// ewpt_brick_v3.hip -- stage-kernel variant 3: 16x8x4 brick, prefetch
// distance 2, XCD-aware block order.
#include "ewpt_brick.h"

namespace ewpt {

bool brick_fits_v3(const ewpt_lattice &L) { return brick_fits<16, 8, 4>(L); }

hipError_t launch_brick_v3(int stage, const StageBufs<double> &b, const ewpt_lattice &L, const Phys &P, double dt,
                            int order, hipStream_t st)
{
    return launch_tile<16, 8, 4, 2>(stage, b, L, P, dt, 1, order, st);
}

}  // namespace ewpt
