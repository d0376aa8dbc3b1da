// ewpt_brick_v16.hip -- stage-kernel variant 16: 16x8x4 brick, prefetch depth
// 4/4/3/2 for stages 1..4, stage 4 reads
// its input as one array written by stage 3, XCD-aware block order.
#include "ewpt_brick.h"

namespace ewpt {

bool brick_fits_v16(const ewpt_lattice &L) { return brick_fits<16, 8, 4>(L); }

hipError_t launch_brick_v16(int stage, const StageBufs<double> &b, const ewpt_lattice &L, const Phys &P, double dt,
                             int order, hipStream_t st)
{
    return launch_tile<16, 8, 4, 404432>(stage, b, L, P, dt, 1, order, st);
}

}  // namespace ewpt
