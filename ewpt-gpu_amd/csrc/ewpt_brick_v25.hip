// ewpt_brick_v25.hip -- stage-kernel variant 25: 16x4x4 brick, 256 threads,
// prefetch depth 4/4/3/2 for stages 1..4, E and Gamma/Sigma centres loaded with
// their passes (two workgroups per CU in FP64), XCD-aware block order.
#include "ewpt_brick.h"

namespace ewpt {

bool brick_fits_v25(const ewpt_lattice &L) { return brick_fits<16, 4, 4>(L); }

hipError_t launch_brick_v25(int stage, const StageBufs<double> &b, const ewpt_lattice &L, const Phys &P, double dt,
                             int order, hipStream_t st)
{
    return launch_tile<16, 4, 4, 1004432>(stage, b, L, P, dt, 1, order, st);
}

}  // namespace ewpt
