// ewpt_brick_v32.hip -- stage-kernel variant 32 (FP64): as 29 with the field
// channels of y' updated in the spread passes from the loads of the partial
// sums (fused_fields), so stage 4 reads each momentum and k value once.
#include "ewpt_brick.h"

namespace ewpt {

bool brick_fits_v32(const ewpt_lattice &L) { return brick_fits<16, 4, 4>(L); }

hipError_t launch_brick_v32(int stage, const StageBufs<double> &b, const ewpt_lattice &L, const Phys &P, double dt,
                             int order, hipStream_t st)
{
    return launch_tile<16, 4, 4, 211704432>(stage, b, L, P, dt, 1, order, st);
}

}  // namespace ewpt
