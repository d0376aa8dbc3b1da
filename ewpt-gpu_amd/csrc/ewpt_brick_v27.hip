// ewpt_brick_v27.hip -- stage-kernel variant 27: as 24 (16x4x4 brick, 256
// threads, late centres, stored stage inputs for stages 2-4) with stages 2-3
// loading the y values of the next stage input during their last passes.
#include "ewpt_brick.h"

namespace ewpt {

bool brick_fits_v27(const ewpt_lattice &L) { return brick_fits<16, 4, 4>(L); }

hipError_t launch_brick_v27(int stage, const StageBufs<double> &b, const ewpt_lattice &L, const Phys &P, double dt,
                             int order, hipStream_t st)
{
    return launch_tile<16, 4, 4, 11704432>(stage, b, L, P, dt, 1, order, st);
}

}  // namespace ewpt
