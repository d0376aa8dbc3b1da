// ewpt_brick_v32.hip -- stage-kernel variant 32 (FP64 default): variant 24 with
// the next stage input's y loaded during the last passes of stages 2-3
// (y_early), stage 4's partial-sum inputs loaded one channel per pass
// (spread_update) and the field channels of y' updated from those loads
// (fused_fields): each momentum and k value is read once.
#include "ewpt_brick.h"

static_assert(ewpt::NDIAG == ewpt::DIAG_SLOTS, "diagnostic row layout");

namespace ewpt {

bool brick_fits_v32(const ewpt_lattice &L) { return brick_fits<16, 4, 4>(L); }

hipError_t launch_brick_v32(int stage, const StageBufs<double> &b, const ewpt_lattice &L, const Phys &P, double dt,
                             int order, hipStream_t st)
{
    return launch_tile<16, 4, 4, 211704432>(stage, b, L, P, dt, 1, order, st);
}

hipError_t launch_brick_v32_diag(const StageBufs<double> &b, const ewpt_lattice &L, const Phys &P, double dt,
                                 int order, double *diag, double sw, double cw, hipStream_t st)
{
    return launch_tile_diag<16, 4, 4, 211704432>(b, L, P, dt, 1, order, diag, sw, cw, st);
}

hipError_t launch_brick_v32_diag_only(const StageBufs<double> &b, const ewpt_lattice &L, const Phys &P, int order,
                                      double *diag, double sw, double cw, hipStream_t st)
{
    return launch_tile_diag_only<16, 4, 4, 211704432>(b, L, P, 1, order, diag, sw, cw, st);
}

}  // namespace ewpt
