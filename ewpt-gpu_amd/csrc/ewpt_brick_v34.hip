// ewpt_brick_v34.hip -- stage-kernel variant 34 (FP64): variant 32 with y
// reconstructed from the stored stage inputs and the stored derivatives
// (recon_y): stage 4 reads u4 and k1..k3 but not y, stage 2 forms u3 from u2
// and k1 instead of y, which cuts 288 + 128 bytes per site of reads per step.
// Stage 4 stages three passes ahead (prefetch depth 4/4/3/3): its update
// inputs, loaded a pass ahead, no longer make it wait for the halo early.
#include "ewpt_brick.h"

namespace ewpt {

bool brick_fits_v34(const ewpt_lattice &L) { return brick_fits<16, 4, 4>(L); }

hipError_t launch_brick_v34(int stage, const StageBufs<double> &b, const ewpt_lattice &L, const Phys &P, double dt,
                             int order, hipStream_t st)
{
    return launch_tile<16, 4, 4, 311704433>(stage, b, L, P, dt, 1, order, st);
}

}  // namespace ewpt
