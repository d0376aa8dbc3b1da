// ewpt_brick.hip -- dispatch over the FP64 LDS-brick stage-kernel variants
// (kernel in ewpt_brick.h, one instantiation per ewpt_brick_v*.hip file so
// they compile in parallel).
//
//   2:  8x8x8 brick, 512 threads, prefetch distance 1 (lattices whose sides
//       are multiples of 8 only)
//   10: 16x8x4 brick, 512 threads, prefetch depth 4/4/3/2 by stage
//   11: 32x4x4 brick, 512 threads, prefetch depth 4/4/3/2 by stage
//   20: 11 with the E and Gamma/Sigma centres loaded with their passes
//       (late_centres)
//   24: 16x4x4 brick of 256 threads (two workgroups per CU), late centres,
//       stored stage inputs for stages 2-4
//   25: 24 without stored stage inputs (every stage input formed on load)
//   32: 24 with the next stage input's y loaded during the last passes of
//       stages 2-3, stage 4's partial-sum inputs loaded one channel per pass
//       and the field channels of y' updated from those same loads; the same
//       arithmetic in the same order as 24 (bitwise equal results)
//   34: 32 with y reconstructed from the stored stage inputs and derivatives
//       (stage 4 reads no y, stage 2 reads k1 instead of y): fewer bytes, one
//       more rounding per reconstructed value; stage-4 prefetch depth 3
// All use the XCD-aware block order; + 100 * o selects brick enumeration
// order o (brick_coords).
#include "ewpt_internal.h"

namespace ewpt {

#define EWPT_DECLARE_BRICK(V)                                                                                     \
    bool brick_fits_v##V(const ewpt_lattice &L);                                                                  \
    hipError_t launch_brick_v##V(int stage, const StageBufs<double> &b, const ewpt_lattice &L, const Phys &P,     \
                                 double dt, int order, hipStream_t st);
EWPT_DECLARE_BRICK(2)
EWPT_DECLARE_BRICK(10)
EWPT_DECLARE_BRICK(11)
EWPT_DECLARE_BRICK(20)
EWPT_DECLARE_BRICK(24)
EWPT_DECLARE_BRICK(25)
EWPT_DECLARE_BRICK(32)
EWPT_DECLARE_BRICK(34)
#undef EWPT_DECLARE_BRICK
hipError_t launch_brick_v32_diag(const StageBufs<double> &b, const ewpt_lattice &L, const Phys &P, double dt,
                                 int order, double *diag, double sw, double cw, hipStream_t st);
hipError_t launch_brick_v32_diag_only(const StageBufs<double> &b, const ewpt_lattice &L, const Phys &P, int order,
                                      double *diag, double sw, double cw, hipStream_t st);

namespace {

struct BrickEntry {
    int v;
    bool (*fits)(const ewpt_lattice &);
    hipError_t (*launch)(int, const StageBufs<double> &, const ewpt_lattice &, const Phys &, double, int,
                         hipStream_t);
};

constexpr BrickEntry kBricks[] = {
    {2, brick_fits_v2, launch_brick_v2},     {10, brick_fits_v10, launch_brick_v10},
    {11, brick_fits_v11, launch_brick_v11},  {20, brick_fits_v20, launch_brick_v20},
    {24, brick_fits_v24, launch_brick_v24},  {25, brick_fits_v25, launch_brick_v25},
    {32, brick_fits_v32, launch_brick_v32},  {34, brick_fits_v34, launch_brick_v34},
};

const BrickEntry *find_brick(int variant)
{
    if (variant < 0 || variant / 100 > 9) return nullptr;
    for (const BrickEntry &e : kBricks)
        if (e.v == variant % 100) return &e;
    return nullptr;
}

}  // namespace

bool brick_exists(int variant) { return find_brick(variant) != nullptr; }

bool brick_stored_inputs(int variant)
{
    const int v = variant % 100;
    return variant >= 0 && (v == 15 || v == 24 || v >= 31);
}

bool brick_supported(int variant, const ewpt_lattice &L)
{
    const BrickEntry *e = find_brick(variant);
    return e && e->fits(L);
}

long diag_rows(const ewpt_lattice &L)
{
    return (long)(L.nx / 16) * (L.ny / 4) * (L.nz / 4);   // the 16x4x4 bricks of variants 32, 34
}

// Variants 32 and 34 share their stage-1 kernel (34 differs in stages 2 and
// 4 only), so both take variant 32's stage 1 with the fused reductions.
bool brick_has_diag(int variant)
{
    return variant >= 0 && (variant % 100 == 32 || variant % 100 == 34) && brick_exists(variant);
}

hipError_t launch_stage1_diag(int variant, const StageBufs<double> &b, const ewpt_lattice &L, const Phys &P, double dt,
                              double *diag, double sw, double cw, hipStream_t st)
{
    if (!brick_has_diag(variant)) return hipErrorInvalidValue;
    return launch_brick_v32_diag(b, L, P, dt, variant / 100, diag, sw, cw, st);
}

bool state_diag_fits(const ewpt_lattice &L) { return brick_supported(32, L); }

hipError_t launch_state_diag(const double *y, long n3, const ewpt_lattice &L, const Phys &P, double *diag, double sw,
                             double cw, hipStream_t st)
{
    if (!state_diag_fits(L)) return hipErrorInvalidValue;
    StageBufs<double> b{};
    b.y = y;
    b.n3 = n3;
    return launch_brick_v32_diag_only(b, L, P, 3, diag, sw, cw, st);   // brick cube order 3, as the stage kernels
}

hipError_t launch_stage_brick(int variant, int stage, const StageBufs<double> &b, const ewpt_lattice &L,
                              const Phys &P, double dt, hipStream_t st)
{
    const BrickEntry *e = find_brick(variant);
    if (!e) return hipErrorInvalidValue;
    return e->launch(stage, b, L, P, dt, variant / 100, st);
}

}  // namespace ewpt
