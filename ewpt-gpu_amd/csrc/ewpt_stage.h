// ewpt_stage.h -- RK4 stage bookkeeping shared by the stage kernels.
//
// The reference (gpu-kernels-rk4/main_rk4.jl:6146-6441) keeps 36 k arrays per
// stage and ~250 broadcast temporaries per step.  Here only the 20 non-trivial
// stage derivatives k_s (state channels 16..35: k_Gamma, k_Sigma, kt) are
// stored; the 16 field derivatives are the stage momenta and are rebuilt
// exactly as the reference forms them:
//   k1_f = pi,  k2_f = pi + kt1*(dt/2),  k3_f = pi + kt2*(dt/2),  k4_f = pi + kt3*dt.
// Stage inputs are y + k1*(dt/2), y + k2*(dt/2), y + k3*dt (6226, 6298, 6370)
// and the update is n = k1*(dt/6) + y; n += k2*(dt/3); n += k3*(dt/3);
// y' = k4*(dt/6) + n (6183-6437).  The stage kernels either form each stage
// input from y and the stored k's when they load it, or (stored stage inputs)
// receive it as one array u written by the previous stage.
#pragma once

#include "ewpt_point.h"

namespace ewpt {

template <typename T>
struct StageBufs {
    const T *y;     // 36 * n3
    const T *k1;    // 20 * n3  (stage-derivative channel m <-> state channel 16+m)
    const T *k2;
    const T *k3;
    T *kout;        // 20 * n3, stages 1-3
    T *ynew;        // 36 * n3, stage 4
    const T *ylo;   // mixed precision: compensation word of y (36 * n3) or null
    T *ylo_new;
    long n3;
    // kernels that pass the stage input on (stored stage inputs): u is this
    // stage's input (36 * n3; stage 1 reads y), unext receives the next one
    const T *u;
    T *unext;
    // z-slab of a lattice split over several contexts: every array holds
    // zghost planes below and above the nz interior planes (the neighbours'
    // halo, filled by the exchange) and is not periodic in z; 0 = periodic
    int zghost = 0;
};

// Value of channel c of the stage-S input at flat site q.
template <int S, typename T, typename R>
__device__ __forceinline__ R stage_in(const StageBufs<T> &b, int c, long q, R dt)
{
    const R yv = R(b.y[(long)c * b.n3 + q]);
    if constexpr (S == 1) {
        return yv;
    } else {
        if (c < 16) {
            const R piv = R(b.y[(long)(c + 20) * b.n3 + q]);
            if constexpr (S == 2) return yv + piv * (dt * R(0.5));
            if constexpr (S == 3) return yv + (piv + R(b.k1[(long)(c + 4) * b.n3 + q]) * (dt * R(0.5))) * (dt * R(0.5));
            if constexpr (S == 4) return yv + (piv + R(b.k2[(long)(c + 4) * b.n3 + q]) * (dt * R(0.5))) * dt;
        }
        const T *kp = (S == 2) ? b.k1 : (S == 3) ? b.k2 : b.k3;
        const R f = (S == 4) ? dt : dt * R(0.5);
        return yv + R(kp[(long)(c - 16) * b.n3 + q]) * f;
    }
}

}  // namespace ewpt
