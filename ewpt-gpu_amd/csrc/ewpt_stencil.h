// ewpt_stencil.h -- 6th-order central stencils (diff_scheme.jl:157-311).
//
// The reference evaluates
//   d1 = (a3 - 9 a2 + 45 a1 - 45 m1 + 9 m2 - m3) / (60 dx)
//   d2 = (a3/90 - 3a2/20 + 3a1/2 - 49a0/18 + 3m1/2 - 3m2/20 + m3/90) / dx^2
// with a_s = A[x+s], m_s = A[x-s] and exact periodic wrap.  The forms below
// pair the symmetric taps first (one subtraction or addition per pair) and
// multiply by precomputed reciprocals; they agree with the reference to
// rounding.
#pragma once

#include <hip/hip_runtime.h>

namespace ewpt {

template <typename R>
struct StencilK {
    R inv60dx;   // 1/(60 dx)
    R invdx2;    // 1/dx^2
};

template <typename R>
__device__ __forceinline__ R sten_d1(R m3, R m2, R m1, R a1, R a2, R a3, R inv60dx)
{
    return (R(45) * (a1 - m1) - R(9) * (a2 - m2) + (a3 - m3)) * inv60dx;
}

template <typename R>
__device__ __forceinline__ R sten_d2(R m3, R m2, R m1, R a0, R a1, R a2, R a3, R invdx2)
{
    if constexpr (sizeof(R) == 4) {
        // FP32: taps relative to the centre (the weights sum to zero), so a
        // smooth field loses no digits to the 49/18 a0 cancellation
        return (R(1.5) * ((a1 - a0) + (m1 - a0)) - R(0.15) * ((a2 - a0) + (m2 - a0)) +
                R(1.0 / 90.0) * ((a3 - a0) + (m3 - a0))) * invdx2;
    } else {
        return (R(1.5) * (a1 + m1) - R(0.15) * (a2 + m2) + R(1.0 / 90.0) * (a3 + m3)
                - R(49.0 / 18.0) * a0) * invdx2;
    }
}

__device__ __forceinline__ int wrap_idx(int i, int n)
{
    return i < 0 ? i + n : (i >= n ? i - n : i);
}

}  // namespace ewpt
