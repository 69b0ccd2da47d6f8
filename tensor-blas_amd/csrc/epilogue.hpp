// Epilogue arithmetic shared by the GEMM kernels (gemm_kernel.hpp, gemm16_kernel.hpp)
// and the persistent batch kernel (batch.hip).
#pragma once
#include <hip/hip_runtime.h>

namespace tb {

// x := (a + b) / d in FP32 with about one rounding: the sum is kept exactly as hi + lo
// (TwoSum), the quotient of hi is corrected by the exact remainder hi - q d (one FMA)
// plus lo.  The leaf epilogues' diagonal division, X = (B + B W^T) D^-1, in FP32 only.
__device__ __forceinline__ float div_sum_f32(float a, float b, float d)
{
    const float hi = a + b;
    const float bb = hi - a;
    const float lo = (a - (hi - bb)) + (b - bb);
    const float q = hi / d;
    const float r = __builtin_fmaf(-q, d, hi) + lo;
    return __builtin_fmaf(r, 1.0f / d, q);
}

}  // namespace tb
