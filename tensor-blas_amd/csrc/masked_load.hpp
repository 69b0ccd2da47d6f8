// Loads that never wait behind a branch: shared by the kernels whose loads are
// predicated per element (conversions, row exponents, epilogues reading C).
#pragma once
#include <hip/hip_runtime.h>

namespace tb {

// x = *p if keep, else 0 (in p's own type), with the load itself unconditional (p
// must be a valid address either way).  The mask goes through an empty (non-volatile)
// asm so that the compiler cannot turn the masking back into a branch around the
// load: a load behind a branch is waited for where the branch rejoins, which
// serializes a row of loads.
template <typename T>
__device__ __forceinline__ T masked_load(const T* p, bool keep)
{
    static_assert(sizeof(T) == 4 || sizeof(T) == 8, "32- or 64-bit element");
    unsigned m = keep ? 0xffffffffu : 0u;
    asm("" : "+v"(m));
    if constexpr (sizeof(T) == 4) {
        return __uint_as_float(*reinterpret_cast<const unsigned*>(p) & m);
    } else {
        const unsigned* w = reinterpret_cast<const unsigned*>(p);
        const unsigned lo = w[0] & m, hi = w[1] & m;
        return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
    }
}



}  // namespace tb
