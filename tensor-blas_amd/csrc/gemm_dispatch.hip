// Instantiations of the grouped GEMM kernels and the host-side dispatcher.
#include <atomic>

#include "gemm16_kernel.hpp"
#include "gemm64_kernel.hpp"
#include "gemm_kernel.hpp"
#include "tb_kernels.hpp"

namespace tb {

namespace {

template <int TIER, int BM, int BN, int BK, int WM, int WN>
hipError_t run(const GemmLaunch& L, hipStream_t s, int max_cus)
{
    auto k = gemm_grouped_kernel<TIER, BM, BN, BK, WM, WN>;
    hipLaunchKernelGGL(k, dim3(capped_grid(reinterpret_cast<const void*>(k), kThreads, L.totalTiles, max_cus)),
                       dim3(kThreads), 0, s, L);
    return hipGetLastError();
}

template <int NPL, int SA, int SB, int BM, int BN, int WM, int WN, int NST>
hipError_t run16(const GemmLaunch& L, hipStream_t s, int max_cus)
{
    auto k = gemm16_kernel<NPL, SA, SB, BM, BN, WM, WN, NST>;
    hipLaunchKernelGGL(k, dim3(capped_grid(reinterpret_cast<const void*>(k), 64 * WM * WN, L.totalTiles, max_cus)),
                       dim3(64 * WM * WN), 0, s, L);
    return hipGetLastError();
}

// compute units of the current device (looked up once per device)
int device_cus()
{
    static std::atomic<int> cache[64];
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return 256;
    int c = cache[dev].load(std::memory_order_relaxed);
    if (c <= 0) {
        if (hipDeviceGetAttribute(&c, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || c <= 0)
            c = 256;
        cache[dev].store(c, std::memory_order_relaxed);
    }
    return c;
}

template <int NPL, int SA, int SB>
hipError_t run16_shape(int shape, const GemmLaunch& L, hipStream_t s, int mc)
{
    // a stage is (BM + BN) x 64 B per plane, and the ring holds as many stages (<= 4)
    // as fit in the 160 KB of LDS.  Eight waves (two per SIMD) where the registers
    // allow.  Square and leaf tiles: 64 x 32 accumulators per wave.  Wide tiles are
    // 256 x 256: eight waves of 128 x 64 in the plain tier (25 % fewer LDS bytes per
    // MFMA than 256 x 128, +14 % at 8192^3), four of 128 x 128 in the split tier, whose
    // MFMAs are issued one at a time with the fragment reads and the next stage's loads
    // between them (+4-5 % at 8192^3 and 16384 x 8192 x 8192 over eight waves of
    // 128 x 64, the same bits; the plain tier measured 6-35 % slower that way).
    if (shape == kShapeLeaf) {
        if constexpr (NPL == 2) return run16<NPL, SA, SB, 64, 256, 1, 8, 3>(L, s, mc);
        return run16<NPL, SA, SB, 64, 256, 1, 4, NPL == 3 ? 2 : 3>(L, s, mc);
    }
    if constexpr (NPL <= 2 && SA == kSrcPlanes && SB == kSrcPlanes) {
        if (shape == kShapeWide) {
            if constexpr (NPL == 1) return run16<NPL, SA, SB, 256, 256, 2, 4, 4>(L, s, 0);
            else return run16<NPL, SA, SB, 256, 256, 2, 2, 2>(L, s, 0);
        }
    }
    if (shape == kShapeWide) return hipErrorInvalidValue;
    // Split-tier square tiles converting an FP32 op(A) in the kernel (the narrow updates
    // of the TRSM recursion) stall on the stage barrier with eight waves per CU; with at
    // least two tiles per CU, two workgroups of two stages each share a CU instead of one
    // of four stages (-15 to -20 % at 16384 x 512 .. 1024, the same bits).
    if constexpr (NPL == 2 && SA == kSrcF32 && SB == kSrcPlanes)
        if (L.totalTiles >= 2 * device_cus()) return run16<NPL, SA, SB, 128, 128, 2, 4, 2>(L, s, mc);
    return run16<NPL, SA, SB, 128, 128, 2, 4, NPL == 3 ? 3 : 4>(L, s, mc);
}

template <int NPL>
hipError_t run16_src(int shape, int srcA, int srcB, const GemmLaunch& L, hipStream_t s, int mc)
{
    if (srcA == kSrcF32 && srcB == kSrcF32) return run16_shape<NPL, kSrcF32, kSrcF32>(shape, L, s, mc);
    if (srcA == kSrcF32 && srcB == kSrcPlanes) return run16_shape<NPL, kSrcF32, kSrcPlanes>(shape, L, s, mc);
    if (srcA == kSrcPlanes && srcB == kSrcPlanes)
        return run16_shape<NPL, kSrcPlanes, kSrcPlanes>(shape, L, s, mc);
    return hipErrorInvalidValue;
}

// The direct-to-LDS FP64 kernel takes launches whose problems all read op(A) = A and
// op(B) = B^T (panel rows contiguous) from 16-byte aligned operands with even leading
// dimensions and batch strides (every 16-byte load then covers two whole elements).
bool gemm64_ok(const GemmLaunch& L)
{
    auto al = [](const void* p) { return (reinterpret_cast<uintptr_t>(p) & 15) == 0; };
    for (int q = 0; q < L.np; ++q) {
        const GemmProb& p = L.p[q];
        if (p.modeA != kOpN || p.modeB != kOpT || p.eOut || !al(p.A) || !al(p.B) ||
            (p.A2 && (!al(p.A2) || !al(p.B2))) || (p.lda | p.ldb | p.sA | p.sB) & 1)
            return false;
    }
    return true;
}

}  // namespace

int capped_grid(const void* kernel, int threads, int total, int max_cus)
{
    if (max_cus <= 0 || total <= 8) return total;
    int per = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, kernel, threads, 0) != hipSuccess || per < 1)
        per = 1;
    long long cap = (long long)max_cus * per / 8 * 8;
    if (cap < 8) cap = 8;
    return cap < total ? (int)cap : total;
}

TileShape gemm_tile_shape(int tier, int shape)
{
    // FP64 leaves: 32-row tiles, so that a solve of 8192 rows fills 256 workgroups
    // rather than 128
    if (shape == kShapeLeaf) return TileShape{tier == kMfmaF64 ? 32 : 64, 256};
    if (shape == kShapeWide) return TileShape{256, 256};
    return TileShape{128, 128};
}

int tier_planes(int tier, int split_planes)
{
    return tier == kSplitF16 ? split_planes : (tier == kPlainF16 ? 1 : 0);
}

hipError_t gemm_launch(int tier, int npl, int shape, int srcA, int srcB, const GemmLaunch& L,
                       hipStream_t s, int max_cus)
{
    if (L.totalTiles <= 0) return hipSuccess;
    switch (tier) {
    case kPlainF16: return run16_src<1>(shape, srcA, srcB, L, s, max_cus);
    case kSplitF16:
        if (npl == 2) return run16_src<2>(shape, srcA, srcB, L, s, max_cus);
        if (npl == 3) return run16_src<3>(shape, srcA, srcB, L, s, max_cus);
        return hipErrorInvalidValue;
    case kFmaF32:
        if (srcA != kSrcF32 || srcB != kSrcF32) return hipErrorInvalidValue;
        return shape == kShapeLeaf ? run<kFmaF32, 64, 256, 32, 1, 4>(L, s, max_cus)
                                   : run<kFmaF32, 128, 128, 32, 2, 2>(L, s, max_cus);
    case kMfmaF64:
        if (srcA != kSrcF32 || srcB != kSrcF32) return hipErrorInvalidValue;
        if (shape == kShapeSquare && gemm64_ok(L)) {
            hipLaunchKernelGGL(gemm64_kernel,
                               dim3(capped_grid(reinterpret_cast<const void*>(gemm64_kernel), g64::NT,
                                                L.totalTiles, max_cus)),
                               dim3(g64::NT), 0, s, L);
            return hipGetLastError();
        }
        return shape == kShapeLeaf ? run<kMfmaF64, 32, 256, 16, 1, 4>(L, s, max_cus)
                                   : run<kMfmaF64, 128, 128, 16, 2, 2>(L, s, max_cus);
    default:
        return hipErrorInvalidValue;
    }
}

}  // namespace tb
