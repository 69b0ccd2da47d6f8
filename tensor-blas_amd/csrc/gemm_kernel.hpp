// Grouped, tier-templated GEMM kernel for gfx950 (CDNA4).
//
// One launch walks the output tiles of up to kMaxProbs strided-batch GEMM problems
// (tb_common.hpp).  Each 256-thread workgroup owns one BM x BN tile of C and loops
// over K in BK steps with two LDS buffers: global loads of step s+1 are in flight
// while the MFMAs of step s run, then the registers are converted and written to
// the other buffer.  One barrier per step.
//
// Operand staging ("fused convert/split prologue"): FP32 operands are read from HBM
// once per tile and converted in registers on their way into LDS:
//   kPlainF16 : fp16 = RNE(x)                        (the reference's s2h_swpipe,
//                                                     include/TensorBLAS.h:185-284)
//   kSplitF16 : xs = x * 2^(15-e_row), hi = RNE(xs), lo = RNE((xs - hi) * 2^11)
//               three products hi*hi + 2^-11 (hi*lo + lo*hi), two FP32 accumulators
//   kFmaF32   : FP32 kept, vector FMA (no matrix cores)
//   kMfmaF64  : FP64 kept, v_mfma_f64_16x16x4
// LDS holds each operand panel as [row][k] (k contiguous, padded) so every MFMA
// fragment is one 16-byte ds_read.  The MFMA's A input is fed from the op(B) panel
// and its B input from the op(A) panel, so the accumulator's lane index runs along
// C's rows: the column-major stores of C are contiguous across lanes.
#pragma once
#include <hip/hip_runtime.h>

#include "epilogue.hpp"
#include "masked_load.hpp"
#include "tb_common.hpp"

namespace tb {

typedef _Float16 half8_t __attribute__((ext_vector_type(8)));
typedef float floatx16_t __attribute__((ext_vector_type(16)));
typedef double doublex4_t __attribute__((ext_vector_type(4)));
typedef float float4_t __attribute__((ext_vector_type(4)));
typedef double double2_t __attribute__((ext_vector_type(2)));

constexpr int kThreads = 256;
constexpr int kExpZeroDev = -100000;   // exponent of an all-zero row (tb_kernels.hpp kExpZero)

template <int TIER>
struct TierTraits;

template <>
struct TierTraits<kPlainF16> {
    using Src = float;
    using Lds = _Float16;
    using Out = float;
    static constexpr int planes = 1;
    static constexpr int pad = 8;
};
template <>
struct TierTraits<kSplitF16> {
    using Src = float;
    using Lds = _Float16;
    using Out = float;
    static constexpr int planes = 2;
    static constexpr int pad = 8;
};
template <>
struct TierTraits<kFmaF32> {
    using Src = float;
    using Lds = float;
    using Out = float;
    static constexpr int planes = 1;
    static constexpr int pad = 4;
};
template <>
struct TierTraits<kMfmaF64> {
    using Src = double;
    using Lds = double;
    using Out = double;
    static constexpr int planes = 1;
    static constexpr int pad = 2;
};

// blockIdx -> tile index so that each XCD (blocks are dealt round-robin to the 8
// XCDs) receives a contiguous run of tiles: neighbouring tiles share operand panels
// in that XCD's L2.
__device__ __forceinline__ int xcd_remap(int b, int total)
{
    constexpr int X = 8;
    int per = total / X, rem = total % X;
    int x = b % X, j = b / X;
    return x * per + (x < rem ? x : rem) + j;
}

// One operand panel (R rows of op(A), or R columns of op(B)) x BK.
//   mn_major: consecutive panel rows are consecutive in memory (A: kOpN, B: kOpT)
//   otherwise consecutive k are consecutive in memory.
// Each thread stages KPT consecutive k of one panel row.
template <int TIER, int R, int BK>
struct Panel {
    using TT = TierTraits<TIER>;
    using Src = typename TT::Src;
    using Lds = typename TT::Lds;
    static constexpr int TPR = kThreads / R;    // threads per panel row
    static constexpr int KPT = BK / TPR;        // k per thread
    static constexpr int LDK = BK + TT::pad;    // LDS row stride (elements)
    static constexpr int PLANE = R * LDK;       // elements of one plane
    static constexpr int VEC = 16 / sizeof(Src);
    static_assert(kThreads % R == 0 && BK % TPR == 0, "panel split");
    static_assert((KPT * sizeof(Lds)) % 16 == 0, "16-byte LDS writes");
    static_assert(KPT % VEC == 0, "vector loads");

    Src v[KPT];
    int row, kseg;
    int sc;  // split tier: power-of-two scale exponent of this panel row

    __device__ __forceinline__ void setup(int tid, bool mn_major)
    {
        if (mn_major) {
            row = tid % R;
            kseg = tid / R;
        } else {
            row = tid / TPR;
            kseg = tid % TPR;
        }
    }

    // Load k-block [k0, k0+BK) of panel rows [r0, r0+R) of op(X).
    //   rows / K: bounds; mode: kOpN/kOpT/kOpSymL/kOpTriL as seen from this panel;
    //   isA: this is the op(A) panel (panel row = row of op(A)); else op(B) column.
    __device__ __forceinline__ void load(const Src* __restrict__ X, long long ld, int mode,
                                         bool isA, int r0, int k0, int rows, int K, bool vec_ok)
    {
        const int r = r0 + row;
        const int kb = k0 + kseg * KPT;
        // A symmetric / triangular operand: blocks strictly below the diagonal are
        // plain column-major reads, blocks strictly above are the mirrored (transposed)
        // reads (SYMM) or zero (TRMM); only blocks crossing the diagonal go elementwise.
        int m = mode;
        bool zero = false, special = false;
        if (isA && (mode == kOpSymL || mode == kOpTriL)) {
            if (k0 + BK <= r0) m = kOpN;
            else if (k0 >= r0 + R) { m = kOpT; zero = (mode == kOpTriL); }
            else special = true;
        }
        if (zero) {
#pragma unroll
            for (int j = 0; j < KPT; ++j) v[j] = Src(0);
            return;
        }
        const bool mn = isA ? (m != kOpT) : (m == kOpT);
        const bool full = (r0 + R <= rows) && (k0 + BK <= K);
        if (!special && full) {
            if (mn) {
                const Src* p = X + (long long)r + (long long)kb * ld;
#pragma unroll
                for (int j = 0; j < KPT; ++j) v[j] = p[(long long)j * ld];
            } else if (vec_ok) {
                const Src* p = X + (long long)r * ld + kb;
#pragma unroll
                for (int j = 0; j < KPT; j += VEC) {
                    if constexpr (VEC == 4) {
                        float4_t t = *reinterpret_cast<const float4_t*>(p + j);
                        v[j] = t[0]; v[j + 1] = t[1]; v[j + 2] = t[2]; v[j + 3] = t[3];
                    } else {
                        double2_t t = *reinterpret_cast<const double2_t*>(p + j);
                        v[j] = t[0]; v[j + 1] = t[1];
                    }
                }
            } else {
                const Src* p = X + (long long)r * ld + kb;
#pragma unroll
                for (int j = 0; j < KPT; ++j) v[j] = p[j];
            }
            return;
        }
        // generic path: edge tiles and blocks crossing the diagonal
#pragma unroll
        for (int j = 0; j < KPT; ++j) {
            const int k = kb + j;
            Src x = Src(0);
            if (r < rows && k < K) {
                if (special) {
                    if (r >= k) x = X[(long long)r + (long long)k * ld];
                    else if (mode == kOpSymL) x = X[(long long)k + (long long)r * ld];
                } else if (mn) {
                    x = X[(long long)r + (long long)k * ld];
                } else {
                    x = X[(long long)r * ld + k];
                }
            }
            v[j] = x;
        }
    }

    // Convert the staged registers and write them to LDS plane(s) at base.
    __device__ __forceinline__ void store(Lds* base) const
    {
        Lds* dst = base + row * LDK + kseg * KPT;
        if constexpr (TIER == kPlainF16) {
#pragma unroll
            for (int j = 0; j < KPT; j += 8) {
                half8_t h;
#pragma unroll
                for (int q = 0; q < 8; ++q) h[q] = (_Float16)v[j + q];
                *reinterpret_cast<half8_t*>(dst + j) = h;
            }
        } else if constexpr (TIER == kSplitF16) {
#pragma unroll
            for (int j = 0; j < KPT; j += 8) {
                half8_t hi, lo;
#pragma unroll
                for (int q = 0; q < 8; ++q) {
                    float xs = __builtin_ldexpf(v[j + q], sc);
                    _Float16 h = (_Float16)xs;
                    hi[q] = h;
                    lo[q] = (_Float16)((xs - (float)h) * 2048.0f);
                }
                *reinterpret_cast<half8_t*>(dst + j) = hi;
                *reinterpret_cast<half8_t*>(dst + PLANE + j) = lo;
            }
        } else if constexpr (TIER == kFmaF32) {
#pragma unroll
            for (int j = 0; j < KPT; j += 4) {
                float4_t t = {v[j], v[j + 1], v[j + 2], v[j + 3]};
                *reinterpret_cast<float4_t*>(dst + j) = t;
            }
        } else {
#pragma unroll
            for (int j = 0; j < KPT; j += 2) {
                double2_t t = {v[j], v[j + 1]};
                *reinterpret_cast<double2_t*>(dst + j) = t;
            }
        }
    }
};

template <int TIER, int BM, int BN, int BK, int WM, int WN>
struct GemmCfg {
    using TT = TierTraits<TIER>;
    using Lds = typename TT::Lds;
    using Out = typename TT::Out;
    using PA = Panel<TIER, BM, BK>;
    using PB = Panel<TIER, BN, BK>;
    static constexpr int WTM = BM / WM;   // wave tile rows (of C)
    static constexpr int WTN = BN / WN;   // wave tile cols
    static_assert(WM * WN == 4, "four waves");
    static constexpr int STAGE = TT::planes * (PA::PLANE + PB::PLANE);
    static constexpr size_t kLdsBytes = 2 * STAGE * sizeof(Lds);
    static_assert(kLdsBytes <= 160 * 1024, "LDS");
};

// ---------------------------------------------------------------- the kernel
// One output tile (flat index g over the launch) per call, computed by the whole
// workgroup.
template <int TIER, int BM, int BN, int BK, int WM, int WN>
__device__ __forceinline__ void gemm_grouped_tile(const GemmLaunch& L, int g)
{
    using Cfg = GemmCfg<TIER, BM, BN, BK, WM, WN>;
    using TT = TierTraits<TIER>;
    using Src = typename TT::Src;
    using Lds = typename TT::Lds;
    using Out = typename TT::Out;
    using PA = typename Cfg::PA;
    using PB = typename Cfg::PB;
    constexpr int WTM = Cfg::WTM, WTN = Cfg::WTN;

    __shared__ __attribute__((aligned(16))) Lds smem[2 * Cfg::STAGE];

    const int tid = threadIdx.x;
    int pi = 0;
#pragma unroll 1
    for (int q = 1; q < L.np; ++q)
        if (g >= L.p[q].tileBegin) pi = q;
    const GemmProb& P = L.p[pi];

    int local = g - P.tileBegin;
    const int bi = local / P.tilesPer;
    local -= bi * P.tilesPer;
    int ti, tj;
    if (P.modeC == kCLower) {
        ti = (int)((__builtin_sqrtf(8.0f * (float)local + 1.0f) - 1.0f) * 0.5f);
        while ((ti + 1) * (ti + 2) / 2 <= local) ++ti;
        while (ti * (ti + 1) / 2 > local) --ti;
        tj = local - ti * (ti + 1) / 2;
    } else {
        ti = local % P.tilesM;
        tj = local / P.tilesM;
    }
    const int i0 = ti * BM, j0 = tj * BN;
    const int M = P.m, N = P.n, K = P.k;

    const Src* Ab = reinterpret_cast<const Src*>(P.A) + (long long)bi * P.sA;
    const Src* Bb = reinterpret_cast<const Src*>(P.B) + (long long)bi * P.sB;
    const Src* A2b = P.A2 ? reinterpret_cast<const Src*>(P.A2) + (long long)bi * P.sA : nullptr;
    const Src* B2b = P.B2 ? reinterpret_cast<const Src*>(P.B2) + (long long)bi * P.sB : nullptr;
    Out* Cb = reinterpret_cast<Out*>(P.C) + (long long)bi * P.sC;

    // K range: a lower-triangular op(A) has no entries right of the tile's last row
    int kend = K;
    if (P.modeA == kOpTriL) kend = min(K, i0 + BM);
    const int kblocks = (kend + BK - 1) / BK;
    const int npairs = A2b ? 2 : 1;
    const int nst = kblocks * npairs;

    PA pa;
    PB pb;
    pa.setup(tid, P.modeA != kOpT);
    pb.setup(tid, P.modeB == kOpT);
    const bool vecA = ((reinterpret_cast<uintptr_t>(Ab) | reinterpret_cast<uintptr_t>(A2b)) % 16 == 0) &&
                      (P.lda % PA::VEC == 0);
    const bool vecB = ((reinterpret_cast<uintptr_t>(Bb) | reinterpret_cast<uintptr_t>(B2b)) % 16 == 0) &&
                      (P.ldb % PB::VEC == 0);

    // Split tier: per-row scale exponents of this tile, read once into LDS so the
    // loaders and the epilogue use the same values even if eA/eB are raised
    // concurrently (eOut may alias them).
    __shared__ int scA_s[(TIER == kSplitF16) ? BM : 1];
    __shared__ int scB_s[(TIER == kSplitF16) ? BN : 1];
    __shared__ int emax_s[BM];
    auto clamp_e = [](int e) { return e < -1000 ? 0 : (e > 1000 ? 1000 : e); };
    if constexpr (TIER == kSplitF16) {
        const int* eA = P.eA + (long long)bi * P.seA;
        const int* eB = P.eB + (long long)bi * P.seB;
        for (int r = tid; r < BM; r += kThreads) scA_s[r] = 15 - (i0 + r < M ? clamp_e(eA[i0 + r]) : 0);
        for (int r = tid; r < BN; r += kThreads) scB_s[r] = 15 - (j0 + r < N ? clamp_e(eB[j0 + r]) : 0);
    }
    for (int r = tid; r < BM; r += kThreads) emax_s[r] = kExpZeroDev;
    __syncthreads();
    if constexpr (TIER == kSplitF16) {
        pa.sc = scA_s[pa.row];
        pb.sc = scB_s[pb.row];
    }

    auto stage_load = [&](int s) {
        const int pair = s / kblocks;
        const int k0 = (s - pair * kblocks) * BK;
        const Src* Xa = pair ? A2b : Ab;
        const Src* Xb = pair ? B2b : Bb;
        pa.load(Xa, P.lda, P.modeA, true, i0, k0, M, K, vecA);
        pb.load(Xb, P.ldb, P.modeB, false, j0, k0, N, K, vecB);
    };
    auto stage_store = [&](int buf) {
        Lds* base = smem + buf * Cfg::STAGE;
        pa.store(base);
        pb.store(base + TT::planes * PA::PLANE);
    };

    const int lane = tid & 63;
    const int wave = tid >> 6;
    const int wm = wave % WM, wn = wave / WM;

    // ------------------------------------------------------------ accumulators
    constexpr bool kMfma16 = (TIER == kPlainF16 || TIER == kSplitF16);
    constexpr int FM16 = WTM / 32, FN16 = WTN / 32;
    constexpr int FM64 = WTM / 16, FN64 = WTN / 16;
    constexpr int RM32 = WTM / 16, CN32 = WTN / 4;   // FMA tier: 16 x 4 lanes
    constexpr int NACC16 = kMfma16 ? FM16 * FN16 : 1;
    constexpr int NACCX = (TIER == kSplitF16) ? FM16 * FN16 : 1;
    constexpr int NACC64 = (TIER == kMfmaF64) ? FM64 * FN64 : 1;
    constexpr int NACC32 = (TIER == kFmaF32) ? RM32 * CN32 : 1;
    floatx16_t acc[NACC16];
    floatx16_t accx[NACCX];
    doublex4_t acc64[NACC64];
    float acc32[NACC32];
#pragma unroll
    for (int q = 0; q < NACC16; ++q) acc[q] = floatx16_t{};
#pragma unroll
    for (int q = 0; q < NACCX; ++q) accx[q] = floatx16_t{};
#pragma unroll
    for (int q = 0; q < NACC64; ++q) acc64[q] = doublex4_t{};
#pragma unroll
    for (int q = 0; q < NACC32; ++q) acc32[q] = 0.0f;

    auto compute = [&](int buf) {
        const Lds* sA = smem + buf * Cfg::STAGE;
        const Lds* sB = sA + TT::planes * PA::PLANE;
        if constexpr (kMfma16) {
            const int fr = lane & 31, fk = (lane >> 5) * 8;
#pragma unroll
            for (int kk = 0; kk < BK; kk += 16) {
                half8_t af[FM16], bf[FN16];
                half8_t afl[TIER == kSplitF16 ? FM16 : 1], bfl[TIER == kSplitF16 ? FN16 : 1];
#pragma unroll
                for (int a = 0; a < FM16; ++a) {
                    const int off = (wm * WTM + a * 32 + fr) * PA::LDK + kk + fk;
                    af[a] = *reinterpret_cast<const half8_t*>(sA + off);
                    if constexpr (TIER == kSplitF16)
                        afl[a] = *reinterpret_cast<const half8_t*>(sA + PA::PLANE + off);
                }
#pragma unroll
                for (int b = 0; b < FN16; ++b) {
                    const int off = (wn * WTN + b * 32 + fr) * PB::LDK + kk + fk;
                    bf[b] = *reinterpret_cast<const half8_t*>(sB + off);
                    if constexpr (TIER == kSplitF16)
                        bfl[b] = *reinterpret_cast<const half8_t*>(sB + PB::PLANE + off);
                }
#pragma unroll
                for (int b = 0; b < FN16; ++b)
#pragma unroll
                    for (int a = 0; a < FM16; ++a) {
                        const int q = b * FM16 + a;
                        acc[q] = __builtin_amdgcn_mfma_f32_32x32x16_f16(bf[b], af[a], acc[q], 0, 0, 0);
                        if constexpr (TIER == kSplitF16) {
                            accx[q] = __builtin_amdgcn_mfma_f32_32x32x16_f16(bf[b], afl[a], accx[q], 0, 0, 0);
                            accx[q] = __builtin_amdgcn_mfma_f32_32x32x16_f16(bfl[b], af[a], accx[q], 0, 0, 0);
                        }
                    }
            }
        } else if constexpr (TIER == kMfmaF64) {
            const int fr = lane & 15, fk = lane >> 4;
#pragma unroll
            for (int kk = 0; kk < BK; kk += 4) {
                double af[FM64], bf[FN64];
#pragma unroll
                for (int a = 0; a < FM64; ++a)
                    af[a] = sA[(wm * WTM + a * 16 + fr) * PA::LDK + kk + fk];
#pragma unroll
                for (int b = 0; b < FN64; ++b)
                    bf[b] = sB[(wn * WTN + b * 16 + fr) * PB::LDK + kk + fk];
#pragma unroll
                for (int b = 0; b < FN64; ++b)
#pragma unroll
                    for (int a = 0; a < FM64; ++a) {
                        const int q = b * FM64 + a;
                        acc64[q] = __builtin_amdgcn_mfma_f64_16x16x4f64(bf[b], af[a], acc64[q], 0, 0, 0);
                    }
            }
        } else {  // kFmaF32: lane (lr, lc) owns rows lr + 16*a, cols lc + 4*b of the wave tile
            const int lr = lane & 15, lc = lane >> 4;
#pragma unroll 2
            for (int kk = 0; kk < BK; kk += 4) {
                float4_t ar[RM32], br[CN32];
#pragma unroll
                for (int a = 0; a < RM32; ++a)
                    ar[a] = *reinterpret_cast<const float4_t*>(sA + (wm * WTM + lr + 16 * a) * PA::LDK + kk);
#pragma unroll
                for (int b = 0; b < CN32; ++b)
                    br[b] = *reinterpret_cast<const float4_t*>(sB + (wn * WTN + lc + 4 * b) * PB::LDK + kk);
#pragma unroll
                for (int q = 0; q < 4; ++q)
#pragma unroll
                    for (int b = 0; b < CN32; ++b)
#pragma unroll
                        for (int a = 0; a < RM32; ++a)
                            acc32[b * RM32 + a] = __builtin_fmaf(ar[a][q], br[b][q], acc32[b * RM32 + a]);
            }
        }
    };

    // ---------------------------------------------------------------- main loop
    if (nst > 0) {
        stage_load(0);
        stage_store(0);
        __syncthreads();
#pragma unroll 1
        for (int s = 0; s < nst; ++s) {
            const bool more = (s + 1 < nst);
            if (more) stage_load(s + 1);
            compute(s & 1);
            if (more) stage_store((s + 1) & 1);
            __syncthreads();
        }
    }

    // ----------------------------------------------------------------- epilogue
    Out alpha, beta;
    if constexpr (sizeof(Out) == 4) {
        alpha = P.alpha32;
        beta = P.beta32;
    } else {
        alpha = P.alpha;
        beta = P.beta;
    }
    const bool lowerC = (P.modeC == kCLower);
    const long long ldc = P.ldc;
    const bool track = (P.eOut != nullptr);
    int emax = kExpZeroDev;   // exponent bound of this lane's current row (FP32 tiers)
    // C elements are written in groups: first every read of the group (C, and the
    // column scales) is issued, unconditionally at clamped addresses, then the values
    // are combined and stored; reads behind per-element branches went out one memory
    // latency apart
    struct CIn {
        bool in;
        Out cold;
        Out cd;
    };
    // the TRSM leaves' diagonal, in the tier's own precision
    const Out* cdiv = static_cast<const Out*>(P.cdiv);
    auto fetch = [&](int i, int j) {
        CIn f;
        f.in = i < M && j < N && (!lowerC || i >= j);
        f.cold = masked_load(Cb + (f.in ? (long long)i + (long long)j * ldc : 0ll),
                             f.in && beta != Out(0));
        f.cd = cdiv ? masked_load(cdiv + (f.in ? j : 0), f.in) : Out(1);
        return f;
    };
    auto put = [&](int i, int j, Out v, const CIn& f) {
        if (f.in) {
            Out* c = Cb + (long long)i + (long long)j * ldc;
            Out w;
            if (cdiv) {
                const Out t = (beta == Out(0)) ? Out(0) : beta * f.cold;
                if constexpr (sizeof(Out) == 4) w = div_sum_f32(alpha * v, t, f.cd);
                else w = (alpha * v + t) / f.cd;
            } else {
                w = (beta == Out(0)) ? alpha * v : alpha * v + beta * f.cold;
            }
            *c = w;
            if constexpr (sizeof(Out) == 4) {
                if (track) {
                    const uint32_t u = __float_as_uint(w) & 0x7fffffffu;
                    // frexp exponent of a normal/inf value, a bound for subnormals
                    const int e = u == 0 ? kExpZeroDev : (int)(u >> 23) - 126;
                    emax = e > emax ? e : emax;
                }
            }
        }
    };
    // flush the lane's row bound into LDS, then one global atomic per row
    auto flush_row = [&](int il) {
        if (track && emax != kExpZeroDev) atomicMax(&emax_s[il], emax);
        emax = kExpZeroDev;
    };
    if constexpr (kMfma16) {
#pragma unroll
        for (int b = 0; b < FN16; ++b)
#pragma unroll
            for (int a = 0; a < FM16; ++a) {
                const int q = b * FM16 + a;
                const int il = wm * WTM + a * 32 + (lane & 31);
                int scA = 0;
                if constexpr (TIER == kSplitF16) scA = scA_s[il];
                CIn f[16];
#pragma unroll
                for (int r = 0; r < 16; ++r)
                    f[r] = fetch(i0 + il, j0 + wn * WTN + b * 32 + (r >> 2) * 8 + (lane >> 5) * 4 + (r & 3));
#pragma unroll
                for (int r = 0; r < 16; ++r) {
                    const int jl = wn * WTN + b * 32 + (r >> 2) * 8 + (lane >> 5) * 4 + (r & 3);
                    float v = acc[q][r];
                    if constexpr (TIER == kSplitF16)
                        v = __builtin_ldexpf(v + accx[q][r] * (1.0f / 2048.0f), -(scA + scB_s[jl]));
                    put(i0 + il, j0 + jl, v, f[r]);
                }
                flush_row(il);
            }
    } else if constexpr (TIER == kMfmaF64) {
#pragma unroll
        for (int b = 0; b < FN64; ++b)
#pragma unroll
            for (int a = 0; a < FM64; ++a) {
                const int q = b * FM64 + a;
                const int i = i0 + wm * WTM + a * 16 + (lane & 15);
                // v_mfma_f64_16x16x4: accumulator r of lane l is row l/16 + 4r, column l%16
                CIn f[4];
#pragma unroll
                for (int r = 0; r < 4; ++r) f[r] = fetch(i, j0 + wn * WTN + b * 16 + (lane >> 4) + 4 * r);
#pragma unroll
                for (int r = 0; r < 4; ++r) put(i, j0 + wn * WTN + b * 16 + (lane >> 4) + 4 * r, acc64[q][r], f[r]);
            }
    } else {
        const int lr = lane & 15, lc = lane >> 4;
#pragma unroll
        for (int a = 0; a < RM32; ++a) {
            const int i = i0 + wm * WTM + lr + 16 * a;
            CIn f[CN32];
#pragma unroll
            for (int b = 0; b < CN32; ++b) f[b] = fetch(i, j0 + wn * WTN + lc + 4 * b);
#pragma unroll
            for (int b = 0; b < CN32; ++b) put(i, j0 + wn * WTN + lc + 4 * b, acc32[b * RM32 + a], f[b]);
            flush_row(wm * WTM + lr + 16 * a);
        }
    }
    if (track) {
        __syncthreads();
        for (int r = tid; r < BM; r += kThreads)
            if (emax_s[r] != kExpZeroDev && i0 + r < M)
                atomicMax(P.eOut + (long long)bi * P.seA + i0 + r, emax_s[r]);
    }
}

// Workgroup b computes tiles b, b + gridDim.x, ... (one tile each when the grid covers
// every tile; a capped grid, a multiple of 8, keeps each workgroup on its XCD's tile
// range through xcd_remap).
template <int TIER, int BM, int BN, int BK, int WM, int WN>
__global__ void __launch_bounds__(kThreads)
gemm_grouped_kernel(GemmLaunch L)
{
#pragma unroll 1
    for (int b = blockIdx.x; b < L.totalTiles; b += gridDim.x) {
        gemm_grouped_tile<TIER, BM, BN, BK, WM, WN>(L, xcd_remap(b, L.totalTiles));
        __syncthreads();      // the next tile reuses the LDS stages
    }
}

}  // namespace tb
