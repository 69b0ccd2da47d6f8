// Grouped FP16-MFMA GEMM for the plain and split tiers (gfx950, v_mfma_f32_32x32x16_f16).
//
// C = alpha * sum_pairs op(A) op(B) + beta * C over a flat list of output tiles of up to
// kMaxProbs strided-batch problems (tb_common.hpp).  512 threads (8 waves) own one
// BM x BN tile; K runs in BK steps through two LDS buffers: the global loads of step
// s+1 are in flight while the MFMAs of step s run.
//
// Each operand arrives in one of two forms (template SRCA / SRCB):
//   kSrcF32    FP32 matrix in any of the kOp* layouts, converted in registers on its way
//              to LDS (used for operands that change during a routine, e.g. TRSM's
//              solved block X);
//   kSrcPlanes NPL pre-converted FP16 planes laid out K-contiguous (plane[r*ldp + k]),
//              copied to LDS with 16-byte loads and no arithmetic.
// NPL = 1 (plain tier): x -> RNE fp16(x), the reference's s2h conversion.
// NPL = 3 (split tier): xs = x * 2^(15 - e_row) is written exactly as
//   xs = h + 2^-11 m + 2^-22 l (h, m, l fp16; 33 significant bits >= FP32's 24)
// and the six products of total weight >= 2^-22 are accumulated in three FP32
// accumulators:  acc0 += hh;  acc1 += hm + mh;  acc2 += hl + mm + lh.
// C = (acc0 + 2^-11 (acc1 + 2^-11 acc2)) * 2^-(scA + scB): products whose weight is
// 2^-33 or less are dropped, so the result carries FP32 accumulation error only.
//
// LDS holds each plane of a panel as [row][BK + 8] halves: every MFMA fragment is one
// 16-byte ds_read and rows are 80 bytes apart (bank-conflict free for 8 lanes).  The
// MFMA's A input comes from the op(B) panel and its B input from the op(A) panel, so
// the accumulator's lane index runs along C's rows and column-major stores coalesce.
#pragma once
#include <hip/hip_runtime.h>

#include "gemm_kernel.hpp"   // vector types, xcd_remap, kExpZeroDev
#include "tb_common.hpp"

namespace tb {

// One operand panel: R rows (of op(A), or columns of op(B)) x BK, staged by NT threads.
template <int NPL, int SRC, int R, int BK, int NT>
struct Panel16 {
    static constexpr int KPT = 8;                 // k per work item (one 16-byte plane row)
    static constexpr int TPR = BK / KPT;          // items per panel row
    static constexpr int ITEMS_TOTAL = R * TPR;
    static constexpr int ITEMS = (ITEMS_TOTAL + NT - 1) / NT;
    static constexpr int LDK = BK + 8;            // LDS row stride (halves)
    static constexpr int PLANE = R * LDK;         // halves per plane
    static_assert(BK % KPT == 0, "BK");
    static_assert(ITEMS_TOTAL % NT == 0 || ITEMS_TOTAL < NT, "items");

    // staging registers
    float vf[SRC == kSrcF32 ? ITEMS : 1][KPT];
    half8_t vh[SRC == kSrcPlanes ? ITEMS : 1][NPL];
    int row[ITEMS], kseg[ITEMS];
    bool live[ITEMS];
    int sc[ITEMS];   // split: scale exponent of the item's row

    __device__ __forceinline__ void setup(int tid, bool mn_major)
    {
#pragma unroll
        for (int q = 0; q < ITEMS; ++q) {
            const int it = tid + q * NT;
            live[q] = it < ITEMS_TOTAL;
            if (mn_major) {
                row[q] = it % R;
                kseg[q] = it / R;
            } else {
                row[q] = it / TPR;
                kseg[q] = it % TPR;
            }
            sc[q] = 0;
        }
    }

    // FP32 source, op(X) rows [r0, r0+R) x k [k0, k0+BK); layouts as in gemm_kernel.hpp
    __device__ __forceinline__ void load_f32(const float* __restrict__ X, long long ld, int mode,
                                             bool isA, int r0, int k0, int rows, int K,
                                             bool vec_ok)
    {
        int m = mode;
        bool zero = false, special = false;
        if (isA && (mode == kOpSymL || mode == kOpTriL)) {
            if (k0 + BK <= r0) m = kOpN;
            else if (k0 >= r0 + R) { m = kOpT; zero = (mode == kOpTriL); }
            else special = true;
        }
        const bool mn = isA ? (m != kOpT) : (m == kOpT);
        const bool full = (r0 + R <= rows) && (k0 + BK <= K);
#pragma unroll
        for (int q = 0; q < ITEMS; ++q) {
            const int r = r0 + row[q];
            const int kb = k0 + kseg[q] * KPT;
            if (zero || !live[q]) {
#pragma unroll
                for (int j = 0; j < KPT; ++j) vf[q][j] = 0.0f;
                continue;
            }
            if (!special && full) {
                if (mn) {
                    const float* p = X + (long long)r + (long long)kb * ld;
#pragma unroll
                    for (int j = 0; j < KPT; ++j) vf[q][j] = p[(long long)j * ld];
                } else if (vec_ok) {
                    const float* p = X + (long long)r * ld + kb;
                    const float4_t a = *reinterpret_cast<const float4_t*>(p);
                    const float4_t b = *reinterpret_cast<const float4_t*>(p + 4);
                    vf[q][0] = a[0]; vf[q][1] = a[1]; vf[q][2] = a[2]; vf[q][3] = a[3];
                    vf[q][4] = b[0]; vf[q][5] = b[1]; vf[q][6] = b[2]; vf[q][7] = b[3];
                } else {
                    const float* p = X + (long long)r * ld + kb;
#pragma unroll
                    for (int j = 0; j < KPT; ++j) vf[q][j] = p[j];
                }
                continue;
            }
#pragma unroll
            for (int j = 0; j < KPT; ++j) {
                const int k = kb + j;
                float x = 0.0f;
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
                vf[q][j] = x;
            }
        }
    }

    // pre-converted planes: plane p of row r at P + p*sp + r*ldp (K-contiguous, zero
    // padded up to a multiple of 8 in k)
    __device__ __forceinline__ void load_planes(const _Float16* __restrict__ P, long long ldp,
                                                long long sp, int r0, int k0, int rows, int K)
    {
#pragma unroll
        for (int q = 0; q < ITEMS; ++q) {
            const int r = r0 + row[q];
            const int kb = k0 + kseg[q] * KPT;
            const bool ok = live[q] && r < rows && kb < K;
#pragma unroll
            for (int p = 0; p < NPL; ++p) {
                if (ok)
                    vh[q][p] = *reinterpret_cast<const half8_t*>(P + p * sp + (long long)r * ldp + kb);
                else
                    vh[q][p] = half8_t{};
            }
            // planes may continue past this problem's K (a sub-block of a larger matrix)
            if (ok && kb + KPT > K) {
#pragma unroll
                for (int j = 0; j < KPT; ++j)
                    if (kb + j >= K)
#pragma unroll
                        for (int p = 0; p < NPL; ++p) vh[q][p][j] = (_Float16)0.0f;
            }
        }
    }

    __device__ __forceinline__ void store(_Float16* base) const
    {
#pragma unroll
        for (int q = 0; q < ITEMS; ++q) {
            if (!live[q]) continue;
            _Float16* dst = base + row[q] * LDK + kseg[q] * KPT;
            if constexpr (SRC == kSrcPlanes) {
#pragma unroll
                for (int p = 0; p < NPL; ++p) *reinterpret_cast<half8_t*>(dst + p * PLANE) = vh[q][p];
            } else if constexpr (NPL == 1) {
                half8_t hh;
#pragma unroll
                for (int j = 0; j < KPT; ++j) hh[j] = (_Float16)vf[q][j];
                *reinterpret_cast<half8_t*>(dst) = hh;
            } else {
                half8_t hh, mm, ll;
#pragma unroll
                for (int j = 0; j < KPT; ++j) {
                    const float xs = __builtin_ldexpf(vf[q][j], sc[q]);
                    const _Float16 h = (_Float16)xs;
                    const float r1 = (xs - (float)h) * 2048.0f;      // exact
                    const _Float16 m = (_Float16)r1;
                    const float r2 = (r1 - (float)m) * 2048.0f;      // exact
                    hh[j] = h;
                    mm[j] = m;
                    ll[j] = (_Float16)r2;
                }
                *reinterpret_cast<half8_t*>(dst) = hh;
                *reinterpret_cast<half8_t*>(dst + PLANE) = mm;
                *reinterpret_cast<half8_t*>(dst + 2 * PLANE) = ll;
            }
        }
    }
};

// tile order inside a full problem: groups of kGroupM tile rows, column by column,
// so the ~32 tiles an XCD runs at once share 8 row panels and 4 column panels
constexpr int kGroupM = 8;

template <int NPL, int SRCA, int SRCB, int BM, int BN, int BK, int WM, int WN>
__global__ void __launch_bounds__(64 * WM * WN)
gemm16_kernel(GemmLaunch L)
{
    constexpr int NT = 64 * WM * WN;
    using PA = Panel16<NPL, SRCA, BM, BK, NT>;
    using PB = Panel16<NPL, SRCB, BN, BK, NT>;
    constexpr int WTM = BM / WM, WTN = BN / WN;
    constexpr int FM = WTM / 32, FN = WTN / 32;
    static_assert(FM >= 1 && FN >= 1, "wave tile");
    constexpr int STAGE = NPL * (PA::PLANE + PB::PLANE);
    constexpr bool kSplit = (NPL == 3);

    __shared__ __attribute__((aligned(16))) _Float16 smem[2 * STAGE];
    __shared__ int scA_s[kSplit ? BM : 1];
    __shared__ int scB_s[kSplit ? BN : 1];
    __shared__ int emax_s[BM];

    const int tid = threadIdx.x;
    const int g = xcd_remap(blockIdx.x, L.totalTiles);
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
        const int per_group = kGroupM * P.tilesN;
        const int grp = local / per_group;
        const int first = grp * kGroupM;
        const int gsz = min(P.tilesM - first, kGroupM);
        const int in = local - grp * per_group;
        ti = first + in % gsz;
        tj = in / gsz;
    }
    const int i0 = ti * BM, j0 = tj * BN;
    const int M = P.m, N = P.n, K = P.k;

    int kend = K;
    if (P.modeA == kOpTriL) kend = min(K, i0 + BM);
    const int kblocks = (kend + BK - 1) / BK;
    const bool pairs = P.A2 != nullptr;
    const int nst = kblocks * (pairs ? 2 : 1);

    // operand bases of this batch element
    const float* Af = reinterpret_cast<const float*>(P.A) + (long long)bi * P.sA;
    const float* A2f = pairs ? reinterpret_cast<const float*>(P.A2) + (long long)bi * P.sA : nullptr;
    const float* Bf = reinterpret_cast<const float*>(P.B) + (long long)bi * P.sB;
    const float* B2f = pairs ? reinterpret_cast<const float*>(P.B2) + (long long)bi * P.sB : nullptr;
    const _Float16* Ah = reinterpret_cast<const _Float16*>(P.A) + (long long)bi * P.sA;
    const _Float16* A2h = pairs ? reinterpret_cast<const _Float16*>(P.A2) + (long long)bi * P.sA : nullptr;
    const _Float16* Bh = reinterpret_cast<const _Float16*>(P.B) + (long long)bi * P.sB;
    const _Float16* B2h = pairs ? reinterpret_cast<const _Float16*>(P.B2) + (long long)bi * P.sB : nullptr;
    float* Cb = reinterpret_cast<float*>(P.C) + (long long)bi * P.sC;

    PA pa;
    PB pb;
    pa.setup(tid, SRCA == kSrcF32 && P.modeA != kOpT);
    pb.setup(tid, SRCB == kSrcF32 && P.modeB == kOpT);
    const bool vecA = SRCA == kSrcF32 &&
                      ((reinterpret_cast<uintptr_t>(Af) | reinterpret_cast<uintptr_t>(A2f)) % 16 == 0) &&
                      (P.lda % 4 == 0);
    const bool vecB = SRCB == kSrcF32 &&
                      ((reinterpret_cast<uintptr_t>(Bf) | reinterpret_cast<uintptr_t>(B2f)) % 16 == 0) &&
                      (P.ldb % 4 == 0);

    auto clamp_e = [](int e) { return e < -1000 ? 0 : (e > 1000 ? 1000 : e); };
    if constexpr (kSplit) {
        const int* eA = P.eA + (long long)bi * P.seA;
        const int* eB = P.eB + (long long)bi * P.seB;
        for (int r = tid; r < BM; r += NT) scA_s[r] = 15 - (i0 + r < M ? clamp_e(eA[i0 + r]) : 0);
        for (int r = tid; r < BN; r += NT) scB_s[r] = 15 - (j0 + r < N ? clamp_e(eB[j0 + r]) : 0);
    }
    for (int r = tid; r < BM; r += NT) emax_s[r] = kExpZeroDev;
    __syncthreads();
    if constexpr (kSplit) {
#pragma unroll
        for (int q = 0; q < PA::ITEMS; ++q) pa.sc[q] = scA_s[pa.row[q]];
#pragma unroll
        for (int q = 0; q < PB::ITEMS; ++q) pb.sc[q] = scB_s[pb.row[q]];
    }

    auto stage_load = [&](int s) {
        const int pair = s >= kblocks;
        const int k0 = (s - pair * kblocks) * BK;
        if constexpr (SRCA == kSrcF32)
            pa.load_f32(pair ? A2f : Af, P.lda, P.modeA, true, i0, k0, M, K, vecA);
        else
            pa.load_planes(pair ? A2h : Ah, P.lda, P.spA, i0, k0, M, K);
        if constexpr (SRCB == kSrcF32)
            pb.load_f32(pair ? B2f : Bf, P.ldb, P.modeB, false, j0, k0, N, K, vecB);
        else
            pb.load_planes(pair ? B2h : Bh, P.ldb, P.spB, j0, k0, N, K);
    };
    auto stage_store = [&](int buf) {
        _Float16* base = smem + buf * STAGE;
        pa.store(base);
        pb.store(base + NPL * PA::PLANE);
    };

    const int lane = tid & 63;
    const int wave = tid >> 6;
    const int wm = wave % WM, wn = wave / WM;

    constexpr int NACC = FM * FN;
    floatx16_t acc0[NACC];
    floatx16_t acc1[kSplit ? NACC : 1];
    floatx16_t acc2[kSplit ? NACC : 1];
#pragma unroll
    for (int q = 0; q < NACC; ++q) acc0[q] = floatx16_t{};
#pragma unroll
    for (int q = 0; q < (kSplit ? NACC : 1); ++q) {
        acc1[q] = floatx16_t{};
        acc2[q] = floatx16_t{};
    }

    auto compute = [&](int buf) {
        const _Float16* sA = smem + buf * STAGE;
        const _Float16* sB = sA + NPL * PA::PLANE;
        const int fr = lane & 31, fk = (lane >> 5) * 8;
#pragma unroll
        for (int kk = 0; kk < BK; kk += 16) {
            half8_t af[FM][NPL], bf[FN][NPL];
#pragma unroll
            for (int a = 0; a < FM; ++a) {
                const int off = (wm * WTM + a * 32 + fr) * PA::LDK + kk + fk;
#pragma unroll
                for (int p = 0; p < NPL; ++p)
                    af[a][p] = *reinterpret_cast<const half8_t*>(sA + p * PA::PLANE + off);
            }
#pragma unroll
            for (int b = 0; b < FN; ++b) {
                const int off = (wn * WTN + b * 32 + fr) * PB::LDK + kk + fk;
#pragma unroll
                for (int p = 0; p < NPL; ++p)
                    bf[b][p] = *reinterpret_cast<const half8_t*>(sB + p * PB::PLANE + off);
            }
            // products ordered so dependent MFMAs on one accumulator are >= NACC apart
#pragma unroll
            for (int q = 0; q < NACC; ++q)
                acc0[q] = __builtin_amdgcn_mfma_f32_32x32x16_f16(bf[q / FM][0], af[q % FM][0], acc0[q], 0, 0, 0);
            if constexpr (kSplit) {
#pragma unroll
                for (int q = 0; q < NACC; ++q)
                    acc1[q] = __builtin_amdgcn_mfma_f32_32x32x16_f16(bf[q / FM][0], af[q % FM][1], acc1[q], 0, 0, 0);
#pragma unroll
                for (int q = 0; q < NACC; ++q)
                    acc2[q] = __builtin_amdgcn_mfma_f32_32x32x16_f16(bf[q / FM][0], af[q % FM][2], acc2[q], 0, 0, 0);
#pragma unroll
                for (int q = 0; q < NACC; ++q)
                    acc1[q] = __builtin_amdgcn_mfma_f32_32x32x16_f16(bf[q / FM][1], af[q % FM][0], acc1[q], 0, 0, 0);
#pragma unroll
                for (int q = 0; q < NACC; ++q)
                    acc2[q] = __builtin_amdgcn_mfma_f32_32x32x16_f16(bf[q / FM][1], af[q % FM][1], acc2[q], 0, 0, 0);
#pragma unroll
                for (int q = 0; q < NACC; ++q)
                    acc2[q] = __builtin_amdgcn_mfma_f32_32x32x16_f16(bf[q / FM][2], af[q % FM][0], acc2[q], 0, 0, 0);
            }
        }
    };

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
    const float alpha = (float)P.alpha, beta = (float)P.beta;
    const bool lowerC = (P.modeC == kCLower);
    const bool track = (P.eOut != nullptr);
    const long long ldc = P.ldc;
#pragma unroll
    for (int a = 0; a < FM; ++a) {
        const int il = wm * WTM + a * 32 + (lane & 31);
        const int i = i0 + il;
        int emax = kExpZeroDev;
        const int scA = kSplit ? scA_s[il] : 0;
#pragma unroll
        for (int b = 0; b < FN; ++b) {
            const int q = b * FM + a;
#pragma unroll
            for (int r = 0; r < 16; ++r) {
                const int jl = wn * WTN + b * 32 + (r >> 2) * 8 + (lane >> 5) * 4 + (r & 3);
                const int j = j0 + jl;
                float v = acc0[q][r];
                if constexpr (kSplit) {
                    const float t = __builtin_fmaf(acc2[q][r], 1.0f / 2048.0f, acc1[q][r]);
                    v = __builtin_ldexpf(__builtin_fmaf(t, 1.0f / 2048.0f, v), -(scA + scB_s[jl]));
                }
                if (i < M && j < N && (!lowerC || i >= j)) {
                    float* c = Cb + (long long)i + (long long)j * ldc;
                    float w;
                    if (P.cscale) {
                        const double cold = (beta == 0.0f) ? 0.0 : (double)*c;
                        w = (float)(((double)alpha * (double)v + (double)beta * cold) * P.cscale[j]);
                    } else {
                        w = (beta == 0.0f) ? alpha * v : alpha * v + beta * *c;
                    }
                    *c = w;
                    if (track) {
                        const uint32_t u = __float_as_uint(w) & 0x7fffffffu;
                        const int e = u == 0 ? kExpZeroDev : (int)(u >> 23) - 126;
                        emax = e > emax ? e : emax;
                    }
                }
            }
        }
        if (track && emax != kExpZeroDev) atomicMax(&emax_s[il], emax);
    }
    if (track) {
        __syncthreads();
        for (int r = tid; r < BM; r += NT)
            if (emax_s[r] != kExpZeroDev && i0 + r < M)
                atomicMax(P.eOut + (long long)bi * P.seA + i0 + r, emax_s[r]);
    }
}

}  // namespace tb
