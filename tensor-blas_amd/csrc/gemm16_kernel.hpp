// Grouped FP16-MFMA GEMM for the plain and split tiers (gfx950, v_mfma_f32_32x32x16_f16).
//
// C = alpha * sum_pairs op(A) op(B) + beta * C over a flat list of output tiles of up to
// kMaxProbs strided-batch problems (tb_common.hpp).  A workgroup of WM x WN waves owns
// one BM x BN tile (128 x 128, 64 x 256 for TRSM leaves, 256 x 256 wide tiles: eight
// waves of 128 x 64 in the plain tier, four of 128 x 128 in the split tier) and walks K
// in BK = 32 steps through a ring of NST LDS stages.
//
// Operands arrive in one of two forms (template SRCA / SRCB):
//   kSrcPlanes NPL pre-converted FP16 planes, K-contiguous (plane[r*ldp + k]).  They are
//              copied global -> LDS by the CDNA4 direct-to-LDS loads
//              (global_load_lds_dwordx4): no registers, no ds_write, and each stage's
//              loads are issued NST-1 stages ahead of their use.
//   kSrcF32    an FP32 matrix in any of the kOp* layouts, loaded one stage ahead into
//              registers and converted on its way to LDS (operands that change during
//              a routine, e.g. TRSM's solved block X).
// NPL = 1 (plain tier): x -> RNE fp16(x), the reference's s2h conversion.
// NPL = 3 (split tier): xs = x * 2^(15 - e_row) is written exactly as
//   xs = h + 2^-11 m + 2^-22 l (h, m, l fp16; 33 significant bits >= FP32's 24)
// and the six products of total weight >= 2^-22 go to three FP32 accumulators:
//   acc0 += hh;  acc1 += hm + mh;  acc2 += hl + mm + lh;
// C = (acc0 + 2^-11 (acc1 + 2^-11 acc2)) * 2^-(scA + scB).
// NPL = 2 (split tier, two planes): xs = h + m + r, m = RNE fp16(xs - h) (so |m| <=
// 2^-11 |xs|) and |r| <= 2^-22 |xs|max (23 significant bits for the row's largest
// entries).  The three products go to one FP32 accumulator, acc0 += hh + hm + mh: every
// product of two fp16 values is exact in FP32 and the accumulation rounds like an FP32
// GEMM's, and one accumulator set leaves the registers for 128 x 64 wave tiles.
// C = acc0 * 2^-(scA + scB).
//
// LDS: every plane of a panel is [row][32 halves] = 64-byte rows, unpadded; the four
// 16-byte units of row r sit at unit position u ^ ((r >> 2) & 3).  Both the 16-lane
// store pattern (4 rows x 4 units) and the MFMA fragment reads (16 rows, one unit) then
// cover all 64 banks once.  The MFMA's A input comes from the op(B) panel and its B
// input from the op(A) panel, so the accumulator's lane index runs along C's rows and
// the column-major stores of C coalesce.
#pragma once
#include <hip/hip_runtime.h>

#include <type_traits>
#include <utility>

#include "epilogue.hpp"
#include "gemm_kernel.hpp"   // vector types, xcd_remap, kExpZeroDev
#include "tb_kernels.hpp"    // plane layouts
#include "tb_common.hpp"

namespace tb {

typedef __attribute__((address_space(3))) void lds_void_t;

constexpr int kBK = 32;          // k per stage
constexpr int kRowBytes = 64;    // one LDS row: 32 halves
// tile rows per group in the tile order of full problems (4 and 16 measured slower at
// 8192^3 and in the TRSM 16384)
constexpr int kGroupM = 8;

__device__ __forceinline__ int swz(int row, int unit) { return unit ^ ((row >> 2) & 3); }

// f(integral_constant<int, 0>), ..., f(integral_constant<int, N-1>) in order
template <class F, int... I>
__device__ __forceinline__ void static_for_impl(F&& f, std::integer_sequence<int, I...>)
{
    (f(std::integral_constant<int, I>{}), ...);
}
template <int N, class F>
__device__ __forceinline__ void static_for(F&& f)
{
    static_for_impl(f, std::make_integer_sequence<int, N>{});
}

// "s_waitcnt vmcnt(n)" then a workgroup barrier.  Written as one asm statement with a
// memory clobber: the compiler neither moves memory operations across it nor adds the
// vmcnt(0) a fenced __syncthreads() would (which would also wait for the prefetched
// stages still in flight).  lgkmcnt(0) completes this wave's ds_writes first.
#define TB_WAIT_VM_BARRIER(n) \
    asm volatile("s_waitcnt vmcnt(%0) lgkmcnt(0)\n\ts_barrier" ::"i"(n) : "memory")

// One direct-to-LDS load: 16 bytes from g (per lane) to LDS[lds_base + 16*lane].
// Issued as asm so that the compiler does not guard every later LDS read with a
// vmcnt(0) for it (it cannot tell the ring slots apart); the kernel's own
// TB_WAIT_VM_BARRIER accounting orders it.  The instruction has no result register;
// its address registers are consumed at issue.  lds_base is wave-uniform; it goes
// through readfirstlane and is bound to m0 by the "{m0}" operand constraint, so the
// compiler itself writes m0 (no clobber of a reserved register).
__device__ __forceinline__ void lds_dma16(const void* g, unsigned lds_base)
{
    asm volatile("global_load_lds_dwordx4 %1, off"
                 :
                 : "{m0}"(__builtin_amdgcn_readfirstlane(lds_base)), "v"(g)
                 : "memory");
}

// One operand panel of R rows x kBK with NPL planes.  LDS layout of a stage:
//   separate planes (NPL 1, 3): plane p at p*PLANE_BYTES, rows of 64 bytes, unit u of row
//     r at swz(r, u) (the 16-lane patterns below cover the 64 banks once);
//   interleaved 2 planes (as in global memory, tb_kernels.hpp): rows of 128 bytes holding
//     the 4 units of plane 0 then the 4 of plane 1, logical unit lu = 4p + u of row r at
//     lu ^ ((r >> 1) & 7): 16 rows of one logical unit fall on 16 distinct 16-byte bank
//     groups, so the fragment reads are conflict-free again.
template <int NPL, int SRC, int R, int NT, int RST = 0>
struct Panel16 {
    static constexpr bool IL = planes_il(NPL);
    static constexpr int UNITS = R * 4;                 // 16-byte units per plane
    static constexpr int PLANE_BYTES = R * kRowBytes;
    static constexpr int BYTES = NPL * PLANE_BYTES;
    // direct-to-LDS: one instruction moves 64 units (1 KB) per wave: 16 rows of one plane
    // (separate planes) or 8 rows of both planes (interleaved)
    static constexpr int DMA_PER_PLANE = UNITS / 64;    // wave-instructions per plane
    static constexpr int NWAVES = NT / 64;
    static constexpr int DMA_PER_WAVE = (NPL * DMA_PER_PLANE + NWAVES - 1) / NWAVES;

    // LDS byte offset of 16-byte unit u (k 8u .. 8u+7) of plane p, row r of a stage
    __device__ static __forceinline__ int lds_off(int r, int p, int u)
    {
        if constexpr (IL) return r * (2 * kRowBytes) + (((p << 2) | u) ^ ((r >> 1) & 7)) * 16;
        else return p * PLANE_BYTES + r * kRowBytes + swz(r, u) * 16;
    }
    // register staging (kSrcF32): ITEMS units of 8 k per thread
    static constexpr int ITEMS = (UNITS + NT - 1) / NT;
    static_assert(UNITS % 64 == 0, "rows");
    // every wave issues the same number of direct-to-LDS loads per stage, so one
    // vmcnt value describes a stage for all of them
    static_assert((NPL * DMA_PER_PLANE) % NWAVES == 0, "even DMA split");

    // register stages of FP32 operands: loaded two iterations before they are stored
    // (one where three planes' accumulators leave no room for a second), or RST (the
    // TRSM leaves: every stage of K <= 256 at once)
    static constexpr int RSTAGES = RST > 0 ? RST : ((SRC == kSrcF32 && NPL < 3) ? 2 : 1);
    float vf[RSTAGES][SRC == kSrcF32 ? ITEMS : 1][8];
    int row[SRC == kSrcF32 ? ITEMS : 1], unit[SRC == kSrcF32 ? ITEMS : 1];
    bool live[SRC == kSrcF32 ? ITEMS : 1];
    int sc[SRC == kSrcF32 ? ITEMS : 1];

    // ---------------------------------------------------------------- planes
    // This wave's share of a stage's direct-to-LDS loads: instruction q moves 1 KB.
    // gp[q] is this lane's source at k = 0 (rows past `rows` clamped onto row rows-1:
    // their C entries are never stored; k past the problem's K must hold zeros in the
    // planes, which the host guarantees), ldo[q] its LDS byte offset.  kofs: the
    // interleaved operand's column offset within the block P points at (a multiple of 8,
    // so no 16-byte unit straddles two blocks).
    const _Float16* gp[SRC == kSrcPlanes ? DMA_PER_WAVE : 1];
    unsigned ldo[SRC == kSrcPlanes ? DMA_PER_WAVE : 1];

    __device__ __forceinline__ void dma_setup(const _Float16* __restrict__ P, long long ldp,
                                             long long sp, int kofs, int r0, int rows, int wave,
                                             int lane)
    {
        if constexpr (SRC == kSrcPlanes) {
#pragma unroll
            for (int q = 0; q < DMA_PER_WAVE; ++q) {
                const int inst = wave + q * NWAVES;
                if constexpr (IL) {
                    // 8 rows x 128 bytes: lane -> row blk*8 + lane/8, physical unit lane%8
                    const int rl = inst * 8 + (lane >> 3);
                    const int lu = (lane & 7) ^ ((rl >> 1) & 7);      // logical unit there
                    const int p = lu >> 2, c = kofs + (lu & 3) * 8;
                    int r = r0 + rl;
                    r = r < rows ? r : rows - 1;
                    gp[q] = P + (long long)r * 2 * ldp + (c >> 5) * 64 + p * 32 + (c & 31);
                    ldo[q] = (unsigned)(inst * 1024);
                } else {
                    const int p = inst / DMA_PER_PLANE;             // 0 .. NPL*DMA_PER_PLANE-1
                    const int blk = inst % DMA_PER_PLANE;
                    const int pos = blk * 64 + lane;            // physical unit in the plane
                    const int rl = pos >> 2;
                    const int u = swz(rl, pos & 3);             // logical unit stored there
                    int r = r0 + rl;
                    r = r < rows ? r : rows - 1;
                    gp[q] = P + p * sp + (long long)r * ldp + u * 8;
                    ldo[q] = (unsigned)(p * PLANE_BYTES + blk * 1024);
                }
            }
        }
    }

    // the loads of the stage at k0 (a multiple of kBK): interleaved rows advance two
    // planes' worth per 32 columns
    __device__ __forceinline__ void dma(int k0, unsigned lds_base) const
    {
        if constexpr (SRC == kSrcPlanes) {
            const int adv = IL ? 2 * k0 : k0;
#pragma unroll
            for (int q = 0; q < DMA_PER_WAVE; ++q) lds_dma16(gp[q] + adv, lds_base + ldo[q]);
        }
    }
    // only load q of the stage at k0 (so that the loads can be spread between MFMAs)
    __device__ __forceinline__ void dma_q(int q, int k0, unsigned lds_base) const
    {
        if constexpr (SRC == kSrcPlanes) {
            const int adv = IL ? 2 * k0 : k0;
            lds_dma16(gp[q] + adv, lds_base + ldo[q]);
        }
    }

    // ------------------------------------------------------------- FP32 source
    __device__ __forceinline__ void setup(int tid, bool mn_major)
    {
        if constexpr (SRC == kSrcF32) {
#pragma unroll
            for (int q = 0; q < ITEMS; ++q) {
                const int it = tid + q * NT;
                live[q] = it < UNITS;
                if (mn_major) {            // consecutive threads: consecutive rows
                    row[q] = it % R;
                    unit[q] = it / R;
                } else {                   // consecutive threads: consecutive k
                    row[q] = it >> 2;
                    unit[q] = it & 3;
                }
                sc[q] = 0;
            }
        }
    }

    template <int BUF>
    __device__ __forceinline__ void load_f32(const float* __restrict__ X, long long ld, int mode,
                                             bool isA, int r0, int k0, int rows, int K,
                                             bool vec_ok)
    {
        float(&vf)[SRC == kSrcF32 ? ITEMS : 1][8] = this->vf[BUF % RSTAGES];
        if constexpr (SRC == kSrcF32) {
            int m = mode;
            bool zero = false, special = false;
            if (isA && (mode == kOpSymL || mode == kOpTriL)) {
                if (k0 + kBK <= r0) m = kOpN;
                else if (k0 >= r0 + R) { m = kOpT; zero = (mode == kOpTriL); }
                else special = true;
            }
            const bool mn = isA ? (m != kOpT) : (m == kOpT);
            const bool full = (r0 + R <= rows) && (k0 + kBK <= K);
#pragma unroll
            for (int q = 0; q < ITEMS; ++q) {
                const int r = r0 + row[q];
                const int kb = k0 + unit[q] * 8;
                if (zero || !live[q]) {
#pragma unroll
                    for (int j = 0; j < 8; ++j) vf[q][j] = 0.0f;
                    continue;
                }
                if (!special && full) {
                    if (mn) {
                        const float* p = X + (long long)r + (long long)kb * ld;
#pragma unroll
                        for (int j = 0; j < 8; ++j) vf[q][j] = p[(long long)j * ld];
                    } else if (vec_ok) {
                        const float* p = X + (long long)r * ld + kb;
                        const float4_t a = *reinterpret_cast<const float4_t*>(p);
                        const float4_t b = *reinterpret_cast<const float4_t*>(p + 4);
                        vf[q][0] = a[0]; vf[q][1] = a[1]; vf[q][2] = a[2]; vf[q][3] = a[3];
                        vf[q][4] = b[0]; vf[q][5] = b[1]; vf[q][6] = b[2]; vf[q][7] = b[3];
                    } else {
                        const float* p = X + (long long)r * ld + kb;
#pragma unroll
                        for (int j = 0; j < 8; ++j) vf[q][j] = p[j];
                    }
                    continue;
                }
#pragma unroll
                for (int j = 0; j < 8; ++j) {
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
    }

    template <int BUF>
    __device__ __forceinline__ void store_f32(char* lds) const
    {
        const float(&vf)[SRC == kSrcF32 ? ITEMS : 1][8] = this->vf[BUF % RSTAGES];
        if constexpr (SRC == kSrcF32) {
#pragma unroll
            for (int q = 0; q < ITEMS; ++q) {
                if (!live[q]) continue;
                char* dst = lds + lds_off(row[q], 0, unit[q]);
                if constexpr (NPL == 1) {
                    half8_t hh;
#pragma unroll
                    for (int j = 0; j < 8; ++j) hh[j] = (_Float16)vf[q][j];
                    *reinterpret_cast<half8_t*>(dst) = hh;
                } else {
                    half8_t hh, mm, ll;
#pragma unroll
                    for (int j = 0; j < 8; ++j) {
                        const float xs = __builtin_ldexpf(vf[q][j], sc[q]);
                        const _Float16 h = (_Float16)xs;
                        hh[j] = h;
                        if constexpr (NPL == 2) {
                            mm[j] = (_Float16)(xs - (float)h);          // xs - h exact
                        } else {
                            const float r1 = (xs - (float)h) * 2048.0f;  // exact
                            const _Float16 m = (_Float16)r1;
                            mm[j] = m;
                            ll[j] = (_Float16)((r1 - (float)m) * 2048.0f);
                        }
                    }
                    *reinterpret_cast<half8_t*>(dst) = hh;
                    *reinterpret_cast<half8_t*>(lds + lds_off(row[q], 1, unit[q])) = mm;
                    if constexpr (NPL == 3) *reinterpret_cast<half8_t*>(lds + lds_off(row[q], 2, unit[q])) = ll;
                }
            }
        }
    }
};

// Compile-time shape of one FP16-MFMA tile configuration.
template <int NPL, int SRCA, int SRCB, int BM, int BN, int WM, int WN, int NST>
struct Gemm16Cfg {
    static constexpr int NT = 64 * WM * WN;
    // TRSM leaf tiles (64 x 256, op(A) the leaf's own block of B in FP32, K <= 256): all
    // of op(A)'s stages are loaded into registers before the first MFMA, so a leaf waits
    // for one memory latency instead of one per stage
    static constexpr int kLeafStages = 8;
    static constexpr bool kLeafAll = SRCA == kSrcF32 && SRCB == kSrcPlanes && BM == 64 && BN == 256 && NPL <= 2;
    using PA = Panel16<NPL, SRCA, BM, NT, kLeafAll ? kLeafStages : 0>;
    using PB = Panel16<NPL, SRCB, BN, NT>;
    static constexpr int STAGE = PA::BYTES + PB::BYTES;
    static constexpr int SMEM = NST * STAGE;        // LDS ring bytes
    // split-tier square tiles with an FP32 op(A) and two stages: two workgroups per CU
    // (66 KB of LDS each), their registers held to 128 per lane (four waves per SIMD)
    static constexpr int kWavesPerEU =
        (NPL == 2 && SRCA == kSrcF32 && SRCB == kSrcPlanes && BM == 128 && BN == 128 && NST == 2) ? 4 : 1;
    static constexpr bool kSplit = (NPL >= 2);     // 2 or 3 planes: split tier
};

// LDS a workgroup needs to run tiles of one configuration.
template <class CFG, int BM, int BN>
struct Gemm16Lds {
    __attribute__((aligned(1024))) char ring[CFG::SMEM];
    int scA[CFG::kSplit ? BM : 1];
    int scB[CFG::kSplit ? BN : 1];
    int emax[BM];
};

// One output tile C(i0:i0+BM, j0:j0+BN) of batch element bi of problem P, computed by
// the whole workgroup.  Ends with a barrier, so a persistent caller may start the next
// tile in the same LDS.
template <int NPL, int SRCA, int SRCB, int BM, int BN, int WM, int WN, int NST>
__device__ __forceinline__ void gemm16_tile(const GemmProb& P, int bi, int i0, int j0,
                                            char* smem, int* scA_s, int* scB_s, int* emax_s)
{
    using CFG = Gemm16Cfg<NPL, SRCA, SRCB, BM, BN, WM, WN, NST>;
    constexpr int NT = CFG::NT;
    using PA = typename CFG::PA;
    using PB = typename CFG::PB;
    constexpr int WTM = BM / WM, WTN = BN / WN;
    constexpr int FM = WTM / 32, FN = WTN / 32;
    static_assert(FM >= 1 && FN >= 1, "wave tile");
    constexpr int STAGE = CFG::STAGE;
    constexpr bool kSplit = CFG::kSplit;
    constexpr bool kLo = (NPL == 3);        // third plane and accumulator
    // direct-to-LDS instructions each wave issues per stage (the vmcnt of one stage)
    constexpr int DMA_W = (SRCA == kSrcPlanes ? PA::DMA_PER_WAVE : 0) +
                          (SRCB == kSrcPlanes ? PB::DMA_PER_WAVE : 0);
    static_assert(NST >= 2 && NST <= 4, "stages");
    // loads that may stay in flight when stage s+1 must be complete: those of stages
    // s+2 .. s+NST-1 (issued later, completing later: vmcnt retires in order)
    constexpr int KEEP = (NST - 2) * DMA_W;
    static_assert(KEEP <= 60, "vmcnt");

    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);   // wave-uniform (SGPR)
    const int M = P.m, N = P.n, K = P.k;

    int kend = K;
    if (P.modeA == kOpTriL) kend = min(K, i0 + BM);
    const int kblocks = (kend + kBK - 1) / kBK;
    const bool pairs = P.A2 != nullptr;
    const int nst = kblocks * (pairs ? 2 : 1);

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
    pa.setup(tid, P.modeA != kOpT);
    pb.setup(tid, P.modeB == kOpT);
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
    if constexpr (kSplit && SRCA == kSrcF32) {
#pragma unroll
        for (int q = 0; q < PA::ITEMS; ++q) pa.sc[q] = scA_s[pa.row[q]];
    }
    if constexpr (kSplit && SRCB == kSrcF32) {
#pragma unroll
        for (int q = 0; q < PB::ITEMS; ++q) pb.sc[q] = scB_s[pb.row[q]];
    }

    auto stage_k0 = [&](int s, int& pair) {
        pair = s >= kblocks;
        return (s - pair * kblocks) * kBK;
    };
    // FP32 operands of stage s into register stage `buf` (an integral constant)
    auto load_regs = [&](auto buf, int s) {
        constexpr int B = decltype(buf)::value;
        int pair;
        const int k0 = stage_k0(s, pair);
        if constexpr (SRCA == kSrcF32)
            pa.template load_f32<B>(pair ? A2f : Af, P.lda, P.modeA, true, i0, k0, M, K, vecA);
        if constexpr (SRCB == kSrcF32)
            pb.template load_f32<B>(pair ? B2f : Bf, P.ldb, P.modeB, false, j0, k0, N, K, vecB);
    };
    auto store_regs = [&](auto buf, int s) {
        constexpr int B = decltype(buf)::value;
        char* base = smem + (s % NST) * STAGE;
        pa.template store_f32<B>(base);
        pb.template store_f32<B>(base + PA::BYTES);
    };
    // plane operands of stage s straight into LDS (sources re-derived when the second
    // product pair starts)
    const unsigned smem_lds = (unsigned)(uintptr_t)(lds_void_t*)smem;
    int dma_pair = -1;
    auto issue_dma = [&](int s) {
        int pair;
        const int k0 = stage_k0(s, pair);
        if (pair != dma_pair) {
            pa.dma_setup(pair ? A2h : Ah, P.lda, P.spA, P.kofsA, i0, M, wave, lane);
            pb.dma_setup(pair ? B2h : Bh, P.ldb, P.spB, P.kofsB, j0, N, wave, lane);
            dma_pair = pair;
        }
        const unsigned base = smem_lds + (unsigned)((s % NST) * STAGE);
        pa.dma(k0, base);
        pb.dma(k0, base + PA::BYTES);
    };

    const int wm = wave % WM, wn = wave / WM;
    constexpr int NACC = FM * FN;
    floatx16_t acc0[NACC];
    floatx16_t acc1[kLo ? NACC : 1];
    floatx16_t acc2[kLo ? NACC : 1];
#pragma unroll
    for (int q = 0; q < NACC; ++q) acc0[q] = floatx16_t{};
#pragma unroll
    for (int q = 0; q < (kLo ? NACC : 1); ++q) acc1[q] = floatx16_t{};
#pragma unroll
    for (int q = 0; q < (kLo ? NACC : 1); ++q) acc2[q] = floatx16_t{};

    // fragment addressing: lane reads row (fragment base + lane%32), 16-byte unit
    // kk/8 + lane/32 (Panel16::lds_off applies the layout's swizzle)
    const int frow = lane & 31, fu = lane >> 5;
    // fragments of the 16-wide k half `kk` of stage s, then their MFMAs
    struct Frags {
        half8_t a[FM][NPL], b[FN][NPL];
    };
    auto load_frags = [&](int s, int kk, Frags& f) {
        const char* sA = smem + (s % NST) * STAGE;
        const char* sB = sA + PA::BYTES;
        // fragment rows are 32-aligned bases + frow: the swizzles only see frow's bits
        const int u = (kk >> 3) + fu;
#pragma unroll
        for (int a = 0; a < FM; ++a) {
            const int r = wm * WTM + a * 32 + frow;
#pragma unroll
            for (int p = 0; p < NPL; ++p)
                f.a[a][p] = *reinterpret_cast<const half8_t*>(sA + PA::lds_off(r, p, u));
        }
#pragma unroll
        for (int b = 0; b < FN; ++b) {
            const int r = wn * WTN + b * 32 + frow;
#pragma unroll
            for (int p = 0; p < NPL; ++p)
                f.b[b][p] = *reinterpret_cast<const half8_t*>(sB + PB::lds_off(r, p, u));
        }
    };
    // dependent MFMAs on one accumulator are NACC issues apart
    auto mma = [&](const Frags& f) {
#pragma unroll
        for (int q = 0; q < NACC; ++q)
            acc0[q] = __builtin_amdgcn_mfma_f32_32x32x16_f16(f.b[q / FM][0], f.a[q % FM][0], acc0[q], 0, 0, 0);
        if constexpr (kLo) {
#pragma unroll
            for (int q = 0; q < NACC; ++q)
                acc1[q] = __builtin_amdgcn_mfma_f32_32x32x16_f16(f.b[q / FM][0], f.a[q % FM][1], acc1[q], 0, 0, 0);
#pragma unroll
            for (int q = 0; q < NACC; ++q)
                acc2[q] = __builtin_amdgcn_mfma_f32_32x32x16_f16(f.b[q / FM][0], f.a[q % FM][2], acc2[q], 0, 0, 0);
#pragma unroll
            for (int q = 0; q < NACC; ++q)
                acc1[q] = __builtin_amdgcn_mfma_f32_32x32x16_f16(f.b[q / FM][1], f.a[q % FM][0], acc1[q], 0, 0, 0);
#pragma unroll
            for (int q = 0; q < NACC; ++q)
                acc2[q] = __builtin_amdgcn_mfma_f32_32x32x16_f16(f.b[q / FM][1], f.a[q % FM][1], acc2[q], 0, 0, 0);
#pragma unroll
            for (int q = 0; q < NACC; ++q)
                acc2[q] = __builtin_amdgcn_mfma_f32_32x32x16_f16(f.b[q / FM][2], f.a[q % FM][0], acc2[q], 0, 0, 0);
        } else if constexpr (kSplit) {
            // two planes (m stored at its own scale, 2^-11 below h): hm + mh into acc0
            // beside hh (the 2^-22 product mm is not formed)
#pragma unroll
            for (int q = 0; q < NACC; ++q)
                acc0[q] = __builtin_amdgcn_mfma_f32_32x32x16_f16(f.b[q / FM][0], f.a[q % FM][1], acc0[q], 0, 0, 0);
#pragma unroll
            for (int q = 0; q < NACC; ++q)
                acc0[q] = __builtin_amdgcn_mfma_f32_32x32x16_f16(f.b[q / FM][1], f.a[q % FM][0], acc0[q], 0, 0, 0);
        }
    };

    // ------------------------------------------------------------------ pipeline
    // Stage s lives in LDS slot s % NST.  Iteration s: FP32 loads of stage s+D (D = 2,
    // into register stage s % 2; D = 1 with three planes), direct-to-LDS loads of s+NST-1 (between the two MFMA halves;
    // after the FP32 stores when there are FP32 operands; first thing with two stages),
    // MFMAs of s, FP32 stores of stage s+1 from its register stage, then wait until
    // only the loads of stages past s+1 are outstanding and barrier: stage s+1 is
    // complete in LDS for every wave, and slot s % NST is free to be reloaded.
    // With FP32 operands the register stores wait for every load issued before the
    // newest register loads (the compiler cannot count the direct-to-LDS loads), so the
    // direct loads go out after the stores and have the whole next iteration to land,
    // and the register loads, issued D iterations ahead, D.
    constexpr bool kRegs = SRCA == kSrcF32 || SRCB == kSrcF32;
    constexpr int D = NPL < 3 ? 2 : 1;
    using RS0 = std::integral_constant<int, 0>;
    using RS1 = std::integral_constant<int, 1>;
    auto iteration = [&](auto rs, int s) __attribute__((always_inline)) {
        constexpr int R = decltype(rs)::value;                  // register stage of s+D
        if constexpr (kRegs)
            if (s + D < nst) load_regs(std::integral_constant<int, D == 2 ? R : 0>{}, s + D);
        const bool ahead = s + NST - 1 < nst;
        // two stages: the other slot is free from the barrier on, so its loads go out
        // first and have the whole stage to land
        if constexpr (NST == 2)
            if (ahead) issue_dma(s + 1);
        // both halves' fragment reads go out before the first MFMA, except where the
        // two halves' fragments and the accumulators would not fit the registers
        // together (2 planes on 128 x 64 wave tiles): there the second half's reads
        // follow the first half's MFMAs
        constexpr bool kSeq = kSplit && WTM * WTN >= 128 * 64;
        Frags f0, f1;
        load_frags(s, 0, f0);
        if constexpr (!kSeq) load_frags(s, 16, f1);
        // split tier: the wave in its MFMA block takes issue priority over the SIMD's
        // other wave (+2-3 % at 8192^3; the plain tier measured 3 % slower with it)
        if constexpr (kSplit) __builtin_amdgcn_s_setprio(1);
        mma(f0);
        if constexpr (kSplit) __builtin_amdgcn_s_setprio(0);
        if constexpr (NST > 2 && !kRegs)
            if (ahead) issue_dma(s + NST - 1);
        if constexpr (kSeq) load_frags(s, 16, f1);
        if constexpr (kSplit) __builtin_amdgcn_s_setprio(1);
        mma(f1);
        if constexpr (kSplit) __builtin_amdgcn_s_setprio(0);
        if constexpr (kRegs)
            if (s + 1 < nst) store_regs(std::integral_constant<int, D == 2 ? 1 - R : 0>{}, s + 1);
        if constexpr (NST > 2 && kRegs)
            if (ahead) issue_dma(s + NST - 1);
        if (KEEP > 0 && ahead) TB_WAIT_VM_BARRIER(KEEP);
        else TB_WAIT_VM_BARRIER(0);
    };

    // Four waves (one per SIMD) of 128 x 128 accumulators on a 256 x 256 tile of plane
    // operands.  With one wave per SIMD nothing else covers a gap in its MFMA stream,
    // so each k half's MFMAs are issued one at a time with the other work placed
    // between them: the other half's fragment reads, and after the stage's one barrier
    // (between the two halves) the direct-to-LDS loads of stage s+NST into stage s's
    // slot.  Every slot gets a stage, the last one again past the end, so every barrier
    // waits for the same count: stage s+1 complete, NST-2 later stages in flight.  The
    // MFMAs of each accumulator run in mma()'s order, so the bits equal the other tiles'.
    constexpr bool kW4 = !kRegs && NT == 256 && BM == 256 && BN == 256;
    if constexpr (kW4) {
        static_assert(!kLo, "one or two planes");
        constexpr int NPROD = kSplit ? 3 : 1;             // hh (, hm, mh)
        constexpr int NM = NPROD * NACC;                  // MFMAs of one k half
        constexpr int NRA = FM * NPL, NRB = FN * NPL;     // fragment reads of one k half
        constexpr int NRD = NRA + NRB;
        static_assert(NRD + DMA_W <= NM && 4 + NRD <= NM, "interleave");
        // one MFMA / one fragment read of the flat lists, the index a compile-time
        // constant (static_for below), so that the accumulators stay in registers
        auto mfma1 = [&](auto jc, const Frags& f) __attribute__((always_inline)) {
            constexpr int j = decltype(jc)::value;
            constexpr int pr = j / NACC, q = j % NACC;
            constexpr int ua = pr == 1 ? 1 : 0, ub = pr == 2 ? 1 : 0;   // hh, hm, mh
            acc0[q] = __builtin_amdgcn_mfma_f32_32x32x16_f16(f.b[q / FM][ub], f.a[q % FM][ua], acc0[q], 0, 0, 0);
        };
        auto read1 = [&](auto ic, int s, int kk, Frags& f) __attribute__((always_inline)) {
            constexpr int i = decltype(ic)::value;
            const char* sA = smem + (s % NST) * STAGE;
            const int u = (kk >> 3) + fu;
            if constexpr (i < NRA) {
                constexpr int a = i / NPL, p = i % NPL;
                f.a[a][p] = *reinterpret_cast<const half8_t*>(sA + PA::lds_off(wm * WTM + a * 32 + frow, p, u));
            } else {
                constexpr int b = (i - NRA) / NPL, p = (i - NRA) % NPL;
                f.b[b][p] = *reinterpret_cast<const half8_t*>(sA + PA::BYTES + PB::lds_off(wn * WTN + b * 32 + frow, p, u));
            }
        };
        // stage s's loads into LDS slot `slot` (sources re-derived when the second
        // product pair starts)
        auto setup_for = [&](int s, int& k0) {
            int pair;
            k0 = stage_k0(s, pair);
            if (pair != dma_pair) {
                pa.dma_setup(pair ? A2h : Ah, P.lda, P.spA, P.kofsA, i0, M, wave, lane);
                pb.dma_setup(pair ? B2h : Bh, P.ldb, P.spB, P.kofsB, j0, N, wave, lane);
                dma_pair = pair;
            }
        };
        if (nst > 0) {
#pragma unroll
            for (int q = 0; q < NST; ++q) {
                int k0;
                setup_for(q < nst ? q : nst - 1, k0);
                const unsigned base = smem_lds + (unsigned)(q * STAGE);
                pa.dma(k0, base);
                pb.dma(k0, base + PA::BYTES);
            }
            TB_WAIT_VM_BARRIER((NST - 1) * DMA_W);
            Frags fa, fb;
            load_frags(0, 0, fa);
#pragma unroll 1
            for (int s = 0; s < nst; ++s) {
                // half 0 (fa); half 1's fragments (fb) read between its MFMAs
                static_for<NM>([&](auto jc) __attribute__((always_inline)) {
                    constexpr int j = decltype(jc)::value;
                    mfma1(jc, fa);
                    if constexpr (j >= 4 && j < 4 + NRD) read1(std::integral_constant<int, j - 4>{}, s, 16, fb);
                    __builtin_amdgcn_sched_barrier(0);
                });
                // stage s+1 complete for every wave, stage s read by every wave
                TB_WAIT_VM_BARRIER(KEEP);
                __builtin_amdgcn_sched_barrier(0);
                // half 1 (fb); the next stage's half-0 fragments and stage s+NST's
                // loads between its MFMAs
                int k0;
                setup_for(s + NST < nst ? s + NST : nst - 1, k0);
                const unsigned base = smem_lds + (unsigned)((s % NST) * STAGE);
                const int s1 = s + 1 < nst ? s + 1 : s;
                static_for<NM>([&](auto jc) __attribute__((always_inline)) {
                    constexpr int j = decltype(jc)::value;
                    mfma1(jc, fb);
                    if constexpr (j < NRD) {
                        read1(jc, s1, 0, fa);
                    } else if constexpr (j < NRD + DMA_W) {
                        constexpr int q = j - NRD;
                        if constexpr (q < PA::DMA_PER_WAVE) pa.dma_q(q, k0, base);
                        else pb.dma_q(q - PA::DMA_PER_WAVE, k0, base + PA::BYTES);
                    }
                    __builtin_amdgcn_sched_barrier(0);
                });
            }
            // the loads past the end land in slots nothing reads, but must land before
            // the workgroup's LDS is handed on
            asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
        }
    } else if (CFG::kLeafAll && nst > 0 && nst <= CFG::kLeafStages) {
        if constexpr (CFG::kLeafAll) {
            static_for<CFG::kLeafStages>([&](auto sc) __attribute__((always_inline)) {
                if (decltype(sc)::value < nst) load_regs(sc, decltype(sc)::value);
            });
#pragma unroll
            for (int s = 0; s < NST - 1; ++s)
                if (s < nst) issue_dma(s);
            store_regs(RS0{}, 0);
            // stage 0 complete (and every register load, issued before the stages)
            const int issued = nst < NST - 1 ? nst : NST - 1;
            if (issued == NST - 1) TB_WAIT_VM_BARRIER(KEEP);
            else if (issued == 3) TB_WAIT_VM_BARRIER(2 * DMA_W);
            else if (issued == 2) TB_WAIT_VM_BARRIER(DMA_W);
            else TB_WAIT_VM_BARRIER(0);
            static_for<CFG::kLeafStages>([&](auto sc) __attribute__((always_inline)) {
                constexpr int s = decltype(sc)::value;
                if (s < nst) {
                    const bool ahead = s + NST - 1 < nst;
                    if constexpr (NST == 2)
                        if (ahead) issue_dma(s + 1);
                    Frags f0, f1;
                    load_frags(s, 0, f0);
                    load_frags(s, 16, f1);
                    if constexpr (kSplit) __builtin_amdgcn_s_setprio(1);
                    mma(f0);
                    mma(f1);
                    if constexpr (kSplit) __builtin_amdgcn_s_setprio(0);
                    if constexpr (s + 1 < CFG::kLeafStages)
                        if (s + 1 < nst) store_regs(std::integral_constant<int, s + 1>{}, s + 1);
                    if constexpr (NST > 2)
                        if (ahead) issue_dma(s + NST - 1);
                    if (KEEP > 0 && ahead) TB_WAIT_VM_BARRIER(KEEP);
                    else TB_WAIT_VM_BARRIER(0);
                }
            });
        }
    } else if (nst > 0) {
        if constexpr (kRegs) {
            load_regs(RS0{}, 0);
            if (D == 2 && nst > 1) load_regs(RS1{}, 1);
        }
#pragma unroll
        for (int s = 0; s < NST - 1; ++s)
            if (s < nst) issue_dma(s);
        if constexpr (kRegs) store_regs(RS0{}, 0);
        // stage 0 complete: only the other issued stages (1 .. issued-1) may still be
        // in flight, DMA_W loads each
        const int issued = nst < NST - 1 ? nst : NST - 1;
        if (issued == NST - 1) TB_WAIT_VM_BARRIER(KEEP);
        else if (issued == 3) TB_WAIT_VM_BARRIER(2 * DMA_W);
        else if (issued == 2) TB_WAIT_VM_BARRIER(DMA_W);
        else TB_WAIT_VM_BARRIER(0);
        if constexpr (kRegs && D == 2) {
            // register stages alternate: two iterations per trip keep their indices
            // compile-time constants (stage s+D goes to register stage s % 2)
#pragma unroll 1
            for (int s = 0; s < nst; s += 2) {
                iteration(RS0{}, s);
                if (s + 1 < nst) iteration(RS1{}, s + 1);
            }
        } else {
#pragma unroll 1
            for (int s = 0; s < nst; ++s) iteration(RS0{}, s);
        }
    }

    // ----------------------------------------------------------------- epilogue
    const float alpha = P.alpha32, beta = P.beta32;
    const bool lowerC = (P.modeC == kCLower);
    const bool track = (P.eOut != nullptr);
    const long long ldc = P.ldc;
    // TRSM leaf tiles read every C value and column divisor of the lane before the first
    // is used (the registers of op(A)'s stages are free by now): one memory latency
    // instead of one per group
    constexpr bool kPre = CFG::kLeafAll;
    float pcold[kPre ? FM : 1][kPre ? FN : 1][16];
    float pcd[kPre ? FN : 1][16];
    if constexpr (kPre) {
        const float* cdiv = static_cast<const float*>(P.cdiv);
        static_for<FM>([&](auto ac) __attribute__((always_inline)) {
            constexpr int a = decltype(ac)::value;
            const int i = i0 + wm * WTM + a * 32 + (lane & 31);
            static_for<FN>([&](auto bc) __attribute__((always_inline)) {
                constexpr int b = decltype(bc)::value;
                static_for<16>([&](auto rc) __attribute__((always_inline)) {
                    constexpr int r = decltype(rc)::value;
                    const int j = j0 + wn * WTN + b * 32 + (r >> 2) * 8 + (lane >> 5) * 4 + (r & 3);
                    const bool in = i < M && j < N && (!lowerC || i >= j);
                    pcold[a][b][r] = masked_load(Cb + (in ? (long long)i + (long long)j * ldc : 0ll),
                                                 in && beta != 0.0f);
                    if constexpr (a == 0)
                        pcd[b][r] = cdiv ? masked_load(cdiv + (j < N ? j : 0), j < N) : 1.0f;
                });
            });
        });
    }
    // every index into the accumulators is a compile-time constant (static_for), so that
    // they stay in registers
    static_for<FM>([&](auto ac) __attribute__((always_inline)) {
        constexpr int a = decltype(ac)::value;
        const int il = wm * WTM + a * 32 + (lane & 31);
        const int i = i0 + il;
        int emax = kExpZeroDev;
        const int scA = kSplit ? scA_s[il] : 0;
        static_for<FN>([&](auto bc) __attribute__((always_inline)) {
            constexpr int b = decltype(bc)::value;
            constexpr int q = b * FM + a;
            // each group's reads of C (and of the column scales) are all issued before
            // the first is used: behind per-element branches they went out one memory
            // latency apart.  Groups of 8; one at a time with three accumulator sets live,
            // whose registers leave no room for a group (the 3-plane batch kernels spill
            // even at 4).
            constexpr int G = kLo ? 1 : 8;
            static_for<16 / G>([&](auto hc) __attribute__((always_inline)) {
                constexpr int hf = decltype(hc)::value;
                bool in[G];
                float cold[G];
                float cd[G];
                const float* cdiv = static_cast<const float*>(P.cdiv);
                static_for<G>([&](auto uc) __attribute__((always_inline)) {
                    constexpr int u = decltype(uc)::value;
                    constexpr int r = hf * G + u;
                    const int j = j0 + wn * WTN + b * 32 + (r >> 2) * 8 + (lane >> 5) * 4 + (r & 3);
                    in[u] = i < M && j < N && (!lowerC || i >= j);
                    if constexpr (kPre) {
                        cold[u] = pcold[a][b][r];
                        cd[u] = pcd[b][r];
                    } else {
                        cold[u] = masked_load(Cb + (in[u] ? (long long)i + (long long)j * ldc : 0ll),
                                              in[u] && beta != 0.0f);
                        cd[u] = cdiv ? masked_load(cdiv + (in[u] ? j : 0), in[u]) : 1.0f;
                    }
                });
                static_for<G>([&](auto uc) __attribute__((always_inline)) {
                    constexpr int u = decltype(uc)::value;
                    constexpr int r = hf * G + u;
                    const int jl = wn * WTN + b * 32 + (r >> 2) * 8 + (lane >> 5) * 4 + (r & 3);
                    const int j = j0 + jl;
                    float v = acc0[q][r];
                    if constexpr (kLo) {
                        const float t = __builtin_fmaf(acc2[q][r], 1.0f / 2048.0f, acc1[q][r]);
                        v = __builtin_ldexpf(__builtin_fmaf(t, 1.0f / 2048.0f, v), -(scA + scB_s[jl]));
                    } else if constexpr (kSplit) {
                        v = __builtin_ldexpf(v, -(scA + scB_s[jl]));
                    }
                    if (in[u]) {
                        float* c = Cb + (long long)i + (long long)j * ldc;
                        float w;
                        if (cdiv) {
                            w = div_sum_f32(alpha * v, beta == 0.0f ? 0.0f : beta * cold[u], cd[u]);
                        } else {
                            w = (beta == 0.0f) ? alpha * v : alpha * v + beta * cold[u];
                        }
                        *c = w;
                        if (track) {
                            const uint32_t bits = __float_as_uint(w) & 0x7fffffffu;
                            const int e = bits == 0 ? kExpZeroDev : (int)(bits >> 23) - 126;
                            emax = e > emax ? e : emax;
                        }
                    }
                });
            });
        });
        if (track && emax != kExpZeroDev) atomicMax(&emax_s[il], emax);
    });
    if (track) {
        __syncthreads();
        for (int r = tid; r < BM; r += NT)
            if (emax_s[r] != kExpZeroDev && i0 + r < M)
                atomicMax(P.eOut + (long long)bi * P.seA + i0 + r, emax_s[r]);
    }
    __syncthreads();
}

template <int NPL, int SRCA, int SRCB, int BM, int BN, int WM, int WN, int NST>
__global__ void __launch_bounds__(64 * WM * WN, (Gemm16Cfg<NPL, SRCA, SRCB, BM, BN, WM, WN, NST>::kWavesPerEU))
gemm16_kernel(GemmLaunch L)
{
    using CFG = Gemm16Cfg<NPL, SRCA, SRCB, BM, BN, WM, WN, NST>;
    __shared__ Gemm16Lds<CFG, BM, BN> lds;
    // workgroup b: tiles b, b + gridDim.x, ... (a capped grid is a multiple of 8, so
    // xcd_remap keeps every workgroup on its XCD's tile range).  Wide tiles (BM = 256)
    // are never capped and run one tile per workgroup: their registers are full.
    constexpr bool kLoop = BM < 256;
    auto tile = [&](int b) {
        const int g = xcd_remap(b, L.totalTiles);
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
        // ends with a barrier: the next tile may reuse the LDS ring
        gemm16_tile<NPL, SRCA, SRCB, BM, BN, WM, WN, NST>(P, bi, ti * BM, tj * BN, lds.ring,
                                                          lds.scA, lds.scB, lds.emax);
    };
    if constexpr (kLoop) {
#pragma unroll 1
        for (int b = blockIdx.x; b < L.totalTiles; b += gridDim.x) tile(b);
    } else {
        tile(blockIdx.x);
    }
}

}  // namespace tb
