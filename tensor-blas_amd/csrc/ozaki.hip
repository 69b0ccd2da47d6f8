// Ozaki-scheme FP64 GEMM on INT8 MFMA (the tc_ozimmu_* tier).
//
// Published method: Ozaki, Ogita, Oishi, Rump (2012) error-free splitting, in the
// INT8 form of Ootomo, Ozaki, Yokota (2024).  Each row of op(A) (column of op(B)) is
// scaled by 2^-e (e = frexp exponent of its max |x|) and cut into S signed 7-bit
// digits:  x = 2^e * sum_s d_s 2^(-7(s+1)) + tail,  d_s = trunc(r_s * 2^7),
// r_{s+1} = r_s * 2^7 - d_s.  Then
//   C(i,j) = 2^(ea_i+eb_j) * sum_g 2^(-7(g+2)) * D_g(i,j),
//   D_g = sum_{s+t=g} A_s * B_t^T     (exact int32 sums, g = 0..S-1)
// The S(S+1)/2 slice-pair products run on v_mfma_i32_32x32x32_i8; all pairs of one g
// share an int32 accumulator (exact while S*kchunk*127^2 < 2^31, K is processed in
// chunks), and each chunk's D_g are added to an FP64 accumulator in increasing g.
// Optionally (G = S + 1 groups) the next anti-diagonal g = S is kept too: S - 1 more
// products on the same slices, which removes the largest dropped term.
// The operation order is exactly oracle/oracle.c orc_ozaki_gemm's, so the result
// is reproducible bit for bit.
#include <hip/hip_runtime.h>

#include <cstdint>

#include "gemm_kernel.hpp"
#include "tb_kernels.hpp"
#include "ozaki_tiles.hpp"

#pragma clang fp contract(off)

namespace tb {

namespace {

typedef int intx4_t __attribute__((ext_vector_type(4)));
typedef int intx16_t __attribute__((ext_vector_type(16)));

constexpr int kSliceTile = 64;
constexpr int kOzBlockK = 32;        // k per block of the slice layout (= the GEMM's stage)

// Slices of a 64 x 64 block of op(X) rows: loaded through LDS so that both the
// source read and the int8 write are coalesced.
__global__ void __launch_bounds__(256)
ozaki_slice_kernel(const double* __restrict__ X, long long ld, int mode, int rows, int K, int kpad,
                   int S, long long sx, const int* __restrict__ exps, int8_t* __restrict__ out,
                   int rblocks)
{
    __shared__ double t[kSliceTile][kSliceTile + 1];   // t[row][k]
    const int q = blockIdx.z;
    const double* Xq = X + (long long)q * sx;
    const int r0 = blockIdx.y * kSliceTile, k0 = blockIdx.x * kSliceTile;
    const int tid = threadIdx.x;
    // load: mode N / TriL / SymL-lower read rows contiguous, mode T reads k contiguous.
    // Every thread issues its 16 loads (unconditionally, at clamped addresses, masked
    // afterwards) before the first LDS store, so they are in flight together.  With
    // K = 0 the operand may be unreferenced and no load is issued.
    constexpr int kPer = kSliceTile * kSliceTile / 256;
    static_assert(kSliceTile == 64, "element e = tid + 256 it sits at (e % 64, e / 64)");
    const bool tr = mode == kOpT;
    const int ea = tid & 63, eb = tid >> 6;
    double v[kPer];
#pragma unroll
    for (int it = 0; it < kPer; ++it) v[it] = 0.0;
    if (K > 0) {
#pragma unroll
        for (int it = 0; it < kPer; ++it) {
            const int rl = tr ? eb + 4 * it : ea, kl = tr ? ea : eb + 4 * it;
            const int r = r0 + rl, k = k0 + kl;
            // X(r, k) stored (N, or the lower triangle of TriL / SymL), else X(k, r)
            const bool direct = mode == kOpN || (!tr && r >= k);
            const bool in = r < rows && k < K && (direct || tr || mode == kOpSymL);
            const long long off = direct ? (long long)r + (long long)k * ld
                                         : (long long)k + (long long)r * ld;
            v[it] = masked_load(Xq + (in ? off : 0ll), in);
        }
    }
#pragma unroll
    for (int it = 0; it < kPer; ++it) {
        const int rl = tr ? eb + 4 * it : ea, kl = tr ? ea : eb + 4 * it;
        t[rl][kl] = v[it];
    }
    __syncthreads();
    // each thread: one row, 16 consecutive k; rows past `rows` (up to the next multiple
    // of 64) get zero digits
    const int rl = tid >> 2, kl0 = (tid & 3) * 16;
    const int r = r0 + rl;
    if (k0 + kl0 >= kpad) return;
    const int e = r < rows ? exps[(long long)q * rows + r] : 0;
    const int sh = (e < -5000) ? 0 : e;   // all-zero row: digits are zero anyway
    double rem[16];
#pragma unroll
    for (int j = 0; j < 16; ++j) rem[j] = __builtin_ldexp(t[rl][kl0 + j], -sh);
    for (int s = 0; s < S; ++s) {
        intx4_t pk;
#pragma unroll
        for (int w = 0; w < 4; ++w) {
            uint32_t word = 0;
#pragma unroll
            for (int b = 0; b < 4; ++b) {
                const int j = w * 4 + b;
                const double tt = rem[j] * 128.0;
                const double d = __builtin_trunc(tt);
                rem[j] = tt - d;
                word |= (uint32_t)(uint8_t)(int8_t)(int)d << (8 * b);
            }
            pk[w] = (int)word;
        }
        // panel-blocked layout: 32 k of all S slices of one 64-row block are contiguous
        const int k = k0 + kl0;
        int8_t* dst = out + (long long)q * S * rblocks * kSliceTile * kpad +
                      (((long long)(k / kOzBlockK) * rblocks + blockIdx.y) * S + s) * (kSliceTile * kOzBlockK) +
                      rl * kOzBlockK + k % kOzBlockK;
        *reinterpret_cast<intx4_t*>(dst) = pk;
    }
}

constexpr int OBM = 64, OBN = 64;   // C tile
constexpr int OBK = kOzBlockK;       // k per stage: one 32-byte LDS row of int8 per slice row
constexpr int kMinSlices = 3, kMaxSlices = 10;

// LDS ring depth for S slices: as many stages (<= 4) of 2*S*64 rows x 32 B as fit
// below 160 KB
template <int S>
constexpr int oz_stages()
{
    return 4 * (2 * S * OBM * OBK) < 160 * 1024 ? 4 : 3;
}

// direct-to-LDS load of 16 bytes per lane (see gemm16_kernel.hpp lds_dma16)
__device__ __forceinline__ void oz_dma16(const void* g, unsigned lds_base)
{
    asm volatile("global_load_lds_dwordx4 %1, off"
                 :
                 : "{m0}"(__builtin_amdgcn_readfirstlane(lds_base)), "v"(g)
                 : "memory");
}
// "s_waitcnt vmcnt(n)" (this wave's direct-to-LDS loads but the newest n are done),
// lgkmcnt(0), then a workgroup barrier
#define TB_OZ_WAIT_BARRIER(n) \
    asm volatile("s_waitcnt vmcnt(%0) lgkmcnt(0)\n\ts_barrier" ::"i"(n) : "memory")


// One 64 x 64 tile of C per workgroup; each wave owns a 32 x 32 block.  A stage holds
// 32 k of all S slices of both 64-row panels (S * 4 KB) in LDS as 32-byte rows whose
// two 16-byte units swap places in every other group of 8 rows (unit ^ ((r >> 3) & 1)),
// so both the fragment reads (16 lanes: 16 rows of one unit) and the 1 KB direct-to-LDS
// writes cover the 64 banks once.  The ring holds NST stages; each stage's loads are
// issued NST-1 stages ahead of its MFMAs.
template <int S, int G>
__global__ void __launch_bounds__(256)
ozaki_gemm_kernel(OzakiProb P, int tilesM, int totalTiles)
{
    constexpr int NST = oz_stages<S>();
    constexpr int PLANE = OBM * OBK;                 // bytes of one slice of one panel
    constexpr int STAGE = 2 * S * PLANE;             // A slices, then B slices
    static_assert(NST * STAGE <= 160 * 1024, "LDS");
    // direct-to-LDS instructions per wave per stage (1 KB each): the vmcnt of a stage
    constexpr int DMA_W = STAGE / 1024 / 4;
    constexpr int KEEP = (NST - 2) * DMA_W;
    static_assert(DMA_W == S && KEEP <= 60, "vmcnt");
    __shared__ __attribute__((aligned(1024))) int8_t smem[NST * STAGE];

    // workgroup blk: tiles blk, blk + gridDim.x, ... (a capped grid is a multiple of 8,
    // so xcd_remap keeps every workgroup on its XCD's tile range).  Each tile's k loop
    // ends with a barrier after the last LDS reads, so the next tile may refill the ring.
#pragma unroll 1
    for (int blk = blockIdx.x; blk < totalTiles; blk += gridDim.x) {
        const int g = xcd_remap(blk, totalTiles);
        int ti, tj;
        if (P.lowerC) {
            lower_tile_banded(g, tilesM, ti, tj);
        } else {
            // groups of kBand tile rows, column by column inside a group
            const int tilesN = totalTiles / tilesM;
            const int per = kBand * tilesN;
            const int first = (g / per) * kBand;
            const int gsz = min(tilesM - first, kBand);
            const int in = g - (g / per) * per;
            ti = first + in % gsz;
            tj = in / gsz;
        }
        const int i0 = ti * OBM, j0 = tj * OBN;
        const int tid = threadIdx.x, lane = tid & 63;
        const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
        const int wm = wave & 1, wn = wave >> 1;

        // this wave's direct-to-LDS loads: instruction q moves 32 rows (1 KB) of one slice of
        // one panel; 4*S instructions per stage, S per wave.  In the panel-blocked slice
        // layout a panel's stage (S slices x 64 rows x 32 k) is one contiguous block laid
        // out as in LDS, and stage kb of a panel is kb * (row blocks) * S * PLANE further.
        const int8_t* gsrc[DMA_W];
        bool isB[DMA_W];
        unsigned ldo[DMA_W];
        const long long kstepA = (long long)((P.m + OBM - 1) / OBM) * S * PLANE;
        const long long kstepB = (long long)((P.n + OBN - 1) / OBN) * S * PLANE;
        int pair_src = -1;
        const unsigned smem_lds = (unsigned)(uintptr_t)(__attribute__((address_space(3))) void*)smem;
        auto dma_setup = [&](int pair) {
            const int8_t* GA = pair ? P.sa2 : P.sa;
            const int8_t* GB = pair ? P.sb2 : P.sb;
#pragma unroll
            for (int q = 0; q < DMA_W; ++q) {
                const int inst = wave + q * 4;                  // 0 .. 4S-1
                const int panel = inst / (2 * S);               // 0: A, 1: B
                const int rest = inst % (2 * S);
                const int sl = rest / 2, blk = rest % 2;
                const int rl = blk * 32 + (lane >> 1);          // LDS row written by this lane
                const int u = (lane & 1) ^ ((rl >> 3) & 1);     // logical unit stored there
                const int tile = panel ? tj : ti;               // 64-row block of the operand
                gsrc[q] = (panel ? GB : GA) + ((long long)tile * S + sl) * PLANE + rl * OBK + u * 16;
                isB[q] = panel != 0;
                ldo[q] = (unsigned)((panel * S + sl) * PLANE + blk * 1024);
            }
        };
        const int npairs = P.sa2 ? 2 : 1;
        const int kblocks = P.kpad / OBK;
        const int nblk = npairs * kblocks;
        auto issue = [&](int b) {
            const int pair = b >= kblocks;
            if (pair != pair_src) {
                dma_setup(pair);
                pair_src = pair;
            }
            const int kb = b - pair * kblocks;
            const unsigned base = smem_lds + (unsigned)((b % NST) * STAGE);
#pragma unroll
            for (int q = 0; q < DMA_W; ++q) oz_dma16(gsrc[q] + kb * (isB[q] ? kstepB : kstepA), base + ldo[q]);
        };

        intx16_t acc[G];
        double facc[16];
#pragma unroll
        for (int r = 0; r < 16; ++r) facc[r] = 0.0;
#pragma unroll
        for (int q = 0; q < G; ++q) acc[q] = intx16_t{};

        const int cblocks = P.kchunk / OBK;          // k blocks per exactness chunk
        // fragment of lane: row (panel base + lane%32), logical unit lane/32
        const int fr = lane & 31;
        const int fo = fr * OBK + (((lane >> 5) ^ ((fr >> 3) & 1)) * 16);
        const int offA = wm * 32 * OBK + fo, offB = S * PLANE + wn * 32 * OBK + fo;

        if (nblk > 0) {
            for (int b = 0; b < NST - 1; ++b)
                if (b < nblk) issue(b);
            // stage 0 complete: only the other issued stages (DMA_W loads each) may be in flight
            const int issued = nblk < NST - 1 ? nblk : NST - 1;
            if (issued == 3) TB_OZ_WAIT_BARRIER(2 * DMA_W);
            else if (issued == 2) TB_OZ_WAIT_BARRIER(DMA_W);
            else TB_OZ_WAIT_BARRIER(0);
        }
#pragma unroll 1
        for (int b = 0; b < nblk; ++b) {
            const int8_t* st = smem + (b % NST) * STAGE;
            intx4_t af[S], bf[S];
#pragma unroll
            for (int s = 0; s < S; ++s) {
                af[s] = *reinterpret_cast<const intx4_t*>(st + s * PLANE + offA);
                bf[s] = *reinterpret_cast<const intx4_t*>(st + s * PLANE + offB);
            }
            // slot (b-1) % NST was read by every wave before the barrier that ended block b-1
            const bool ahead = b + NST - 1 < nblk;
            if (ahead) issue(b + NST - 1);
            // D_g += A_s B_{g-s} (s, g-s < S; g < G): consecutive MFMAs update different D_g
#pragma unroll
            for (int s = 0; s < S; ++s) {
#pragma unroll
                for (int gg = s; gg < G; ++gg)
                    if (gg - s < S)
                        acc[gg] = __builtin_amdgcn_mfma_i32_32x32x32_i8(bf[gg - s], af[s], acc[gg], 0, 0, 0);
            }
            const int kb = b % kblocks;
            if ((kb + 1) % cblocks == 0 || kb + 1 == kblocks) {
                // exactness chunk done: D_g into FP64, increasing g
#pragma unroll
                for (int gg = 0; gg < G; ++gg) {
#pragma unroll
                    for (int r = 0; r < 16; ++r)
                        facc[r] = facc[r] + __builtin_ldexp((double)acc[gg][r], -7 * (gg + 2));
                    acc[gg] = intx16_t{};
                }
            }
            // block b+1 complete in LDS (later stages may still be in flight)
            if (ahead) TB_OZ_WAIT_BARRIER(KEEP);
            else TB_OZ_WAIT_BARRIER(0);
        }

        // epilogue: D[jj][ii], ii = lane%32 runs along C's rows
        const int i = i0 + wm * 32 + (lane & 31);
        if (i < P.m) {
            const int ea = P.ea[i];
            const int sea = ea < -5000 ? 0 : ea;
#pragma unroll
            for (int r = 0; r < 16; ++r) {
                const int j = j0 + wn * 32 + (r >> 2) * 8 + (lane >> 5) * 4 + (r & 3);
                if (j >= P.n || (P.lowerC && i < j)) continue;
                const int eb = P.eb[j];
                const int seb = eb < -5000 ? 0 : eb;
                const double v = __builtin_ldexp(facc[r], sea + seb);
                double* c = P.C + (long long)i + (long long)j * P.ldc;
                *c = (P.beta == 0.0) ? P.alpha * v : __builtin_fma(P.alpha, v, P.beta * *c);
            }
        }
    }
}

// ---------------------------------------------------------------- wide tiles
// The same products on 128 x 64 tiles of C with eight waves (two per SIMD, each a 32 x 32
// block: one wave's stage bubbles -- fragment reads, the barrier -- are the other's MFMA
// time) for K within one exactness chunk (kpad <= kchunk), where no FP64 accumulation
// across chunks is kept in registers: the G int32 sums D_g are converted once, in
// increasing g, as the chunked kernel's single chunk is, so the results are the same
// bits.  Longer K and SYR2K's second product run as one launch per exactness chunk of
// each product, in the chunked kernel's order: the first stores its unscaled FP64 sums
// in C (partial = 1), middle ones add their D_g onto them (partial = 3), and the last
// adds its D_g and applies the scale, alpha and beta = 0 (partial = 2), so the launches
// give the chunked kernel's bits.  A stage holds
// 32 k of all S slices of the two 64-row blocks of A (adjacent in the panel-blocked
// layout: A's slices are laid out with an even number of row blocks) and of one 64-row
// block of B: 6 S KB, in the LDS layout above.  Per wave, the S A fragments stay in
// registers and the B fragments are read one slice at a time.
constexpr int OWM = 128;             // rows of a wide tile (two 64-row slice blocks)

template <int S>
constexpr int ozw_stages()
{
    return 4 * (6 * S * 1024) <= 160 * 1024 ? 4 : (3 * (6 * S * 1024) <= 160 * 1024 ? 3 : 2);
}


template <int S, int G>
__global__ void __launch_bounds__(512)
ozaki_gemm_wide_kernel(OzakiProb P, int tilesM, int tilesN, int totalTiles)
{
    constexpr int NST = ozw_stages<S>();
    constexpr int PLANE = OBM * OBK;                 // one slice of one 64-row block: 2 KB
    constexpr int A_BYTES = 2 * S * PLANE;
    constexpr int STAGE = 3 * S * PLANE;
    constexpr int NINST = STAGE / 1024;              // 1 KB direct-to-LDS instructions
    constexpr int DMA_W = (NINST + 7) / 8;           // per wave (the last ones repeat)
    constexpr int KEEP = (NST - 2) * DMA_W;
    static_assert(NST >= 2 && NST * STAGE <= 160 * 1024 && KEEP <= 60, "stages");
    __shared__ __attribute__((aligned(1024))) int8_t smem[NST * STAGE];

#pragma unroll 1
    for (int blk = blockIdx.x; blk < totalTiles; blk += gridDim.x) {
        const int g = xcd_remap(blk, totalTiles);
        int ti, tj;
        if (P.lowerC) {
            lower_wide_banded(g, tilesM, tilesN, ti, tj);
        } else {
            const int per = kBandW * tilesN;
            const int first = (g / per) * kBandW;
            const int gsz = min(tilesM - first, kBandW);
            const int in = g - (g / per) * per;
            ti = first + in % gsz;
            tj = in / gsz;
        }
        const int i0 = ti * OWM, j0 = tj * OBN;
        const int tid = threadIdx.x, lane = tid & 63;
        const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
        const int wm = wave & 3, wn = wave >> 2;

        // instruction inst: A (inst < 4S): row block inst / 2S, slice, 1 KB half; B: slice,
        // half.  Waves with fewer real instructions repeat their previous one (the same
        // bytes to the same place), so every wave counts DMA_W per stage.
        const int8_t* gsrc[DMA_W];
        bool isB[DMA_W];
        unsigned ldo[DMA_W];
        const long long rbA = (long long)tilesM * 2;                 // A's row blocks (even)
        const long long rbB = (long long)(P.n + OBN - 1) / OBN;
        const long long kstepA = rbA * S * PLANE, kstepB = rbB * S * PLANE;
        const unsigned smem_lds = (unsigned)(uintptr_t)(__attribute__((address_space(3))) void*)smem;
        {
            const int8_t* GA = P.sa;
            const int8_t* GB = P.sb;
#pragma unroll
            for (int q = 0; q < DMA_W; ++q) {
                int inst = wave + q * 8;
                if (inst >= NINST) inst -= 8;
                const bool b = inst >= 4 * S;
                const int rest = b ? inst - 4 * S : inst;
                const int rb = b ? 0 : rest / (2 * S);                  // A row block in the tile
                const int sl = (rest % (2 * S)) / 2, half = rest % 2;
                const int rl = half * 32 + (lane >> 1);                 // row within the 64-row block
                const int u = (lane & 1) ^ ((rl >> 3) & 1);             // logical unit stored there
                const long long blk64 = b ? tj : 2ll * ti + rb;         // 64-row block of the operand
                gsrc[q] = (b ? GB : GA) + (blk64 * S + sl) * PLANE + rl * OBK + u * 16;
                isB[q] = b;
                ldo[q] = (unsigned)((b ? A_BYTES + sl * PLANE : (rb * S + sl) * PLANE) + half * 1024);
            }
        }
        const int nblk = P.kpad / OBK;
        auto issue = [&](int kb) {
            const unsigned base = smem_lds + (unsigned)((kb % NST) * STAGE);
#pragma unroll
            for (int q = 0; q < DMA_W; ++q) oz_dma16(gsrc[q] + kb * (isB[q] ? kstepB : kstepA), base + ldo[q]);
        };

        intx16_t acc[G];
#pragma unroll
        for (int q = 0; q < G; ++q) acc[q] = intx16_t{};
        // fragment of lane: row (block base + lane%32), logical unit lane/32
        const int fr = lane & 31;
        const int fo = fr * OBK + (((lane >> 5) ^ ((fr >> 3) & 1)) * 16);
        const int offA = (wm >> 1) * S * PLANE + (wm & 1) * 32 * OBK + fo;
        const int offB = A_BYTES + wn * 32 * OBK + fo;

        if (nblk > 0) {
            for (int b = 0; b < NST - 1; ++b)
                if (b < nblk) issue(b);
            const int issued = nblk < NST - 1 ? nblk : NST - 1;
            if (issued == 3) TB_OZ_WAIT_BARRIER(2 * DMA_W);
            else if (issued == 2) TB_OZ_WAIT_BARRIER(DMA_W);
            else TB_OZ_WAIT_BARRIER(0);
        }
        // D_{s+t} += A_s B_t (s, t < S; s + t < G): consecutive MFMAs update different D_g;
        // int32 sums are exact, so their order does not matter
        auto mfma_t = [&](const int8_t* st, int t, const intx4_t(&a)[S]) __attribute__((always_inline)) {
            const intx4_t bt = *reinterpret_cast<const intx4_t*>(st + t * PLANE + offB);
#pragma unroll
            for (int s = 0; s < S; ++s)
                if (s + t < G)
                    acc[s + t] = __builtin_amdgcn_mfma_i32_32x32x32_i8(bt, a[s], acc[s + t], 0, 0, 0);
        };
        if constexpr (NST >= 3) {
            // The block's barrier sits halfway through its MFMAs: after it (block b+1
            // complete, block b-1 read by every wave) block b+NST-1's loads go into
            // block b-1's slot and block b+1's A fragments are read while the second
            // half's MFMAs run, so no wave waits for its A fragments after a barrier.
            // first half: the slices t < TH of B, chosen so that the halves hold about
            // the same number of MFMAs (slice t has min(S, G - t) of them)
            constexpr int TH = [] {
                int total = 0;
                for (int t = 0; t < S; ++t) total += (G - t < S ? G - t : S);
                int c = 0;
                for (int t = 0; t < S; ++t) {
                    c += (G - t < S ? G - t : S);
                    if (2 * c >= total) return t + 1;
                }
                return S;
            }();
            constexpr int KEEPM = (NST - 3) * DMA_W;
            intx4_t af[S], an[S];
            if (nblk > 0) {
#pragma unroll
                for (int s = 0; s < S; ++s) af[s] = *reinterpret_cast<const intx4_t*>(smem + s * PLANE + offA);
            }
#pragma unroll 1
            for (int b = 0; b < nblk; ++b) {
                const int8_t* st = smem + (b % NST) * STAGE;
#pragma unroll
                for (int t = 0; t < TH; ++t) mfma_t(st, t, af);
                // (the MFMAs stay on their side of the barrier and of the reads after it)
                __builtin_amdgcn_sched_barrier(0);
                // block b+1's loads went out after the barrier of block b+2-NST; those of
                // blocks b+3-NST .. b-1 may stay in flight when block b-1 issued its own
                if (KEEPM > 0 && b + NST - 2 < nblk) TB_OZ_WAIT_BARRIER(KEEPM);
                else TB_OZ_WAIT_BARRIER(0);
                if (b + NST - 1 < nblk) issue(b + NST - 1);
                // block b+1's A fragments (past the last block: block b's again, unused)
                const int8_t* sn = smem + ((b + 1 < nblk ? b + 1 : b) % NST) * STAGE;
#pragma unroll
                for (int s = 0; s < S; ++s) an[s] = *reinterpret_cast<const intx4_t*>(sn + s * PLANE + offA);
                __builtin_amdgcn_sched_barrier(0);
#pragma unroll
                for (int t = TH; t < S; ++t) mfma_t(st, t, af);
#pragma unroll
                for (int s = 0; s < S; ++s) af[s] = an[s];
            }
        } else {
#pragma unroll 1
            for (int b = 0; b < nblk; ++b) {
                const int8_t* st = smem + (b % NST) * STAGE;
                intx4_t af[S];
#pragma unroll
                for (int s = 0; s < S; ++s) af[s] = *reinterpret_cast<const intx4_t*>(st + s * PLANE + offA);
                // slot (b-1) % NST was read by every wave before the barrier that ended block b-1
                const bool ahead = b + NST - 1 < nblk;
                if (ahead) issue(b + NST - 1);
#pragma unroll
                for (int t = 0; t < S; ++t) mfma_t(st, t, af);
                if (ahead) TB_OZ_WAIT_BARRIER(KEEP);
                else TB_OZ_WAIT_BARRIER(0);
            }
        }

        // epilogue: the exactness chunk's D_g into FP64 in increasing g (on top of the
        // first product's sum for partial = 2: the chunked kernel's order), then the
        // scale, alpha and beta
        const int i = i0 + wm * 32 + (lane & 31);
        if (i < P.m) {
            const int ea = P.ea[i];
            const int sea = ea < -5000 ? 0 : ea;
#pragma unroll
            for (int r = 0; r < 16; ++r) {
                const int j = j0 + wn * 32 + (r >> 2) * 8 + (lane >> 5) * 4 + (r & 3);
                if (j >= P.n || (P.lowerC && i < j)) continue;
                double* c = P.C + (long long)i + (long long)j * P.ldc;
                double f = P.partial >= 2 ? *c : 0.0;
#pragma unroll
                for (int gg = 0; gg < G; ++gg) f = f + __builtin_ldexp((double)acc[gg][r], -7 * (gg + 2));
                if (P.partial == 1 || P.partial == 3) {
                    *c = f;
                    continue;
                }
                const int eb = P.eb[j];
                const int seb = eb < -5000 ? 0 : eb;
                const double v = __builtin_ldexp(f, sea + seb);
                *c = (P.beta == 0.0) ? P.alpha * v : __builtin_fma(P.alpha, v, P.beta * *c);
            }
        }
        __syncthreads();      // the next tile refills the ring
    }
}

}  // namespace

int ozaki_wide_rows() { return OWM; }

bool ozaki_wide_ok(int nslices, int ngroups)
{
    if (nslices < 3 || nslices > 10 || ngroups < nslices || ngroups > nslices + 1) return false;
    return ngroups == nslices || (nslices >= 6 && nslices <= 9);
}

hipError_t ozaki_slice(const double* X, long long ld, int mode, int rows, int K, int kpad,
                       int nslices, int nb, long long sx, const int* exps, int8_t* slices,
                       hipStream_t s, int rblocks)
{
    if (rows <= 0 || kpad <= 0 || nb <= 0) return hipSuccess;
    const int need = (rows + kSliceTile - 1) / kSliceTile;
    if (rblocks <= 0) rblocks = need;
    if (kpad % kSliceTile || nslices < kMinSlices || nslices > kMaxSlices || rblocks < need)
        return hipErrorInvalidValue;
    // every row block of the layout is written (rows past `rows` as zero digits)
    dim3 g(kpad / kSliceTile, rblocks, nb);
    hipLaunchKernelGGL(ozaki_slice_kernel, g, dim3(256), 0, s, X, ld, mode, rows, K, kpad, nslices,
                       sx, exps, slices, rblocks);
    return hipGetLastError();
}

hipError_t ozaki_gemm(const OzakiProb& p, hipStream_t s, int max_cus)
{
    if (p.m <= 0 || p.n <= 0) return hipSuccess;
    if (p.nslices < kMinSlices || p.nslices > kMaxSlices || p.kpad % OBK || p.kchunk < OBK ||
        p.kchunk % OBK || (p.ngroups != p.nslices && p.ngroups != p.nslices + 1))
        return hipErrorInvalidValue;
    if (p.lowerC && (p.m != p.n)) return hipErrorInvalidValue;
    if (p.wide) {
        // one exactness chunk, one product (pairs are two launches, the partial sum kept
        // in C); A's slices laid out with an even number of row blocks
        if (p.kpad > p.kchunk || p.sa2 || p.partial < 0 || p.partial > 3 ||
            (p.partial && p.beta != 0.0))
            return hipErrorInvalidValue;
        const int tm = (p.m + OWM - 1) / OWM, tn = (p.n + OBN - 1) / OBN;
        int total = 0;
        if (p.lowerC) {
            for (int t = 0; t < tm; ++t) total += min(2 * t + 2, tn);
        } else {
            total = tm * tn;
        }
        const int key = p.nslices * 2 + (p.ngroups - p.nslices);
        switch (key) {
#define TB_OZW_CASE(S_, E_)                                                                       \
    case S_ * 2 + E_:                                                                             \
        hipLaunchKernelGGL((ozaki_gemm_wide_kernel<S_, S_ + E_>),                                 \
                           dim3(capped_grid(reinterpret_cast<const void*>(ozaki_gemm_wide_kernel<S_, S_ + E_>), \
                                            512, total, max_cus)),                                \
                           dim3(512), 0, s, p, tm, tn, total);                                    \
        break;
            TB_OZW_CASE(3, 0) TB_OZW_CASE(4, 0) TB_OZW_CASE(5, 0) TB_OZW_CASE(6, 0) TB_OZW_CASE(7, 0)
            TB_OZW_CASE(8, 0) TB_OZW_CASE(9, 0) TB_OZW_CASE(10, 0)
            TB_OZW_CASE(6, 1) TB_OZW_CASE(7, 1) TB_OZW_CASE(8, 1) TB_OZW_CASE(9, 1)
#undef TB_OZW_CASE
        default: return hipErrorInvalidValue;
        }
        return hipGetLastError();
    }
    const int tm = (p.m + OBM - 1) / OBM, tn = (p.n + OBN - 1) / OBN;
    const int total = p.lowerC ? tm * (tm + 1) / 2 : tm * tn;
    const int key = p.nslices * 2 + (p.ngroups - p.nslices);
    switch (key) {
#define TB_OZ_CASE(S_, E_)                                                                        \
    case S_ * 2 + E_:                                                                             \
        hipLaunchKernelGGL((ozaki_gemm_kernel<S_, S_ + E_>),                                      \
                           dim3(capped_grid(reinterpret_cast<const void*>(ozaki_gemm_kernel<S_, S_ + E_>), \
                                            256, total, max_cus)),                                \
                           dim3(256), 0, s, p, tm, total);                                        \
        break;
        TB_OZ_CASE(3, 0) TB_OZ_CASE(4, 0) TB_OZ_CASE(5, 0) TB_OZ_CASE(6, 0) TB_OZ_CASE(7, 0)
        TB_OZ_CASE(8, 0) TB_OZ_CASE(9, 0) TB_OZ_CASE(10, 0)
        TB_OZ_CASE(6, 1) TB_OZ_CASE(7, 1) TB_OZ_CASE(8, 1) TB_OZ_CASE(9, 1)
#undef TB_OZ_CASE
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

}  // namespace tb
