// Grouped FP64-MFMA GEMM for the FP64 tier (gfx950, v_mfma_f64_16x16x4_f64), for
// operands whose panel rows are contiguous in memory: op(A) = A (kOpN) and op(B) = B^T
// (kOpT), the form of SYRK, SYR2K, GEMM-NT and the TRSM updates.
//
// 512 threads (8 waves, two per SIMD, a 64 x 32 accumulator block each) own one
// 128 x 128 tile of C and walk K in BK = 16 steps through a ring of NST LDS stages.
// A stage holds each panel as 16 k-rows of 128 doubles: one k-row of a panel is 1 KB
// contiguous in global memory (128 consecutive rows of one column) and is copied by one
// direct-to-LDS load (global_load_lds_dwordx4, 16 bytes per lane), issued NST-1 stages
// ahead of its use, with no registers or ds_writes on the way.  k-rows are 1152 bytes
// apart in LDS: the fragment reads (16 lanes of one k-row, 128 contiguous bytes, four
// k-rows per read) then cover the banks without conflict.
//
// Bounds: rows past the problem (in a 128-row panel) are read at clamped addresses and
// only reach C entries that are never stored.  The last k-block of each product is
// loaded by the threads themselves with masks (k past K as zeros, rows past the matrix
// never read), so every direct-to-LDS load stays inside the operand's columns
// 0 .. K-2 (plus at most one element past the end of such a column).
#pragma once
#include <hip/hip_runtime.h>

#include "gemm_kernel.hpp"     // vector types, xcd_remap
#include "gemm16_kernel.hpp"   // lds_dma16, TB_WAIT_VM_BARRIER, lds_void_t
#include "masked_load.hpp"
#include "tb_common.hpp"

namespace tb {

namespace g64 {
constexpr int BM = 128, BN = 128, BK = 16;
constexpr int NT = 512;
constexpr int WM = 2, WN = 4;                 // waves along rows / columns of C
constexpr int WTM = BM / WM, WTN = BN / WN;   // 64 x 32 per wave
constexpr int FM = WTM / 16, FN = WTN / 16;   // 16 x 16 MFMA blocks per wave
constexpr int KROW = 128 * 8 + 128;           // bytes of one k-row of a panel (+ pad)
constexpr int PANEL = BK * KROW;
constexpr int STAGE = 2 * PANEL;
constexpr int NST = 4;
constexpr int DMA_W = 2 * BK / (NT / 64);     // direct-to-LDS loads per wave per stage
static_assert(NST * STAGE <= 160 * 1024, "LDS");
}  // namespace g64

// Launchable: every problem NT with 16-byte aligned operands and even leading
// dimensions and batch strides (checked by the dispatcher).
__device__ __forceinline__ void gemm64_tile(const GemmLaunch& L, int g, char* smem)
{
    using namespace g64;
    const int tid = threadIdx.x, lane = tid & 63;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
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
    const long long lda = P.lda, ldb = P.ldb;
    const bool pairs = P.A2 != nullptr;
    const double* Ab = reinterpret_cast<const double*>(P.A) + (long long)bi * P.sA;
    const double* Bb = reinterpret_cast<const double*>(P.B) + (long long)bi * P.sB;
    const double* A2b = pairs ? reinterpret_cast<const double*>(P.A2) + (long long)bi * P.sA : nullptr;
    const double* B2b = pairs ? reinterpret_cast<const double*>(P.B2) + (long long)bi * P.sB : nullptr;
    double* Cb = reinterpret_cast<double*>(P.C) + (long long)bi * P.sC;

    const int kblocks = (K + BK - 1) / BK;
    const int nst = kblocks * (pairs ? 2 : 1);
    // stage s: product s / kblocks, k-block s % kblocks; the last k-block of a product
    // goes through registers
    auto is_dma = [&](int s) __attribute__((always_inline)) { return s < nst && (s % kblocks) != kblocks - 1; };

    // this wave's direct-to-LDS loads of a stage: k-rows `wave` and `wave + 8` of both
    // panels (op(A) rows, op(B) columns); lane l brings panel rows 2l, 2l + 1, clamped
    // into the matrix (past its last row they only feed C entries never stored)
    const int ra = min(i0 + 2 * lane, (M - 1) & ~1), rb = min(j0 + 2 * lane, (N - 1) & ~1);
    const long long oA = (long long)wave * lda + ra, oB = (long long)wave * ldb + rb;
    const unsigned smem_lds = (unsigned)(uintptr_t)(lds_void_t*)smem;
    auto issue = [&](int s) __attribute__((always_inline)) {
        const int pair = s >= kblocks;
        const long long k0 = (long long)(s - pair * kblocks) * BK;
        const double* xa = (pair ? A2b : Ab) + (oA + k0 * lda);
        const double* xb = (pair ? B2b : Bb) + (oB + k0 * ldb);
        const unsigned base = smem_lds + (unsigned)((s % NST) * STAGE) + wave * KROW;
        lds_dma16(xa, base);
        lds_dma16(xa + 8 * lda, base + 8 * KROW);
        lds_dma16(xb, base + PANEL);
        lds_dma16(xb + 8 * ldb, base + PANEL + 8 * KROW);
    };
    static_assert(DMA_W == 4 && NT / 64 == 8 && BK == 16, "two k-rows of each panel per wave");
    // the last k-block of a product by the threads: 2 x 2048 doubles, 8 per thread
    // (k-row idx / 128 of the panel, row idx % 128), zeros for k >= K and for rows past
    // the matrix
    constexpr int RV = 2 * BK * 128 / NT;
    double rv[RV];
    auto load_regs = [&](int s) __attribute__((always_inline)) {
        const int pair = s >= kblocks;
        const int k0 = (s - pair * kblocks) * BK;
        const double* xa = pair ? A2b : Ab;
        const double* xb = pair ? B2b : Bb;
#pragma unroll
        for (int i = 0; i < RV; ++i) {
            const int idx = tid + NT * (i % (RV / 2));
            const int kr = idx / 128, r = idx % 128;
            const int k = k0 + kr;
            if (i < RV / 2) {
                const bool in = i0 + r < M && k < K;
                rv[i] = masked_load(xa + (in ? (long long)k * lda + i0 + r : 0ll), in);
            } else {
                const bool in = j0 + r < N && k < K;
                rv[i] = masked_load(xb + (in ? (long long)k * ldb + j0 + r : 0ll), in);
            }
        }
    };
    auto store_regs = [&](int s) __attribute__((always_inline)) {
        char* base = smem + (s % NST) * STAGE;
#pragma unroll
        for (int i = 0; i < RV; ++i) {
            const int idx = tid + NT * (i % (RV / 2));
            const int kr = idx / 128, r = idx % 128;
            *reinterpret_cast<double*>(base + (i < RV / 2 ? 0 : PANEL) + kr * KROW + r * 8) = rv[i];
        }
    };

    const int wm = wave % WM, wn = wave / WM;
    doublex4_t acc[FM * FN];
#pragma unroll
    for (int q = 0; q < FM * FN; ++q) acc[q] = doublex4_t{};
    // fragment of lane l for k-step kk: row (block base + l % 16), k-row kk + l / 16
    const int fo = (lane >> 4) * KROW + (lane & 15) * 8;
    const int offA = wm * WTM * 8 + fo;
    const int offB = PANEL + wn * WTN * 8 + fo;

    // ---------------------------------------------------------------- pipeline
    // Iteration s: fragment reads of stage s, loads of stage s + NST - 1 (direct or
    // through registers), the MFMAs of stage s, then wait until stage s + 1 is in LDS
    // and barrier.  A register stage waits for its loads (and so for every direct load
    // issued before them), so only the direct stages issued after the last register
    // stage may still be in flight at a barrier.
    if (nst > 0) {
        int last_reg = -1;
        for (int s = 0; s < NST - 1 && s < nst; ++s) {
            if (is_dma(s)) {
                issue(s);
            } else {
                load_regs(s);
                store_regs(s);
                last_reg = s;
            }
        }
        {
            // stage 0 complete: the direct stages after it (and after the last register
            // stage) may stay in flight
            int after = 0;
            for (int t = (last_reg > 0 ? last_reg : 0) + 1; t < NST - 1 && t < nst; ++t) after += is_dma(t);
            if (after >= 2) TB_WAIT_VM_BARRIER(2 * DMA_W);
            else if (after == 1) TB_WAIT_VM_BARRIER(DMA_W);
            else TB_WAIT_VM_BARRIER(0);
        }
#pragma unroll 1
        for (int s = 0; s < nst; ++s) {
            const char* st = smem + (s % NST) * STAGE;
            double af[BK / 4][FM], bf[BK / 4][FN];
#pragma unroll
            for (int kq = 0; kq < BK / 4; ++kq) {
#pragma unroll
                for (int a = 0; a < FM; ++a)
                    af[kq][a] = *reinterpret_cast<const double*>(st + offA + kq * 4 * KROW + a * 128);
#pragma unroll
                for (int b = 0; b < FN; ++b)
                    bf[kq][b] = *reinterpret_cast<const double*>(st + offB + kq * 4 * KROW + b * 128);
            }
            const int nx = s + NST - 1;
            const bool reg_next = nx < nst && !is_dma(nx);
            if (nx < nst) {
                if (reg_next) load_regs(nx);
                else issue(nx);
            }
#pragma unroll
            for (int kq = 0; kq < BK / 4; ++kq)
#pragma unroll
                for (int b = 0; b < FN; ++b)
#pragma unroll
                    for (int a = 0; a < FM; ++a)
                        acc[b * FM + a] = __builtin_amdgcn_mfma_f64_16x16x4f64(bf[kq][b], af[kq][a],
                                                                                acc[b * FM + a], 0, 0, 0);
            if (reg_next) {
                store_regs(nx);
                last_reg = nx;
            }
            // stages issued after s + 1 that may stay in flight: direct ones after the
            // last register stage
            int after = 0;
#pragma unroll
            for (int t = 2; t < NST; ++t) {
                const int u = s + t;
                after += (u < nst && u > last_reg && is_dma(u)) ? 1 : 0;
            }
            if (after >= 2) TB_WAIT_VM_BARRIER(2 * DMA_W);
            else if (after == 1) TB_WAIT_VM_BARRIER(DMA_W);
            else TB_WAIT_VM_BARRIER(0);
        }
    }

    // ----------------------------------------------------------------- epilogue
    const double alpha = P.alpha, beta = P.beta;
    const bool lowerC = P.modeC == kCLower;
    const long long ldc = P.ldc;
    const double* cdiv = static_cast<const double*>(P.cdiv);
#pragma unroll
    for (int b = 0; b < FN; ++b)
#pragma unroll
        for (int a = 0; a < FM; ++a) {
            const int i = i0 + wm * WTM + a * 16 + (lane & 15);
            // v_mfma_f64_16x16x4: accumulator r of lane l is row l % 16 (of C), column
            // l / 16 + 4 r
            double cold[4], cd[4];
            bool in[4];
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int j = j0 + wn * WTN + b * 16 + (lane >> 4) + 4 * r;
                in[r] = i < M && j < N && (!lowerC || i >= j);
                cold[r] = masked_load(Cb + (in[r] ? (long long)i + (long long)j * ldc : 0ll), in[r] && beta != 0.0);
                cd[r] = cdiv ? masked_load(cdiv + (in[r] ? j : 0), in[r]) : 1.0;
            }
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                if (!in[r]) continue;
                const int j = j0 + wn * WTN + b * 16 + (lane >> 4) + 4 * r;
                const double v = acc[b * FM + a][r];
                double w;
                if (cdiv) w = (alpha * v + (beta == 0.0 ? 0.0 : beta * cold[r])) / cd[r];
                else w = beta == 0.0 ? alpha * v : alpha * v + beta * cold[r];
                Cb[(long long)i + (long long)j * ldc] = w;
            }
        }
}

__global__ void __launch_bounds__(g64::NT) gemm64_kernel(GemmLaunch L)
{
    __shared__ __attribute__((aligned(1024))) char smem[g64::NST * g64::STAGE];
#pragma unroll 1
    for (int b = blockIdx.x; b < L.totalTiles; b += gridDim.x) {
        gemm64_tile(L, xcd_remap(b, L.totalTiles), smem);
        __syncthreads();      // the next tile refills the ring
    }
}

}  // namespace tb
