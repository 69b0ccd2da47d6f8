// Device bodies shared by the single-call kernels (aux_kernels.hip) and the batched
// kernels (batch.hip): row exponents and the inverses of TRSM diagonal blocks.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

#include "tb_kernels.hpp"

namespace tb {

__device__ __forceinline__ int expo(float x)
{
    uint32_t u = __float_as_uint(x);
    int b = (u >> 23) & 0xff;
    if (b) return b - 126;
    if ((u & 0x7fffffffu) == 0) return kExpZero;
    int e;
    frexpf(x, &e);
    return e;
}

__device__ __forceinline__ int expo(double x)
{
    uint64_t u = (uint64_t)__double_as_longlong(x);
    int b = (int)((u >> 52) & 0x7ff);
    if (b) return b - 1022;
    if ((u & 0x7fffffffffffffffull) == 0) return kExpZero;
    int e;
    frexp(x, &e);
    return e;
}

constexpr int kRowChunk = 256;    // k per block of the row-major pass
constexpr int kColChunk = 1024;   // k per block of the k-contiguous pass

// rows contiguous across threads: one thread per row of rows rb*256.., k chunk kb
template <typename T>
__device__ __forceinline__ void rowexp_mn_block(const T* __restrict__ X, long long ld, int rows,
                                                int K, bool lower_only, int* out, int rb, int kb)
{
    const int r = rb * 256 + threadIdx.x;
    if (r >= rows) return;
    const int k0 = kb * kRowChunk;
    int k1 = min(K, k0 + kRowChunk);
    if (lower_only) k1 = min(k1, r + 1);
    int e = kExpZero;
    for (int k = k0; k < k1; ++k) e = max(e, expo(X[(long long)r + (long long)k * ld]));
    if (e != kExpZero) atomicMax(out + r, e);
}

// k contiguous: one wave per row (rows rb*4..), lanes across k chunk kb; kfrom_diag
// starts row r at k = r+1
template <typename T>
__device__ __forceinline__ void rowexp_k_block(const T* __restrict__ X, long long ld, int rows,
                                               int K, bool kfrom_diag, int* out, int rb, int kb)
{
    const int r = rb * 4 + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    if (r >= rows) return;
    const T* Xr = X + (long long)r * ld;
    int k0 = kb * kColChunk;
    const int k1 = min(K, k0 + kColChunk);
    if (kfrom_diag) k0 = max(k0, r + 1);
    int e = kExpZero;
    for (int k = k0 + lane; k < k1; k += 64) e = max(e, expo(Xr[k]));
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) e = max(e, __shfl_xor(e, o));
    if (lane == 0 && e != kExpZero) atomicMax(out + r, e);
}

// Diagonal block L (s x s, lower) = (I + N) D with D = diag(L): inv(L) = D^-1 (I+N)^-1.
// Writes W = D inv(L) - I = (I+N)^-1 - I (strictly lower) as Z = W^T (Z[c + r*s] =
// W(r, c)) and dinv[r] = 1/L(r,r), FP64 arithmetic.  One 64-lane block owns the
// kTrtriCols columns cb.. of inv(L): four lanes per column compute it by forward
// substitution, each summing every fourth term of a row's dot product (combined by
// shuffles); the columns live in LDS (y: 256 x kTrtriCols doubles) and row r of L is
// staged in LDS (lrow: 256) so the inner products read it as a broadcast.  The rows of
// L do not depend on the results, so each lane loads them kTrtriAhead rows early into
// a register ring: the row loads' latency overlaps the substitution (s <= 256).  The
// block is one wave, so LDS results are shared by waiting for the wave's own LDS
// operations (no barrier, whose fence would also wait for the prefetched rows).  A
// staged-chunk variant (32 rows of L per column, line-contiguous) measured slower: its
// larger LDS footprint halves the blocks resident per CU.
constexpr int kTrtriAhead = 8;
#define TB_WAVE_LDS_SYNC() asm volatile("s_waitcnt lgkmcnt(0)" ::: "memory")

template <typename Tin, typename Tout>
__device__ __forceinline__ void trtri_block(const Tin* __restrict__ Lq, long long lda, int s,
                                            Tout* __restrict__ Zq, double* __restrict__ dinvq,
                                            int cb, double* y, double* lrow)
{
    const int lane = threadIdx.x;
    const int cl = lane >> 2, part = lane & 3;
    const int col = cb + cl;
    const bool live = col < s;
    if (cb == 0)
        for (int r = lane; r < s; r += 64) dinvq[r] = 1.0 / (double)Lq[r + (long long)r * lda];
    for (int r = part; r < cb; r += 4)
        if (live) Zq[col + (long long)r * s] = Tout(0);
    // row r of L, columns cb + lane + 64 j (j < 4) with column <= r
    Tin pre[kTrtriAhead][4];
    auto fetch = [&](int r, Tin* v) {
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const int c = cb + lane + 64 * j;
            v[j] = (r < s && c <= r) ? Lq[r + (long long)c * lda] : Tin(0);
        }
    };
#pragma unroll
    for (int d = 0; d < kTrtriAhead; ++d) fetch(cb + d, pre[d]);
    for (int r0 = cb; r0 < s; r0 += kTrtriAhead) {
#pragma unroll
        for (int d = 0; d < kTrtriAhead; ++d) {
            const int r = r0 + d;
            if (r >= s) break;
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                const int c = cb + lane + 64 * j;
                if (c <= r) lrow[c - cb] = (double)pre[d][j];
            }
            fetch(r + kTrtriAhead, pre[d]);
            TB_WAVE_LDS_SYNC();
            // every fourth term of the row's dot product, in four independent partial
            // sums so that several LDS reads are in flight at once
            double s0 = 0.0, s1 = 0.0, s2 = 0.0, s3 = 0.0;
            int p = cb + part;
            for (; p + 12 < r; p += 16) {
                s0 += lrow[p - cb] * y[(p - cb) * kTrtriCols + cl];
                s1 += lrow[p + 4 - cb] * y[(p + 4 - cb) * kTrtriCols + cl];
                s2 += lrow[p + 8 - cb] * y[(p + 8 - cb) * kTrtriCols + cl];
                s3 += lrow[p + 12 - cb] * y[(p + 12 - cb) * kTrtriCols + cl];
            }
            for (; p < r; p += 4) s0 += lrow[p - cb] * y[(p - cb) * kTrtriCols + cl];
            double sum = (s0 + s1) + (s2 + s3);
            sum += __shfl_xor(sum, 1);
            sum += __shfl_xor(sum, 2);
            const double dg = lrow[r - cb];
            const double v = (r < col || !live) ? 0.0 : (((r == col) ? 1.0 : 0.0) - sum) / dg;
            if (part == 0) {
                y[(r - cb) * kTrtriCols + cl] = v;
                // W(r, col) = L(r,r) * inv(L)(r, col) for r > col
                if (live) Zq[col + (long long)r * s] = (r > col) ? (Tout)(dg * v) : Tout(0);
            }
            TB_WAVE_LDS_SYNC();
        }
    }
}

// 64 rows x 64 plane columns per block through an LDS tile, so both the source read
// (rows or k contiguous) and the 16-byte plane writes are coalesced.  e_used (split
// tier, optional) receives the row exponents the planes were scaled with, for a GEMM
// that runs while e itself keeps growing.
// Block (bx, by) of the conversion: rows by*64.., plane columns bx*64.. (t: a 64 x 65
// float LDS tile).  e_used is written by the blocks with bx == 0.
template <int NPL>
__device__ __forceinline__ void split_planes_block(const float* __restrict__ X, long long ld, int mode,
                                                   int rows, int K, const int* __restrict__ e,
                                                   int* __restrict__ e_used,
                                                   _Float16* __restrict__ out, long long ldp,
                                                   long long sp, int kshift, int bx, int by,
                                                   float (*t)[65])
{
    const int r0 = by * 64, c0 = bx * 64;   // c: plane column = kshift + k
    const int tid = threadIdx.x;
    for (int q = tid; q < 64 * 64; q += 256) {
        int rl, cl;
        if (mode == kOpT) { cl = q % 64; rl = q / 64; }
        else { rl = q % 64; cl = q / 64; }
        const int r = r0 + rl, k = c0 + cl - kshift;
        float v = 0.0f;
        if (r < rows && k >= 0 && k < K) {
            if (mode == kOpN) v = X[(long long)r + (long long)k * ld];
            else if (mode == kOpT) v = X[(long long)k + (long long)r * ld];
            else if (r >= k) v = X[(long long)r + (long long)k * ld];
            else if (mode == kOpSymL) v = X[(long long)k + (long long)r * ld];
        }
        t[rl][cl] = v;
    }
    __syncthreads();
    const int rl = tid >> 2, cl0 = (tid & 3) * 16;
    const int r = r0 + rl;
    if (r >= rows) return;
    int sc = 0;
    if (NPL >= 2) {
        const int ex = e[r];
        sc = 15 - (ex < -1000 ? 0 : (ex > 1000 ? 1000 : ex));
        if (e_used && bx == 0 && (tid & 3) == 0) e_used[r] = ex;
    }
    typedef _Float16 half8 __attribute__((ext_vector_type(8)));
    for (int half = 0; half < 2; ++half) {
        half8 hh, mm, ll;
#pragma unroll
        for (int j = 0; j < 8; ++j) {
            const float x = t[rl][cl0 + half * 8 + j];
            if (NPL == 1) {
                hh[j] = (_Float16)x;
            } else {
                const float xs = __builtin_ldexpf(x, sc);
                const _Float16 h = (_Float16)xs;
                const float r1 = (xs - (float)h) * 2048.0f;
                const _Float16 m = (_Float16)r1;
                hh[j] = h;
                mm[j] = m;
                if (NPL == 3) ll[j] = (_Float16)((r1 - (float)m) * 2048.0f);
            }
        }
        _Float16* dst = out + (long long)r * ldp + c0 + cl0 + half * 8;
        *reinterpret_cast<half8*>(dst) = hh;
        if (NPL >= 2) *reinterpret_cast<half8*>(dst + sp) = mm;
        if (NPL == 3) *reinterpret_cast<half8*>(dst + 2 * sp) = ll;
    }
}

}  // namespace tb
