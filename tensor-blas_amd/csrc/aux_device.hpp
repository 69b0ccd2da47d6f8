// Device bodies shared by the single-call kernels (aux_kernels.hip) and the batched
// kernels (batch.hip): row exponents and the inverses of TRSM diagonal blocks.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

#include "masked_load.hpp"
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

// rows contiguous across threads: one thread per row of rows rb*256.., k chunk kb of
// `chunk` columns
template <typename T>
__device__ __forceinline__ void rowexp_mn_block(const T* __restrict__ X, long long ld, int rows,
                                                int K, bool lower_only, int* out, int rb, int kb,
                                                int chunk = kRowChunk)
{
    const int r = rb * 256 + threadIdx.x;
    if (r >= rows) return;
    const int k0 = kb * chunk;
    int k1 = min(K, k0 + chunk);
    if (lower_only) k1 = min(k1, r + 1);
    int e = kExpZero;
    int k = k0;
    // eight independent loads in flight per thread: one at a time, the pass waited a
    // full memory latency per element
    for (; k + 8 <= k1; k += 8) {
        T v[8];
#pragma unroll
        for (int q = 0; q < 8; ++q) v[q] = X[(long long)r + (long long)(k + q) * ld];
#pragma unroll
        for (int q = 0; q < 8; ++q) e = max(e, expo(v[q]));
    }
    for (; k < k1; ++k) e = max(e, expo(X[(long long)r + (long long)k * ld]));
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
    int k = k0 + lane;
    for (; k + 192 < k1; k += 256) {          // four independent loads in flight
        const T v0 = Xr[k], v1 = Xr[k + 64], v2 = Xr[k + 128], v3 = Xr[k + 192];
        e = max(max(e, expo(v0)), max(expo(v1), max(expo(v2), expo(v3))));
    }
    for (; k < k1; k += 64) e = max(e, expo(Xr[k]));
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) e = max(e, __shfl_xor(e, o));
    if (lane == 0 && e != kExpZero) atomicMax(out + r, e);
}

// Diagonal block L (s x s, lower) = (I + N) D with D = diag(L): inv(L) = D^-1 (I+N)^-1.
// Writes W = D inv(L) - I = (I+N)^-1 - I (strictly lower) as Z = W^T (Z[c + r*s] =
// W(r, c)) and ldiag[r] = L(r,r), in the arithmetic of T: FP32 for the FP32-class tiers
// (plain, split, FP32-FMA), FP64 for the FP64 tiers.  One wave owns the kTrtriCols
// columns cb.. of W and solves them right-looking: lane l holds rows l, l+64, l+128,
// l+192 of those columns in registers (acc, starting as the identity's columns); at
// step r the finished row r is broadcast by readlane (a wavefront shuffle), v = acc_r /
// L(r,r), and every later row q subtracts L(q,r) v.  Row q's acc is final once step q
// is reached and is W(q, .) itself (W = D inv(L) - I needs no final scale), so the wave
// stores it at the end.  Column r of L is contiguous in memory and is loaded
// kTrtriAhead steps early into a register ring.  No LDS and no barriers: the chain per
// step is one readlane broadcast, one division and the row updates (s <= 256).
constexpr int kTrtriAhead = 8;

__device__ __forceinline__ double trtri_readlane(double x, int lane)
{
    const long long v = __double_as_longlong(x);
    const int lo = __builtin_amdgcn_readlane((int)(v & 0xffffffff), lane);
    const int hi = __builtin_amdgcn_readlane((int)(v >> 32), lane);
    return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}

__device__ __forceinline__ float trtri_readlane(float x, int lane)
{
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(x), lane));
}

template <typename T>
__device__ __forceinline__ void trtri_block(const T* __restrict__ Lq, long long lda, int s,
                                            T* __restrict__ Zq, T* __restrict__ ldiagq, int cb)
{
    constexpr int C = kTrtriCols;
    const int lane = threadIdx.x & 63;
    if (cb == 0)
        for (int r = lane; r < s; r += 64) ldiagq[r] = Lq[r + (long long)r * lda];
    T acc[4][C];
#pragma unroll
    for (int u = 0; u < 4; ++u)
#pragma unroll
        for (int c = 0; c < C; ++c) acc[u][c] = (u * 64 + lane == cb + c) ? T(1) : T(0);
    // column r of L at rows u*64 + lane >= r (others read as zero), and 1 / L(r, r)
    // (broadcast from the lane holding the diagonal; the division is off the chain)
    T pre[kTrtriAhead][4], rcpr[kTrtriAhead];
    // unconditional loads at clamped addresses (masked_load), so a step's loads never
    // wait for each other
    auto fetch = [&](int r, T* v) {
        const int rc = min(r, s - 1);
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const int q = u * 64 + lane, qc = min(q, s - 1);
            v[u] = masked_load(Lq + qc + (long long)rc * lda, r < s && q < s && q >= r);
        }
    };
    auto diag_rcp = [&](int r, const T* v) {
        const int ur = (r >> 6) & 3, lr = r & 63;
        T dg = trtri_readlane(v[0], lr);
#pragma unroll
        for (int u = 1; u < 4; ++u) dg = ur == u ? trtri_readlane(v[u], lr) : dg;
        return r < s ? T(1) / dg : T(0);
    };
#pragma unroll
    for (int d = 0; d < kTrtriAhead; ++d) {
        fetch(cb + d, pre[d]);
        rcpr[d] = diag_rcp(cb + d, pre[d]);
    }
#pragma unroll
    for (int u = 0; u < 4; ++u) {
        // steps r = u*64 + lr owned by slot u (cb is a multiple of kTrtriCols and of 8)
        const int first = max(cb - u * 64, 0), last = min(64, s - u * 64);
        for (int lr0 = first; lr0 < last; lr0 += kTrtriAhead) {
#pragma unroll
            for (int d = 0; d < kTrtriAhead; ++d) {
                const int lr = lr0 + d;
                const int r = u * 64 + lr;
                if (lr < last) {
                    const T rcp = rcpr[d];                                       // 1 / L(r, r)
                    T v[C];
#pragma unroll
                    for (int c = 0; c < C; ++c) v[c] = trtri_readlane(acc[u][c], lr) * rcp;
#pragma unroll
                    for (int w = u; w < 4; ++w) {
                        const T lq = pre[d][w];                  // L(w*64 + lane, r), 0 unless below r
                        if (w > u || lane > lr) {
#pragma unroll
                            for (int c = 0; c < C; ++c) acc[w][c] -= lq * v[c];
                        }
                    }
                }
                // every lane issues every load (past the end: masked, clamped), and the
                // reciprocal of a column's diagonal is formed half the prefetch distance
                // after its loads, so no step waits for the loads it has just issued
                fetch(r + kTrtriAhead, pre[d]);
                const int dh = (d + kTrtriAhead / 2) % kTrtriAhead;
                rcpr[dh] = diag_rcp(r + kTrtriAhead / 2, pre[dh]);
            }
        }
    }
    // row q of W: acc (q > column), zero on and above the diagonal
#pragma unroll
    for (int u = 0; u < 4; ++u) {
        const int q = u * 64 + lane;
        if (q < s) {
#pragma unroll
            for (int c = 0; c < C; ++c)
                if (cb + c < s) Zq[(cb + c) + (long long)q * s] = (q > cb + c) ? acc[u][c] : T(0);
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
                                                   float (*t)[65], bool lower_only = false)
{
    const int r0 = by * 64, c0 = bx * 64;   // c: plane column = kshift + k
    const int tid = threadIdx.x;
    // the 16 elements of this thread: all loads issued before the first LDS store
    // (unconditional loads at a clamped address, then masked: loads behind per-element
    // branches went out one memory latency apart).  A lower-stored operand's blocks
    // wholly below (above) the diagonal read as mode N (as the mirror, mode T; TriL:
    // zeros), with rows (k) contiguous across threads, so every read coalesces: with
    // all loads in flight at once, strided reads would fetch each line many times.
    int em = mode;
    if (mode == kOpTriL || mode == kOpSymL) {
        if (c0 + 63 - kshift <= r0) em = kOpN;
        else if (c0 - kshift > r0 + 63) em = mode == kOpSymL ? kOpT : -1;
    }
    if (em < 0 && lower_only) return;      // wholly above the diagonal: nobody reads it
    const bool kmaj = em == kOpT;
    float v[16];
#pragma unroll
    for (int it = 0; it < 16; ++it) {
        const int q = tid + it * 256;
        const int rl = kmaj ? q / 64 : q % 64, cl = kmaj ? q % 64 : q / 64;
        const int r = r0 + rl, k = c0 + cl - kshift;
        bool keep = r < rows && k >= 0 && k < K && em >= 0;
        long long idx;
        if (em == kOpN) {
            idx = (long long)r + (long long)k * ld;
        } else if (em == kOpT) {
            idx = (long long)k + (long long)r * ld;
        } else {                               // across the diagonal: the other triangle mirrored
            keep = keep && (r >= k || mode == kOpSymL);
            idx = r >= k ? (long long)r + (long long)k * ld : (long long)k + (long long)r * ld;
        }
        v[it] = (float)masked_load(X + (keep ? idx : 0ll), keep);
    }
#pragma unroll
    for (int it = 0; it < 16; ++it) {
        const int q = tid + it * 256;
        t[kmaj ? q / 64 : q % 64][kmaj ? q % 64 : q / 64] = v[it];
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
                hh[j] = h;
                if (NPL == 2) {
                    mm[j] = (_Float16)(xs - (float)h);          // xs - h exact in FP32
                } else {
                    const float r1 = (xs - (float)h) * 2048.0f;
                    const _Float16 m = (_Float16)r1;
                    mm[j] = m;
                    ll[j] = (_Float16)((r1 - (float)m) * 2048.0f);
                }
            }
        }
        const int c = c0 + cl0 + half * 8;
        *reinterpret_cast<half8*>(out + plane_off(NPL, ldp, sp, 0, r, c)) = hh;
        if (NPL >= 2) *reinterpret_cast<half8*>(out + plane_off(NPL, ldp, sp, 1, r, c)) = mm;
        if (NPL == 3) *reinterpret_cast<half8*>(out + plane_off(NPL, ldp, sp, 2, r, c)) = ll;
    }
}

}  // namespace tb
