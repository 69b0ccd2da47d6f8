// Diagonal-block Cholesky of the recursive POTRF (host/tb_potrf.cpp).
//
// One workgroup (256 threads) factors one s x s diagonal block (s <= kPotrfLeaf) in
// place in FP64, A_jj = L_jj L_jj^T, lower triangle only.  The block is loaded into
// LDS as doubles (column-major, padded) and factored right-looking in panels of
// kPanel = 16 columns (32 made the fully unrolled code larger than the 64 KB
// instruction cache, and the kernel instruction-fetch bound):
//   1. the panel's 16 x 16 diagonal part, by one wave: lane i holds row i in registers
//      and the pivots and column values travel by readlane (every index is a
//      compile-time constant, so nothing is spilled or indexed dynamically);
//   2. the rows below it, L21 = A21 L11^-T: one thread per row, forward substitution
//      in registers with L11 read from LDS (all threads read the same entry);
//   3. the trailing lower triangle, A22 -= L21 L21^T: two threads per row, each row's
//      panel entries in registers, the other row's read from LDS.
// A non-positive (or NaN) pivot at local column j stops the factorization and records
// off + j + 1 in *info if it is still 0 (blocks run in order, so the first failure in
// the matrix is kept): LAPACK dpotrf's info convention.  A block whose *info is
// already set returns at once, so the factor stays the partial one.
#include <hip/hip_runtime.h>

#include "tb_kernels.hpp"

namespace tb {

namespace {

constexpr int kLd = kPotrfLeaf + 1;     // padded LDS column
constexpr int kPanel = 16;     // keeps the unrolled code inside the instruction cache

__device__ __forceinline__ double readlane64(double x, int lane)
{
    const long long v = __double_as_longlong(x);
    const int lo = __builtin_amdgcn_readlane((int)(v & 0xffffffff), lane);
    const int hi = __builtin_amdgcn_readlane((int)(v >> 32), lane);
    return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}

template <typename T>
__global__ void __launch_bounds__(256)
potrf_leaf_kernel(T* __restrict__ A, long long lda, int s, long long off, int* __restrict__ info)
{
    __shared__ double a[kPotrfLeaf * kLd];
    __shared__ int fail;                          // local column of a failing pivot, or -1
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    if (*info != 0) return;                       // an earlier block failed
    // thread (ri, kc): row ri of columns kc, kc + 2, ...; eight loads in flight per
    // thread, so the block arrives in s/16 memory latencies, not s/2
    const int ri = tid & 127, kc = tid >> 7;
    for (int k0 = 0; k0 < s; k0 += 16) {
        T v[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) {
            const int k = k0 + 2 * u + kc;
            v[u] = (k < s && ri < s && ri >= k) ? A[ri + (long long)k * lda] : T(0);
        }
#pragma unroll
        for (int u = 0; u < 8; ++u) {
            const int k = k0 + 2 * u + kc;
            if (k < s && ri < s && ri >= k) a[k * kLd + ri] = (double)v[u];
        }
    }
    if (tid == 0) fail = -1;
    __syncthreads();
    for (int p = 0; p < s; p += kPanel) {
        const int pw = min(kPanel, s - p);
        // ---- 1. diagonal part of the panel (wave 0; lane i < pw owns row p + i)
        if (wave == 0) {
            double r[kPanel];
#pragma unroll
            for (int c = 0; c < kPanel; ++c)
                r[c] = (lane < pw && c <= lane) ? a[(p + c) * kLd + p + lane] : 0.0;
            int bad = -1;
#pragma unroll
            for (int j = 0; j < kPanel; ++j) {
                if (j < pw && bad < 0) {
                    const double piv = readlane64(r[j], j);
                    if (!(piv > 0.0)) {
                        bad = j;
                    } else {
                        const double d = __builtin_sqrt(piv);
                        r[j] = lane == j ? d : r[j] / d;
#pragma unroll
                        for (int k = j + 1; k < kPanel; ++k) {
                            const double lkj = readlane64(r[j], k);
                            if (lane >= k) r[k] -= r[j] * lkj;
                        }
                    }
                }
            }
            if (lane < pw) {
#pragma unroll
                for (int c = 0; c < kPanel; ++c)
                    if (c <= lane) a[(p + c) * kLd + p + lane] = r[c];
            }
            if (lane == 0 && bad >= 0) {
                fail = p + bad;
                atomicCAS(info, 0, (int)(off + p + bad + 1));
            }
        }
        __syncthreads();
        if (fail >= 0) break;
        const int m = s - p - pw;                 // rows below the panel
        if (m <= 0) break;
        // ---- 2. L21 = A21 L11^-T, one thread per row (m <= 112 < 256)
        if (tid < m) {
            const int row = p + pw + tid;
            double x[kPanel];
#pragma unroll
            for (int c = 0; c < kPanel; ++c) x[c] = c < pw ? a[(p + c) * kLd + row] : 0.0;
#pragma unroll
            for (int j = 0; j < kPanel; ++j) {
                if (j < pw) {
                    x[j] /= a[(p + j) * kLd + p + j];
#pragma unroll
                    for (int k = j + 1; k < kPanel; ++k)
                        if (k < pw) x[k] -= x[j] * a[(p + j) * kLd + p + k];
                }
            }
#pragma unroll
            for (int c = 0; c < kPanel; ++c)
                if (c < pw) a[(p + c) * kLd + row] = x[c];
        }
        __syncthreads();
        // ---- 3. A22 -= L21 L21^T (lower), two threads per row: columns of parity h
        if (tid < 2 * m) {
            const int il = tid >> 1, h = tid & 1;
            const int i = p + pw + il;
            double li[kPanel];
#pragma unroll
            for (int c = 0; c < kPanel; ++c) li[c] = c < pw ? a[(p + c) * kLd + i] : 0.0;
            for (int kl = h; kl <= il; kl += 2) {
                const int k = p + pw + kl;
                double sum = 0.0;
#pragma unroll
                for (int c = 0; c < kPanel; ++c)
                    if (c < pw) sum += li[c] * a[(p + c) * kLd + k];
                a[k * kLd + i] -= sum;
            }
        }
        __syncthreads();
    }
    __syncthreads();
    for (int k = kc; k < s; k += 2)
        if (ri < s && ri >= k) A[ri + (long long)k * lda] = (T)a[k * kLd + ri];
}

}  // namespace

hipError_t potrf_leaf(bool f64, void* A, long long lda, int s, long long off, int* info,
                      hipStream_t st)
{
    if (s <= 0) return hipSuccess;
    if (s > kPotrfLeaf) return hipErrorInvalidValue;
    if (f64)
        hipLaunchKernelGGL(potrf_leaf_kernel<double>, dim3(1), dim3(256), 0, st,
                           static_cast<double*>(A), lda, s, off, info);
    else
        hipLaunchKernelGGL(potrf_leaf_kernel<float>, dim3(1), dim3(256), 0, st,
                           static_cast<float*>(A), lda, s, off, info);
    return hipGetLastError();
}

}  // namespace tb
