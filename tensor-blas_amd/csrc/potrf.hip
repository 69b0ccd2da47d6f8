// Diagonal-block Cholesky of the recursive POTRF (host/tb_potrf.cpp).
//
// One workgroup factors one s x s diagonal block (s <= kPotrfLeaf = 128) in place in
// the arithmetic of the matrix (FP32 for the FP32-class tiers, FP64 for the FP64
// tiers), A_jj = L_jj L_jj^T, lower triangle only, right-looking with the block in
// registers.  Thread tiles are 4 x 4; a wave holds an 8 x 8 block of tiles (32 x 32
// elements), and the ten waves hold the ten lower-triangular such blocks, so a wave
// whose rows or columns are all finished skips a step as a whole.
// Each step finalizes two columns j, j+1 behind one barrier: every thread reads both
// (current) columns from a double-buffered LDS pair, forms
//   d0 = sqrt(a_jj), l_i0 = a_ij / d0, d1 = sqrt(a_j+1,j+1 - l_j+1,0^2),
//   l_i1 = (a_i,j+1 - l_i0 l_j+1,0) / d1
// for its rows and columns -- the operations the one-column recursion does, in the
// same order -- finalizes columns j, j+1 and subtracts l_i0 l_k0, then l_i1 l_k1 from
// every later element; the threads holding columns j+2, j+3 then publish them.
// A non-positive (or NaN) pivot at column c stops the factorization and records
// off + c + 1 in *info if it is still 0 (blocks run in order, so the first failure in
// the matrix is kept): LAPACK dpotrf's info convention; the columns before c are
// factored.  A block whose *info is already set returns at once.
#include <hip/hip_runtime.h>

#include "aux_device.hpp"
#include "tb_kernels.hpp"

namespace tb {

namespace {

static_assert(kPotrfLeaf == 128, "4 x 4 blocks of 8 x 8 tiles of 4 x 4");
constexpr int kLeafWaves = 10;                   // lower-triangular 32 x 32 blocks

__device__ __forceinline__ float leaf_sqrt(float x) { return __builtin_sqrtf(x); }
__device__ __forceinline__ double leaf_sqrt(double x) { return __builtin_sqrt(x); }

template <typename T>
__global__ void __launch_bounds__(64 * kLeafWaves)
potrf_leaf_kernel(T* __restrict__ A, long long lda, int s, long long off, int* __restrict__ info)
{
    __shared__ T colA[2][kPotrfLeaf];             // column j of the current step
    __shared__ T colB[2][kPotrfLeaf];             // column j + 1
    const int tid = threadIdx.x;
    if (*info != 0) return;                       // an earlier block failed
    // wave w -> lower block (bi, bj): w = bi (bi + 1) / 2 + bj
    const int w = tid >> 6, lane = tid & 63;
    const int bi = w >= 6 ? 3 : (w >= 3 ? 2 : (w >= 1 ? 1 : 0));
    const int bj = w - bi * (bi + 1) / 2;
    const int i0 = 32 * bi + 4 * (lane >> 3), k0 = 32 * bj + 4 * (lane & 7);
    const bool live = k0 <= i0 + 3 && i0 < s && k0 < s;   // tile touches the lower block
    T a[4][4];
#pragma unroll
    for (int ii = 0; ii < 4; ++ii) {
        const int i = i0 + ii, ic = min(i, s - 1);
#pragma unroll
        for (int kk = 0; kk < 4; ++kk) {
            const int k = k0 + kk, kc = min(k, s - 1);
            a[ii][kk] = masked_load(A + ic + (long long)kc * lda, live && i < s && k <= i);
        }
    }
    if (k0 == 0) {                                // columns 0 and 1 for step 0
#pragma unroll
        for (int ii = 0; ii < 4; ++ii) {
            colA[0][i0 + ii] = a[ii][0];
            colB[0][i0 + ii] = a[ii][1];
        }
    }
    for (int j = 0; j < s; j += 2) {
        __syncthreads();
        const int b = (j >> 1) & 1;
        const T* ca = colA[b];
        const T* cb = colB[b];
        const bool has1 = j + 1 < s;              // uniform
        const T piv0 = ca[j];
        if (!(piv0 > T(0))) {                      // uniform: every thread reads the same pivot
            if (tid == 0) atomicCAS(info, 0, (int)(off + j + 1));
            break;
        }
        const T d0 = leaf_sqrt(piv0);
        const T r0 = T(1) / d0;
        const T l1 = has1 ? ca[j + 1] * r0 : T(0);
        const T piv1 = has1 ? cb[j + 1] - l1 * l1 : T(1);
        const bool ok1 = piv1 > T(0);              // uniform; false for NaN
        const T d1 = leaf_sqrt(ok1 ? piv1 : T(1));
        const T r1 = T(1) / d1;
        const bool two = has1 && ok1;             // columns j and j+1 this step
        if (live && i0 + 3 >= j && k0 + 3 >= j) {   // rows or columns left to finish
            T li0[4], li1[4], lk0[4], lk1[4];
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                li0[q] = ca[i0 + q] * r0;
                li1[q] = (cb[i0 + q] - li0[q] * l1) * r1;
                lk0[q] = ca[k0 + q] * r0;
                lk1[q] = (cb[k0 + q] - lk0[q] * l1) * r1;
            }
            if (two && k0 > j + 1 && i0 > j + 1) {
                // the whole tile is trailing: the rank-2 update alone, without the
                // per-element selects (the same operations in the same order)
#pragma unroll
                for (int ii = 0; ii < 4; ++ii) {
#pragma unroll
                    for (int kk = 0; kk < 4; ++kk) {
                        T v = a[ii][kk];
                        v = v - li0[ii] * lk0[kk];
                        v = v - li1[ii] * lk1[kk];
                        a[ii][kk] = v;
                    }
                }
            } else {
#pragma unroll
                for (int ii = 0; ii < 4; ++ii) {
                    const int i = i0 + ii;
#pragma unroll
                    for (int kk = 0; kk < 4; ++kk) {
                        const int k = k0 + kk;
                        T v = a[ii][kk];
                        if (k == j) {
                            v = i == j ? d0 : (i > j ? li0[ii] : v);
                        } else if (two && k == j + 1) {
                            v = i == j + 1 ? d1 : (i > j + 1 ? li1[ii] : v);
                        } else if (two && k > j + 1 && i > j + 1) {
                            v = v - li0[ii] * lk0[kk];
                            v = v - li1[ii] * lk1[kk];
                        }
                        a[ii][kk] = v;
                    }
                }
            }
            // the tiles holding columns j+2, j+3 publish them (static indices only: a
            // select on a run-time column index would put the tile in scratch memory)
            if (two) {
                T* na = colA[b ^ 1];
                T* nb = colB[b ^ 1];
#pragma unroll
                for (int kk = 0; kk < 4; ++kk) {
                    if (k0 + kk == j + 2) {
#pragma unroll
                        for (int ii = 0; ii < 4; ++ii)
                            if (i0 + ii >= j + 2) na[i0 + ii] = a[ii][kk];
                    }
                    if (k0 + kk == j + 3) {
#pragma unroll
                        for (int ii = 0; ii < 4; ++ii)
                            if (i0 + ii >= j + 3) nb[i0 + ii] = a[ii][kk];
                    }
                }
            }
        }
        if (has1 && !ok1) {                       // column j is final; j+1 is not positive
            if (tid == 0) atomicCAS(info, 0, (int)(off + j + 2));
            break;
        }
    }
#pragma unroll
    for (int ii = 0; ii < 4; ++ii) {
        const int i = i0 + ii;
#pragma unroll
        for (int kk = 0; kk < 4; ++kk) {
            const int k = k0 + kk;
            if (live && i < s && k <= i) A[i + (long long)k * lda] = a[ii][kk];
        }
    }
}

}  // namespace

hipError_t potrf_leaf(bool f64, void* A, long long lda, int s, long long off, int* info,
                      hipStream_t st)
{
    if (s <= 0) return hipSuccess;
    if (s > kPotrfLeaf) return hipErrorInvalidValue;
    if (f64)
        hipLaunchKernelGGL(potrf_leaf_kernel<double>, dim3(1), dim3(64 * kLeafWaves), 0, st,
                           static_cast<double*>(A), lda, s, off, info);
    else
        hipLaunchKernelGGL(potrf_leaf_kernel<float>, dim3(1), dim3(64 * kLeafWaves), 0, st,
                           static_cast<float*>(A), lda, s, off, info);
    return hipGetLastError();
}

}  // namespace tb
