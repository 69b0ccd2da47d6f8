// Diagonal-block Cholesky of the recursive POTRF (host/tb_potrf.cpp).
//
// One workgroup of 1024 threads factors one s x s diagonal block (s <= kPotrfLeaf =
// 128) in place in FP64, A_jj = L_jj L_jj^T, lower triangle only, right-looking with
// the block in registers: thread (rt, ct) holds the 4 x 4 tile of rows 4rt.. and
// columns 4ct.. (tiles above the diagonal stay idle).  Step j reads the pivot column j
// from a double-buffered LDS column, forms l = a / sqrt(a_jj) for the tile's rows and
// columns, finalizes column j and subtracts l_i l_k from every later element; the
// threads holding column j + 1 then publish it.  One barrier per step; 4 waves per
// SIMD hide the LDS and FP64 latencies.
// A non-positive (or NaN) pivot at column j stops the factorization and records
// off + j + 1 in *info if it is still 0 (blocks run in order, so the first failure in
// the matrix is kept): LAPACK dpotrf's info convention.  A block whose *info is
// already set returns at once, so the factor stays the partial one.
#include <hip/hip_runtime.h>

#include "aux_device.hpp"
#include "tb_kernels.hpp"

namespace tb {

namespace {

static_assert(kPotrfLeaf == 128, "32 x 32 tiles of 4 x 4");

template <typename T>
__global__ void __launch_bounds__(1024) __attribute__((amdgpu_waves_per_eu(4, 4)))
potrf_leaf_kernel(T* __restrict__ A, long long lda, int s, long long off, int* __restrict__ info)
{
    __shared__ double col[2][kPotrfLeaf];         // pivot column of the current step
    const int tid = threadIdx.x;
    if (*info != 0) return;                       // an earlier block failed
    const int ct = tid & 31, rt = tid >> 5;
    const int i0 = 4 * rt, k0 = 4 * ct;
    const bool live = k0 <= i0 + 3 && i0 < s && k0 < s;   // tile touches the lower block
    double a[4][4];
#pragma unroll
    for (int ii = 0; ii < 4; ++ii) {
        const int i = i0 + ii, ic = min(i, s - 1);
#pragma unroll
        for (int kk = 0; kk < 4; ++kk) {
            const int k = k0 + kk, kc = min(k, s - 1);
            a[ii][kk] = masked_load(A + ic + (long long)kc * lda, live && i < s && k <= i);
        }
    }
    if (ct == 0) {                                // column 0 for step 0
#pragma unroll
        for (int ii = 0; ii < 4; ++ii) col[0][i0 + ii] = a[ii][0];
    }
    for (int j = 0; j < s; ++j) {
        __syncthreads();
        const double* cj = col[j & 1];
        const double piv = cj[j];
        if (!(piv > 0.0)) {                       // uniform: every thread reads the same pivot
            if (tid == 0) atomicCAS(info, 0, (int)(off + j + 1));
            break;
        }
        if (!live || i0 + 3 < j) continue;       // every row of the tile finished (no barrier
                                                  // skipped: the loop's barrier is at its top)
        const double d = __builtin_sqrt(piv);
        const double rcp = 1.0 / d;
        double li[4], lk[4];
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            li[q] = cj[i0 + q] * rcp;
            lk[q] = cj[k0 + q] * rcp;
        }
#pragma unroll
        for (int ii = 0; ii < 4; ++ii) {
            const int i = i0 + ii;
#pragma unroll
            for (int kk = 0; kk < 4; ++kk) {
                const int k = k0 + kk;
                const double upd = a[ii][kk] - li[ii] * lk[kk];
                const double fin = i == j ? d : (i > j ? li[ii] : a[ii][kk]);
                a[ii][kk] = k == j ? fin : ((k > j && i > j) ? upd : a[ii][kk]);
            }
        }
        // the tile holding column j + 1 publishes it (static indices only: a select
        // on a run-time column index would put the tile in scratch memory)
        double* cn = col[(j + 1) & 1];
#pragma unroll
        for (int kk = 0; kk < 4; ++kk) {
            if (k0 + kk == j + 1) {
#pragma unroll
                for (int ii = 0; ii < 4; ++ii)
                    if (i0 + ii > j) cn[i0 + ii] = a[ii][kk];
            }
        }
    }
#pragma unroll
    for (int ii = 0; ii < 4; ++ii) {
        const int i = i0 + ii;
#pragma unroll
        for (int kk = 0; kk < 4; ++kk) {
            const int k = k0 + kk;
            if (live && i < s && k <= i) A[i + (long long)k * lda] = (T)a[ii][kk];
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
        hipLaunchKernelGGL(potrf_leaf_kernel<double>, dim3(1), dim3(1024), 0, st,
                           static_cast<double*>(A), lda, s, off, info);
    else
        hipLaunchKernelGGL(potrf_leaf_kernel<float>, dim3(1), dim3(1024), 0, st,
                           static_cast<float*>(A), lda, s, off, info);
    return hipGetLastError();
}

}  // namespace tb
