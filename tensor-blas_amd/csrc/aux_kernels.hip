// Auxiliary device kernels: row exponents for the split and Ozaki tiers, inverses
// of the TRSM diagonal blocks, and the FP32->FP16 / mirror / fill / copy kernels of
// the reference's util/util.cu:127-280 and include/TensorBLAS.h:185-284.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>

#include "aux_device.hpp"
#include "tb_kernels.hpp"

namespace tb {

namespace {

template <typename T>
__global__ void __launch_bounds__(256)
rowexp_mn_kernel(const T* __restrict__ X, long long ld, int rows, int K, bool lower_only,
                 long long sx, int* out, long long so, int chunk)
{
    rowexp_mn_block(X + (long long)blockIdx.z * sx, ld, rows, K, lower_only,
                    out + (long long)blockIdx.z * so, blockIdx.x, blockIdx.y, chunk);
}

template <typename T>
__global__ void __launch_bounds__(256)
rowexp_k_kernel(const T* __restrict__ X, long long ld, int rows, int K, bool kfrom_diag,
                long long sx, int* out, long long so)
{
    rowexp_k_block(X + (long long)blockIdx.z * sx, ld, rows, K, kfrom_diag,
                   out + (long long)blockIdx.z * so, blockIdx.x, blockIdx.y);
}

__global__ void fill_i32_kernel(int* p, long long n, int v)
{
    for (long long i = blockIdx.x * 256ll + threadIdx.x; i < n; i += 256ll * gridDim.x) p[i] = v;
}

template <typename T>
__global__ void __launch_bounds__(64)
trtri_lower_t_kernel(const T* __restrict__ L, long long lda, long long bstride, int s,
                     T* __restrict__ Z, T* __restrict__ ldiag)
{
    trtri_block<T>(L + (long long)blockIdx.y * bstride, lda, s, Z + (long long)blockIdx.y * s * s,
                   ldiag + (long long)blockIdx.y * s, blockIdx.x * kTrtriCols);
}

// one column c of the block per workgroup: A(r, c) = (delta(r, c) + W(r, c)) / L(r, r),
// W(r, c) = Z[c + r*s] (Z = W^T), in T's arithmetic
template <typename T>
__global__ void __launch_bounds__(256)
trtri_store_kernel(const T* __restrict__ Z, const T* __restrict__ ldiag, int s,
                   T* __restrict__ A, long long lda, int unit)
{
    const int c = blockIdx.x;
    for (int r = c + unit + (int)threadIdx.x; r < s; r += 256) {
        const T w = r == c ? T(1) : Z[c + (long long)r * s];
        A[r + (long long)c * lda] = w / ldiag[r];
    }
}

template <typename T>
__global__ void __launch_bounds__(256)
diag_zero_kernel(const T* __restrict__ A, long long lda, long long n, int* __restrict__ info)
{
    for (long long i = blockIdx.x * 256ll + threadIdx.x; i < n; i += 256ll * gridDim.x)
        if (A[i + i * lda] == T(0)) atomicMin(info, (int)(i + 1));
}

// FP32 -> FP16, round to nearest even
__global__ void __launch_bounds__(256)
s2h_kernel(long long m, long long n, const float* __restrict__ as, long long ldas,
           _Float16* __restrict__ ah, long long ldah)
{
    for (long long j = blockIdx.y; j < n; j += gridDim.y)
        for (long long i = blockIdx.x * 256ll + threadIdx.x; i < m; i += 256ll * gridDim.x)
            ah[i + j * ldah] = (_Float16)as[i + j * ldas];
}

// row-major m x n FP32 -> column-major m x n FP16 through a 32x32 LDS tile
__global__ void __launch_bounds__(256)
s2h_transpose_kernel(long long m, long long n, const float* __restrict__ as,
                     _Float16* __restrict__ ah)
{
    __shared__ float t[32][33];
    const long long i0 = blockIdx.y * 32ll, j0 = blockIdx.x * 32ll;
    const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
    for (int q = ty; q < 32; q += 8) {
        long long i = i0 + q, j = j0 + tx;                 // read row i, contiguous j
        if (i < m && j < n) t[q][tx] = as[i * n + j];
    }
    __syncthreads();
    for (int q = ty; q < 32; q += 8) {
        long long j = j0 + q, i = i0 + tx;                 // write column j, contiguous i
        if (i < m && j < n) ah[i + j * m] = (_Float16)t[tx][q];
    }
}

template <typename T>
__global__ void __launch_bounds__(256)
mirror_lower_kernel(long long n, T* A, long long lda)
{
    __shared__ T t[32][33];
    const long long bi = blockIdx.y, bj = blockIdx.x;      // destination tile (bi, bj), bi <= bj
    if (bi > bj) return;
    const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
    // source tile (bj, bi) in the lower triangle
    for (int q = ty; q < 32; q += 8) {
        long long i = bj * 32 + tx, j = bi * 32 + q;
        if (i < n && j < n) t[q][tx] = A[i + j * lda];
    }
    __syncthreads();
    for (int q = ty; q < 32; q += 8) {
        long long i = bi * 32 + tx, j = bj * 32 + q;      // A(i, j) := A(j, i) for i < j
        if (i < n && j < n && i < j) A[i + j * lda] = t[tx][q];
    }
}

template <typename T>
__global__ void fill_matrix_kernel(long long m, long long n, T* a, long long lda, T v)
{
    for (long long j = blockIdx.y; j < n; j += gridDim.y)
        for (long long i = blockIdx.x * 256ll + threadIdx.x; i < m; i += 256ll * gridDim.x)
            a[i + j * lda] = v;
}

template <typename TA, typename TB>
__global__ void copy_matrix_kernel(long long m, long long n, const TA* a, long long lda, TB* b,
                                   long long ldb)
{
    for (long long j = blockIdx.y; j < n; j += gridDim.y)
        for (long long i = blockIdx.x * 256ll + threadIdx.x; i < m; i += 256ll * gridDim.x)
            b[i + j * ldb] = (TB)a[i + j * lda];
}

template <typename TA, typename TB>
__global__ void copy_lower_kernel(long long n, const TA* a, long long lda, TB* b, long long ldb)
{
    for (long long j = blockIdx.y; j < n; j += gridDim.y)
        for (long long i = j + blockIdx.x * 256ll + threadIdx.x; i < n; i += 256ll * gridDim.x)
            b[i + j * ldb] = (TB)a[i + j * lda];
}

__global__ void add_matrix_f32_kernel(long long m, long long n, float* x, long long ldx,
                                      const float* y, long long ldy)
{
    for (long long j = blockIdx.y; j < n; j += gridDim.y)
        for (long long i = blockIdx.x * 256ll + threadIdx.x; i < m; i += 256ll * gridDim.x)
            x[i + j * ldx] += y[i + j * ldy];
}

template <int NPL>
__global__ void __launch_bounds__(256)
split_planes_kernel(const float* __restrict__ X, long long ld, int mode, int rows, int K,
                    const int* __restrict__ e, int* __restrict__ e_used,
                    _Float16* __restrict__ out, long long ldp, long long sp, int kshift,
                    bool lower_only)
{
    __shared__ float t[64][65];
    split_planes_block<NPL>(X, ld, mode, rows, K, e, e_used, out, ldp, sp, kshift, blockIdx.x,
                            blockIdx.y, t, lower_only);
}

dim3 grid2d(long long m, long long n)
{
    long long gx = (m + 255) / 256;
    if (gx > 65535) gx = 65535;
    long long gy = n < 65535 ? n : 65535;
    if (gx < 1) gx = 1;
    if (gy < 1) gy = 1;
    return dim3((unsigned)gx, (unsigned)gy);
}

}  // namespace

hipError_t split_planes(int npl, const float* X, long long ld, int mode, int rows, int K,
                        const int* e, _Float16* out, long long ldp, long long sp, int kshift,
                        hipStream_t s, int* e_used, bool lower_only)
{
    if (rows <= 0) return hipSuccess;
    if (ldp % 64 || kshift < 0 || ldp < kshift + K || npl < 1 || npl > 3 ||
        (reinterpret_cast<uintptr_t>(out) % 16) || (npl > 1 && sp % 8))
        return hipErrorInvalidValue;
    dim3 g((unsigned)(ldp / 64), (unsigned)((rows + 63) / 64));
    if (npl == 1)
        hipLaunchKernelGGL(split_planes_kernel<1>, g, dim3(256), 0, s, X, ld, mode, rows, K, e,
                           e_used, out, ldp, sp, kshift, lower_only && mode == kOpTriL);
    else if (npl == 2)
        hipLaunchKernelGGL(split_planes_kernel<2>, g, dim3(256), 0, s, X, ld, mode, rows, K, e,
                           e_used, out, ldp, sp, kshift, lower_only && mode == kOpTriL);
    else
        hipLaunchKernelGGL(split_planes_kernel<3>, g, dim3(256), 0, s, X, ld, mode, rows, K, e,
                           e_used, out, ldp, sp, kshift, lower_only && mode == kOpTriL);
    return hipGetLastError();
}

hipError_t fill_i32(int* p, long long n, int v, hipStream_t s)
{
    if (n <= 0) return hipSuccess;
    long long g = (n + 255) / 256;
    if (g > 4096) g = 4096;
    hipLaunchKernelGGL(fill_i32_kernel, dim3((unsigned)g), dim3(256), 0, s, p, n, v);
    return hipGetLastError();
}

hipError_t rowexp(bool f64, const void* X, long long ld, int mode, int rows, int K, int nb,
                  long long sx, int* out, long long so, hipStream_t s)
{
    if (rows <= 0 || K <= 0 || nb <= 0) return hipSuccess;
    if (mode == kOpN || mode == kOpTriL || mode == kOpSymL) {
        // k chunks of 256 columns, halved (down to 32) while the grid has fewer than 1024
        // workgroups: a thread's chunk is a chain of dependent rounds of 8 loads, so
        // small passes (the 256-column blocks of a TRSM recursion) need the shorter chains
        const long long rb = (rows + 255) / 256;
        int chunk = kRowChunk;
        while (chunk > 32 && rb * ((K + chunk - 1) / chunk) * nb < 1024) chunk /= 2;
        dim3 g((unsigned)rb, (K + chunk - 1) / chunk, nb);
        const bool lower = mode != kOpN;
        if (f64)
            hipLaunchKernelGGL(rowexp_mn_kernel<double>, g, dim3(256), 0, s,
                               (const double*)X, ld, rows, K, lower, sx, out, so, chunk);
        else
            hipLaunchKernelGGL(rowexp_mn_kernel<float>, g, dim3(256), 0, s,
                               (const float*)X, ld, rows, K, lower, sx, out, so, chunk);
        hipError_t e = hipGetLastError();
        if (e != hipSuccess || mode != kOpSymL) return e;
    }
    if (mode == kOpT || mode == kOpSymL) {
        // kOpSymL: the entries right of the diagonal of row r are column r below it
        dim3 g((rows + 3) / 4, (K + kColChunk - 1) / kColChunk, nb);
        const bool from_diag = mode == kOpSymL;
        if (f64)
            hipLaunchKernelGGL(rowexp_k_kernel<double>, g, dim3(256), 0, s,
                               (const double*)X, ld, rows, K, from_diag, sx, out, so);
        else
            hipLaunchKernelGGL(rowexp_k_kernel<float>, g, dim3(256), 0, s,
                               (const float*)X, ld, rows, K, from_diag, sx, out, so);
        return hipGetLastError();
    }
    return hipSuccess;
}

hipError_t trtri_lower_t(bool f64, const void* L, long long lda, long long bstride, int s,
                         int nblocks, void* Z, void* ldiag, hipStream_t st)
{
    if (s <= 0 || nblocks <= 0) return hipSuccess;
    if (s > 256) return hipErrorInvalidValue;
    dim3 g((s + kTrtriCols - 1) / kTrtriCols, nblocks);
    if (f64)
        hipLaunchKernelGGL((trtri_lower_t_kernel<double>), g, dim3(64), 0, st,
                           (const double*)L, lda, bstride, s, (double*)Z, (double*)ldiag);
    else
        hipLaunchKernelGGL((trtri_lower_t_kernel<float>), g, dim3(64), 0, st,
                           (const float*)L, lda, bstride, s, (float*)Z, (float*)ldiag);
    return hipGetLastError();
}

hipError_t trtri_store(bool f64, const void* Z, const void* ldiag, int s, void* A, long long lda,
                       bool unit, hipStream_t st)
{
    if (s <= 0) return hipSuccess;
    if (f64)
        hipLaunchKernelGGL(trtri_store_kernel<double>, dim3(s), dim3(256), 0, st,
                           (const double*)Z, (const double*)ldiag, s, (double*)A, lda, (int)unit);
    else
        hipLaunchKernelGGL(trtri_store_kernel<float>, dim3(s), dim3(256), 0, st,
                           (const float*)Z, (const float*)ldiag, s, (float*)A, lda, (int)unit);
    return hipGetLastError();
}

hipError_t diag_zero_scan(bool f64, const void* A, long long lda, long long n, int* info,
                          hipStream_t st)
{
    if (n <= 0) return hipSuccess;
    const dim3 g((unsigned)std::min<long long>((n + 255) / 256, 1024));
    if (f64)
        hipLaunchKernelGGL(diag_zero_kernel<double>, g, dim3(256), 0, st, (const double*)A, lda, n, info);
    else
        hipLaunchKernelGGL(diag_zero_kernel<float>, g, dim3(256), 0, st, (const float*)A, lda, n, info);
    return hipGetLastError();
}

hipError_t s2h(long long m, long long n, const float* as, long long ldas, void* ah, long long ldah,
               hipStream_t s)
{
    if (m <= 0 || n <= 0) return hipSuccess;
    hipLaunchKernelGGL(s2h_kernel, grid2d(m, n), dim3(256), 0, s, m, n, as, ldas, (_Float16*)ah,
                       ldah);
    return hipGetLastError();
}

namespace {

// one 32 x 32 tile of dst per block (32 x 8 threads): the src block it maps from is
// read column by column into LDS, then dst is written column by column
template <typename T>
__global__ void __launch_bounds__(256)
xform_kernel(const T* __restrict__ src, long long lds, T* __restrict__ dst, long long ldd, int rows,
             int cols, int flags, T alpha)
{
    __shared__ T t[32][33];
    const bool tr = flags & kXformTrans, rr = flags & kXformRevRows, rc = flags & kXformRevCols;
    const int bi = blockIdx.x * 32, bj = blockIdx.y * 32;
    const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
    // dst rows bi.. map to i' in [a0, a0+31], dst cols bj.. to j' in [b0, b0+31]
    const int a0 = rr ? rows - bi - 32 : bi, b0 = rc ? cols - bj - 32 : bj;
    const int sr0 = tr ? b0 : a0, sc0 = tr ? a0 : b0;             // src block origin
    const int srows = tr ? cols : rows, scols = tr ? rows : cols;
    T v[4];                                   // all four loads in flight (clamped, masked)
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        const int si = sr0 + tx, sj = sc0 + ty + 8 * q;
        const bool in = si >= 0 && si < srows && sj >= 0 && sj < scols;
        v[q] = (T)masked_load(src + (in ? si + (long long)sj * lds : 0ll), in);
    }
#pragma unroll
    for (int q = 0; q < 4; ++q) t[ty + 8 * q][tx] = v[q];
    __syncthreads();
    for (int y = ty; y < 32; y += 8) {
        const int i = bi + tx, j = bj + y;
        if (i >= rows || j >= cols) continue;
        const int ip = (rr ? rows - 1 - i : i) - a0, jp = (rc ? cols - 1 - j : j) - b0;   // in [0, 32)
        const T v = tr ? t[ip][jp] : t[jp][ip];
        dst[i + (long long)j * ldd] = ((flags & kXformUnitDiag) && i == j) ? T(1) : alpha * v;
    }
}

}  // namespace

hipError_t xform(bool f64, const void* src, long long lds, void* dst, long long ldd, int rows, int cols,
                 int flags, double alpha, hipStream_t s)
{
    if (rows <= 0 || cols <= 0) return hipSuccess;
    const dim3 g((rows + 31) / 32, (cols + 31) / 32);
    if (f64)
        hipLaunchKernelGGL(xform_kernel<double>, g, dim3(256), 0, s, (const double*)src, lds,
                           (double*)dst, ldd, rows, cols, flags, alpha);
    else
        hipLaunchKernelGGL(xform_kernel<float>, g, dim3(256), 0, s, (const float*)src, lds,
                           (float*)dst, ldd, rows, cols, flags, (float)alpha);
    return hipGetLastError();
}

hipError_t s2h_transpose(long long m, long long n, const float* as, void* ah, hipStream_t s)
{
    if (m <= 0 || n <= 0) return hipSuccess;
    dim3 g((unsigned)((n + 31) / 32), (unsigned)((m + 31) / 32));
    hipLaunchKernelGGL(s2h_transpose_kernel, g, dim3(256), 0, s, m, n, as, (_Float16*)ah);
    return hipGetLastError();
}

hipError_t mirror_lower(bool f64, long long n, void* A, long long lda, hipStream_t s)
{
    if (n <= 0) return hipSuccess;
    dim3 g((unsigned)((n + 31) / 32), (unsigned)((n + 31) / 32));
    if (f64)
        hipLaunchKernelGGL(mirror_lower_kernel<double>, g, dim3(256), 0, s, n, (double*)A, lda);
    else
        hipLaunchKernelGGL(mirror_lower_kernel<float>, g, dim3(256), 0, s, n, (float*)A, lda);
    return hipGetLastError();
}

hipError_t fill_matrix(bool f64, long long m, long long n, void* a, long long lda, double v,
                       hipStream_t s)
{
    if (m <= 0 || n <= 0) return hipSuccess;
    if (f64)
        hipLaunchKernelGGL(fill_matrix_kernel<double>, grid2d(m, n), dim3(256), 0, s, m, n,
                           (double*)a, lda, v);
    else
        hipLaunchKernelGGL(fill_matrix_kernel<float>, grid2d(m, n), dim3(256), 0, s, m, n,
                           (float*)a, lda, (float)v);
    return hipGetLastError();
}

hipError_t copy_matrix(bool a_f64, bool b_f64, long long m, long long n, const void* a,
                       long long lda, void* b, long long ldb, hipStream_t s)
{
    if (m <= 0 || n <= 0) return hipSuccess;
    dim3 g = grid2d(m, n);
    if (a_f64 && b_f64)
        hipLaunchKernelGGL((copy_matrix_kernel<double, double>), g, dim3(256), 0, s, m, n,
                           (const double*)a, lda, (double*)b, ldb);
    else if (a_f64)
        hipLaunchKernelGGL((copy_matrix_kernel<double, float>), g, dim3(256), 0, s, m, n,
                           (const double*)a, lda, (float*)b, ldb);
    else if (b_f64)
        hipLaunchKernelGGL((copy_matrix_kernel<float, double>), g, dim3(256), 0, s, m, n,
                           (const float*)a, lda, (double*)b, ldb);
    else
        hipLaunchKernelGGL((copy_matrix_kernel<float, float>), g, dim3(256), 0, s, m, n,
                           (const float*)a, lda, (float*)b, ldb);
    return hipGetLastError();
}

hipError_t copy_lower(bool a_f64, bool b_f64, long long n, const void* a, long long lda, void* b,
                      long long ldb, hipStream_t s)
{
    if (n <= 0) return hipSuccess;
    dim3 g = grid2d(n, n);
    if (a_f64 && b_f64)
        hipLaunchKernelGGL((copy_lower_kernel<double, double>), g, dim3(256), 0, s, n,
                           (const double*)a, lda, (double*)b, ldb);
    else if (a_f64)
        hipLaunchKernelGGL((copy_lower_kernel<double, float>), g, dim3(256), 0, s, n,
                           (const double*)a, lda, (float*)b, ldb);
    else if (b_f64)
        hipLaunchKernelGGL((copy_lower_kernel<float, double>), g, dim3(256), 0, s, n,
                           (const float*)a, lda, (double*)b, ldb);
    else
        hipLaunchKernelGGL((copy_lower_kernel<float, float>), g, dim3(256), 0, s, n,
                           (const float*)a, lda, (float*)b, ldb);
    return hipGetLastError();
}

hipError_t add_matrix_f32(long long m, long long n, void* x, long long ldx, const float* y,
                          long long ldy, hipStream_t s)
{
    if (m <= 0 || n <= 0) return hipSuccess;
    hipLaunchKernelGGL(add_matrix_f32_kernel, grid2d(m, n), dim3(256), 0, s, m, n, (float*)x, ldx,
                       y, ldy);
    return hipGetLastError();
}

}  // namespace tb
