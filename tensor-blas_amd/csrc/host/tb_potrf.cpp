// Recursive Cholesky factorization A = L L^T (lower), composed from the library's own
// SYRK and TRSM: the downstream caller these BLAS-3 routines exist for.
//
//   potrf(A):  n <= 128: factor the diagonal block on the device in the matrix's
//              precision, FP32 or FP64 (potrf.hip)
//              else, with A = [A11 . ; A21 A22] split at n1 (a multiple of 64 near n/2):
//                potrf(A11)
//                A21 := A21 L11^-T          tb_trsm  (X L^T = B, right/lower/T)
//                A22 := A22 - A21 A21^T     tb_syrk  (alpha = -1, beta = 1, lower)
//                potrf(A22)
//
// Every step runs in the caller's tier, so the split tier factors at FP32 accuracy on
// FP16 MFMA and the Ozaki tier at FP64 accuracy on INT8 MFMA.  Work: n^3 / 3 flops.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "tb_internal.hpp"

namespace tb {

namespace {

struct PotrfCtx {
    tb_handle_s* h;
    int tier;
    char* A;          // byte pointer to A(0, 0)
    long lda;
    size_t es;        // element size
    long nb;

    void* at(long i, long j) const { return A + (size_t)(i + j * lda) * es; }

    int leaf(long off, long s)
    {
        hipEvent_t e0;
        TB_TRY(prof_begin(h, &e0));
        TB_HIP(potrf_leaf(es == 8, at(off, off), lda, (int)s, off, h->dev_info, h->stream));
        return prof_end(h, kProfPotrfLeaf, e0, (double)s * s * s / 3.0);
    }

    int rec(long off, long s)
    {
        if (s <= kPotrfLeaf) return leaf(off, s);
        const long n1 = std::min(s - 1, (s / 2 + 63) / 64 * 64);
        const long n2 = s - n1;
        TB_TRY(rec(off, n1));
        TB_TRY(trsm(h, tier, n2, n1, at(off, off), lda, at(off + n1, off), lda, nb));
        TB_TRY(syrk(h, tier, n2, n1, -1.0, at(off + n1, off), lda, false, 1.0, at(off + n1, off + n1),
                    lda, nb));
        return rec(off + n1, n2);
    }
};

}  // namespace

int potrf(tb_handle_s* h, int tier, long n, void* A, long lda, long nb, long* info)
{
    if (!h || !info || tier < kPlainF16 || tier > kMfmaF64 || n < 0 || n >= (1l << 31) ||
        nb <= 0 || lda < std::max(1l, n) || (n > 0 && !A))
        return TB_ERR_INVALID;
    *info = 0;
    if (n == 0) return TB_SUCCESS;
    if (!h->dev_info) TB_HIP(hipMalloc(&h->dev_info, sizeof(int)));
    TB_HIP(fill_i32(h->dev_info, 1, 0, h->stream));
    const bool f64 = (tier == kOzakiI8 || tier == kMfmaF64);
    PotrfCtx c{h, tier, static_cast<char*>(A), lda, f64 ? 8u : 4u, nb};
    TB_TRY(c.rec(0, n));
    int v = 0;
    TB_HIP(hipMemcpyAsync(&v, h->dev_info, sizeof(int), hipMemcpyDeviceToHost, h->stream));
    TB_HIP(hipStreamSynchronize(h->stream));
    *info = v;
    return TB_SUCCESS;
}

}  // namespace tb
