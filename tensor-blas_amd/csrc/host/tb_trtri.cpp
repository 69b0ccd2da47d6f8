// Triangular inverse in place, A := inv(A) (LAPACK dtrtri's contract), composed from the
// library's TRSM and the diagonal-block inversion kernel of the TRSM leaves.
//
//   lower, order s:  s <= 256: invert the diagonal block on the device in the matrix's
//                    precision (FP32 for the FP32-class tiers; trtri_lower_t, then
//                    trtri_store writes D^-1 (I + W))
//                    else, with A = [L11 . ; L21 L22] split at n1 (a multiple of 64 near
//                    s/2), inv(A) = [X11 . ; X21 X22] with X21 = -X22 L21 X11:
//                      X11, X22 recursively, in place
//                      T := L21 X11            tb_trmm_ex (right, lower, N) into workspace
//                      L21 := -X22 T           tb_trmm (left, lower, N: the native form)
//   upper:           inv(U) = inv(U^T)^T: the lower inverse of a transposed copy
//                    (workspace slot 3), copied back whole: its upper triangle holds A's
//                    lower triangle transposed, so that part of A is rewritten unchanged.
// The products are single triangular GEMM launches (no inner solves), in the caller's
// tier.  Work: n^3 / 3 flops.
// info: LAPACK's, i if A(i, i) is exactly zero (non-unit only); A is then unchanged.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cctype>
#include <climits>

#include "tb_internal.hpp"

namespace tb {

namespace {

constexpr long kTrtriLeaf = 256;    // largest block trtri_lower_t inverts

struct TrtriCtx {
    tb_handle_s* h;
    int tier;
    char* A;          // byte pointer to A(0, 0) of the lower-triangular matrix
    long lda;
    size_t es;        // element size
    long nb;
    char diag;
    char* T;          // n2 x n1 product buffer (ld n2) in workspace slot 3

    bool f64() const { return es == 8; }
    void* at(long i, long j) const { return A + (size_t)(i + j * lda) * es; }

    int leaf(long off, long s)
    {
        const bool unit = diag == 'U';
        const size_t zb = (size_t)s * s * es, db = (size_t)s * es;
        const size_t ob = unit ? ((size_t)s * s * es + 255) / 256 * 256 : 0;
        char* ws = nullptr;
        TB_TRY(workspace_get(h, ob + zb + db + 256, &ws, 2));
        const void* src = at(off, off);
        long ld = lda;
        if (unit) {      // a copy with ones on the diagonal: A's diagonal is not referenced
            TB_HIP(xform(f64(), src, lda, ws, s, (int)s, (int)s, kXformUnitDiag, 1.0, h->stream));
            src = ws;
            ld = s;
        }
        // in the matrix's own precision: FP32 blocks for the FP32-class tiers
        char* Z = ws + ob;
        char* ldiag = ws + ob + zb;
        hipEvent_t e0;
        TB_TRY(prof_begin(h, &e0));
        TB_HIP(trtri_lower_t(f64(), src, ld, 0, (int)s, 1, Z, ldiag, h->stream));
        TB_HIP(trtri_store(f64(), Z, ldiag, (int)s, at(off, off), lda, unit, h->stream));
        return prof_end(h, kProfTrtri, e0, (double)s * s * s / 3.0);
    }

    int rec(long off, long s)
    {
        if (s <= kTrtriLeaf) return leaf(off, s);
        const long n1 = std::min(s - 1, (s / 2 + 63) / 64 * 64);
        const long n2 = s - n1;
        TB_TRY(rec(off, n1));
        TB_TRY(rec(off + n1, n2));
        void* L21 = at(off + n1, off);
        TB_TRY(trmm_ex(h, tier, 'C', 'R', 'L', 'N', diag, n2, n1, 1.0, at(off, off), lda, L21, lda, T, n2,
                       nb));
        return trmm_ex(h, tier, 'C', 'L', 'L', 'N', diag, n2, n1, -1.0, at(off + n1, off + n1), lda, T, n2,
                       L21, lda, nb);
    }
};

}  // namespace

int trtri(tb_handle_s* h, int tier, char uplo, char diag, long n, void* A, long lda, long nb, long* info)
{
    const char ul = (char)std::toupper((unsigned char)uplo), dg = (char)std::toupper((unsigned char)diag);
    if (!h || !info || tier < kPlainF16 || tier > kMfmaF64 || (ul != 'L' && ul != 'U') ||
        (dg != 'N' && dg != 'U') || n < 0 || n >= (1l << 31) || nb <= 0 || lda < std::max(1l, n) ||
        (n > 0 && !A))
        return TB_ERR_INVALID;
    *info = 0;
    if (n == 0) return TB_SUCCESS;
    const bool f64 = (tier == kOzakiI8 || tier == kMfmaF64);
    const size_t es = f64 ? 8 : 4;
    if (dg == 'N') {     // singular: report the first zero on the diagonal, change nothing
        if (!h->dev_info) TB_HIP(hipMalloc(&h->dev_info, sizeof(int)));
        TB_HIP(fill_i32(h->dev_info, 1, INT_MAX, h->stream));
        TB_HIP(diag_zero_scan(f64, A, lda, n, h->dev_info, h->stream));
        int v = 0;
        TB_HIP(hipMemcpyAsync(&v, h->dev_info, sizeof(int), hipMemcpyDeviceToHost, h->stream));
        TB_HIP(hipStreamSynchronize(h->stream));
        if (v != INT_MAX) {
            *info = v;
            return TB_SUCCESS;
        }
    }
    // slot 3: [transposed copy of A (upper only)] [product buffer, <= n/2 x (n/2 + 64)]
    const size_t cb = ul == 'U' ? ((size_t)n * n * es + 255) / 256 * 256 : 0;
    const size_t tb = n > kTrtriLeaf ? (size_t)(n / 2 + 64) * (n / 2 + 64) * es : 0;
    char* ws = nullptr;
    TB_TRY(workspace_get(h, cb + tb + 256, &ws, 3));
    if (ul == 'L') {
        TrtriCtx c{h, tier, static_cast<char*>(A), lda, es, nb, dg, ws};
        return c.rec(0, n);
    }
    TB_HIP(xform(f64, A, lda, ws, n, (int)n, (int)n, kXformTrans, 1.0, h->stream));
    TrtriCtx c{h, tier, ws, n, es, nb, dg, ws + cb};
    TB_TRY(c.rec(0, n));
    return hip_status(xform(f64, ws, n, A, lda, (int)n, (int)n, kXformTrans, 1.0, h->stream));
}

}  // namespace tb
