// The BLAS option sets around the reference's fixed forms (a "next" step of the
// hot-path table: uplo = U, trans = T, side = L/R, diag = U).  The native routines
// solve or multiply with a lower-triangular (or lower-stored symmetric) operand in one
// orientation:
//   TRSM  X M^T = B        (right, lower, T, non-unit; BLAS/tc_trsm.cu:31-137)
//   TRMM  C = alpha M B    (left, lower, N, non-unit, out of place; BLAS/tc_trmm.cu)
//   SYMM  C = alpha A B + beta C, A lower-stored (left; BLAS/tc_symm.cu)
//   SYRK / SYR2K  lower triangle of C (trans N; BLAS/tc_syrk.cu, tc_syr2k.cu)
// Every other option set is one of these on transformed operands, by three identities
// (J reverses the order of rows or columns):
//   * transposition: op(A) X = B  <=>  X^T op(A)^T = B^T, and C^T of a symmetric
//     product is its other triangle;
//   * reversal: J U J is lower for U upper, so X U = B <=> (X J)(J U J) = B J and
//     U B = J (J U J)(J B);
//   * a unit diagonal is written into the copy of the triangular operand.
// The copies (xform, one pass over O(n^2) or O(mn) elements) cost little against the
// O(mn^2) routines; they live in the handle's workspace slot 2.  Results are those of
// the native routine on the transformed problem, so accuracy is the tier's.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cctype>

#include "tb_internal.hpp"

namespace tb {

namespace {

bool f64_of(int tier) { return tier == kOzakiI8 || tier == kMfmaF64; }

int norm_opt(char c, const char* allowed, char* out)
{
    const char u = (char)std::toupper((unsigned char)c);
    for (const char* a = allowed; *a; ++a)
        if (u == *a) {
            *out = u == 'C' ? 'T' : u;      // real data: conjugate transpose = transpose
            return TB_SUCCESS;
        }
    return TB_ERR_INVALID;
}

bool dims_fit(long v) { return v >= 0 && v < (1l << 31); }

// storage order 'C' (column-major) or 'R' (row-major).  A row-major matrix is the
// column-major storage of its transpose, so a row-major call is the column-major call
// with the other side and the other triangle, m and n exchanged (TRSM, TRMM, SYMM), and
// the other trans (SYRK, SYR2K: C's transpose is C): the CBLAS mapping.
int row_major(char order, char* side, char* uplo, char* trans, long* m, long* n)
{
    const char o = (char)std::toupper((unsigned char)order);
    if (o == 'C') return TB_SUCCESS;
    if (o != 'R') return TB_ERR_INVALID;
    auto flip = [](char* c, char a, char b) {
        if (c) *c = *c == a ? b : a;
    };
    flip(side, 'L', 'R');
    flip(uplo, 'L', 'U');
    flip(trans, 'N', 'T');
    if (m && n) std::swap(*m, *n);
    return TB_SUCCESS;
}

// carves workspace slot 2 into consecutive buffers
struct ExWs {
    tb_handle_s* h;
    size_t es, total = 0;
    size_t add(long rows, long cols)
    {
        const size_t o = total;
        total += ((size_t)std::max(rows, 1l) * std::max(cols, 1l) * es + 255) / 256 * 256;
        return o;
    }
    int get(char** base) { return workspace_get(h, std::max<size_t>(total, 256), base, 2); }
};

int xf(tb_handle_s* h, bool f64, const void* src, long lds, void* dst, long ldd, long rows, long cols,
       int flags, double alpha = 1.0)
{
    return hip_status(xform(f64, src, lds, dst, ldd, (int)rows, (int)cols, flags, alpha, h->stream));
}

int zero_fill(tb_handle_s* h, size_t es, void* B, long ldb, long m, long n)
{
    return hip_status(hipMemset2DAsync(B, (size_t)ldb * es, 0, (size_t)m * es, (size_t)n, h->stream));
}

}  // namespace

int trsm_ex(tb_handle_s* h, int tier, char order, char side, char uplo, char trans, char diag, long m, long n,
            double alpha, const void* A, long lda, void* B, long ldb, long nb)
{
    char sd, ul, tr, dg;
    if (!h || tier < kPlainF16 || tier > kMfmaF64 || norm_opt(side, "LR", &sd) || norm_opt(uplo, "LU", &ul) ||
        norm_opt(trans, "NTC", &tr) || norm_opt(diag, "NU", &dg) || !dims_fit(m) || !dims_fit(n) || nb <= 0 ||
        row_major(order, &sd, &ul, nullptr, &m, &n))
        return TB_ERR_INVALID;
    const long na = sd == 'L' ? m : n;                     // order of A
    if (lda < std::max(1l, na) || ldb < std::max(1l, m)) return TB_ERR_INVALID;
    if (m == 0 || n == 0) return TB_SUCCESS;
    if (!B || (alpha != 0.0 && !A)) return TB_ERR_INVALID;
    const bool f64 = f64_of(tier);
    const size_t es = f64 ? 8 : 4;
    if (alpha == 0.0) return zero_fill(h, es, B, ldb, m, n);
    // right-side problem X' op'(A) = alpha B' with B' mm x nn: the left side works on B^T
    const bool left = sd == 'L';
    const long mm = left ? n : m, nn = left ? m : n;
    const char tp = left ? (tr == 'N' ? 'T' : 'N') : tr;
    // M lower with X' M^T = B'' (B'' = B' J when the form needs reversal)
    int mflags = 0;
    bool rev = false, copyM = false;
    if (ul == 'L' && tp == 'T') {                          // native
        copyM = dg == 'U';
    } else if (ul == 'U' && tp == 'N') {                   // X U = B: M = U^T
        mflags = kXformTrans;
        copyM = true;
    } else if (ul == 'L' && tp == 'N') {                   // X L = B: M = J L^T J, B J
        mflags = kXformTrans | kXformRevRows | kXformRevCols;
        rev = copyM = true;
    } else {                                               // X U^T = B: M = J U J, B J
        mflags = kXformRevRows | kXformRevCols;
        rev = copyM = true;
    }
    if (dg == 'U') mflags |= kXformUnitDiag;
    const bool copyB = left || rev || alpha != 1.0;
    ExWs w{h, es};
    const size_t oM = copyM ? w.add(na, na) : 0;
    const size_t oB = copyB ? w.add(mm, nn) : 0;
    char* ws = nullptr;
    TB_TRY(w.get(&ws));
    const void* M = A;
    long ldm = lda;
    if (copyM) {
        TB_TRY(xf(h, f64, A, lda, ws + oM, na, na, na, mflags));
        M = ws + oM;
        ldm = na;
    }
    void* Bp = B;
    long ldbp = ldb;
    // B -> B' and back: left: transpose (and reverse the columns of B'); right: reverse
    const int fwd = (left ? kXformTrans : 0) | (rev ? kXformRevCols : 0);
    const int back = left ? (kXformTrans | (rev ? kXformRevRows : 0)) : (rev ? kXformRevCols : 0);
    if (copyB) {
        TB_TRY(xf(h, f64, B, ldb, ws + oB, mm, mm, nn, fwd, alpha));
        Bp = ws + oB;
        ldbp = mm;
    }
    TB_TRY(trsm(h, tier, mm, nn, M, ldm, Bp, ldbp, nb));
    if (copyB) TB_TRY(xf(h, f64, Bp, ldbp, B, ldb, m, n, back));
    return TB_SUCCESS;
}

int trmm_ex(tb_handle_s* h, int tier, char order, char side, char uplo, char trans, char diag, long m, long n,
            double alpha, const void* A, long lda, const void* B, long ldb, void* C, long ldc, long nb)
{
    char sd, ul, tr, dg;
    if (!h || tier < kPlainF16 || tier > kMfmaF64 || norm_opt(side, "LR", &sd) || norm_opt(uplo, "LU", &ul) ||
        norm_opt(trans, "NTC", &tr) || norm_opt(diag, "NU", &dg) || !dims_fit(m) || !dims_fit(n) || nb <= 0 ||
        row_major(order, &sd, &ul, nullptr, &m, &n))
        return TB_ERR_INVALID;
    const long na = sd == 'L' ? m : n;
    if (lda < std::max(1l, na) || ldb < std::max(1l, m) || ldc < std::max(1l, m)) return TB_ERR_INVALID;
    if (m == 0 || n == 0) return TB_SUCCESS;
    if (!C || (alpha != 0.0 && (!A || !B))) return TB_ERR_INVALID;
    const bool f64 = f64_of(tier);
    const size_t es = f64 ? 8 : 4;
    if (alpha == 0.0) return zero_fill(h, es, C, ldc, m, n);
    // left-side problem C' = alpha op'(A) B' with B', C' mm x nn: the right side works on
    // the transposes (C^T = alpha op(A)^T B^T)
    const bool right = sd == 'R';
    const long mm = right ? n : m, nn = right ? m : n;
    const char tp = right ? (tr == 'N' ? 'T' : 'N') : tr;
    int mflags = 0;
    bool rev = false, copyM = false;
    if (ul == 'L' && tp == 'N') {                          // native
        copyM = dg == 'U';
    } else if (ul == 'U' && tp == 'T') {                   // U^T B: M = U^T
        mflags = kXformTrans;
        copyM = true;
    } else if (ul == 'U' && tp == 'N') {                   // U B = J (J U J)(J B)
        mflags = kXformRevRows | kXformRevCols;
        rev = copyM = true;
    } else {                                               // L^T B = J (J L^T J)(J B)
        mflags = kXformTrans | kXformRevRows | kXformRevCols;
        rev = copyM = true;
    }
    if (dg == 'U') mflags |= kXformUnitDiag;
    const bool copyBC = right || rev;
    ExWs w{h, es};
    const size_t oM = copyM ? w.add(na, na) : 0;
    const size_t oB = copyBC ? w.add(mm, nn) : 0;
    const size_t oC = copyBC ? w.add(mm, nn) : 0;
    char* ws = nullptr;
    TB_TRY(w.get(&ws));
    const void* M = A;
    long ldm = lda;
    if (copyM) {
        TB_TRY(xf(h, f64, A, lda, ws + oM, na, na, na, mflags));
        M = ws + oM;
        ldm = na;
    }
    if (!copyBC) return trmm(h, tier, m, n, alpha, M, ldm, B, ldb, C, ldc, nb);
    // B' = (right ? B^T : B) with rows reversed when rev; C = the inverse map of C'
    const int fwd = (right ? kXformTrans : 0) | (rev ? kXformRevRows : 0);
    const int back = right ? (kXformTrans | (rev ? kXformRevCols : 0)) : (rev ? kXformRevRows : 0);
    TB_TRY(xf(h, f64, B, ldb, ws + oB, mm, mm, nn, fwd));
    TB_TRY(trmm(h, tier, mm, nn, alpha, M, ldm, ws + oB, mm, ws + oC, mm, nb));
    return xf(h, f64, ws + oC, mm, C, ldc, m, n, back);
}

int symm_ex(tb_handle_s* h, int tier, char order, char side, char uplo, long m, long n, double alpha, const void* A,
            long lda, const void* B, long ldb, double beta, void* C, long ldc)
{
    char sd, ul;
    if (!h || tier < kPlainF16 || tier > kMfmaF64 || norm_opt(side, "LR", &sd) || norm_opt(uplo, "LU", &ul) ||
        !dims_fit(m) || !dims_fit(n) || row_major(order, &sd, &ul, nullptr, &m, &n))
        return TB_ERR_INVALID;
    const long na = sd == 'L' ? m : n;
    if (lda < std::max(1l, na) || ldb < std::max(1l, m) || ldc < std::max(1l, m)) return TB_ERR_INVALID;
    if (m == 0 || n == 0) return TB_SUCCESS;
    // null operands are rejected before any operand copy launches
    if (!A || !B || !C) return TB_ERR_INVALID;
    if (sd == 'L' && ul == 'L') return symm(h, tier, m, n, alpha, A, lda, B, ldb, beta, C, ldc);
    const bool f64 = f64_of(tier);
    const size_t es = f64 ? 8 : 4;
    // left problem C' = alpha A B' + beta C' (C^T = alpha A B^T + beta C^T on the right)
    const bool right = sd == 'R';
    const long mm = right ? n : m, nn = right ? m : n;
    ExWs w{h, es};
    const size_t oA = ul == 'U' ? w.add(na, na) : 0;
    const size_t oB = right ? w.add(mm, nn) : 0;
    const size_t oC = right ? w.add(mm, nn) : 0;
    char* ws = nullptr;
    TB_TRY(w.get(&ws));
    const void* Al = A;
    long ldal = lda;
    if (ul == 'U') {                                       // the lower triangle of A^T
        TB_TRY(xf(h, f64, A, lda, ws + oA, na, na, na, kXformTrans));
        Al = ws + oA;
        ldal = na;
    }
    if (!right) return symm(h, tier, m, n, alpha, Al, ldal, B, ldb, beta, C, ldc);
    TB_TRY(xf(h, f64, B, ldb, ws + oB, mm, mm, nn, kXformTrans));
    if (beta != 0.0) TB_TRY(xf(h, f64, C, ldc, ws + oC, mm, mm, nn, kXformTrans));
    TB_TRY(symm(h, tier, mm, nn, alpha, Al, ldal, ws + oB, mm, beta, ws + oC, mm));
    return xf(h, f64, ws + oC, mm, C, ldc, m, n, kXformTrans);
}

namespace {

// lower-triangle routine on C (uplo L) or on a transposed copy of C (uplo U): C's
// upper triangle is the lower triangle of C^T; copying the whole C^T back rewrites
// C's lower triangle with its own values
template <class F>
int on_triangle(tb_handle_s* h, bool f64, char ul, long n, void* C, long ldc, size_t oC, char* ws, F&& run)
{
    if (ul == 'L') return run(C, ldc);
    TB_TRY(xf(h, f64, C, ldc, ws + oC, n, n, n, kXformTrans));
    TB_TRY(run(ws + oC, n));
    return xf(h, f64, ws + oC, n, C, ldc, n, n, kXformTrans);
}

}  // namespace

int syrk_ex(tb_handle_s* h, int tier, char order, char uplo, char trans, long n, long k, double alpha, const void* A,
            long lda, double beta, void* C, long ldc, long nb)
{
    char ul, tr;
    if (!h || tier < kPlainF16 || tier > kMfmaF64 || norm_opt(uplo, "LU", &ul) || norm_opt(trans, "NTC", &tr) ||
        !dims_fit(n) || !dims_fit(k) || nb <= 0 || row_major(order, nullptr, &ul, &tr, nullptr, nullptr))
        return TB_ERR_INVALID;
    if (ldc < std::max(1l, n) || lda < std::max(1l, tr == 'N' ? n : k)) return TB_ERR_INVALID;
    if (n == 0) return TB_SUCCESS;
    if (!C || (k > 0 && !A)) return TB_ERR_INVALID;
    const bool f64 = f64_of(tier);
    ExWs w{h, f64 ? 8u : 4u};
    const size_t oC = ul == 'U' ? w.add(n, n) : 0;
    char* ws = nullptr;
    if (ul == 'U') TB_TRY(w.get(&ws));
    return on_triangle(h, f64, ul, n, C, ldc, oC, ws, [&](void* Ct, long ldt) {
        return syrk(h, tier, n, k, alpha, A, lda, tr == 'T', beta, Ct, ldt, nb);
    });
}

int syr2k_ex(tb_handle_s* h, int tier, char order, char uplo, char trans, long n, long k, double alpha, const void* A,
             long lda, const void* B, long ldb, double beta, void* C, long ldc, long nb)
{
    char ul, tr;
    if (!h || tier < kPlainF16 || tier > kMfmaF64 || norm_opt(uplo, "LU", &ul) || norm_opt(trans, "NTC", &tr) ||
        !dims_fit(n) || !dims_fit(k) || nb <= 0 || row_major(order, nullptr, &ul, &tr, nullptr, nullptr))
        return TB_ERR_INVALID;
    const long ar = tr == 'N' ? n : k;
    if (ldc < std::max(1l, n) || lda < std::max(1l, ar) || ldb < std::max(1l, ar)) return TB_ERR_INVALID;
    if (n == 0) return TB_SUCCESS;
    if (!C || (k > 0 && (!A || !B))) return TB_ERR_INVALID;
    const bool f64 = f64_of(tier);
    ExWs w{h, f64 ? 8u : 4u};
    const size_t oA = tr == 'T' ? w.add(n, k) : 0;
    const size_t oB = tr == 'T' ? w.add(n, k) : 0;
    const size_t oC = ul == 'U' ? w.add(n, n) : 0;
    char* ws = nullptr;
    if (tr == 'T' || ul == 'U') TB_TRY(w.get(&ws));
    const void* An = A;
    const void* Bn = B;
    long ldan = lda, ldbn = ldb;
    if (tr == 'T') {                                       // n x k operands A^T, B^T
        if (k > 0) {
            TB_TRY(xf(h, f64, A, lda, ws + oA, n, n, k, kXformTrans));
            TB_TRY(xf(h, f64, B, ldb, ws + oB, n, n, k, kXformTrans));
            An = ws + oA;
            Bn = ws + oB;
        }
        ldan = ldbn = n;
    }
    return on_triangle(h, f64, ul, n, C, ldc, oC, ws, [&](void* Ct, long ldt) {
        return syr2k(h, tier, n, k, alpha, An, ldan, Bn, ldbn, beta, Ct, ldt, nb);
    });
}

}  // namespace tb
