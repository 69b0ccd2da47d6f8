/*
 * oracle.c -- CPU restatement of Tensor-BLAS's hot path, for TESTS ONLY.
 *
 * This file is test infrastructure.  Only tests/, __graft_entry__.smoke() and
 * bench.py's cpu_baseline leg may load it, and only as the checker.  The
 * product library (tensor-blas_amd/) never links or calls it.
 *
 * Contents, each citing the reference file:line it follows:
 *   1. chunk planners   find_mat_size_syrk / find_mat_size_trsm  (util/mat_size.cu:5-55)
 *   2. FP32->FP16 round-to-nearest-even and FP16->FP32          (__float2half in
 *      include/TensorBLAS.h:185-284 s2h_swpipe, util/util.cu:127-145 s2h/s2hTranspose)
 *      plus FP32->BF16 RNE (split-tier helper, no reference counterpart)
 *   3. FP64 BLAS-3 in the reference's fixed options (lower, N; TRSM right/lower/T/non-unit;
 *      TRMM left/lower/N/non-unit out of place; SYMM left, A stored lower):
 *      BLAS/tc_syrk.cu:41-125, tc_syr2k.cu:3-115, tc_trmm.cu:4-98, tc_trsm.cu:31-137,
 *      tc_symm.cu:3-20
 *   4. emulation of the plain tier's TRSM recursion (FP16-rounded update operands,
 *      exact leaves): BLAS/tc_trsm.cu:31-137
 *   5. Ozaki INT8 slicing and the slice-pair product (published method: Ozaki, Ogita,
 *      Oishi, Rump 2012; Ootomo, Ozaki, Yokota 2024 "DGEMM on integer matrix
 *      multiplication unit").  No reference counterpart: the reference calls cuBLAS
 *      FP64 (BLAS/tc_ozimmu_syrk.cu:33-99); parity for this tier is pinned by exact
 *      integer arithmetic and by the FP64 routines above.
 *   6. error metrics (normwise as testing/testing_syrk.cu:152, max-relative, componentwise)
 *   7. Cholesky factorization A = L L^T (lower), the downstream caller composed from
 *      SYRK and TRSM.  No reference counterpart (the reference stops at the BLAS-3
 *      routines); LAPACK dpotrf semantics, pinned by tests against numpy.linalg.cholesky
 *      and by the reconstruction L L^T = A.
 *   8. sampled columns of the FP64 products in extended precision (x87 long double,
 *      64-bit significand): the reference for the Ozaki tier's 1e-14 target at full size
 *      (BLAS/tc_ozimmu_syr2k.cu:3-31 computes them in FP64).  Pinned by exact rational
 *      arithmetic on small inputs.
 *
 * All matrices are column-major, leading dimensions in elements.
 */
#include <math.h>
#include <omp.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

/* ------------------------------------------------------------------ 1. planners */

/* util/mat_size.cu:5-29.  Greedy decomposition of n into 16384s, then one 8192,
 * then the remainder (< 8192).  Writes the chunks to out[] (largest first) and
 * returns the count.  The reference returns `length = count-1` and walks the
 * array last-first, so out[count-1] is the chunk at offset 0. */
static int plan_chunks(int64_t n, const int64_t *powers, int np, int64_t floor_,
                       int64_t *out)
{
    int cnt = 0;
    for (int i = np - 1; i >= -1; i--) {
        if (n < floor_) {
            if (n == 0) break;
            out[cnt++] = n;
            break;
        }
        /* the reference reads powers[-1] here when n exceeds the sum of all
         * powers (util/mat_size.cu:12, 20): outside its domain, reported as -1. */
        if (i < 0) return -1;
        if (n >= powers[i]) {
            out[cnt++] = powers[i];
            n -= powers[i];
        }
    }
    return cnt;
}

int orc_find_mat_size_syrk(int64_t n, int64_t *out)
{
    static const int64_t powers[] = {8192, 16384, 16384, 16384, 16384, 16384,
                                     16384, 16384, 16384, 16384, 16384, 16384};
    return plan_chunks(n, powers, 12, 8192, out);
}

/* util/mat_size.cu:31-55: powers {2048,4096,8192,16384 x9}, floor 2048. */
int orc_find_mat_size_trsm(int64_t n, int64_t *out)
{
    static const int64_t powers[] = {2048, 4096, 8192, 16384, 16384, 16384,
                                     16384, 16384, 16384, 16384, 16384, 16384};
    return plan_chunks(n, powers, 12, 2048, out);
}

/* ------------------------------------------------------------ 2. conversions */

static uint32_t f2u(float x) { uint32_t u; memcpy(&u, &x, 4); return u; }
static float u2f(uint32_t u) { float x; memcpy(&x, &u, 4); return x; }

/* IEEE-754 binary32 -> binary16, round to nearest, ties to even.  Written from
 * the IEEE-754 definition: overflow to +-inf at |x| >= 65520, gradual underflow
 * below 2^-14, NaN stays NaN (quiet), -0 stays -0. */
uint16_t orc_f32_to_f16(float x)
{
    uint32_t u = f2u(x);
    uint16_t sign = (uint16_t)((u >> 16) & 0x8000u);
    uint32_t ex = (u >> 23) & 0xffu;
    uint32_t man = u & 0x7fffffu;
    if (ex == 0xffu) {                           /* inf / nan */
        if (man == 0) return sign | 0x7c00u;
        return sign | 0x7e00u | (uint16_t)(man >> 13);
    }
    int32_t e = (int32_t)ex - 127;               /* unbiased exponent */
    if (e > 15) return sign | 0x7c00u;           /* certainly overflows */
    if (e >= -14) {                              /* normal half range */
        uint32_t hm = man >> 13;                 /* 10 mantissa bits */
        uint32_t rem = man & 0x1fffu;            /* 13 dropped bits */
        uint32_t h = ((uint32_t)(e + 15) << 10) | hm;
        if (rem > 0x1000u || (rem == 0x1000u && (hm & 1u))) h += 1u; /* carries into exponent ok */
        return sign | (uint16_t)h;               /* h == 0x7c00 means overflow to inf */
    }
    if (e < -25) return sign;                    /* below half of the smallest subnormal */
    /* subnormal: value = m * 2^-24, m = full significand >> shift */
    uint32_t full = man | 0x800000u;             /* 24-bit significand, value full*2^(e-23) */
    int shift = -e - 1;                          /* (e-23) - (-24) = e+1  -> shift right by -(e+1) */
    uint32_t hm = full >> shift;
    uint32_t rem = full & ((1u << shift) - 1u);
    uint32_t half = 1u << (shift - 1);
    if (rem > half || (rem == half && (hm & 1u))) hm += 1u;
    return sign | (uint16_t)hm;
}

float orc_f16_to_f32(uint16_t h)
{
    uint32_t sign = (uint32_t)(h & 0x8000u) << 16;
    uint32_t ex = (h >> 10) & 0x1fu;
    uint32_t man = h & 0x3ffu;
    if (ex == 0x1fu) return u2f(sign | 0x7f800000u | (man << 13));
    if (ex == 0) {
        if (man == 0) return u2f(sign);
        float v = (float)man * 5.9604644775390625e-08f; /* 2^-24, exact */
        return sign ? -v : v;
    }
    return u2f(sign | ((ex + 112u) << 23) | (man << 13));
}

/* binary32 -> bfloat16, round to nearest even (NaN kept quiet). */
uint16_t orc_f32_to_bf16(float x)
{
    uint32_t u = f2u(x);
    if ((u & 0x7f800000u) == 0x7f800000u && (u & 0x7fffffu))
        return (uint16_t)((u >> 16) | 0x40u);
    uint32_t lsb = (u >> 16) & 1u;
    u += 0x7fffu + lsb;
    return (uint16_t)(u >> 16);
}

void orc_f32_to_f16_array(int64_t n, const float *x, uint16_t *h)
{
    for (int64_t i = 0; i < n; i++) h[i] = orc_f32_to_f16(x[i]);
}

/* Round an FP32 array through FP16 and back (what s2h + the FP16 GEMM operand sees). */
void orc_round_f16_array(int64_t n, const float *x, double *out)
{
    for (int64_t i = 0; i < n; i++) out[i] = (double)orc_f16_to_f32(orc_f32_to_f16(x[i]));
}

/* ------------------------------------------------------------ 3. FP64 BLAS-3 */

/* General C = alpha*op(A)*op(B) + beta*C, m x n, inner k. transX: 0 = N, 1 = T. */
void orc_dgemm(int transa, int transb, int64_t m, int64_t n, int64_t k, double alpha,
               const double *A, int64_t lda, const double *B, int64_t ldb, double beta,
               double *C, int64_t ldc)
{
    for (int64_t j = 0; j < n; j++) {
        for (int64_t i = 0; i < m; i++) {
            double s = 0.0;
            for (int64_t p = 0; p < k; p++) {
                double a = transa ? A[p + i * lda] : A[i + p * lda];
                double b = transb ? B[j + p * ldb] : B[p + j * ldb];
                s += a * b;
            }
            double c = beta == 0.0 ? 0.0 : beta * C[i + j * ldc];
            C[i + j * ldc] = alpha * s + c;
        }
    }
}

/* SYRK lower, N: C := alpha*A*A^T + beta*C on i >= j only (BLAS/tc_syrk.cu:41-125
 * computes the same lower triangle; the strictly-upper part is not referenced). */
void orc_dsyrk_ln(int64_t n, int64_t k, double alpha, const double *A, int64_t lda,
                  double beta, double *C, int64_t ldc)
{
    for (int64_t j = 0; j < n; j++)
        for (int64_t i = j; i < n; i++) {
            double s = 0.0;
            for (int64_t p = 0; p < k; p++) s += A[i + p * lda] * A[j + p * lda];
            double c = beta == 0.0 ? 0.0 : beta * C[i + j * ldc];
            C[i + j * ldc] = alpha * s + c;
        }
}

/* SYR2K lower, N: C := alpha*(A*B^T + B*A^T) + beta*C (BLAS/tc_syr2k.cu:3-115). */
void orc_dsyr2k_ln(int64_t n, int64_t k, double alpha, const double *A, int64_t lda,
                   const double *B, int64_t ldb, double beta, double *C, int64_t ldc)
{
    for (int64_t j = 0; j < n; j++)
        for (int64_t i = j; i < n; i++) {
            double s = 0.0;
            for (int64_t p = 0; p < k; p++)
                s += A[i + p * lda] * B[j + p * ldb] + B[i + p * ldb] * A[j + p * lda];
            double c = beta == 0.0 ? 0.0 : beta * C[i + j * ldc];
            C[i + j * ldc] = alpha * s + c;
        }
}

/* TRMM left, lower, N, non-unit, out of place: C := alpha*L*B, L m x m (upper
 * triangle ignored), B and C m x n (BLAS/tc_trmm.cu:29-98, beta 0). */
void orc_dtrmm_llnn(int64_t m, int64_t n, double alpha, const double *L, int64_t lda,
                    const double *B, int64_t ldb, double *C, int64_t ldc)
{
    for (int64_t j = 0; j < n; j++)
        for (int64_t i = 0; i < m; i++) {
            double s = 0.0;
            for (int64_t p = 0; p <= i; p++) s += L[i + p * lda] * B[p + j * ldb];
            C[i + j * ldc] = alpha * s;
        }
}

/* TRSM right, lower, T, non-unit, in place: solve X*L^T = B, B m x n overwritten
 * with X, L n x n lower (BLAS/tc_trsm.cu:31-137).  Row i of X solves
 * L x^T = b^T by forward substitution. */
void orc_dtrsm_rltn(int64_t m, int64_t n, const double *L, int64_t lda, double *B,
                    int64_t ldb)
{
    for (int64_t i = 0; i < m; i++)
        for (int64_t j = 0; j < n; j++) {
            double s = B[i + j * ldb];
            for (int64_t p = 0; p < j; p++) s -= L[j + p * lda] * B[i + p * ldb];
            B[i + j * ldb] = s / L[j + j * lda];
        }
}

/* TRSM left, lower, N, non-unit, in place: solve L*X = B, B m x n (the form named
 * by the TRSM benchmark; it equals the right/T form applied to B^T). */
void orc_dtrsm_llnn(int64_t m, int64_t n, const double *L, int64_t lda, double *B,
                    int64_t ldb)
{
    for (int64_t j = 0; j < n; j++)
        for (int64_t i = 0; i < m; i++) {
            double s = B[i + j * ldb];
            for (int64_t p = 0; p < i; p++) s -= L[i + p * lda] * B[p + j * ldb];
            B[i + j * ldb] = s / L[i + i * lda];
        }
}

/* SYMM left, A symmetric stored lower: C := alpha*A*B + beta*C, A m x m,
 * B and C m x n (BLAS/tc_symm.cu:3-20 mirrors A then runs one GEMM; here the
 * upper triangle of A is never read). */
void orc_dsymm_ll(int64_t m, int64_t n, double alpha, const double *A, int64_t lda,
                  const double *B, int64_t ldb, double beta, double *C, int64_t ldc)
{
    for (int64_t j = 0; j < n; j++)
        for (int64_t i = 0; i < m; i++) {
            double s = 0.0;
            for (int64_t p = 0; p < m; p++) {
                double a = i >= p ? A[i + p * lda] : A[p + i * lda];
                s += a * B[p + j * ldb];
            }
            double c = beta == 0.0 ? 0.0 : beta * C[i + j * ldc];
            C[i + j * ldc] = alpha * s + c;
        }
}

/* ------------------------------------- 4. plain-tier TRSM recursion, emulated */

/* BLAS/tc_trsm.cu:31-80 (tc_rtrsm_p2): n <= nb -> exact leaf (cublasStrsm there,
 * exact FP64 here); else recurse on the left half, then
 *     B2 -= fp16(B1) * fp16(A21)^T         (s2h_swpipe + cublasGemmEx, FP32 accumulate)
 * and recurse on the right half.  The update is evaluated in FP64 on the
 * FP16-rounded operands, so this isolates exactly the rounding the plain tier
 * introduces; accumulation-order effects are the only remaining difference. */
static void trsm_p2_emul(int64_t m, int64_t n, const double *A, int64_t lda, double *B,
                         int64_t ldb, int64_t nb, double *wA, double *wB)
{
    if (n <= nb) {
        orc_dtrsm_rltn(m, n, A, lda, B, ldb);
        return;
    }
    int64_t h = n / 2;
    trsm_p2_emul(m, h, A, lda, B, ldb, nb, wA, wB);
    for (int64_t j = 0; j < h; j++)
        for (int64_t i = 0; i < h; i++)
            wA[i + j * h] = (double)orc_f16_to_f32(orc_f32_to_f16((float)A[h + i + j * lda]));
    for (int64_t j = 0; j < h; j++)
        for (int64_t i = 0; i < m; i++)
            wB[i + j * m] = (double)orc_f16_to_f32(orc_f32_to_f16((float)B[i + j * ldb]));
    orc_dgemm(0, 1, m, n - h, h, -1.0, wB, m, wA, h, 1.0, B + h * ldb, ldb);
    trsm_p2_emul(m, n - h, A + h * lda + h, lda, B + h * ldb, ldb, nb, wA, wB);
}

/* BLAS/tc_trsm.cu:83-137 (tc_trsm_p3): chunks from find_mat_size_trsm, processed
 * from offset 0; multiples of 2048 recurse, others are one leaf; between chunks a
 * panel update B(:, rest) -= fp16(B(:,chunk)) * fp16(A(rest, chunk))^T.
 * Requires m*n + n*n doubles of workspace in w. */
void orc_tc_trsm_plain_emul(int64_t m, int64_t n, const double *A, int64_t lda, double *B,
                            int64_t ldb, int64_t nb, double *w)
{
    int64_t chunks[16];
    int cnt = orc_find_mat_size_trsm(n, chunks);
    int64_t off = 0, rest = n;
    double *wA = w, *wB = w + n * n;
    for (int c = cnt - 1; c >= 0; c--) {
        int64_t nn = chunks[c];
        if (nn % 2048 == 0)
            trsm_p2_emul(m, nn, A + off + off * lda, lda, B + off * ldb, ldb, nb, wA, wB);
        else
            orc_dtrsm_rltn(m, nn, A + off + off * lda, lda, B + off * ldb, ldb);
        if (c != 0) {
            rest -= nn;
            for (int64_t j = 0; j < nn; j++)
                for (int64_t i = 0; i < rest; i++)
                    wA[i + j * rest] = (double)orc_f16_to_f32(
                        orc_f32_to_f16((float)A[off + nn + i + (off + j) * lda]));
            for (int64_t j = 0; j < nn; j++)
                for (int64_t i = 0; i < m; i++)
                    wB[i + j * m] = (double)orc_f16_to_f32(
                        orc_f32_to_f16((float)B[i + (off + j) * ldb]));
            orc_dgemm(0, 1, m, rest, nn, -1.0, wB, m, wA, rest, 1.0, B + (off + nn) * ldb, ldb);
        }
        off += nn;
    }
}

/* ------------------------------------------------------------- 5. Ozaki slices */

/* Row exponent as used by the split and Ozaki tiers: the e with
 * max_k |x_k| = f * 2^e, f in [0.5, 1) (frexp); 0 for an all-zero row. */
static int row_exponent(const double *x, int64_t stride, int64_t k)
{
    double mx = 0.0;
    for (int64_t p = 0; p < k; p++) {
        double a = fabs(x[p * stride]);
        if (a > mx) mx = a;
    }
    if (mx == 0.0) return 0;
    int e;
    frexp(mx, &e);
    return e;
}

/* Ozaki INT8 slicing of op(X) rows (X m x k, row i strided by 1, k by ldx when
 * trans == 0; row i strided by ldx, k by 1 when trans == 1).  Each entry is written
 * as  x = 2^(e_i) * sum_{s=0}^{S-1} d_s * 2^(-7(s+1)) + tail,  d_s integer in
 * [-127,127]: with r_0 = x * 2^(-e_i) in (-1,1), t_s = r_s * 2^7, d_s = trunc(t_s)
 * (toward zero, so |d_s| <= 127 always), r_{s+1} = t_s - d_s (exact in binary64,
 * |r_{s+1}| < 1).  Each slice carries 7 bits of the row-scaled mantissa: the
 * digit splitting of Ootomo-Ozaki-Yokota's INT8 DGEMM with a 7-bit digit.
 * slices[(s*m + i)*k + p] (int8, row-major per slice), exps[i] = e_i. */
void orc_ozaki_slice(int trans, int64_t m, int64_t k, const double *X, int64_t ldx,
                     int nslices, int8_t *slices, int32_t *exps)
{
    for (int64_t i = 0; i < m; i++) {
        const double *row = trans ? X + i * ldx : X + i;
        int64_t stride = trans ? 1 : ldx;
        int e = row_exponent(row, stride, k);
        exps[i] = e;
        for (int64_t p = 0; p < k; p++) {
            double r = ldexp(row[p * stride], -e);
            for (int s = 0; s < nslices; s++) {
                double t = r * 128.0;
                double d = trunc(t);              /* |t| < 128 -> |d| <= 127 */
                slices[((int64_t)s * m + i) * k + p] = (int8_t)d;
                r = t - d;                        /* exact */
            }
        }
    }
}

/* Ozaki product from slices.  With g = s+t, the slice pairs of one g share the
 * weight 2^(-7(g+2)), so their integer products are summed together:
 *     D_g(i,j) = sum_{s+t=g} sum_{p in chunk} a_s(i,p) * b_t(j,p)     (exact integer)
 * for g = 0..ngroups-1 with s, t < nslices (ngroups = nslices: pairs with
 * s+t >= nslices are dropped, the triangular truncation of the Ozaki scheme; the
 * library's option ngroups = nslices + 1 also keeps the group g = nslices, which has
 * no counterpart in the published scheme).  For each k chunk in order, and g in increasing
 * order, acc += D_g * 2^(-7(g+2)) in binary64.  Finally v = acc * 2^(ea_i+eb_j)
 * and C = fma(alpha, v, beta*C) (alpha*v when beta == 0).  This is the exact
 * operation order of the device kernel, so results agree bit for bit.  kchunk
 * keeps each D_g inside int32 on the device: nslices*kchunk*127^2 < 2^31. */
void orc_ozaki_gemm(int64_t m, int64_t n, int64_t k, int nslices, const int8_t *sa,
                    const int32_t *ea, const int8_t *sb, const int32_t *eb, double alpha,
                    double beta, double *C, int64_t ldc, int64_t kchunk, int ngroups)
{
    if (kchunk <= 0) kchunk = k;
    if (ngroups <= 0) ngroups = nslices;
    for (int64_t j = 0; j < n; j++)
        for (int64_t i = 0; i < m; i++) {
            double acc = 0.0;
            for (int64_t k0 = 0; k0 < k; k0 += kchunk) {
                int64_t k1 = k0 + kchunk < k ? k0 + kchunk : k;
                for (int g = 0; g < ngroups; g++) {
                    int64_t dot = 0;
                    for (int s = 0; s <= g; s++) {
                        int t = g - s;
                        if (s >= nslices || t >= nslices) continue;
                        const int8_t *a = sa + ((int64_t)s * m + i) * k;
                        const int8_t *b = sb + ((int64_t)t * n + j) * k;
                        for (int64_t p = k0; p < k1; p++) dot += (int64_t)a[p] * b[p];
                    }
                    acc += ldexp((double)dot, -7 * (g + 2));
                }
            }
            double v = ldexp(acc, ea[i] + eb[j]);
            C[i + j * ldc] = beta == 0.0 ? alpha * v : fma(alpha, v, beta * C[i + j * ldc]);
        }
}

/* ---------------------------------------------------------------- 6. metrics */

/* out[0] normwise ||X-R||_F/||R||_F; out[1] max|X-R| / max|R|;
 * out[2] max_ij |X-R|_ij / S_ij where S (may be NULL) is the |A||B| magnitude
 * bound matrix.  Only the lower triangle when lower != 0. */
void orc_errors(int64_t m, int64_t n, const double *X, int64_t ldx, const double *R,
                int64_t ldr, const double *S, int64_t lds, int lower, double *out)
{
    double num = 0.0, den = 0.0, mxd = 0.0, mxr = 0.0, comp = 0.0;
    for (int64_t j = 0; j < n; j++)
        for (int64_t i = lower ? j : 0; i < m; i++) {
            double r = R[i + j * ldr];
            double d = fabs(X[i + j * ldx] - r);
            num += d * d;
            den += r * r;
            if (d > mxd || d != d) mxd = d;
            if (fabs(r) > mxr) mxr = fabs(r);
            if (S) {
                double s = S[i + j * lds];
                double c = s > 0 ? d / s : (d > 0 ? INFINITY : 0.0);
                if (c > comp || c != c) comp = c;
            }
        }
    out[0] = den > 0 ? sqrt(num / den) : sqrt(num);
    out[1] = mxr > 0 ? mxd / mxr : mxd;
    out[2] = comp;
}

/* ------------------------------------------------------------- 7. Cholesky */

/* Lower Cholesky in place, right-looking, unblocked: for j = 0..n-1,
 * l_jj = sqrt(a_jj), l_ij = a_ij / l_jj (i > j), a_ik -= l_ij l_kj (j < k <= i).
 * Only the lower triangle is read or written.  Returns 0, or j+1 when the leading
 * minor of order j+1 is not positive definite (a_jj <= 0 or NaN at step j; A then
 * holds the partial factorization), as LAPACK dpotrf's info. */
int orc_dpotrf_lower(int64_t n, double *A, int64_t lda)
{
    for (int64_t j = 0; j < n; j++) {
        double d = A[j + j * lda];
        if (!(d > 0.0)) return (int)(j + 1);
        d = sqrt(d);
        A[j + j * lda] = d;
        for (int64_t i = j + 1; i < n; i++) A[i + j * lda] /= d;
        for (int64_t k = j + 1; k < n; k++) {
            const double lkj = A[k + j * lda];
            for (int64_t i = k; i < n; i++) A[i + k * lda] -= A[i + j * lda] * lkj;
        }
    }
    return 0;
}

/* ------------------------------------------- 8. sampled columns, extended precision */

/* out[c*n + i] = sum_p X[i,p] Y[cols[c],p] (+ X2[i,p] Y2[cols[c],p] when X2 != NULL)
 * for i < n and c < ncols, with the rows of X, Y, X2, Y2 contiguous (row r of X at
 * X + r*ldx, k elements: row-major storage).  Products and sums are formed in long
 * double (64-bit significand on x86-64), then rounded once to double.
 *   SYR2K column j (lower part, i >= j): X = A, Y = B, X2 = B, Y2 = A, cols = {j}.
 *   SYMM / GEMM column j of A*B: X = rows of A, Y = rows of B^T.
 * Rows are shared among OpenMP threads; the sampled rows of Y, Y2 stay in cache. */
void orc_rowdot_xp(int64_t n, int64_t k, const double *X, int64_t ldx, const double *Y,
                   int64_t ldy, const double *X2, int64_t ldx2, const double *Y2, int64_t ldy2,
                   int64_t ncols, const int64_t *cols, double *out)
{
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; i++) {
        const double *x = X + i * ldx;
        const double *x2 = X2 ? X2 + i * ldx2 : NULL;
        for (int64_t c = 0; c < ncols; c++) {
            const double *y = Y + cols[c] * ldy;
            long double s0 = 0.0L, s1 = 0.0L, s2 = 0.0L, s3 = 0.0L;
            int64_t p = 0;
            for (; p + 4 <= k; p += 4) {
                s0 += (long double)x[p] * y[p];
                s1 += (long double)x[p + 1] * y[p + 1];
                s2 += (long double)x[p + 2] * y[p + 2];
                s3 += (long double)x[p + 3] * y[p + 3];
            }
            for (; p < k; p++) s0 += (long double)x[p] * y[p];
            if (x2) {
                const double *y2 = Y2 + cols[c] * ldy2;
                for (p = 0; p + 4 <= k; p += 4) {
                    s0 += (long double)x2[p] * y2[p];
                    s1 += (long double)x2[p + 1] * y2[p + 1];
                    s2 += (long double)x2[p + 2] * y2[p + 2];
                    s3 += (long double)x2[p + 3] * y2[p + 3];
                }
                for (; p < k; p++) s0 += (long double)x2[p] * y2[p];
            }
            out[c * n + i] = (double)((s0 + s1) + (s2 + s3));
        }
    }
}
