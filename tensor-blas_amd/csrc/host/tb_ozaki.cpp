// Host side of the Ozaki INT8 tier: exponents, slicing, the slice GEMM.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "tb_internal.hpp"

namespace tb {

namespace {
int flip_mode(int mode) { return mode == kOpT ? kOpN : (mode == kOpN ? kOpT : mode); }
}  // namespace

// K is processed in chunks so that every int32 sum D_g stays exact:
// nslices * kchunk * 127^2 < 2^31 (all pairs of one g share an accumulator).
long ozaki_kchunk(int nslices)
{
    long kc = ((1l << 31) - 1) / ((long)nslices * 127 * 127);
    return kc / 64 * 64;
}

int ozaki_gemm_host(tb_handle_s* h, long m, long n, long k, double alpha, const double* A,
                    long lda, int modeA, const double* B, long ldb, int modeB, const double* A2,
                    const double* B2, double beta, double* C, long ldc, bool lowerC)
{
    if (m <= 0 || n <= 0) return TB_SUCCESS;
    const int S = h->ozaki_slices;
    if (h->ozaki_extra && (S < 6 || S > 9)) return TB_ERR_INVALID;   // kernels instantiated
    const long kpad = std::max(64l, (k + 63) / 64 * 64);
    const bool pairs = A2 != nullptr;
    Arena a;
    const size_t oEA = a.reserve(sizeof(int) * m);
    const size_t oEB = a.reserve(sizeof(int) * n);
    // slices hold rows padded to a multiple of 64; the wide kernel (128 x 64 tiles, eight
    // waves) takes one exactness chunk of one product per launch (more than one with
    // beta = 0: the FP64 sum of the chunks so far kept in C), and its A slices cover a
    // multiple of 128 rows
    const long kchunk = ozaki_kchunk(S);
    const long nchunks = (kpad + kchunk - 1) / kchunk;
    const bool wide = ((nchunks == 1 && !pairs) || beta == 0.0) &&
                      ozaki_wide_ok(S, S + h->ozaki_extra);
    const long arows = wide ? (m + ozaki_wide_rows() - 1) / ozaki_wide_rows() * ozaki_wide_rows() : m;
    const int rbA = (int)((arows + 63) / 64);
    const size_t mp = (size_t)rbA * 64, np = (size_t)(n + 63) / 64 * 64;
    const size_t oSA = a.reserve((size_t)S * mp * kpad);
    const size_t oSB = a.reserve((size_t)S * np * kpad);
    const size_t oSA2 = pairs ? a.reserve((size_t)S * mp * kpad) : 0;
    const size_t oSB2 = pairs ? a.reserve((size_t)S * np * kpad) : 0;
    char* ws;
    TB_TRY(workspace_get(h, a.total(), &ws, 1));
    hipStream_t st = h->stream;
    int* eA = reinterpret_cast<int*>(ws + oEA);
    int* eB = reinterpret_cast<int*>(ws + oEB);
    int8_t* sA = reinterpret_cast<int8_t*>(ws + oSA);
    int8_t* sB = reinterpret_cast<int8_t*>(ws + oSB);
    // op(B) columns are the rows of op(B)^T
    const int rmB = flip_mode(modeB);
    // eB follows eA in the workspace (reserved in that order): one fill covers both and
    // the alignment gap between them
    TB_HIP(fill_i32(eA, (long)(eB - eA) + n, kExpZero, st));
    if (k > 0) {
        TB_HIP(rowexp(true, A, lda, modeA, (int)m, (int)k, 1, 0, eA, 0, st));
        TB_HIP(rowexp(true, B, ldb, rmB, (int)n, (int)k, 1, 0, eB, 0, st));
        if (pairs) {
            // both products accumulate together: one scale per row and per column
            TB_HIP(rowexp(true, A2, ldb, modeA, (int)m, (int)k, 1, 0, eA, 0, st));
            TB_HIP(rowexp(true, B2, lda, rmB, (int)n, (int)k, 1, 0, eB, 0, st));
        }
    }
    TB_HIP(ozaki_slice(A, lda, modeA, (int)m, (int)k, (int)kpad, S, 1, 0, eA, sA, st, rbA));
    TB_HIP(ozaki_slice(B, ldb, rmB, (int)n, (int)k, (int)kpad, S, 1, 0, eB, sB, st));
    OzakiProb p{};
    p.sa = sA;
    p.sb = sB;
    if (pairs) {
        p.sa2 = reinterpret_cast<int8_t*>(ws + oSA2);
        p.sb2 = reinterpret_cast<int8_t*>(ws + oSB2);
        TB_HIP(ozaki_slice(A2, ldb, modeA, (int)m, (int)k, (int)kpad, S, 1, 0, eA,
                           const_cast<int8_t*>(p.sa2), st, rbA));
        TB_HIP(ozaki_slice(B2, lda, rmB, (int)n, (int)k, (int)kpad, S, 1, 0, eB,
                           const_cast<int8_t*>(p.sb2), st));
    }
    p.ea = eA;
    p.eb = eB;
    p.C = C;
    p.ldc = ldc;
    p.alpha = alpha;
    p.beta = beta;
    p.m = (int)m;
    p.n = (int)n;
    p.kpad = (int)kpad;
    p.nslices = S;
    p.ngroups = S + h->ozaki_extra;
    p.kchunk = (int)kchunk;
    p.lowerC = lowerC ? 1 : 0;
    p.wide = wide ? 1 : 0;
    const double flops = (lowerC ? (double)n * (n + 1) * k : 2.0 * m * n * k) * (pairs ? 2 : 1);
    hipEvent_t e0;
    TB_TRY(prof_begin(h, &e0));
    if (wide && (pairs || nchunks > 1)) {
        // the same sums in the same order as one chunked launch over both products: each
        // chunk of each product one launch, the FP64 sum so far kept in C, scaled by the
        // last.  The slices are blocked by 32 columns: chunk c of a product starts
        // c * kchunk / 32 blocks in.
        const long kstepA = (long)rbA * S * 64 * 32, kstepB = (long)(np / 64) * S * 64 * 32;
        const int nl = (pairs ? 2 : 1) * (int)nchunks;
        int li = 0;
        for (int pp = 0; pp < (pairs ? 2 : 1); ++pp) {
            for (long c = 0; c < nchunks; ++c, ++li) {
                OzakiProb q = p;
                q.sa = (pp ? p.sa2 : p.sa) + c * (kchunk / 32) * kstepA;
                q.sb = (pp ? p.sb2 : p.sb) + c * (kchunk / 32) * kstepB;
                q.sa2 = q.sb2 = nullptr;
                q.kpad = (int)std::min(kchunk, kpad - c * kchunk);
                q.partial = li == 0 ? 1 : (li == nl - 1 ? 2 : 3);
                TB_HIP(ozaki_gemm(q, st, h->max_cus));
            }
        }
    } else {
        TB_HIP(ozaki_gemm(p, st, h->max_cus));
    }
    return prof_end(h, kProfOzakiGemm, e0, flops);
}

}  // namespace tb
