// The BLAS-3 routines: argument checks, the reference's schedules, and the launches.
//
// SYRK / SYR2K / TRMM / SYMM.  The reference decomposes each of these into
// diagonal-block leaves and off-diagonal GEMM updates (BLAS/tc_syrk.cu:41-125:
// chunks from find_mat_size_syrk, then per chunk the nb x nb diagonal batch and
// the i*nb x i*nb level batches, then the panel GEMM).  Every piece writes a
// disjoint block of C with the full k sum, so the union is exactly the set of C
// tiles on and below the diagonal.  Here that whole set is one grouped launch:
// a kCLower problem enumerates the lower tiles directly, diagonal tiles mask their
// strictly upper half (the reference computed full squares there), and TRMM/SYMM
// tiles clip or mirror their K range at the diagonal instead of requiring a zeroed
// or mirrored A (BLAS/tc_trmm.cu:29-98, tc_symm.cu:3-20).  Each C element is the
// same k-sum either way, so the result does not depend on nb.
//
// In the FP16 tiers every operand is first converted once into K-contiguous FP16
// planes (the reference's s2h pass, BLAS/tc_syrk.cu:70-86; three planes hi/mid/lo
// in the split tier), so the GEMM tiles copy planes instead of re-converting the
// same panel for every tile.
//
// TRSM keeps the reference's recursion (BLAS/tc_trsm.cu:31-137) because its order
// matters: chunks from find_mat_size_trsm, each 2048-multiple chunk halved down to
// nb with the update B2 -= X1*A21^T between the halves, a panel update between
// chunks.  Leaves (the reference's cublasStrsm) are solved as X := B * inv(L_jj)^T,
// written as (B + B W^T) D^-1 with L_jj = (I + N) D and W = (I + N)^-1 - I: all the
// W are formed up front by one batched kernel (in FP32 for the FP32-class tiers, FP64
// for the FP64 tiers), and each leaf is one in-place 64 x 256-tile GEMM at FP32
// accuracy (split MFMA) whose epilogue divides by the diagonal.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>

#include "tb_internal.hpp"

namespace tb {

namespace {

bool bad_tier(int t) { return t < kPlainF16 || t > kMfmaF64; }
int flip(int mode) { return mode == kOpT ? kOpN : kOpT; }
bool fp16_tier(int t) { return t == kPlainF16 || t == kSplitF16; }
long round_up(long v, long q) { return (v + q - 1) / q * q; }

}  // namespace

int workspace_get(tb_handle_s* h, size_t bytes, char** base, int slot)
{
    if (bytes > h->ws_bytes[slot]) {
        if (h->ws[slot]) {
            TB_HIP(hipStreamSynchronize(h->stream));
            TB_HIP(hipFree(h->ws[slot]));
            h->ws[slot] = nullptr;
            h->ws_bytes[slot] = 0;
        }
        size_t want = (bytes + (size_t(2) << 20) - 1) & ~((size_t(2) << 20) - 1);
        if (hipMalloc(&h->ws[slot], want) != hipSuccess) {
            (void)hipGetLastError();
            h->ws[slot] = nullptr;
            return TB_ERR_NOMEM;
        }
        h->ws_bytes[slot] = want;
    }
    *base = static_cast<char*>(h->ws[slot]);
    return TB_SUCCESS;
}

// ------------------------------------------------------------------ profiling
int prof_begin(tb_handle_s* h, hipEvent_t* e0)
{
    *e0 = nullptr;
    if (!h->prof) return TB_SUCCESS;
    TB_HIP(hipEventCreate(e0));
    return hip_status(hipEventRecord(*e0, h->stream));
}

int prof_end(tb_handle_s* h, int cat, hipEvent_t e0, double flops)
{
    if (!h->prof || !e0) return TB_SUCCESS;
    hipEvent_t e1;
    TB_HIP(hipEventCreate(&e1));
    TB_HIP(hipEventRecord(e1, h->stream));
    h->prof_recs.push_back({cat, e0, e1, flops});
    return TB_SUCCESS;
}

namespace {
// algorithmic flops of one problem (useful multiply-adds x 2)
double prob_flops(const GemmProb& p)
{
    const double m = p.m, n = p.n, k = p.k, pairs = p.A2 ? 2.0 : 1.0;
    double f;
    if (p.modeC == kCLower) f = n * (n + 1) * k;
    else if (p.modeA == kOpTriL) f = m * (m + 1) * n;
    else f = 2.0 * m * n * k;
    return f * pairs * p.batch;
}
}  // namespace

// ------------------------------------------------------------------ GEMM queue
GemmQueue::GemmQueue(tb_handle_s* h, int tier, int shape, int srcA, int srcB)
    : h_(h), tier_(tier), shape_(shape), srcA_(srcA), srcB_(srcB),
      ts_(gemm_tile_shape(tier, shape))
{
    std::memset(&L_, 0, sizeof(L_));
}

int GemmQueue::add(const GemmProb& p0)
{
    if (p0.m <= 0 || p0.n <= 0 || p0.batch <= 0) return TB_SUCCESS;
    if (p0.srcA != srcA_ || p0.srcB != srcB_) return TB_ERR_INVALID;
    GemmProb p = p0;
    p.tilesM = (p.m + ts_.bm - 1) / ts_.bm;
    p.tilesN = (p.n + ts_.bn - 1) / ts_.bn;
    if (p.modeC == kCLower) {
        if (ts_.bm != ts_.bn || p.m != p.n) return TB_ERR_INVALID;
        p.tilesPer = p.tilesM * (p.tilesM + 1) / 2;
    } else {
        p.tilesPer = p.tilesM * p.tilesN;
    }
    const long long tiles = (long long)p.tilesPer * p.batch;
    if (tiles > (1ll << 30)) return TB_ERR_INVALID;
    if (L_.np == kMaxProbs || (long long)L_.totalTiles + tiles > (1ll << 30)) TB_TRY(flush());
    p.tileBegin = L_.totalTiles;
    L_.p[L_.np++] = p;
    L_.totalTiles += (int)tiles;
    return TB_SUCCESS;
}

// Square launches of plane products (full or lower-triangular C) with at least one wide
// tile per CU run on 256 x 256 tiles instead (fewer operand bytes per MFMA); smaller
// launches keep 128 x 128, whose more numerous tiles fill more of the device.
bool GemmQueue::retile_wide()
{
    const int npl = planes_of(h_, tier_);
    if (shape_ != kShapeSquare || srcA_ != kSrcPlanes || srcB_ != kSrcPlanes || npl < 1 || npl > 2 ||
        h_->max_cus > 0)          // wide tiles launch one workgroup per tile: never capped
        return false;
    const TileShape w = gemm_tile_shape(tier_, kShapeWide);
    long long tiles = 0;
    // lower-triangular C (SYRK, SYR2K) takes the tiles on and below the diagonal
    auto per = [&](const GemmProb& p) -> long long {
        const long long tm = (p.m + w.bm - 1) / w.bm, tn = (p.n + w.bn - 1) / w.bn;
        return p.modeC == kCLower ? tm * (tm + 1) / 2 : tm * tn;
    };
    for (int q = 0; q < L_.np; ++q) tiles += per(L_.p[q]) * L_.p[q].batch;
    if (tiles < kWideMinTiles) return false;
    int begin = 0;
    for (int q = 0; q < L_.np; ++q) {
        GemmProb& p = L_.p[q];
        p.tilesM = (p.m + w.bm - 1) / w.bm;
        p.tilesN = (p.n + w.bn - 1) / w.bn;
        p.tilesPer = (int)per(p);
        p.tileBegin = begin;
        begin += p.tilesPer * p.batch;
    }
    L_.totalTiles = begin;
    return true;
}

int GemmQueue::flush()
{
    if (L_.np == 0) return TB_SUCCESS;
    double flops = 0.0;
    for (int q = 0; q < L_.np; ++q) flops += prob_flops(L_.p[q]);
    const int shape = retile_wide() ? kShapeWide : shape_;
    hipEvent_t e0;
    TB_TRY(prof_begin(h_, &e0));
    hipError_t e = gemm_launch(tier_, planes_of(h_, tier_), shape, srcA_, srcB_, L_, h_->stream,
                               h_->max_cus);
    std::memset(&L_, 0, sizeof(L_));
    TB_HIP(e);
    return prof_end(h_, kProfGemm0 + tier_ * 2 + shape_, e0, flops);
}

GemmProb make_prob(long m, long n, long k, const void* A, long lda, int modeA, const void* B,
                   long ldb, int modeB, void* C, long ldc, double alpha, double beta, int modeC)
{
    GemmProb p;
    std::memset(&p, 0, sizeof(p));
    p.A = A;
    p.B = B;
    p.C = C;
    p.lda = lda;
    p.ldb = ldb;
    p.ldc = ldc;
    p.alpha = alpha;
    p.beta = beta;
    p.alpha32 = (float)alpha;
    p.beta32 = (float)beta;
    p.m = (int)m;
    p.n = (int)n;
    p.k = (int)k;
    p.batch = 1;
    p.modeA = modeA;
    p.modeB = modeB;
    p.modeC = modeC;
    p.srcA = kSrcF32;
    p.srcB = kSrcF32;
    return p;
}

// ---------------------------------------------------------------- planners
namespace {
std::vector<long> plan_chunks(long n, const long* powers, int np, long floor_)
{
    std::vector<long> out;
    for (int i = np - 1; i >= 0 && n > 0; --i) {
        if (n < floor_) break;
        if (n >= powers[i]) {
            out.push_back(powers[i]);
            n -= powers[i];
        }
    }
    // past the end of the table (where the reference reads powers[-1]): 16384s,
    // then the remainder as one last chunk
    while (n >= 16384 && out.size() >= (size_t)np) {
        out.push_back(16384);
        n -= 16384;
    }
    if (n > 0) out.push_back(n);
    return out;
}
}  // namespace

std::vector<long> plan_chunks_syrk(long n)
{
    static const long powers[] = {8192, 16384, 16384, 16384, 16384, 16384,
                                  16384, 16384, 16384, 16384, 16384, 16384};
    return plan_chunks(n, powers, 12, 8192);
}

std::vector<long> plan_chunks_trsm(long n)
{
    static const long powers[] = {2048, 4096, 8192, 16384, 16384, 16384,
                                  16384, 16384, 16384, 16384, 16384, 16384};
    return plan_chunks(n, powers, 12, 2048);
}

// ----------------------------------------------------- exponents and FP16 planes
namespace {

int exps_of(tb_handle_s* h, bool f64, const void* X, long ld, int mode, long rows, long K, int* e,
            bool fill = true)
{
    if (fill) TB_HIP(fill_i32(e, rows, kExpZero, h->stream));
    return hip_status(rowexp(f64, X, ld, mode, (int)rows, (int)K, 1, 0, e, 0, h->stream));
}

bool dims_ok(long v) { return v >= 0 && v < (1l << 31); }

// K-contiguous FP16 planes of the rows of op(X)
struct Planes {
    _Float16* p = nullptr;
    long ldp = 0;
    long long sp = 0;
};

size_t planes_bytes(int npl, long rows, long K, int kshift = 0)
{
    const long ldp = round_up(std::max(1l, K + kshift), 64);
    return (size_t)npl * rows * ldp * sizeof(_Float16);
}

int make_planes(tb_handle_s* h, int npl, const float* X, long ld, int mode, long rows, long K,
                const int* e, char* mem, Planes* out, int kshift = 0, int* e_used = nullptr,
                bool lower_only = false)
{
    out->p = reinterpret_cast<_Float16*>(mem);
    out->ldp = round_up(std::max(1l, K + kshift), 64);
    out->sp = (long long)rows * out->ldp;
    hipEvent_t e0;
    TB_TRY(prof_begin(h, &e0));
    TB_HIP(split_planes(npl, X, ld, mode, (int)rows, (int)K, e, out->p, out->ldp, out->sp,
                        kshift, h->stream, e_used, lower_only));
    return prof_end(h, kProfPlanes, e0, (double)rows * (4.0 * K + 2.0 * npl * out->ldp));
}

// One FP16-tier product  C = alpha * op(A) op(B) [+ alpha * op(B-rows) op(A-rows)^T]
// + beta * C.  a: the rows of op(A) (m x k, rowexp mode); b: the rows of op(B)^T
// (n x k).  same_ab: op(B)^T == op(A) (SYRK).  pairs: SYR2K's second product with
// the roles of a and b swapped.
struct RowsOf {
    const float* X;
    long ld;
    int mode;
};

// fixed_ws: when set, the planes and exponents go there (at most fixed_bytes) instead of
// the handle's workspace, and nothing is allocated (TB_ERR_SCRATCH when they do not fit)
int fp16_product(tb_handle_s* h, int tier, long m, long n, long k, double alpha, RowsOf a,
                 RowsOf b, bool same_ab, bool pairs, double beta, float* C, long ldc, int modeC,
                 char* fixed_ws = nullptr, size_t fixed_bytes = 0)
{
    const int npl = planes_of(h, tier);
    const bool split = tier == kSplitF16;
    const bool sep_b = !same_ab;
    Arena ar;
    const size_t oEA = ar.reserve(sizeof(int) * m);
    const size_t oEB = ar.reserve(sizeof(int) * n);
    const size_t oPA = ar.reserve(planes_bytes(npl, m, k));
    const size_t oPB = sep_b ? ar.reserve(planes_bytes(npl, n, k)) : 0;
    char* ws;
    if (fixed_ws) {
        if (ar.total() > fixed_bytes) return TB_ERR_SCRATCH;
        ws = fixed_ws;
    } else {
        TB_TRY(workspace_get(h, ar.total(), &ws));
    }
    int* eA = reinterpret_cast<int*>(ws + oEA);
    int* eB = reinterpret_cast<int*>(ws + oEB);
    if (split && k > 0) {
        TB_TRY(exps_of(h, false, a.X, a.ld, a.mode, m, k, eA));
        if (pairs) {
            // both products share one accumulator: one scale per row of [A B]
            TB_TRY(exps_of(h, false, b.X, b.ld, b.mode, n, k, eA, false));
            eB = eA;
        } else if (sep_b) {
            TB_TRY(exps_of(h, false, b.X, b.ld, b.mode, n, k, eB));
        } else {
            eB = eA;
        }
    }
    Planes pa, pb;
    TB_TRY(make_planes(h, npl, a.X, a.ld, a.mode, m, k, eA, ws + oPA, &pa));
    if (sep_b) TB_TRY(make_planes(h, npl, b.X, b.ld, b.mode, n, k, eB, ws + oPB, &pb));
    else pb = pa;
    GemmProb p = make_prob(m, n, k, pa.p, pa.ldp, a.mode == kOpTriL ? kOpTriL : kOpN, pb.p, pb.ldp,
                           kOpN, C, ldc, alpha, beta, modeC);
    p.srcA = p.srcB = kSrcPlanes;
    p.spA = pa.sp;
    p.spB = pb.sp;
    if (pairs) {
        p.A2 = pb.p;
        p.B2 = pa.p;
    }
    if (split) {
        p.eA = eA;
        p.eB = eB;
    }
    GemmQueue q(h, tier, kShapeSquare, kSrcPlanes, kSrcPlanes);
    TB_TRY(q.add(p));
    return q.flush();
}

}  // namespace

// -------------------------------------------------------------------- GEMM
int gemm(tb_handle_s* h, int tier, int transa, int transb, long m, long n, long k, double alpha,
         const void* A, long lda, const void* B, long ldb, double beta, void* C, long ldc)
{
    if (!h || bad_tier(tier) || !dims_ok(m) || !dims_ok(n) || !dims_ok(k)) return TB_ERR_INVALID;
    const long arows = transa ? k : m, brows = transb ? n : k;
    if (lda < std::max(1l, arows) || ldb < std::max(1l, brows) || ldc < std::max(1l, m))
        return TB_ERR_INVALID;
    if (m == 0 || n == 0) return TB_SUCCESS;
    if (!C || (k > 0 && (!A || !B))) return TB_ERR_INVALID;
    const int mA = transa ? kOpT : kOpN, mB = transb ? kOpT : kOpN;
    if (tier == kOzakiI8)
        return ozaki_gemm_host(h, m, n, k, alpha, (const double*)A, lda, mA, (const double*)B, ldb,
                               mB, nullptr, nullptr, beta, (double*)C, ldc, false);
    if (fp16_tier(tier))
        return fp16_product(h, tier, m, n, k, alpha, {(const float*)A, lda, mA},
                            {(const float*)B, ldb, flip(mB)}, false, false, beta, (float*)C, ldc,
                            kCFull);
    GemmQueue q(h, tier, kShapeSquare);
    TB_TRY(q.add(make_prob(m, n, k, A, lda, mA, B, ldb, mB, C, ldc, alpha, beta, kCFull)));
    return q.flush();
}

// -------------------------------------------------------------------- SYRK
int syrk(tb_handle_s* h, int tier, long n, long k, double alpha, const void* A, long lda,
         bool trans_a, double beta, void* C, long ldc, long nb)
{
    if (!h || bad_tier(tier) || !dims_ok(n) || !dims_ok(k) || nb <= 0) return TB_ERR_INVALID;
    const long arows = trans_a ? k : n;
    if (lda < std::max(1l, arows) || ldc < std::max(1l, n)) return TB_ERR_INVALID;
    if (n == 0) return TB_SUCCESS;
    if (!C || (k > 0 && !A)) return TB_ERR_INVALID;
    const int mA = trans_a ? kOpT : kOpN, mB = flip(mA);
    if (tier == kOzakiI8)
        return ozaki_gemm_host(h, n, n, k, alpha, (const double*)A, lda, mA, (const double*)A, lda,
                               mB, nullptr, nullptr, beta, (double*)C, ldc, true);
    if (fp16_tier(tier))
        return fp16_product(h, tier, n, n, k, alpha, {(const float*)A, lda, mA},
                            {(const float*)A, lda, mA}, true, false, beta, (float*)C, ldc, kCLower);
    GemmQueue q(h, tier, kShapeSquare);
    TB_TRY(q.add(make_prob(n, n, k, A, lda, mA, A, lda, mB, C, ldc, alpha, beta, kCLower)));
    return q.flush();
}

// Same as syrk with trans_a = false, except that the FP16 tiers' operand planes and row
// exponents live in the handle's fixed no-allocation scratch (kNoAllocScratchBytes per
// device, shared by every call on that device) instead of the handle's workspace.  The
// FP32 and FP64 tiers need no scratch; the Ozaki tier's slice workspace is not offered
// here.  Profiling events are not recorded (creating them allocates).
int syrk_noalloc(tb_handle_s* h, int tier, long n, long k, double alpha, const void* A,
                 long lda, double beta, void* C, long ldc, long nb)
{
    if (!h || bad_tier(tier) || tier == kOzakiI8 || !dims_ok(n) || !dims_ok(k) || nb <= 0)
        return TB_ERR_INVALID;
    if (lda < std::max(1l, n) || ldc < std::max(1l, n)) return TB_ERR_INVALID;
    if (n == 0) return TB_SUCCESS;
    if (!C || (k > 0 && !A)) return TB_ERR_INVALID;
    static_assert(kNoAllocScratchBytes == (size_t)TB_NOALLOC_SCRATCH_BYTES, "scratch size");
    if (!h->scratch) return TB_ERR_DEVICE;
    const bool prof = h->prof;
    h->prof = false;
    int st;
    if (fp16_tier(tier)) {
        st = fp16_product(h, tier, n, n, k, alpha, {(const float*)A, lda, kOpN},
                          {(const float*)A, lda, kOpN}, true, false, beta, (float*)C, ldc, kCLower,
                          static_cast<char*>(h->scratch), kNoAllocScratchBytes);
    } else {
        GemmQueue q(h, tier, kShapeSquare);
        st = q.add(make_prob(n, n, k, A, lda, kOpN, A, lda, kOpT, C, ldc, alpha, beta, kCLower));
        if (st == TB_SUCCESS) st = q.flush();
    }
    h->prof = prof;
    return st;
}

// ------------------------------------------------------------------- SYR2K
int syr2k(tb_handle_s* h, int tier, long n, long k, double alpha, const void* A, long lda,
          const void* B, long ldb, double beta, void* C, long ldc, long nb)
{
    if (!h || bad_tier(tier) || !dims_ok(n) || !dims_ok(k) || nb <= 0) return TB_ERR_INVALID;
    if (lda < std::max(1l, n) || ldb < std::max(1l, n) || ldc < std::max(1l, n))
        return TB_ERR_INVALID;
    if (n == 0) return TB_SUCCESS;
    if (!C || (k > 0 && (!A || !B))) return TB_ERR_INVALID;
    if (tier == kOzakiI8)
        return ozaki_gemm_host(h, n, n, k, alpha, (const double*)A, lda, kOpN, (const double*)B, ldb,
                               kOpT, (const double*)B, (const double*)A, beta, (double*)C, ldc,
                               true);
    if (fp16_tier(tier))
        return fp16_product(h, tier, n, n, k, alpha, {(const float*)A, lda, kOpN},
                            {(const float*)B, ldb, kOpN}, false, true, beta, (float*)C, ldc,
                            kCLower);
    GemmQueue q(h, tier, kShapeSquare);
    if (ldb != lda && k > 0) {
        // the descriptor's second pair shares the first pair's leading dimensions
        TB_TRY(q.add(make_prob(n, n, k, A, lda, kOpN, B, ldb, kOpT, C, ldc, alpha, beta, kCLower)));
        TB_TRY(q.flush());   // the second product accumulates onto the first
        TB_TRY(q.add(make_prob(n, n, k, B, ldb, kOpN, A, lda, kOpT, C, ldc, alpha, 1.0, kCLower)));
        return q.flush();
    }
    GemmProb p = make_prob(n, n, k, A, lda, kOpN, B, ldb, kOpT, C, ldc, alpha, beta, kCLower);
    p.A2 = B;
    p.B2 = A;
    TB_TRY(q.add(p));
    return q.flush();
}

// -------------------------------------------------------------------- TRMM
int trmm(tb_handle_s* h, int tier, long m, long n, double alpha, const void* A, long lda,
         const void* B, long ldb, void* C, long ldc, long nb)
{
    if (!h || bad_tier(tier) || !dims_ok(m) || !dims_ok(n) || nb <= 0) return TB_ERR_INVALID;
    if (lda < std::max(1l, m) || ldb < std::max(1l, m) || ldc < std::max(1l, m))
        return TB_ERR_INVALID;
    if (m == 0 || n == 0) return TB_SUCCESS;
    if (!A || !B || !C) return TB_ERR_INVALID;
    if (tier == kOzakiI8)
        return ozaki_gemm_host(h, m, n, m, alpha, (const double*)A, lda, kOpTriL, (const double*)B,
                               ldb, kOpN, nullptr, nullptr, 0.0, (double*)C, ldc, false);
    if (fp16_tier(tier))
        return fp16_product(h, tier, m, n, m, alpha, {(const float*)A, lda, kOpTriL},
                            {(const float*)B, ldb, kOpT}, false, false, 0.0, (float*)C, ldc,
                            kCFull);
    GemmQueue q(h, tier, kShapeSquare);
    TB_TRY(q.add(make_prob(m, n, m, A, lda, kOpTriL, B, ldb, kOpN, C, ldc, alpha, 0.0, kCFull)));
    return q.flush();
}

// -------------------------------------------------------------------- SYMM
int symm(tb_handle_s* h, int tier, long m, long n, double alpha, const void* A, long lda,
         const void* B, long ldb, double beta, void* C, long ldc)
{
    if (!h || bad_tier(tier) || !dims_ok(m) || !dims_ok(n)) return TB_ERR_INVALID;
    if (lda < std::max(1l, m) || ldb < std::max(1l, m) || ldc < std::max(1l, m))
        return TB_ERR_INVALID;
    if (m == 0 || n == 0) return TB_SUCCESS;
    if (!A || !B || !C) return TB_ERR_INVALID;
    if (tier == kOzakiI8)
        return ozaki_gemm_host(h, m, n, m, alpha, (const double*)A, lda, kOpSymL, (const double*)B,
                               ldb, kOpN, nullptr, nullptr, beta, (double*)C, ldc, false);
    if (fp16_tier(tier))
        return fp16_product(h, tier, m, n, m, alpha, {(const float*)A, lda, kOpSymL},
                            {(const float*)B, ldb, kOpT}, false, false, beta, (float*)C, ldc,
                            kCFull);
    GemmQueue q(h, tier, kShapeSquare);
    TB_TRY(q.add(make_prob(m, n, m, A, lda, kOpSymL, B, ldb, kOpN, C, ldc, alpha, beta, kCFull)));
    return q.flush();
}

// -------------------------------------------------------------------- TRSM
namespace {

constexpr long kLeafMax = 256;   // largest diagonal block the leaf kernels take
// Updates whose solved block X1 is at least this wide read it from FP16 planes
// converted once for the update; narrower ones read X1 in FP32 and convert it on its
// way to LDS in every tile column (at most eight of them at m x 1024): a separate
// conversion launch costs more than that there (at n = m = 16384, 57 of the 63
// updates, 0.77 ms of conversions per solve on MI355X)
constexpr long kXPlanesMinK = 2048;

struct Leaf {
    long off, s;
};

struct PlanOp {
    long type, a, b, c, d;   // leaf: 0, off, s;  update: 1, c0, w, k0, kw
    bool acc = false;        // update inside a wide leaf block: runs in the leaf tier
};

struct TrsmCtx {
    tb_handle_s* h;
    int tier;        // the routine's tier
    int upd_tier;    // tier of the recursion's update GEMMs
    int acc_tier;    // tier of leaf applications and of updates inside large leaves
    bool f64;
    long m, n, nb;
    const char* A;
    long lda;
    char* B;
    long ldb;
    size_t es;       // element size
    // workspace
    char* Z;         // inverses of the leaves, transposed, each s x s dense
    Planes Zp;       // their FP16 planes (acc tier), rows of all leaves stacked
    Planes Lp;       // FP16 planes of the rows of L (update tier), shifted by kshift
    int kshift = 0;  // column shift of Lp so that the usual update offsets are 16B aligned
    int* eBX;        // running per-row exponent bound of B / X (split leaves and updates)
    int* eL;         // exponent of each row of L (lower part)
    int* eZ;         // exponents of the rows of each W_jj (columns of Z)
    int* eX;         // the eBX values the planes in xmem were scaled with
    char* xmem;      // FP16 planes of the solved block read by the current update / leaf
    char* ldiag;     // L(j, j) for every column j (element size es)
    std::vector<Leaf> leaves;
    std::vector<PlanOp> ops;    // the schedule, recorded by the collecting pass
    std::vector<size_t> zoff;   // element offset of each leaf's Z block (= s * first row)
    std::vector<size_t> eoff;   // offset of each leaf's eZ block (= its first stacked row)
    size_t next_leaf = 0;
    bool collect = true;        // first pass: collect the leaves only

    const void* Aat(long i, long j) const { return A + (i + j * lda) * es; }
    char* Bcol(long j) const { return B + j * ldb * es; }
    bool split_exps() const { return upd_tier == kSplitF16 || acc_tier == kSplitF16; }
    bool l_planes() const { return fp16_tier(upd_tier); }

    // op(A) of p := FP16 planes of B(:, k0:k0+kw), scaled with the current eBX, which
    // the product then reads through its snapshot eX (its own epilogue raises eBX)
    int x_planes(int tier, long k0, long kw, GemmProb* p)
    {
        const bool split = tier == kSplitF16;
        Planes xp;
        TB_TRY(make_planes(h, planes_of(h, tier), (const float*)Bcol(k0), ldb, kOpN, m, kw,
                           split ? eBX : nullptr, xmem, &xp, 0, split ? eX : nullptr));
        p->A = xp.p;
        p->lda = xp.ldp;
        p->spA = xp.sp;
        p->srcA = kSrcPlanes;
        if (split) p->eA = eX;
        return TB_SUCCESS;
    }

    // B(:, c0:c0+w) -= B(:, k0:k0+kw) * A(c0:c0+w, k0:k0+kw)^T
    // in_leaf: an update inside a leaf block wider than the leaf kernel (acc tier)
    int update(long c0, long w, long k0, long kw, int tier, bool in_leaf = false)
    {
        if (w <= 0 || kw <= 0) return TB_SUCCESS;
        if (collect) {
            ops.push_back({1, c0, w, k0, kw, in_leaf});
            return TB_SUCCESS;
        }
        if (tier == kOzakiI8)
            return ozaki_gemm_host(h, m, w, kw, -1.0, (const double*)Bcol(k0), ldb, kOpN,
                                   (const double*)Aat(c0, k0), lda, kOpT, nullptr, nullptr, 1.0,
                                   (double*)Bcol(c0), ldb, false);
        GemmProb p = make_prob(m, w, kw, Bcol(k0), ldb, kOpN, Aat(c0, k0), lda, kOpT, Bcol(c0), ldb,
                               -1.0, 1.0, kCFull);
        if (tier == kSplitF16) {
            p.eA = eBX;
            p.eB = eL + c0;
        }
        // the L planes serve updates of the update tier whose column offset is 16-byte
        // aligned and whose width is whole 32-column stages (the kernel copies planes
        // by stages and the planes continue past the block)
        if (l_planes() && tier == upd_tier && (k0 + kshift) % 8 == 0 && kw % 32 == 0) {
            const PlaneRef lr = plane_ref(planes_of(h, tier), Lp.p, Lp.ldp, c0, kshift + k0);
            p.B = lr.p;
            p.kofsB = lr.kofs;
            p.ldb = Lp.ldp;
            p.spB = Lp.sp;
            p.srcB = kSrcPlanes;
            if (kw >= kXPlanesMinK) {
                // the solved block X1 is final: convert it once here rather than in
                // every tile column of the product
                TB_TRY(x_planes(tier, k0, kw, &p));
            } else if (tier == kSplitF16) {
                // X1 in FP32, split in the tiles with a snapshot of its rows' exponents
                // (this product's epilogues raise eBX meanwhile)
                TB_HIP(hipMemcpyAsync(eX, eBX, sizeof(int) * m, hipMemcpyDeviceToDevice, h->stream));
                p.eA = eX;
            }
        } else if (tier == kSplitF16) {
            // FP32 operands: every tile scales X1's rows by eA, which this product's
            // epilogues raise meanwhile; read a snapshot so all tiles agree
            TB_HIP(hipMemcpyAsync(eX, eBX, sizeof(int) * m, hipMemcpyDeviceToDevice, h->stream));
            p.eA = eX;
        }
        if (split_exps() && !f64) p.eOut = eBX;
        GemmQueue q(h, tier, kShapeSquare, p.srcA, p.srcB);
        TB_TRY(q.add(p));
        return q.flush();
    }

    // leaf: B(:, off:off+s) := B * inv(L_jj)^T = (B + B W^T) D^-1, s <= kLeafMax.  The
    // GEMM only forms the small correction B W^T (|W| ~ |offdiag L| / |diag L|) on top
    // of B, and the epilogue divides B + B W^T by the diagonal in the tier's precision
    // (FP32: TwoSum-corrected, about one rounding): the rounding of the dominant
    // diagonal term happens once, as in substitution.
    int apply_leaf(long off, long s)
    {
        if (collect) {
            leaves.push_back({off, s});
            ops.push_back({0, off, s, 0, 0, false});
            return TB_SUCCESS;
        }
        const size_t q = next_leaf++;
        GemmProb p = make_prob(m, s, s, Bcol(off), ldb, kOpN, Z + zoff[q] * es, s, kOpN, Bcol(off),
                               ldb, 1.0, 1.0, kCFull);
        p.cdiv = ldiag + off * es;   // X = (B + B W^T) D^-1
        if (acc_tier == kSplitF16) {
            p.B = plane_ref(planes_of(h, acc_tier), Zp.p, Zp.ldp, eoff[q], 0).p;
            p.ldb = Zp.ldp;
            p.spB = Zp.sp;
            p.srcB = kSrcPlanes;
            p.eB = eZ + eoff[q];
            p.eOut = eBX;
            // op(A) = the leaf's own block of B, read in FP32 and split on its way to
            // LDS (no separate plane conversion): a leaf tile owns all s columns of its
            // rows, so each element is converted once, and no other tile of the launch
            // raises those rows' eBX while the tile reads it
            p.srcA = kSrcF32;
            p.eA = eBX;
        }
        GemmQueue g(h, acc_tier, kShapeLeaf, p.srcA, p.srcB);
        TB_TRY(g.add(p));
        return g.flush();
    }

    // a diagonal block solved at FP32 (FP64) accuracy: the reference's cublasStrsm.
    // Blocks wider than the leaf kernels are halved at multiples of 8 columns.
    int leaf_block(long off, long s)
    {
        if (s <= kLeafMax) return apply_leaf(off, s);
        const long hh = round_up(s / 2, 8);
        TB_TRY(leaf_block(off, hh));
        TB_TRY(update(off + hh, s - hh, off, hh, acc_tier, true));
        return leaf_block(off + hh, s - hh);
    }

    // BLAS/tc_trsm.cu:31-80 tc_rtrsm_p2
    int p2(long off, long s)
    {
        if (s <= nb) return leaf_block(off, s);
        const long hh = s / 2;
        TB_TRY(p2(off, hh));
        TB_TRY(update(off + hh, s - hh, off, hh, upd_tier));
        return p2(off + hh, s - hh);
    }

    // BLAS/tc_trsm.cu:83-137 tc_trsm_p3
    int p3()
    {
        std::vector<long> ch = plan_chunks_trsm(n);
        long off = 0, rest = n;
        for (int c = (int)ch.size() - 1; c >= 0; --c) {
            const long nn = ch[c];
            if (nn % 2048 == 0) TB_TRY(p2(off, nn));
            else TB_TRY(leaf_block(off, nn));
            if (c != 0) {
                rest -= nn;
                TB_TRY(update(off + nn, rest, off, nn, upd_tier));
            }
            off += nn;
        }
        return TB_SUCCESS;
    }
};

int tiers_of(int tier, int* upd, int* acc)
{
    switch (tier) {
    case kPlainF16: *upd = kPlainF16; *acc = kSplitF16; return TB_SUCCESS;
    case kSplitF16: *upd = kSplitF16; *acc = kSplitF16; return TB_SUCCESS;
    case kFmaF32: *upd = kFmaF32; *acc = kFmaF32; return TB_SUCCESS;
    case kMfmaF64: *upd = kMfmaF64; *acc = kMfmaF64; return TB_SUCCESS;
    case kOzakiI8: *upd = kOzakiI8; *acc = kMfmaF64; return TB_SUCCESS;
    default: return TB_ERR_INVALID;
    }
}

int trsm_check(tb_handle_s* h, int tier, long m, long n, const void* A, long lda, void* B,
               long ldb, long nb)
{
    if (!h || bad_tier(tier) || !dims_ok(m) || !dims_ok(n) || nb <= 0) return TB_ERR_INVALID;
    if (lda < std::max(1l, n) || ldb < std::max(1l, m)) return TB_ERR_INVALID;
    if (m > 0 && n > 0 && (!A || !B)) return TB_ERR_INVALID;
    return TB_SUCCESS;
}

struct TrsmLayout {
    size_t oZ, oE, oD, oZp, oLp, oXp, total;
    size_t zsum, esum;
};

// plans the leaves and the workspace layout of trsm_run
TrsmLayout trsm_plan(TrsmCtx& c)
{
    c.collect = true;
    c.leaves.clear();
    c.ops.clear();
    c.p3();
    c.zoff.clear();
    c.eoff.clear();
    size_t z = 0, e = 0;
    for (const Leaf& l : c.leaves) {
        c.zoff.push_back(z);
        c.eoff.push_back(e);
        z += (size_t)l.s * l.s;
        e += (size_t)l.s;
    }
    // the first chunk (offset 0) is the one not a multiple of 2048; every later
    // block offset is congruent to its size modulo 8
    std::vector<long> ch = plan_chunks_trsm(c.n);
    const long first = ch.empty() ? 0 : ch.back();
    c.kshift = (int)((8 - first % 8) % 8);
    TrsmLayout t;
    t.zsum = z;
    t.esum = e;
    Arena a;
    t.oZ = a.reserve(z * c.es);
    t.oE = a.reserve(sizeof(int) * (2 * c.m + c.n + e));
    t.oD = a.reserve(c.es * c.n);
    t.oZp = c.acc_tier == kSplitF16 ? a.reserve(planes_bytes(planes_of(c.h, kSplitF16), (long)e, kLeafMax)) : 0;
    t.oLp = c.l_planes() ? a.reserve(planes_bytes(planes_of(c.h, c.upd_tier), c.n, c.n, c.kshift)) : 0;
    // planes of the widest solved block an update or a leaf reads
    long xk = 0;
    for (const PlanOp& o : c.ops) xk = std::max(xk, o.type == 0 ? o.b : o.d);
    const int xnpl = std::max(planes_of(c.h, c.upd_tier), planes_of(c.h, c.acc_tier));
    t.oXp = xnpl > 0 ? a.reserve(planes_bytes(xnpl, c.m, xk)) : 0;
    t.total = a.total() + 1024;
    return t;
}

int trsm_run(TrsmCtx& c, const TrsmLayout& t, char* ws)
{
    c.Z = ws + t.oZ;
    c.eBX = reinterpret_cast<int*>(ws + t.oE);
    c.eL = c.eBX + c.m;
    c.eZ = c.eL + c.n;
    c.eX = c.eZ + t.esum;
    c.xmem = ws + t.oXp;
    c.ldiag = ws + t.oD;
    hipStream_t st = c.h->stream;
    if (c.split_exps()) {
        TB_HIP(fill_i32(c.eBX, c.m + c.n + (long long)t.esum, kExpZero, st));
        TB_HIP(rowexp(false, c.B, c.ldb, kOpN, (int)c.m, (int)c.n, 1, 0, c.eBX, 0, st));
        TB_HIP(rowexp(false, c.A, c.lda, kOpTriL, (int)c.n, (int)c.n, 1, 0, c.eL, 0, st));
    }
    if (c.l_planes()) {
        const int npl = planes_of(c.h, c.upd_tier);
        // every update reads L below its diagonal blocks (rows c0 .. c0+w-1 at columns
        // k0 .. k0+kw-1 with k0 + kw <= c0): blocks above the diagonal stay unwritten
        TB_TRY(make_planes(c.h, npl, (const float*)c.A, c.lda, kOpTriL, c.n, c.n, c.eL,
                           ws + t.oLp, &c.Lp, c.kshift, nullptr, true));
    }
    if (c.acc_tier == kSplitF16) {
        c.Zp.p = reinterpret_cast<_Float16*>(ws + t.oZp);
        c.Zp.ldp = kLeafMax;
        c.Zp.sp = (long long)t.esum * kLeafMax;
    }

    // inverses of all diagonal blocks: one launch per run of equal, adjacent leaves
    for (size_t q = 0; q < c.leaves.size();) {
        size_t r = q + 1;
        while (r < c.leaves.size() && c.leaves[r].s == c.leaves[q].s &&
               c.leaves[r].off == c.leaves[r - 1].off + c.leaves[q].s)
            ++r;
        const long s = c.leaves[q].s, off = c.leaves[q].off;
        const int cnt = (int)(r - q);
        hipEvent_t e0;
        TB_TRY(prof_begin(c.h, &e0));
        TB_HIP(trtri_lower_t(c.f64, c.Aat(off, off), c.lda, s * (c.lda + 1), (int)s, cnt,
                             c.Z + c.zoff[q] * c.es, c.ldiag + off * c.es, st));
        TB_TRY(prof_end(c.h, kProfTrtri, e0, (double)s * s * s / 3.0 * cnt));
        if (c.acc_tier == kSplitF16) {
            // the blocks of the run stacked: row R = q'*s + j at Z + R*s, K = s
            int* ez = c.eZ + c.eoff[q];
            const float* zq = reinterpret_cast<const float*>(c.Z + c.zoff[q] * c.es);
            TB_HIP(rowexp(false, zq, s, kOpT, (int)(s * cnt), (int)s, 1, 0, ez, 0, st));
            TB_HIP(split_planes(planes_of(c.h, kSplitF16), zq, s, kOpT, (int)(s * cnt), (int)s, ez,
                                c.Zp.p + (long long)c.eoff[q] *
                                             plane_row_stride(planes_of(c.h, kSplitF16), c.Zp.ldp),
                                c.Zp.ldp, c.Zp.sp, 0, st));
        }
        q = r;
    }
    c.collect = false;
    c.next_leaf = 0;
    return c.p3();
}

TrsmCtx make_ctx(tb_handle_s* h, int tier, long m, long n, const void* A, long lda, void* B,
                 long ldb, long nb)
{
    TrsmCtx c;
    c.h = h;
    c.tier = tier;
    tiers_of(tier, &c.upd_tier, &c.acc_tier);
    c.f64 = tier_is_f64(tier);
    c.m = m;
    c.n = n;
    c.nb = nb;
    c.A = static_cast<const char*>(A);
    c.lda = lda;
    c.B = static_cast<char*>(B);
    c.ldb = ldb;
    c.es = c.f64 ? 8 : 4;
    return c;
}

}  // namespace

int trsm(tb_handle_s* h, int tier, long m, long n, const void* A, long lda, void* B, long ldb,
         long nb)
{
    TB_TRY(trsm_check(h, tier, m, n, A, lda, B, ldb, nb));
    if (m == 0 || n == 0) return TB_SUCCESS;
    TrsmCtx c = make_ctx(h, tier, m, n, A, lda, B, ldb, nb);
    const TrsmLayout t = trsm_plan(c);
    char* ws;
    TB_TRY(workspace_get(h, t.total, &ws));
    return trsm_run(c, t, ws);
}

void trsm_schedule(long n, long nb, std::vector<TrsmStep>* steps, std::vector<TrsmLeaf>* leaves)
{
    TrsmCtx c = make_ctx(nullptr, kSplitF16, 1, n, nullptr, n, nullptr, 1, nb);
    trsm_plan(c);
    steps->clear();
    leaves->clear();
    for (const PlanOp& o : c.ops) steps->push_back({o.type, o.a, o.b, o.c, o.d, o.acc});
    for (const Leaf& l : c.leaves) leaves->push_back({l.off, l.s});
}

int plan_trsm(long n, long nb, long* out, int cap)
{
    if (n < 0 || nb <= 0 || cap < 0 || (!out && cap > 0)) return -TB_ERR_INVALID;
    TrsmCtx c = make_ctx(nullptr, kSplitF16, 1, n, nullptr, n, nullptr, 1, nb);
    trsm_plan(c);
    if ((long)c.ops.size() > cap) return -TB_ERR_INVALID;
    for (size_t q = 0; q < c.ops.size(); ++q) {
        const PlanOp& o = c.ops[q];
        long* w = out + 5 * q;
        w[0] = o.type; w[1] = o.a; w[2] = o.b; w[3] = o.c; w[4] = o.d;
    }
    return (int)c.ops.size();
}

int trsm_refined(tb_handle_s* h, int tier, long m, long n, const void* A, long lda, void* B,
                 long ldb, long nb)
{
    TB_TRY(trsm_check(h, tier, m, n, A, lda, B, ldb, nb));
    if (tier_is_f64(tier)) return TB_ERR_INVALID;
    if (m == 0 || n == 0) return TB_SUCCESS;
    TrsmCtx c = make_ctx(h, tier, m, n, A, lda, B, ldb, nb);
    const TrsmLayout t = trsm_plan(c);
    Arena a;
    const size_t oR = a.reserve(sizeof(double) * m * n);   // B0, then the residual
    const size_t oX = a.reserve(sizeof(double) * m * n);   // X in FP64
    const size_t oL = a.reserve(sizeof(double) * n * n);   // lower(L) in FP64
    const size_t oD = a.reserve(sizeof(float) * m * n);    // correction
    const size_t oT = a.reserve(t.total);
    char* ws;
    TB_TRY(workspace_get(h, a.total(), &ws));
    hipStream_t st = h->stream;
    double* R = reinterpret_cast<double*>(ws + oR);
    double* X64 = reinterpret_cast<double*>(ws + oX);
    double* L64 = reinterpret_cast<double*>(ws + oL);
    float* D = reinterpret_cast<float*>(ws + oD);
    TB_HIP(copy_matrix(false, true, m, n, B, ldb, R, m, st));
    TB_TRY(trsm_run(c, t, ws + oT));
    TB_HIP(copy_matrix(false, true, m, n, B, ldb, X64, m, st));
    TB_HIP(fill_matrix(true, n, n, L64, n, 0.0, st));
    TB_HIP(copy_lower(false, true, n, A, lda, L64, n, st));
    // R = B0 - X * L^T on FP64 MFMA
    GemmQueue q(h, kMfmaF64, kShapeSquare);
    TB_TRY(q.add(make_prob(m, n, n, X64, m, kOpN, L64, n, kOpT, R, m, -1.0, 1.0, kCFull)));
    TB_TRY(q.flush());
    TB_HIP(copy_matrix(true, false, m, n, R, m, D, m, st));
    TrsmCtx c2 = make_ctx(h, tier, m, n, A, lda, D, m, nb);
    const TrsmLayout t2 = trsm_plan(c2);
    TB_TRY(trsm_run(c2, t2, ws + oT));
    return hip_status(add_matrix_f32(m, n, B, ldb, D, m, st));
}

}  // namespace tb
