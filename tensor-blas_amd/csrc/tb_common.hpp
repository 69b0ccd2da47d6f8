// Shared host/device definitions of the grouped GEMM work description.
//
// Every BLAS-3 routine of the library is planned on the host into a short list of
// GEMM problems (the reference's off-diagonal GEMM updates and diagonal-block
// leaves, BLAS/tc_syrk.cu:41-125, tc_trsm.cu:31-137, ...).  One kernel launch runs
// up to kMaxProbs problems, each a strided batch, as one flat list of output tiles.
#pragma once
#include <cstdint>

namespace tb {

// Precision tiers.  The reference's families map to the first three:
//   tc_*            -> kPlainF16   (FP32 -> FP16 operands, FP16 MFMA, FP32 accumulate)
//   tc_cumpsgemm_*  -> kSplitF16   (FP32 operands split into scaled FP16 hi/lo, three
//                                   MFMA products, FP32-accurate)
//   tc_ozimmu_*     -> kOzakiI8    (FP64 operands sliced to INT8, INT8 MFMA, FP64 sum)
// plus two comparators that have no reference family:
//   kFmaF32 -- plain FP32 vector-FMA kernel of the same recursion ("vs plain")
//   kMfmaF64 -- native FP64 MFMA
enum Tier : int {
    kPlainF16 = 0,
    kSplitF16 = 1,
    kOzakiI8 = 2,
    kFmaF32 = 3,
    kMfmaF64 = 4,
};

inline bool tier_is_f64(int t) { return t == kOzakiI8 || t == kMfmaF64; }

// How a GEMM operand is read.  op(A) is m x k, op(B) is k x n; matrices are
// column-major with leading dimension ld.
enum OpMode : int {
    kOpN = 0,     // op(X) = X
    kOpT = 1,     // op(X) = X^T
    kOpSymL = 2,  // op(A) = A, A symmetric with only its lower triangle stored (SYMM)
    kOpTriL = 3,  // op(A) = A, A lower triangular, strict upper part treated as zero (TRMM)
};

// How an operand of the FP16 tiers is stored: an FP32 matrix converted while it is
// staged, or pre-converted K-contiguous FP16 planes (1 plane plain, 3 planes split).
enum SrcKind : int { kSrcF32 = 0, kSrcPlanes = 1 };

// Which part of C a problem writes.
enum CMode : int {
    kCFull = 0,
    kCLower = 1,  // only C(i,j) with i >= j; square C, tiles above the diagonal skipped
};

struct GemmProb {
    const void* A;
    const void* B;
    const void* A2;   // second product pair (SYR2K): C += alpha*op(A2)*op(B2); may be null
    const void* B2;
    void* C;
    const int* eA;    // split tier: power-of-two exponent of each row of op(A) (and op(A2)),
                      // any upper bound of the row's max exponent is valid
    const int* eB;    // split tier: exponent of each column of op(B) (and op(B2))
    int* eOut;        // if set: atomicMax of the exponents of the values written to each
                      // row of C (keeps a running per-row bound for later split GEMMs)
    const void* cdiv;      // if set: C(i,j) = (alpha*AB + beta*C)(i,j) / cdiv[j] (the TRSM
                           // leaves' diagonal), in the tier's own precision: float* for the
                           // FP32-class tiers (a TwoSum-corrected FP32 division), double*
                           // for the FP64 tier
    long long lda, ldb, ldc;
    long long sA, sB, sC;      // batch strides (elements)
    long long seA, seB;        // batch strides of eA / eB
    long long spA, spB;        // plane strides of pre-converted FP16 operands (elements)
    double alpha, beta;
    float alpha32, beta32;     // the same, rounded to FP32 on the host (FP32-class tiers)
    int m, n, k, batch;
    int modeA, modeB, modeC;
    int srcA, srcB;            // FP16 tiers: 0 = FP32 matrix, 1 = K-contiguous FP16 planes
                               // (then A/B point at plane 0 and lda/ldb are the plane ld)
    int kofsA, kofsB;          // interleaved 2-plane operands: column offset (< 32) of k = 0
                               // from the 32-column block A/B point at (plane_ref)
    int tilesM, tilesN;        // tile grid of one batch element (tilesM == tilesN for kCLower)
    int tilesPer;              // tiles of one batch element
    int tileBegin;             // index of this problem's first tile in the launch
};

constexpr int kMaxProbs = 12;

struct GemmLaunch {
    GemmProb p[kMaxProbs];
    int np;
    int totalTiles;
};

}  // namespace tb
