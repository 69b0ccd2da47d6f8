// Device tables of a batch plan: many independent SYRK and TRSM instances of mixed
// sizes run by one persistent kernel (batch.hip), planned on the host
// (host/tb_batch.cpp).
//
// Work items:
//   SYRK  one lower output tile C(i0:, j0:) of one instance;
//   TRSM  one panel of 64 rows of X*L^T = B of one instance.  Rows of X depend only
//         on the same rows of B, so a panel walks the instance's whole recursion
//         (BLAS/tc_trsm.cu:31-137: leaves and updates, in schedule order) by itself
//         and no workgroup ever waits for another.
// Items are sorted by decreasing work and handed out through an atomic counter, so
// the largest panels start first and the small ones fill the tail.
#pragma once
#include <cstdint>

namespace tb {

enum BatchOp : int { kBatchSyrk = 0, kBatchTrsm = 1 };

constexpr int kPanelRows = 64;     // rows of a TRSM panel (= BM of the panel tiles)
constexpr int kPanelCols = 256;    // columns of a panel tile (>= every leaf width)

struct BatchInst {
    int op;
    int m, n, k;                 // SYRK: n x k A; TRSM: m x n B, n x n L
    float alpha, beta;           // SYRK
    const float* A;              // SYRK: A; TRSM: L (lower part read)
    long long lda;
    float* C;                    // SYRK: C (lower part written); TRSM: B, overwritten by X
    long long ldc;
    int* eRow;                   // split exponents: SYRK rows of A; TRSM running bound per row of B
    int* eSnap;                  // TRSM: eRow of a panel's rows as of the start of a step
    const int* eL;               // TRSM: exponents of the rows of L
    const int* eZ;               // TRSM: exponents of the columns of each leaf's W^T, stacked
    const float* Z;              // TRSM: W^T = (D inv(L_jj) - I)^T of each leaf, s x s, stacked
    const double* dinv;          // TRSM: 1 / L(j, j)
    int step0, nsteps;           // TRSM: its steps in the step table
};

struct BatchStep {
    int type;                    // 0 leaf (a = off, b = s); 1 update (a = c0, b = w, c = k0,
                                 // d = kw); 2 update in the leaf tier (inside a wide leaf block)
    int a, b, c, d;
    long long zoff;              // leaf: element offset of its W^T block in Z
    int eoff;                    // leaf: offset of its exponents in eZ
};

struct BatchItem {
    int inst;
    int i0, j0;                  // SYRK: tile origin; TRSM: first row of the panel (j0 = 0)
    int pad;
};

// Batched prologue jobs: row exponents and diagonal-block inverses of all instances,
// each a contiguous range of blocks of one launch.
struct RowexpJob {
    const float* X;
    long long ld;
    int rows, K;
    int kcontig;                 // 0: rows contiguous (kOpN / kOpTriL), 1: k contiguous (kOpT)
    int lower;                   // kOpTriL: only k <= r
    int* out;
    int blockBegin, rowBlocks, kBlocks;
};

struct TrtriJob {
    const float* L;              // first diagonal block (L + off*(lda+1))
    long long lda, bstride;      // stride between consecutive blocks of the run
    int s, count;                // block size, blocks in the run
    float* Z;                    // count s x s W^T blocks
    double* dinv;                // count*s
    int blockBegin, colBlocks;   // colBlocks = ceil(s / 64)
};

}  // namespace tb
