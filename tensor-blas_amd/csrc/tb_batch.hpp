// Device tables of a batch plan: many independent SYRK, TRMM and TRSM instances of
// mixed sizes run by one persistent kernel (batch.hip), planned on the host
// (host/tb_batch.cpp).
//
// Work items:
//   SYRK  one lower output tile C(i0:, j0:) of one instance;
//   TRMM  one output tile C(i0:, j0:) of C = alpha L B (L lower, its strict upper part
//         read as zero: the k range of a tile stops at its last row);
//   TRSM  one panel of 64 rows of X*L^T = B of one instance.  Rows of X depend only
//         on the same rows of B, so a panel walks the instance's whole recursion
//         (BLAS/tc_trsm.cu:31-137: leaves and updates, in schedule order) by itself
//         and no workgroup ever waits for another.
// Items are sorted by decreasing work and handed out through an atomic counter, so
// the largest panels start first and the small ones fill the tail.
#pragma once
#include <cstdint>

namespace tb {

enum BatchOp : int { kBatchSyrk = 0, kBatchTrsm = 1, kBatchTrmm = 2 };

// tile shape of a SYRK / TRMM work item
enum BatchShape : int { kShapePanel = 0,    // 64 x 256 (the TRSM panel tile)
                        kShapeSq = 1 };     // 128 x 128

constexpr int kPanelRows = 64;     // rows of a TRSM panel (= BM of the panel tiles)
constexpr int kPanelCols = 256;    // columns of a panel tile (>= every leaf width)

struct BatchInst {
    int op;
    int m, n, k;                 // SYRK: n x k A; TRSM: m x n B, n x n L; TRMM: m x m L, m x n B
    float alpha, beta;           // SYRK, TRMM (beta 0)
    const float* A;              // SYRK: A; TRSM, TRMM: L (lower part read)
    long long lda;
    float* C;                    // SYRK: C (lower part written); TRSM: B, overwritten by X;
                                 // TRMM: C (m x n)
    long long ldc;
    int* eRow;                   // split exponents: SYRK rows of A; TRSM running bound per row
                                 // of B; TRMM rows of L
    int* eSnap;                  // TRSM: eRow of a panel's rows as of the start of a step
    const int* eL;               // TRSM: exponents of the rows of L
    const int* eZ;               // TRSM: exponents of the columns of each leaf's W^T, stacked
    const float* Z;              // TRSM: W^T = (D inv(L_jj) - I)^T of each leaf, s x s, stacked
    const float* ldiag;          // TRSM: L(j, j), the leaves' FP32 divisors
    int step0, nsteps;           // TRSM: its steps in the step table
    // FP16 planes (K-contiguous rows, plane stride sp) made by the prologue:
    const void* LpU;             // TRSM: rows of L in the update tier's planes (ldpL)
    const void* LpA;             // TRSM: rows of L in the leaf tier's planes
    const void* LpU1;            // TRSM: the same shifted right by kshift1 columns, for
    const void* LpA1;            //   updates whose column offset is not a multiple of 8
    long long ldpL, spLU, spLA;
    int kshift1;
    const void* Zp;              // TRSM: rows of every leaf's W^T (leaf tier), ld kZpLd
    long long spZ;
    const void* Ap;              // SYRK: rows of A; TRMM: rows of L (zero above the diagonal)
    long long ldpA, spA;
    const void* Bp;              // TRMM: columns of B (rows of B^T)
    long long ldpB, spB;
    const int* eCol;             // TRMM split exponents of B's columns
};

constexpr int kZpLd = 256;       // plane row length of the leaves' W^T (>= every leaf width)

// per-workgroup scratch of the persistent kernel: a panel's solved block in planes
struct BatchScratch {
    void* base;                  // workgroup b uses base + b * bytes_per_wg
    long long bytes_per_wg;
    long long ldp, sp;           // plane row length and plane stride (64 rows)
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
    int i0, j0;                  // SYRK / TRMM: tile origin; TRSM: first row of the panel (j0 = 0)
    int shape;                   // SYRK / TRMM: BatchShape of the tile
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

struct PlanesJob {               // one FP32 -> FP16 plane conversion (split_planes)
    const float* X;
    long long ld;
    int mode, rows, K, npl, kshift;
    const int* e;                // row exponents (split planes), else null
    void* out;                   // plane 0; planes at out + p * sp (halves)
    long long ldp, sp;
    int blockBegin, gx, gy;      // gx = ldp / 64 plane-column blocks, gy = ceil(rows / 64)
};

struct TrtriJob {
    const float* L;              // first diagonal block (L + off*(lda+1))
    long long lda, bstride;      // stride between consecutive blocks of the run
    int s, count;                // block size, blocks in the run
    float* Z;                    // count s x s W^T blocks
    float* ldiag;                // count*s: L(j, j) of each block
    int blockBegin, colBlocks;   // colBlocks = ceil(s / 64)
};

}  // namespace tb
