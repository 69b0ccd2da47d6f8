// Host-callable launchers of the library's device kernels (all asynchronous on the
// given stream).  Kernels live in gemm_dispatch.hip, aux_kernels.hip, ozaki.hip.
#pragma once
#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>

#include "tb_common.hpp"

namespace tb {

// ------------------------------------------------------------- grouped GEMM
enum GemmShape : int {
    kShapeSquare = 0,  // 128 x 128 tiles: off-diagonal updates and whole routines
    kShapeLeaf = 1,    // 64 x 256 tiles (FP64: 32 x 256): a workgroup owns all n <= 256
                       // columns of its rows, so C may alias op(A) (in-place TRSM leaf
                       // X := B * inv(L)^T)
    kShapeWide = 2,    // 256 x 256 tiles (plain and 2-plane split): full or lower C,
                       // both operands in FP16 planes; chosen by GemmQueue::flush for
                       // launches with enough tiles to fill the device
};
struct TileShape {
    int bm, bn;
};
TileShape gemm_tile_shape(int tier, int shape);
// FP16 planes per operand element of a tier: plain 1, split `split_planes` (2 or 3),
// others 0
int tier_planes(int tier, int split_planes);
// all problems of L share the operand source kinds srcA/srcB (GemmProb::srcA/srcB);
// npl: planes of the split tier (2 or 3); max_cus > 0 caps the grid at the workgroups
// that fit on that many CUs (the kernels loop over the remaining tiles)
hipError_t gemm_launch(int tier, int npl, int shape, int srcA, int srcB, const GemmLaunch& L,
                       hipStream_t s, int max_cus = 0);
// Grid of a tile-looping kernel: every tile (total) when max_cus <= 0, else at most the
// workgroups of `kernel` resident on max_cus CUs, rounded down to a multiple of 8 (the
// XCD count, which the kernels' xcd_remap tile order relies on).
int capped_grid(const void* kernel, int threads, int total, int max_cus);

// FP16 plane layouts, element (plane p, row r, column c), c = kshift + k counted from
// the row start:
//   separate planes (plain: 1; split with 3):  p*sp + r*ldp + c
//   the two split planes, interleaved per 32 columns:  r*2*ldp + (c/32)*64 + p*32 + c%32
// so that one 32-column stage of a 2-plane row is one contiguous 128-byte segment: the
// direct-to-LDS loads stream 128-byte segments about half again as fast as 64-byte ones
// (MI355X, measured).  ldp (a multiple of 64) is the columns of one plane's row either
// way, so both layouts take npl * rows * ldp halves.
__host__ __device__ constexpr bool planes_il(int npl) { return npl == 2; }
__host__ __device__ inline long long plane_row_stride(int npl, long long ldp)
{
    return planes_il(npl) ? 2 * ldp : ldp;
}
__host__ __device__ inline long long plane_off(int npl, long long ldp, long long sp, int p, long long r,
                                               long long c)
{
    return planes_il(npl) ? r * 2 * ldp + (c >> 5) * 64 + p * 32 + (c & 31) : p * sp + r * ldp + c;
}
// A GEMM operand that starts at row r0 and column c0 of planes at base: the pointer the
// GemmProb takes, and the column offset kofs (< 32, a multiple of 8) it takes with it
// (interleaved planes only: their pointer stays on a 32-column block boundary).
struct PlaneRef {
    const _Float16* p;
    int kofs;
};
__host__ __device__ inline PlaneRef plane_ref(int npl, const _Float16* base, long long ldp, long long r0,
                                              long long c0)
{
    if (planes_il(npl)) return {base + r0 * 2 * ldp + (c0 >> 5) * 64, (int)(c0 & 31)};
    return {base + r0 * ldp + c0, 0};
}

// Pre-conversion of the rows of op(X) (FP32, rowexp modes kOpN/kOpT/kOpSymL/kOpTriL,
// rows x K) into npl K-contiguous FP16 planes: element (r, k) of plane p at
// out[plane_off(npl, ldp, sp, p, r, kshift + k)].  npl = 1: RNE fp16(x).  With
// xs = x * 2^(15 - e[r]) (e: row exponents, kExpZero -> 0) and h = RNE fp16(xs):
// npl = 2: m = RNE fp16(xs - h) (xs ~ h + m);  npl = 3: xs = h + 2^-11 m + 2^-22 l.
// ldp must be a multiple of 64 with ldp >= kshift + K; entries of a row outside
// [kshift, kshift + K) are written as zero.  e_used (optional, split tier): receives
// e[r] as read, the exponents a GEMM must use with these planes when e keeps changing.
// lower_only (mode kOpTriL): 64 x 64 blocks wholly above the diagonal are left as they
// are instead of written as zeros, for consumers that read only below the diagonal.
hipError_t split_planes(int npl, const float* X, long long ld, int mode, int rows, int K,
                        const int* e, _Float16* out, long long ldp, long long sp, int kshift,
                        hipStream_t s, int* e_used = nullptr, bool lower_only = false);

// -------------------------------------------------------------- row exponents
// Sentinel exponent of an all-zero row.
constexpr int kExpZero = -100000;

// out[b*so + r] := max(out[...], max_k expo(op(X)(r, k))) for r < rows, where
// expo(x) is the e of x = f*2^e, f in [0.5, 1) (kExpZero for x == 0) and
//   mode kOpN   : op(X)(r, k) = X[r + k*ld], k < K
//   mode kOpTriL: same, k <= r only
//   mode kOpSymL: row r of the symmetric matrix stored lower: X[r + k*ld] for k <= r
//                 and X[k + r*ld] for r < k < K
//   mode kOpT   : op(X)(r, k) = X[k + r*ld]
// out must hold kExpZero (fill_i32) or earlier maxima.  Batched over nb items with
// element stride sx and output stride so.
hipError_t rowexp(bool f64, const void* X, long long ld, int mode, int rows, int K, int nb,
                  long long sx, int* out, long long so, hipStream_t s);
hipError_t fill_i32(int* p, long long n, int v, hipStream_t s);

// ------------------------------------------------------ diagonal-block inverses
// For q < nblocks: L_q = the s x s lower triangle at L + q*bstride (ld lda), factored
// as L_q = (I + N) D with D = diag(L_q).  Writes Z_q = W^T with W = (I+N)^-1 - I =
// D inv(L_q) - I (strictly lower; Z_q dense s x s, ld s, at Z + q*s*s) and
// ldiag[q*s + r] = L_q(r, r).  Then X L_q^T = B  <=>  X = (B + B W^T) D^-1.
// Arithmetic, L, Z and ldiag in FP64 when f64, else FP32 (the FP32-class tiers keep
// their leaves free of FP64).  s <= 256.
// columns of inv(L) per 64-lane block (one wave) of the diagonal-block inversion
constexpr int kTrtriCols = 8;
hipError_t trtri_lower_t(bool f64, const void* L, long long lda, long long bstride, int s,
                         int nblocks, void* Z, void* ldiag, hipStream_t st);
// Writes inv(L_q) = D^-1 (I + W) into the lower triangle of the s x s block A (ld lda)
// from trtri_lower_t's outputs of one block (Z s x s, ldiag s), in the same precision;
// unit: the diagonal of A is left as it is.  The strict upper triangle is not written.
hipError_t trtri_store(bool f64, const void* Z, const void* ldiag, int s, void* A, long long lda,
                       bool unit, hipStream_t st);
// *info = min(*info, i + 1) over the i < n with A(i, i) == 0
hipError_t diag_zero_scan(bool f64, const void* A, long long lda, long long n, int* info,
                          hipStream_t st);

// ------------------------------------------------------------- conversions etc.
// ah(i,j) = RNE fp16(as(i,j)), m x n (util/util.cu:127-135 s2h; include/TensorBLAS.h:185-284)
hipError_t s2h(long long m, long long n, const float* as, long long ldas, void* ah, long long ldah,
               hipStream_t s);
// row-major m x n -> column-major m x n fp16 (util/util.cu:137-145 s2hTranspose)
hipError_t s2h_transpose(long long m, long long n, const float* as, void* ah, hipStream_t s);
// A(i,j) := A(j,i) for i < j (util/util.cu:148-167)
hipError_t mirror_lower(bool f64, long long n, void* A, long long lda, hipStream_t s);

// dst (rows x cols, ld ldd) := alpha * src under an index map: with
// i' = kXformRevRows ? rows-1-i : i and j' = kXformRevCols ? cols-1-j : j, dst(i, j) =
// src(j', i') if kXformTrans else src(i', j'); kXformUnitDiag makes dst's diagonal 1.
// src and dst must not overlap.  Float or double (f64).  Tiled through LDS so that
// both sides are read and written in columns.
enum XformFlags : int {
    kXformTrans = 1,
    kXformRevRows = 2,
    kXformRevCols = 4,
    kXformUnitDiag = 8,
};
hipError_t xform(bool f64, const void* src, long long lds, void* dst, long long ldd, int rows, int cols,
                 int flags, double alpha, hipStream_t s);
// a(i,j) := v (util/util.cu:240-257)
hipError_t fill_matrix(bool f64, long long m, long long n, void* a, long long lda, double v,
                       hipStream_t s);
// b(i,j) := a(i,j) with type conversion (util/util.cu:263-280 is the same-type case)
hipError_t copy_matrix(bool a_f64, bool b_f64, long long m, long long n, const void* a,
                       long long lda, void* b, long long ldb, hipStream_t s);
// b(i,j) := a(i,j) for i >= j only (lower triangle), with type conversion
hipError_t copy_lower(bool a_f64, bool b_f64, long long n, const void* a, long long lda, void* b,
                      long long ldb, hipStream_t s);
// x(i,j) += y(i,j), FP32
hipError_t add_matrix_f32(long long m, long long n, void* x, long long ldx, const float* y,
                          long long ldy, hipStream_t s);

// ---------------------------------------------------------------------- Ozaki
// Slices of the rows of op(X) (FP64), 7-bit digits (oracle/oracle.c orc_ozaki_slice):
// out layout, blocked by 32 k and 64 rows (R = ceil(rows / 64) row blocks, rows padded
// with zero digits to 64 R, k to kpad, a multiple of 64):
//   slices[q*S*64R*kpad + (((p/32)*R + r/64)*S + s)*2048 + (r%64)*32 + p%32]  int8,
// so that one 32-k stage of a 64-row panel (all S slices) is 2048 S contiguous bytes.
// exps[q*rows + r].  mode kOpN/kOpT/kOpSymL/kOpTriL as rowexp.  exps must be filled by
// rowexp first.
// rblocks: 64-row blocks of the layout (0: ceil(rows / 64)); blocks past the rows hold
// zero digits.
hipError_t ozaki_slice(const double* X, long long ld, int mode, int rows, int K, int kpad,
                       int nslices, int nb, long long sx, const int* exps, int8_t* slices,
                       hipStream_t s, int rblocks = 0);
// rows of a wide Ozaki tile: with OzakiProb::wide, op(A)'s slices must be laid out with
// a multiple of this many rows (an even number of 64-row blocks)
int ozaki_wide_rows();
// the wide kernel exists for these slices / groups
bool ozaki_wide_ok(int nslices, int ngroups);
struct OzakiProb {
    const int8_t* sa;   // slices of op(A) rows   [kpad/32][ceil(m/64)][S][64][32]
    const int8_t* sb;   // slices of op(B) columns [kpad/32][ceil(n/64)][S][64][32]
    const int8_t* sa2;  // second pair (SYR2K) or null
    const int8_t* sb2;
    const int* ea;      // exponents of op(A) rows
    const int* eb;      // exponents of op(B) columns
    double* C;
    long long ldc;
    double alpha, beta;
    int m, n, kpad, nslices, kchunk;
    int ngroups;        // slice-pair groups g = 0 .. ngroups-1 kept: nslices or nslices + 1
    int lowerC;
    int wide;           // 128 x 64 tiles on eight waves (needs kpad <= kchunk and op(A)'s
                        // slices over a multiple of ozaki_wide_rows() rows); longer K and a
                        // SYR2K pair are one launch per chunk of each product, sa2 / sb2
                        // unused:
    int partial;        //   1: store the unscaled FP64 sum into C (the first chunk)
                        //   3: add this chunk onto that sum (a middle chunk)
                        //   2: add this chunk onto that sum, then scale (beta must be 0)
};
hipError_t ozaki_gemm(const OzakiProb& p, hipStream_t s, int max_cus = 0);

// Cholesky of one s x s diagonal block (s <= kPotrfLeaf) in place, lower triangle, in
// the arithmetic of A: FP32 (f64 = false, the FP32-class tiers) or FP64.  A failing
// pivot at local column j sets *info = off + j + 1 if *info is 0; a block with
// *info != 0 does nothing.
constexpr int kPotrfLeaf = 128;
hipError_t potrf_leaf(bool f64, void* A, long long lda, int s, long long off, int* info,
                      hipStream_t st);

// ------------------------------------------------------ no-allocation scratch
// Bytes of the fixed device scratch behind tb_syrk_noalloc (one array per GPU;
// TB_NOALLOC_SCRATCH_BYTES of the public header).
constexpr size_t kNoAllocScratchBytes = size_t(64) << 20;
// Device address of that array on `device`, which must be the current device: read
// with hipGetSymbolAddress on the first call for the device, then cached.
hipError_t noalloc_scratch(int device, void** out);

// ---------------------------------------------------------------------- batches
struct RowexpJob;
struct TrtriJob;
struct PlanesJob;
struct BatchItem;
struct BatchInst;
struct BatchStep;
struct BatchScratch;
// row exponents / diagonal-block inverses of many operands in one launch each
// (nblocks = the jobs' summed block counts)
hipError_t batch_rowexp(const RowexpJob* jobs, int njobs, int nblocks, hipStream_t s);
hipError_t batch_trtri(const TrtriJob* jobs, int njobs, int nblocks, hipStream_t s);
hipError_t batch_planes(const PlanesJob* jobs, int njobs, int nblocks, hipStream_t s);
// the persistent batch kernel: unpl / anpl = FP16 planes of update (and SYRK / TRMM)
// tiles / leaf tiles, waves = 4, or 8 (2-plane plans only) per workgroup; each SYRK /
// TRMM item names its own tile shape; grid = workgroups (each with its scratch in S);
// counter must hold 0
hipError_t batch_run(int unpl, int anpl, int waves, const BatchItem* items, int nitems,
                     const BatchInst* insts, const BatchStep* steps, int* counter,
                     const BatchScratch& S, int grid, hipStream_t s);
// workgroups of the persistent batch kernel that one device runs at once
int batch_grid(int device);

}  // namespace tb
