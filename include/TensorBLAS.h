/*
 * TensorBLAS.h -- public interface of the MI355X (gfx950) Tensor-BLAS library.
 *
 * Drop-in for the reference's include/TensorBLAS.h:135-167.  Two layers:
 *
 *  1. The C ABI (extern "C", plain pointers and sizes).  Everything a foreign
 *     binding (ctypes, cgo, JNI, N-API) needs:
 *       - print_env, tc_syrk_wrapper, tc_trsm_wrapper: the reference's own C-ABI
 *         exports, same signatures and meaning;
 *       - tb_*: one entry point per routine with an explicit precision tier, which
 *         replace the reference's C++-only tc_*, tc_cumpsgemm_*, tc_ozimmu_* families.
 *  2. The reference's C++ entry points (tc_syrk, tc_cumpsgemm_syrk, ..., 14 of them,
 *     plus tc_ozimmu_symm which the reference lacks) with the same argument order and
 *     meaning.  The cuBLAS / cuMpSGEMM handle arguments become this library's handle
 *     (tb_handle_t); the FP16 workspace arguments are accepted and may be NULL (the
 *     library stages operands itself).
 *
 * Fixed options, as in the reference:
 *   SYRK / SYR2K  C := alpha*A*A^T + beta*C (resp. alpha*(A*B^T + B*A^T) + beta*C),
 *                 lower triangle written, A (and B) n x k, not transposed.
 *   TRMM          C := alpha*L*B, left, lower, not transposed, non-unit, out of place.
 *   TRSM          X*L^T = B, right, lower, transposed, non-unit, B overwritten by X.
 *   SYMM          C := alpha*A*B + beta*C, left, A symmetric with its lower triangle stored.
 * All matrices are column-major device pointers; dimensions are long.  Deliberate
 * differences from the reference (documented at each function): the strictly upper
 * triangle of a SYRK/SYR2K result is not written; TRMM ignores the upper triangle
 * of L instead of requiring zeros; SYMM never modifies A; odd sizes run natively.
 *
 * Errors: every tb_* function returns TB_SUCCESS (0) or a TB_ERR_* code and
 * launches nothing when it rejects its arguments.  The reference's tc_* returned
 * void; here they return the same int status.  The library never prints.
 *
 * Streams: a handle owns one HIP stream (default: the null stream).  tb_* and tc_*
 * calls are asynchronous on it; tc_syrk_wrapper / tc_trsm_wrapper synchronize before
 * returning, as the reference's did (BLAS/tc_syrk.cu:29, tc_trsm.cu:18).
 */
#ifndef TENSORBLAS_MI355X_H
#define TENSORBLAS_MI355X_H

/* exported symbols; everything else in the library has hidden visibility */
#define TB_API __attribute__((visibility("default")))

#ifdef __cplusplus
extern "C" {
#endif

typedef struct tb_handle_s* tb_handle_t;

/* Precision tiers.  TB_TIER_PLAIN is the reference's tc_* family (FP32 operands
 * converted to FP16 round-to-nearest-even, FP16 MFMA, FP32 accumulation);
 * TB_TIER_SPLIT is the tc_cumpsgemm_* family (FP32-accurate: each FP32 operand is
 * scaled per row by a power of two and split into FP16 hi + 2^-11 lo; the three
 * products hh | hm+mh run on FP16 MFMA into two FP32 accumulators; see
 * tb_set_split_planes for the exact three-plane form); TB_TIER_OZAKI is the tc_ozimmu_* family (FP64 operands sliced into
 * INT8 digits, INT8 MFMA with exact int32 sums, FP64 accumulation).  The two
 * comparators have no reference family: TB_TIER_FP32 runs the same recursion on
 * FP32 vector FMA, TB_TIER_FP64 on native FP64 MFMA. */
enum {
    TB_TIER_PLAIN = 0,
    TB_TIER_SPLIT = 1,
    TB_TIER_OZAKI = 2,
    TB_TIER_FP32 = 3,
    TB_TIER_FP64 = 4
};

enum {
    TB_SUCCESS = 0,
    TB_ERR_INVALID = 1,   /* bad dimension, leading dimension, nb, tier or null pointer */
    TB_ERR_DEVICE = 2,    /* a HIP call failed (launch, allocation, copy) */
    TB_ERR_NOMEM = 3,     /* device workspace could not be allocated */
    TB_ERR_SCRATCH = 4    /* the problem needs more than the no-allocation scratch holds */
};

/* ------------------------------------------------------------- handle */
TB_API int tb_create(tb_handle_t* handle);
TB_API int tb_destroy(tb_handle_t handle);
/* stream is a hipStream_t (NULL = null stream); streams keep the default priority */
TB_API int tb_set_stream(tb_handle_t handle, void* stream);
TB_API int tb_synchronize(tb_handle_t handle);
/* Ozaki tier: number of INT8 slices (3..10, default 7) */
TB_API int tb_set_ozaki_slices(tb_handle_t handle, int nslices);
/* Ozaki tier: 1 also sums the slice pairs (s, t) with s + t = nslices (0-based), the
 * first group the triangular scheme drops: nslices - 1 more INT8 products on the same
 * slices, for about the accuracy of one more slice without its memory traffic.
 * 0 (default) is the published scheme.  With extra = 1, nslices must be 6..9.  No
 * counterpart in the reference. */
TB_API int tb_set_ozaki_extra(tb_handle_t handle, int extra);
/* Split tier: FP16 planes per FP32 operand (default 2).  2: x*2^s = h + 2^-11 m
 * (23 significant bits for each row's largest entries), three MFMA products
 * (hh | hm+mh): the published FP16 error-correction scheme.  3: h + 2^-11 m + 2^-22 l,
 * exact for FP32 inputs, six products (hh | hm+mh | hl+mm+lh) at twice the MFMA work,
 * for inputs whose rows span many binades. */
TB_API int tb_set_split_planes(tb_handle_t handle, int nplanes);
/* Occupancy cap (no counterpart in the reference; for callers that run several
 * handles concurrently, such as tb_serve): the handle's GEMM kernels launch at most
 * the workgroups that fit on max_cus compute units and loop over the remaining tiles,
 * so that the rest of the device stays free for other streams.  0 (default): one
 * workgroup per tile.  Results do not depend on it. */
TB_API int tb_set_max_cus(tb_handle_t handle, int max_cus);
/* bytes of device workspace the handle currently holds */
TB_API long tb_workspace_bytes(tb_handle_t handle);
/* Per-launch profiling with HIP events on the handle's stream (off by default;
 * switching it clears earlier records).  tb_profile_read synchronizes and writes,
 * for each category c < ncat, out[3c] = launches, out[3c+1] = summed milliseconds,
 * out[3c+2] = summed work: algorithmic flops, except category 13 (bytes read and
 * written).  Categories: 0..9 grouped GEMM (tier*2 + tile shape, shape 0 = 128x128
 * updates, 1 = 64x256 TRSM leaves), 10 Ozaki slice GEMM, 11 Ozaki slicing,
 * 12 diagonal-block inversion, 13 FP32 -> FP16 plane conversion of operands, 14 one
 * run of a batch plan (tb_batch_run: prologue and persistent kernel), 15 POTRF
 * diagonal-block Cholesky. */
#define TB_PROF_CATEGORIES 16
TB_API int tb_set_profiling(tb_handle_t handle, int on);
TB_API int tb_profile_read(tb_handle_t handle, double* out, int ncat);
/* human-readable text of a status code */
TB_API const char* tb_status_string(int status);
/* library version: major*10000 + minor*100 + patch */
TB_API int tb_version(void);

/* ------------------------------------------------ routines (C ABI, tiered)
 * Pointer types follow the tier: float* for PLAIN, SPLIT and FP32; double* for
 * OZAKI and FP64.  alpha and beta are given in double and used at the tier's
 * precision.  nb is the recursion / diagonal-block size of the reference (its
 * README suggests 256 or 512); it changes the schedule, not the result, except
 * for TRSM where it sets the leaf size. */

/* replaces tc_syrk / tc_cumpsgemm_syrk / tc_ozimmu_syrk (include/TensorBLAS.h:139-143) */
TB_API int tb_syrk(tb_handle_t h, int tier, long n, long k, double alpha, const void* A, long lda,
            double beta, void* C, long ldc, long nb);
/* No counterpart in the reference.  For callers that cannot allocate device memory
 * (code under HIP graph capture, tight serving loops, harnesses that forbid hipMalloc):
 * tb_syrk for PLAIN, SPLIT, FP32 and FP64 that allocates nothing.  The FP16 tiers'
 * operand planes and row exponents go into one fixed device array of
 * TB_NOALLOC_SCRATCH_BYTES (64 MiB) per GPU, held from library load until the process
 * ends, instead of the handle's workspace (which it never touches).  A split-tier
 * problem needs about 4 n round_up(k, 64) + 4 n bytes of it (plain: half of the planes);
 * larger problems return TB_ERR_SCRATCH before launching anything.  OZAKI returns
 * TB_ERR_INVALID.  Every call on one GPU shares that GPU's array and leaves its contents
 * behind, so calls that overlap on one GPU (two streams or threads) race.  No profiling
 * records are taken.  Results equal tb_syrk's bit for bit. */
#define TB_NOALLOC_SCRATCH_BYTES (64L << 20)
TB_API int tb_syrk_noalloc(tb_handle_t h, int tier, long n, long k, double alpha, const void* A,
                           long lda, double beta, void* C, long ldc, long nb);
/* replaces tc_syr2k / tc_cumpsgemm_syr2k / tc_ozimmu_syr2k (include/TensorBLAS.h:159-163) */
TB_API int tb_syr2k(tb_handle_t h, int tier, long n, long k, double alpha, const void* A, long lda,
             const void* B, long ldb, double beta, void* C, long ldc, long nb);
/* replaces tc_trmm / tc_cumpsgemm_trmm / tc_ozimmu_trmm (include/TensorBLAS.h:153-157) */
TB_API int tb_trmm(tb_handle_t h, int tier, long m, long n, double alpha, const void* A, long lda,
            const void* B, long ldb, void* C, long ldc, long nb);
/* replaces tc_trsm / tc_cumpsgemm_trsm / tc_ozimmu_trsm (include/TensorBLAS.h:147-151) */
TB_API int tb_trsm(tb_handle_t h, int tier, long m, long n, const void* A, long lda, void* B, long ldb,
            long nb);
/* TRSM followed by one step of iterative refinement with an FP64 residual
 * R = B - X*L^T computed on FP64 MFMA; the correction is solved in the same tier.
 * FP32 tiers only.  No reference counterpart (added by the FP32-accuracy target). */
TB_API int tb_trsm_refined(tb_handle_t h, int tier, long m, long n, const void* A, long lda, void* B,
                    long ldb, long nb);
/* Cholesky factorization A = L L^T of a symmetric positive definite n x n matrix, lower
 * triangle in place (the upper triangle is neither read nor written).  No counterpart
 * in the reference, which stops at the BLAS-3 routines; this is their canonical
 * caller, composed from them: recursive halving, A21 := A21 L11^-T by tb_trsm and
 * A22 -= A21 A21^T by tb_syrk in the given tier, diagonal blocks of at most 128
 * factored on the device in the matrix's precision (FP32 for plain/split/fp32, FP64 for
 * ozaki/fp64).  *info = 0, or j > 0 when the leading minor of order
 * j is not positive definite (LAPACK dpotrf's info; A then holds a partial factor).
 * Synchronizes the handle's stream to read info.  float* A for plain/split/fp32,
 * double* A for ozaki/fp64; nb is passed to tb_trsm and tb_syrk. */
TB_API int tb_potrf(tb_handle_t h, int tier, long n, void* A, long lda, long nb, long* info);
/* Triangular inverse in place, A := inv(A), for the uplo triangle of A (no counterpart in
 * the reference; LAPACK dtrtri's contract): uplo 'L'/'U', diag 'N' or 'U' (unit diagonal,
 * not referenced).  Composed from this library's TRMM in the chosen tier, with the
 * diagonal blocks inverted in the matrix's precision; the other strict triangle is not
 * changed.  *info = 0,
 * or i when A(i, i) is exactly zero (A is then left unchanged).  n^3/3 flops.
 * Synchronizes the handle's stream (diag 'N': to read the zero test). */
TB_API int tb_trtri(tb_handle_t h, int tier, char uplo, char diag, long n, void* A, long lda,
                    long nb, long* info);
/* The BLAS option sets around the fixed forms above (no counterpart in the reference,
 * whose routines take one option set each).  Options are BLAS characters: order 'C'
 * (column-major) or 'R' (row-major, as CBLAS: every matrix argument and its leading
 * dimension in row-major storage), side 'L'/'R', uplo 'L'/'U', transa 'N'/'T' ('C' = 'T'
 * for real data), diag 'N'/'U'.  Semantics:
 *   tb_trsm_ex:  op(A) X = alpha B (side L) or X op(A) = alpha B (side R); B := X
 *   tb_trmm_ex:  C = alpha op(A) B (side L) or alpha B op(A) (side R), out of place
 *                (as tc_trmm); A triangular with uplo/diag
 *   tb_symm_ex:  C = alpha A B + beta C (side L) or alpha B A + beta C (side R), A
 *                symmetric, only its uplo triangle read
 *   tb_syrk_ex:  C = alpha op(A) op(A)^T + beta C, op(A) = A (n x k) or A^T (A k x n);
 *                only the uplo triangle of C written
 *   tb_syr2k_ex: C = alpha (op(A) op(B)^T + op(B) op(A)^T) + beta C, likewise
 * Each maps onto the native form by transposed / reversed operand copies in the
 * handle's workspace (O(n^2) or O(mn) extra traffic), so results carry the tier's
 * accuracy. */
TB_API int tb_trsm_ex(tb_handle_t h, int tier, char order, char side, char uplo, char transa,
                      char diag, long m, long n, double alpha, const void* A, long lda, void* B,
                      long ldb, long nb);
TB_API int tb_trmm_ex(tb_handle_t h, int tier, char order, char side, char uplo, char transa,
                      char diag, long m, long n, double alpha, const void* A, long lda,
                      const void* B, long ldb, void* C, long ldc, long nb);
TB_API int tb_symm_ex(tb_handle_t h, int tier, char order, char side, char uplo, long m, long n,
                      double alpha, const void* A, long lda, const void* B, long ldb, double beta,
                      void* C, long ldc);
TB_API int tb_syrk_ex(tb_handle_t h, int tier, char order, char uplo, char trans, long n, long k,
                      double alpha, const void* A, long lda, double beta, void* C, long ldc,
                      long nb);
TB_API int tb_syr2k_ex(tb_handle_t h, int tier, char order, char uplo, char trans, long n, long k,
                       double alpha, const void* A, long lda, const void* B, long ldb,
                       double beta, void* C, long ldc, long nb);
/* replaces tc_symm / tc_cumpsgemm_symm (include/TensorBLAS.h:165-167); the OZAKI and
 * FP64 tiers add the FP64 SYMM the reference does not have */
TB_API int tb_symm(tb_handle_t h, int tier, long m, long n, double alpha, const void* A, long lda,
            const void* B, long ldb, double beta, void* C, long ldc);
/* General C := alpha*op(A)*op(B) + beta*C (transa/transb: 0 = N, 1 = T); the update
 * GEMM every routine above is built from (the reference's cublasGemmEx call sites,
 * BLAS/tc_syrk.cu:106-118). */
TB_API int tb_gemm(tb_handle_t h, int tier, int transa, int transb, long m, long n, long k, double alpha,
            const void* A, long lda, const void* B, long ldb, double beta, void* C, long ldc);

/* ------------------------------------------------------------ batched plans
 * Many independent SYRK, TRMM and TRSM problems of mixed sizes, in one FP16 tier (PLAIN
 * or SPLIT), planned once and then run as one HIP graph: the row exponents, the FP16
 * planes and the diagonal-block inverses of every instance (one launch each), then one
 * persistent kernel whose workgroups take work items -- SYRK lower tiles, TRMM tiles,
 * and 64-row TRSM panels that walk their instance's whole recursion -- from a queue
 * sorted by decreasing work.  Each batched instance gives the same bits as its single
 * tb_syrk / tb_trmm / tb_trsm call in the same tier with the same nb.  The plan keeps the
 * pointers it was given; the data behind them may change between runs (TRSM solves B in
 * place on every run).  Tile choices (64 x 256 or 128 x 128 tiles per SYRK / TRMM
 * instance, 4 or 8 waves per workgroup) come from the library's tuning table, measured on
 * MI355X and keyed by (tier, op, size bucket), or are timed on the plan's own instances
 * (TB_BATCH_AUTOTUNE); they never change results.
 * No reference counterpart: the reference runs one problem per call (its batched TRMM
 * levels, BLAS/tc_trmm.cu:4-27, are within one problem). */
typedef struct tb_batch_s* tb_batch_t;
enum { TB_OP_SYRK = 0, TB_OP_TRSM = 1, TB_OP_TRMM = 2 };
enum {
    TB_BATCH_AUTOTUNE = 1,      /* time the tile choices on scratch copies of the outputs
                                   (per op and size bucket) instead of using the table */
    TB_BATCH_NO_GRAPH = 2,      /* launch the kernels directly on each run (no graph) */
    TB_BATCH_TILES_PANEL = 4,   /* every SYRK / TRMM instance on 64 x 256 tiles */
    TB_BATCH_TILES_SQUARE = 8,  /* every SYRK / TRMM instance on 128 x 128 tiles */
    TB_BATCH_WAVES_4 = 16,      /* four waves per workgroup */
    TB_BATCH_WAVES_8 = 32       /* eight waves per workgroup (SPLIT with 2 planes only) */
};   /* the TILES and WAVES flags override the table or the autotuner for that choice */
typedef struct {
    int op;              /* TB_OP_SYRK: C := alpha*A*A^T + beta*C, A n x k, lower C n x n
                            TB_OP_TRSM: X*A^T = B in place, A = L n x n lower, B m x n
                            TB_OP_TRMM: C := alpha*A*B, A = L m x m lower (upper part not
                                        read), B and C m x n (beta unused) */
    long m, n, k;        /* SYRK: n, k (m unused);  TRSM, TRMM: m, n (k unused) */
    double alpha, beta;  /* SYRK; TRMM alpha */
    const float* A;      /* SYRK: A;  TRSM, TRMM: L */
    long lda;
    float* C;            /* SYRK, TRMM: C;  TRSM: B */
    long ldc;
    const float* B;      /* TRMM: B (unused otherwise) */
    long ldb;
} tb_batch_entry;
TB_API int tb_batch_create(tb_handle_t handle, int tier, const tb_batch_entry* entries, int count,
                    long nb, int flags, tb_batch_t* batch);
/* one run of the plan on the handle's stream (asynchronous) */
TB_API int tb_batch_run(tb_batch_t batch);
/* out[0] waves per workgroup, out[1] work items, out[2] instances, out[3] SYRK / TRMM
 * instances on 128 x 128 tiles, out[4] autotuning time in microseconds (0: the table's
 * choices), out[5] device bytes of the plan */
TB_API int tb_batch_info(tb_batch_t batch, long* out, int n);
/* the plan's tile choices, in the tuning table's layout: out[b] the SYRK tile shape of
 * size bucket b (0: 64 x 256, 1: 128 x 128), out[6 + b] the TRMM one, out[12] waves per
 * workgroup; buckets: n < 128, < 256, < 512, < 1024, < 2048, >= 2048 */
TB_API int tb_batch_tiles(tb_batch_t batch, int* out, int n);
TB_API int tb_batch_destroy(tb_batch_t batch);

/* ------------------------------------------------------------- request loop
 * tb_serve runs a stream of requests -- any routine, any tier, each with its own
 * arrival time -- on nslots handles (each with its own stream and workspace) within
 * an HBM budget.  Requests go to one of three lanes by their estimated time alone
 * (a per-routine, per-tier model): those above cfg->large_us form the large lane, of
 * which at most cfg->max_large run at once; those below cfg->small_us the small lane,
 * for which cfg->small_slots slots are kept (the other requests never hold more than
 * nslots - small_slots slots); the rest the middle lane.  The lanes are served small
 * first, then middle, then large, and no request waits behind one of a lane served
 * after its own.  Within a lane a request is admitted first-come first-served as soon as a
 * slot is free and its bytes fit next to those already admitted; admitted requests
 * run concurrently on the device.  Operands are device pointers the caller keeps
 * resident.  Times are device times (HIP events) in microseconds from the start of
 * the stream; the call returns when every request has finished.  A request whose
 * bytes exceed the budget is refused with TB_ERR_NOMEM.  Each request computes what
 * the same call made alone computes (the kernels are deterministic).
 * No reference counterpart. */
enum { TB_REQ_SYRK = 0, TB_REQ_SYR2K = 1, TB_REQ_TRMM = 2, TB_REQ_TRSM = 3, TB_REQ_SYMM = 4 };
typedef struct {
    int op, tier;
    long m, n, k, nb;      /* the routine's dimensions (SYRK/SYR2K: n, k; TRMM/TRSM/SYMM: m, n) */
    double alpha, beta;
    const void* A;         /* TRMM/TRSM: L */
    long lda;
    const void* B;         /* SYR2K, TRMM, SYMM */
    long ldb;
    void* C;               /* TRSM: B, solved in place */
    long ldc;
    double arrival_us;     /* from the start of the stream */
    long bytes;            /* HBM held while admitted; 0: twice the operand bytes */
} tb_request;
typedef struct {
    double admit_us, done_us;
    int slot, status;      /* status of the routine call (TB_ERR_NOMEM: refused) */
    double est_us;         /* the estimate that chose the lane */
} tb_request_result;
typedef struct {
    long hbm_budget;       /* bytes admitted requests may hold together */
    double large_us;       /* estimated time alone above which a request is large (0: one lane) */
    int max_large;         /* large requests running at once (>= 1) */
    int order;             /* within a lane: TB_SERVE_FIFO (oldest waiting request first),
                              TB_SERVE_SHORTEST (shortest estimated time alone first) or
                              TB_SERVE_SHORTEST_SMALL (shortest first in the small and
                              middle lanes, oldest first in the large lane) */
    int reserve_cus;       /* compute units the upper lanes leave free, so that the
                              lower ones always find free CUs; 0: none.  Upper: the large
                              lane, or, with small_slots > 0, every lane but the small one */
    int partition;         /* how they are left free: 0, upper-lane requests' GEMMs run
                              under tb_set_max_cus (the CUs left, shared by as many of them
                              as may run at once); 1, upper-lane requests run on CU-masked
                              streams (every kernel; needs GPU_MAX_HW_QUEUES >= nslots + as
                              many); 2, the other way round: the lower lanes run on
                              CU-masked streams and the reserved CUs stay free for the
                              upper ones (GPU_MAX_HW_QUEUES >= 2 nslots) */
    double small_us;       /* estimated time alone below which a request is small (0: no
                              small lane) */
    int small_slots;       /* slots kept for small requests (0 .. nslots - 1) */
} tb_serve_config;
enum { TB_SERVE_FIFO = 0, TB_SERVE_SHORTEST = 1, TB_SERVE_SHORTEST_SMALL = 2 };
TB_API int tb_serve(tb_handle_t* slots, int nslots, const tb_serve_config* cfg, const tb_request* reqs,
             int count, tb_request_result* results);
/* Grow a handle's two device workspaces (routines; Ozaki products) to at least these
 * sizes, so that later calls never reallocate (which synchronizes the handle's stream);
 * tb_workspace_sizes reads the current sizes into out[0], out[1]. */
TB_API int tb_reserve_workspace(tb_handle_t handle, long bytes, long ozaki_bytes);
TB_API int tb_workspace_sizes(tb_handle_t handle, long* out);

/* ---------------------------------- conversion kernels (util/util.cu:127-167) */
/* ah = fp16(as), m x n (RNE); the reference's s2h / s2h_swpipe */
TB_API int tb_s2h(tb_handle_t h, long m, long n, const float* as, long ldas, void* ah, long ldah);
/* row-major m x n FP32 -> column-major FP16; the reference's s2hTranspose */
TB_API int tb_s2h_transpose(tb_handle_t h, long m, long n, const float* as, void* ah);
/* A(i,j) := A(j,i), i < j; the reference's copy_lower_to_upper{,Double} */
TB_API int tb_copy_lower_to_upper(tb_handle_t h, int f64, long n, void* A, long lda);

/* ------------------------------------------------ schedule inspection (host only)
 * Chunks of n in processing order as util/mat_size.cu:5-55 computes them
 * (find_mat_size_syrk when trsm_planner == 0, find_mat_size_trsm otherwise),
 * continued with 16384s past the reference's table.  Returns the count, or
 * -TB_ERR_INVALID when cap is too small.  Launches nothing. */
TB_API int tb_plan_chunks(int trsm_planner, long n, long* out, int cap);
/* The TRSM schedule of BLAS/tc_trsm.cu:31-137 for an n-column problem with block nb:
 * 5 longs per step, in execution order: (0, off, s, 0, 0) a diagonal leaf of size s,
 * (1, c0, w, k0, kw) an update B(:,c0:c0+w) -= B(:,k0:k0+kw)*L(c0:c0+w,k0:k0+kw)^T.
 * Returns the step count or -TB_ERR_INVALID.  Launches nothing. */
TB_API int tb_plan_trsm(long n, long nb, long* out, int cap);

/* ---------------------------------------- the reference's own C-ABI exports */
/* include/TensorBLAS.h:135 -- prints the device properties */
TB_API void print_env(void);
/* include/TensorBLAS.h:137, BLAS/tc_syrk.cu:4-39: A is n x k ROW-major (a torch
 * tensor), C (n x n) receives the full symmetric A*A^T (plain tier).  Returns 0 on
 * success, 1 on error; synchronous. */
TB_API int tc_syrk_wrapper(long n, long k, float* A, float* C, long nb);
/* include/TensorBLAS.h:145, BLAS/tc_trsm.cu:4-28: solves X*A^T = B in place for
 * column-major A (n x n, lda = n) and B (m x n, ldb = m), plain tier.  Returns 0/1;
 * synchronous. */
TB_API int tc_trsm_wrapper(long m, long n, float* A, float* B, long nb);

#ifdef __cplusplus
} /* extern "C" */

/* ----------------------------------------- the reference's C++ entry points
 * include/TensorBLAS.h:139-167 with tb_handle_t in place of cublasHandle_t and
 * cumpsgemm::handle_t.  __half* workspaces are accepted (NULL allowed) and unused. */
struct __half;

TB_API int tc_syrk(tb_handle_t handle, long n, long k, float alpha, float* A, long lda, float beta,
            float* C, long ldc, __half* Ah, long nb);
TB_API int tc_cumpsgemm_syrk(tb_handle_t handle, long n, long k, float alpha, float* A, long lda,
                      float beta, float* C, long ldc, long nb);
TB_API int tc_ozimmu_syrk(tb_handle_t handle, long n, long k, double alpha, double* A, long lda,
                   double beta, double* C, long ldc, long nb);
TB_API int tc_trsm(tb_handle_t handle, long m, long n, float* A, long lda, float* B, long ldb,
            __half* hwork, long nb);
/* the reference takes both a cuBLAS and a cuMpSGEMM handle here */
TB_API int tc_cumpsgemm_trsm(tb_handle_t handle, tb_handle_t handle2, long m, long n, float* A, long lda,
                      float* B, long ldb, long nb);
TB_API int tc_ozimmu_trsm(tb_handle_t handle, long m, long n, double* A, long lda, double* B, long ldb,
                   long nb);
TB_API int tc_trmm(tb_handle_t handle, long m, long n, float alpha, float* A, long lda, float* B,
            long ldb, float* C, long ldc, __half* hwork, long nb);
TB_API int tc_cumpsgemm_trmm(tb_handle_t handle, long m, long n, float alpha, float* A, long lda,
                      float* B, long ldb, float* C, long ldc, long nb);
TB_API int tc_ozimmu_trmm(tb_handle_t handle, long m, long n, double alpha, double* A, long lda,
                   double* B, long ldb, double* C, long ldc, long nb);
TB_API int tc_syr2k(tb_handle_t handle, long n, long k, float alpha, float* A, long lda, float* B,
             long ldb, float beta, float* C, long ldc, __half* Ah, long nb);
TB_API int tc_cumpsgemm_syr2k(tb_handle_t handle, long n, long k, float alpha, float* A, long lda,
                       float* B, long ldb, float beta, float* C, long ldc, long nb);
TB_API int tc_ozimmu_syr2k(tb_handle_t handle, long n, long k, double alpha, double* A, long lda,
                    double* B, long ldb, double beta, double* C, long ldc, long nb);
TB_API int tc_symm(tb_handle_t handle, long m, long n, float alpha, float* A, long lda, float* B,
            int ldb, float beta, float* C, long ldc, __half* Ah);
TB_API int tc_cumpsgemm_symm(tb_handle_t handle, long m, long n, float alpha, float* A, long lda,
                      float* B, int ldb, float beta, float* C, long ldc);
/* no reference counterpart: the FP64 SYMM named by the Ozaki benchmark */
TB_API int tc_ozimmu_symm(tb_handle_t handle, long m, long n, double alpha, double* A, long lda,
                   double* B, long ldb, double beta, double* C, long ldc);
#endif

#endif /* TENSORBLAS_MI355X_H */
