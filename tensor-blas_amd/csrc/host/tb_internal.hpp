// Host-side internals: the handle, its device workspace, the GEMM problem queue and
// the routine planners (tb_blas3.cpp).
#pragma once
#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <vector>

#include "../../../include/TensorBLAS.h"
#include "../tb_kernels.hpp"

// One profiled launch: events recorded on the handle's stream around it.
struct tb_prof_rec {
    int cat;          // tb::prof_category
    hipEvent_t e0, e1;
    double flops;     // algorithmic flops of the launch
};

struct tb_handle_s {
    hipStream_t stream = nullptr;
    bool prof = false;
    std::vector<tb_prof_rec> prof_recs;
    int device = 0;
    int ozaki_slices = 7;
    int ozaki_extra = 0;    // 1: also the slice-pair group g = nslices (6 <= nslices <= 9)
    int split_planes = 2;   // FP16 planes per operand of the split tier (2 or 3)
    int max_cus = 0;        // > 0: GEMM launches occupy at most this many CUs
    // device workspaces, grown on demand, freed by tb_destroy: slot 0 for routines,
    // slot 1 for the Ozaki GEMM (which routines call with their own slot live), slot 2
    // for the BLAS-option layer's transformed operands (live across a routine call),
    // slot 3 for a factorization's own copies (live across option-layer calls)
    static constexpr int kWsSlots = 4;
    void* ws[kWsSlots] = {};
    size_t ws_bytes[kWsSlots] = {};
    int* dev_info = nullptr;   // POTRF / TRTRI status word (device), allocated on first use
    void* scratch = nullptr;   // the fixed no-allocation scratch of this handle's device
};

namespace tb {

// Profiling categories: grouped GEMM (tier*2 + shape), then the other kernels.
enum ProfCat : int {
    kProfGemm0 = 0,            // .. kProfGemm0 + 9: tier * 2 + shape
    kProfOzakiGemm = 10,
    kProfOzakiSlice = 11,
    kProfTrtri = 12,
    kProfPlanes = 13,          // FP32 -> FP16 plane conversion (work = bytes moved)
    kProfBatch = 14,           // one run of a batch plan (work = its algorithmic flops)
    kProfPotrfLeaf = 15,       // diagonal-block Cholesky of POTRF (work = s^3/3 flops)
    kProfCats = 16,
};
// Wraps one launch in events when the handle profiles.
int prof_begin(tb_handle_s* h, hipEvent_t* e0);
int prof_end(tb_handle_s* h, int cat, hipEvent_t e0, double flops);

// status helpers
inline int hip_status(hipError_t e) { return e == hipSuccess ? TB_SUCCESS : TB_ERR_DEVICE; }
#define TB_TRY(expr)                          \
    do {                                      \
        int _st = (expr);                     \
        if (_st != TB_SUCCESS) return _st;    \
    } while (0)
#define TB_HIP(expr) TB_TRY(::tb::hip_status(expr))

// Workspace carving: sizes are reserved first, then the handle's buffer is grown
// once (synchronizing its stream before freeing a smaller one).
class Arena {
public:
    size_t reserve(size_t bytes)
    {
        size_t off = (total_ + 255) & ~size_t(255);
        total_ = off + bytes;
        return off;
    }
    size_t total() const { return total_; }

private:
    size_t total_ = 0;
};
int workspace_get(tb_handle_s* h, size_t bytes, char** base, int slot = 0);

// Collects GEMM problems of one tier, tile shape and operand source kinds (all
// problems must match them); launches up to kMaxProbs problems per kernel launch,
// in order.
// FP16 planes an operand of `tier` is converted to on handle h (h may be null: default)
inline int planes_of(const tb_handle_s* h, int tier)
{
    return tier_planes(tier, h ? h->split_planes : 2);
}

class GemmQueue {
public:
    GemmQueue(tb_handle_s* h, int tier, int shape, int srcA = 0, int srcB = 0);
    int add(const GemmProb& p);
    int flush();
    static constexpr long long kWideMinTiles = 256;   // one wide tile per CU

private:
    bool retile_wide();
    tb_handle_s* h_;
    int tier_, shape_, srcA_, srcB_;
    TileShape ts_;
    GemmLaunch L_;
};

GemmProb make_prob(long m, long n, long k, const void* A, long lda, int modeA, const void* B,
                   long ldb, int modeB, void* C, long ldc, double alpha, double beta, int modeC);

// Chunk planners: the reference's util/mat_size.cu:5-55, continued with 16384s past
// the end of its power table (where the reference reads out of bounds).
std::vector<long> plan_chunks_syrk(long n);
std::vector<long> plan_chunks_trsm(long n);

// routines (argument validation included); trans_a: op(A) = A^T with A stored k x n
int syrk(tb_handle_s* h, int tier, long n, long k, double alpha, const void* A, long lda,
         bool trans_a, double beta, void* C, long ldc, long nb);
int syr2k(tb_handle_s* h, int tier, long n, long k, double alpha, const void* A, long lda,
          const void* B, long ldb, double beta, void* C, long ldc, long nb);
int trmm(tb_handle_s* h, int tier, long m, long n, double alpha, const void* A, long lda,
         const void* B, long ldb, void* C, long ldc, long nb);
int potrf(tb_handle_s* h, int tier, long n, void* A, long lda, long nb, long* info);
int trtri(tb_handle_s* h, int tier, char uplo, char diag, long n, void* A, long lda, long nb,
          long* info);
// BLAS option sets (side 'L'/'R', uplo 'L'/'U', trans 'N'/'T'/'C', diag 'N'/'U'),
// mapped onto the native forms above by transposed and reversed operand copies
int trsm_ex(tb_handle_s* h, int tier, char order, char side, char uplo, char trans, char diag, long m, long n,
            double alpha, const void* A, long lda, void* B, long ldb, long nb);
int trmm_ex(tb_handle_s* h, int tier, char order, char side, char uplo, char trans, char diag, long m, long n,
            double alpha, const void* A, long lda, const void* B, long ldb, void* C, long ldc, long nb);
int symm_ex(tb_handle_s* h, int tier, char order, char side, char uplo, long m, long n, double alpha, const void* A,
            long lda, const void* B, long ldb, double beta, void* C, long ldc);
int syrk_ex(tb_handle_s* h, int tier, char order, char uplo, char trans, long n, long k, double alpha, const void* A,
            long lda, double beta, void* C, long ldc, long nb);
int syr2k_ex(tb_handle_s* h, int tier, char order, char uplo, char trans, long n, long k, double alpha, const void* A,
             long lda, const void* B, long ldb, double beta, void* C, long ldc, long nb);
int trsm(tb_handle_s* h, int tier, long m, long n, const void* A, long lda, void* B, long ldb,
         long nb);
// syrk (trans_a = false) with its operand planes and exponents in the handle's fixed
// no-allocation scratch instead of the handle's workspace: no device allocation
int syrk_noalloc(tb_handle_s* h, int tier, long n, long k, double alpha, const void* A,
                 long lda, double beta, void* C, long ldc, long nb);
int trsm_refined(tb_handle_s* h, int tier, long m, long n, const void* A, long lda, void* B,
                 long ldb, long nb);
// schedule of TRSM (leaves and updates in execution order), for inspection
int plan_trsm(long n, long nb, long* out, int cap);
// The TRSM schedule of an n-column solve with block nb as the single call runs it
// (BLAS/tc_trsm.cu:31-137): steps in execution order (type 0 leaf: a = off, b = s;
// type 1 update: a = c0, b = w, c = k0, d = kw; acc: an update inside a leaf block
// wider than the leaf kernel, run in the leaf tier) and the leaves in order.
struct TrsmStep {
    long type, a, b, c, d;
    bool acc;
};
struct TrsmLeaf {
    long off, s;
};
void trsm_schedule(long n, long nb, std::vector<TrsmStep>* steps, std::vector<TrsmLeaf>* leaves);
int symm(tb_handle_s* h, int tier, long m, long n, double alpha, const void* A, long lda,
         const void* B, long ldb, double beta, void* C, long ldc);
int gemm(tb_handle_s* h, int tier, int transa, int transb, long m, long n, long k, double alpha,
         const void* A, long lda, const void* B, long ldb, double beta, void* C, long ldc);

// Ozaki tier (tb_ozaki.cpp): C = alpha*op(A)*op(B) (+ op(A2)*op(B2)) + beta*C in FP64.
// The second pair is SYR2K's swap: A2 is read with ldb and B2 with lda.
int ozaki_gemm_host(tb_handle_s* h, long m, long n, long k, double alpha, const double* A,
                    long lda, int modeA, const double* B, long ldb, int modeB, const double* A2,
                    const double* B2, double beta, double* C, long ldc, bool lowerC);

}  // namespace tb
