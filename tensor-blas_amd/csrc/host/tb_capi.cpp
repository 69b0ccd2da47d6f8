// The C ABI and the reference's C++ entry points (include/TensorBLAS.h:135-167).
#include <hip/hip_runtime.h>

#include <cstdio>
#include <mutex>

#include "tb_internal.hpp"

using namespace tb;

extern "C" {

int tb_create(tb_handle_t* handle)
{
    if (!handle) return TB_ERR_INVALID;
    tb_handle_s* h = new (std::nothrow) tb_handle_s();
    if (!h) return TB_ERR_NOMEM;
    if (hipGetDevice(&h->device) != hipSuccess) {
        delete h;
        return TB_ERR_DEVICE;
    }
    // set-up of the no-allocation scratch: its device address on this (current) device
    if (noalloc_scratch(h->device, &h->scratch) != hipSuccess) {
        (void)hipGetLastError();
        delete h;
        return TB_ERR_DEVICE;
    }
    *handle = h;
    return TB_SUCCESS;
}

static void prof_clear(tb_handle_t h)
{
    for (tb_prof_rec& r : h->prof_recs) {
        (void)hipEventDestroy(r.e0);
        (void)hipEventDestroy(r.e1);
    }
    h->prof_recs.clear();
}

int tb_set_profiling(tb_handle_t h, int on)
{
    if (!h) return TB_ERR_INVALID;
    int st = hip_status(hipStreamSynchronize(h->stream));
    prof_clear(h);
    h->prof = on != 0;
    return st;
}

int tb_profile_read(tb_handle_t h, double* out, int ncat)
{
    if (!h || ncat < 0 || (!out && ncat > 0)) return TB_ERR_INVALID;
    TB_HIP(hipStreamSynchronize(h->stream));
    for (int c = 0; c < 3 * ncat; ++c) out[c] = 0.0;
    for (const tb_prof_rec& r : h->prof_recs) {
        if (r.cat >= ncat) continue;
        float ms = 0.0f;
        TB_HIP(hipEventElapsedTime(&ms, r.e0, r.e1));
        out[3 * r.cat] += 1.0;
        out[3 * r.cat + 1] += ms;
        out[3 * r.cat + 2] += r.flops;
    }
    return TB_SUCCESS;
}

int tb_destroy(tb_handle_t h)
{
    if (!h) return TB_ERR_INVALID;
    int st = hip_status(hipStreamSynchronize(h->stream));
    prof_clear(h);
    for (int s = 0; s < tb_handle_s::kWsSlots; ++s)
        if (h->ws[s]) (void)hipFree(h->ws[s]);
    if (h->dev_info) (void)hipFree(h->dev_info);
    delete h;
    return st;
}

int tb_set_stream(tb_handle_t h, void* stream)
{
    if (!h) return TB_ERR_INVALID;
    h->stream = static_cast<hipStream_t>(stream);
    return TB_SUCCESS;
}

int tb_synchronize(tb_handle_t h)
{
    if (!h) return TB_ERR_INVALID;
    return hip_status(hipStreamSynchronize(h->stream));
}

int tb_set_ozaki_slices(tb_handle_t h, int nslices)
{
    if (!h || nslices < 3 || nslices > 10) return TB_ERR_INVALID;
    h->ozaki_slices = nslices;
    return TB_SUCCESS;
}

int tb_set_ozaki_extra(tb_handle_t h, int extra)
{
    if (!h || extra < 0 || extra > 1) return TB_ERR_INVALID;
    h->ozaki_extra = extra;
    return TB_SUCCESS;
}

int tb_set_max_cus(tb_handle_t h, int max_cus)
{
    if (!h || max_cus < 0) return TB_ERR_INVALID;
    h->max_cus = max_cus;
    return TB_SUCCESS;
}

int tb_set_split_planes(tb_handle_t h, int nplanes)
{
    if (!h || nplanes < 2 || nplanes > 3) return TB_ERR_INVALID;
    h->split_planes = nplanes;
    return TB_SUCCESS;
}

long tb_workspace_bytes(tb_handle_t h)
{
    if (!h) return 0;
    size_t b = 0;
    for (int s = 0; s < tb_handle_s::kWsSlots; ++s) b += h->ws_bytes[s];
    return (long)b;
}

const char* tb_status_string(int status)
{
    switch (status) {
    case TB_SUCCESS: return "success";
    case TB_ERR_INVALID: return "invalid argument";
    case TB_ERR_DEVICE: return "device error";
    case TB_ERR_NOMEM: return "out of device memory";
    case TB_ERR_SCRATCH: return "problem exceeds the no-allocation scratch";
    default: return "unknown status";
    }
}

int tb_version(void) { return 100; }

int tb_syrk(tb_handle_t h, int tier, long n, long k, double alpha, const void* A, long lda,
            double beta, void* C, long ldc, long nb)
{
    return syrk(h, tier, n, k, alpha, A, lda, false, beta, C, ldc, nb);
}

int tb_syrk_noalloc(tb_handle_t h, int tier, long n, long k, double alpha, const void* A, long lda,
                    double beta, void* C, long ldc, long nb)
{
    return syrk_noalloc(h, tier, n, k, alpha, A, lda, beta, C, ldc, nb);
}

int tb_syr2k(tb_handle_t h, int tier, long n, long k, double alpha, const void* A, long lda,
             const void* B, long ldb, double beta, void* C, long ldc, long nb)
{
    return syr2k(h, tier, n, k, alpha, A, lda, B, ldb, beta, C, ldc, nb);
}

int tb_trmm(tb_handle_t h, int tier, long m, long n, double alpha, const void* A, long lda,
            const void* B, long ldb, void* C, long ldc, long nb)
{
    return trmm(h, tier, m, n, alpha, A, lda, B, ldb, C, ldc, nb);
}

int tb_trsm(tb_handle_t h, int tier, long m, long n, const void* A, long lda, void* B, long ldb,
            long nb)
{
    return trsm(h, tier, m, n, A, lda, B, ldb, nb);
}

int tb_potrf(tb_handle_t h, int tier, long n, void* A, long lda, long nb, long* info)
{
    return potrf(h, tier, n, A, lda, nb, info);
}

int tb_trtri(tb_handle_t h, int tier, char uplo, char diag, long n, void* A, long lda, long nb,
             long* info)
{
    return trtri(h, tier, uplo, diag, n, A, lda, nb, info);
}

int tb_trsm_ex(tb_handle_t h, int tier, char order, char side, char uplo, char transa, char diag, long m, long n,
               double alpha, const void* A, long lda, void* B, long ldb, long nb)
{
    return trsm_ex(h, tier, order, side, uplo, transa, diag, m, n, alpha, A, lda, B, ldb, nb);
}

int tb_trmm_ex(tb_handle_t h, int tier, char order, char side, char uplo, char transa, char diag, long m, long n,
               double alpha, const void* A, long lda, const void* B, long ldb, void* C, long ldc, long nb)
{
    return trmm_ex(h, tier, order, side, uplo, transa, diag, m, n, alpha, A, lda, B, ldb, C, ldc, nb);
}

int tb_symm_ex(tb_handle_t h, int tier, char order, char side, char uplo, long m, long n, double alpha, const void* A,
               long lda, const void* B, long ldb, double beta, void* C, long ldc)
{
    return symm_ex(h, tier, order, side, uplo, m, n, alpha, A, lda, B, ldb, beta, C, ldc);
}

int tb_syrk_ex(tb_handle_t h, int tier, char order, char uplo, char trans, long n, long k, double alpha, const void* A,
               long lda, double beta, void* C, long ldc, long nb)
{
    return syrk_ex(h, tier, order, uplo, trans, n, k, alpha, A, lda, beta, C, ldc, nb);
}

int tb_syr2k_ex(tb_handle_t h, int tier, char order, char uplo, char trans, long n, long k, double alpha, const void* A,
                long lda, const void* B, long ldb, double beta, void* C, long ldc, long nb)
{
    return syr2k_ex(h, tier, order, uplo, trans, n, k, alpha, A, lda, B, ldb, beta, C, ldc, nb);
}

int tb_trsm_refined(tb_handle_t h, int tier, long m, long n, const void* A, long lda, void* B,
                    long ldb, long nb)
{
    return trsm_refined(h, tier, m, n, A, lda, B, ldb, nb);
}

int tb_symm(tb_handle_t h, int tier, long m, long n, double alpha, const void* A, long lda,
            const void* B, long ldb, double beta, void* C, long ldc)
{
    return symm(h, tier, m, n, alpha, A, lda, B, ldb, beta, C, ldc);
}

int tb_gemm(tb_handle_t h, int tier, int transa, int transb, long m, long n, long k, double alpha,
            const void* A, long lda, const void* B, long ldb, double beta, void* C, long ldc)
{
    if (transa < 0 || transa > 1 || transb < 0 || transb > 1) return TB_ERR_INVALID;
    return gemm(h, tier, transa, transb, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc);
}

int tb_s2h(tb_handle_t h, long m, long n, const float* as, long ldas, void* ah, long ldah)
{
    if (!h || m < 0 || n < 0 || ldas < (m > 1 ? m : 1) || ldah < (m > 1 ? m : 1))
        return TB_ERR_INVALID;
    if (m == 0 || n == 0) return TB_SUCCESS;
    if (!as || !ah) return TB_ERR_INVALID;
    return hip_status(s2h(m, n, as, ldas, ah, ldah, h->stream));
}

int tb_s2h_transpose(tb_handle_t h, long m, long n, const float* as, void* ah)
{
    if (!h || m < 0 || n < 0) return TB_ERR_INVALID;
    if (m == 0 || n == 0) return TB_SUCCESS;
    if (!as || !ah) return TB_ERR_INVALID;
    return hip_status(s2h_transpose(m, n, as, ah, h->stream));
}

int tb_copy_lower_to_upper(tb_handle_t h, int f64, long n, void* A, long lda)
{
    if (!h || n < 0 || lda < (n > 1 ? n : 1)) return TB_ERR_INVALID;
    if (n == 0) return TB_SUCCESS;
    if (!A) return TB_ERR_INVALID;
    return hip_status(mirror_lower(f64 != 0, n, A, lda, h->stream));
}

int tb_plan_chunks(int trsm_planner, long n, long* out, int cap)
{
    if (n < 0 || cap < 0 || (!out && cap > 0)) return -TB_ERR_INVALID;
    std::vector<long> ch = trsm_planner ? plan_chunks_trsm(n) : plan_chunks_syrk(n);
    if ((long)ch.size() > cap) return -TB_ERR_INVALID;
    for (size_t q = 0; q < ch.size(); ++q) out[q] = ch[q];
    return (int)ch.size();
}

int tb_plan_trsm(long n, long nb, long* out, int cap) { return plan_trsm(n, nb, out, cap); }

// ------------------------------------------------------ reference C-ABI exports
void print_env(void)
{
    int dev = 0;
    hipDeviceProp_t p;
    if (hipGetDevice(&dev) != hipSuccess || hipGetDeviceProperties(&p, dev) != hipSuccess) {
        std::printf("no HIP device\n");
        return;
    }
    std::printf("Device %d: %s (%s)\n", dev, p.name, p.gcnArchName);
    std::printf("  compute units: %d, wavefront: %d, clock: %d kHz\n", p.multiProcessorCount,
                p.warpSize, p.clockRate);
    std::printf("  global memory: %.1f GiB, LDS per workgroup: %zu KiB, L2: %d KiB\n",
                p.totalGlobalMem / 1073741824.0, p.sharedMemPerBlock / 1024, p.l2CacheSize / 1024);
    std::printf("  memory clock: %d kHz, bus width: %d bits\n", p.memoryClockRate,
                p.memoryBusWidth);
}

// The reference created (and leaked) a cuBLAS handle and a workspace per call
// (BLAS/tc_syrk.cu:7-11).  Here the wrappers share one handle, serialized.
static std::mutex g_wrap_mu;
static tb_handle_t g_wrap_handle = nullptr;

static tb_handle_t wrap_handle()
{
    if (!g_wrap_handle && tb_create(&g_wrap_handle) != TB_SUCCESS) g_wrap_handle = nullptr;
    return g_wrap_handle;
}

int tc_syrk_wrapper(long n, long k, float* A, float* C, long nb)
{
    std::lock_guard<std::mutex> lock(g_wrap_mu);
    tb_handle_t h = wrap_handle();
    if (!h) return 1;
    // A is n x k row-major == the k x n column-major matrix X, op(A) = X^T
    // (the reference transposes it into its FP16 workspace, BLAS/tc_syrk.cu:83-85)
    int st = syrk(h, kPlainF16, n, k, 1.0, A, k > 1 ? k : 1, true, 0.0, C, n > 1 ? n : 1, nb);
    if (st == TB_SUCCESS && n > 0) st = hip_status(mirror_lower(false, n, C, n, h->stream));
    int sy = tb_synchronize(h);
    return (st == TB_SUCCESS && sy == TB_SUCCESS) ? 0 : 1;
}

int tc_trsm_wrapper(long m, long n, float* A, float* B, long nb)
{
    std::lock_guard<std::mutex> lock(g_wrap_mu);
    tb_handle_t h = wrap_handle();
    if (!h) return 1;
    int st = trsm(h, kPlainF16, m, n, A, n > 1 ? n : 1, B, m > 1 ? m : 1, nb);
    int sy = tb_synchronize(h);
    return (st == TB_SUCCESS && sy == TB_SUCCESS) ? 0 : 1;
}

}  // extern "C"

// ------------------------------------------------ reference C++ entry points
int tc_syrk(tb_handle_t handle, long n, long k, float alpha, float* A, long lda, float beta,
            float* C, long ldc, __half*, long nb)
{
    return syrk(handle, kPlainF16, n, k, alpha, A, lda, false, beta, C, ldc, nb);
}
int tc_cumpsgemm_syrk(tb_handle_t handle, long n, long k, float alpha, float* A, long lda,
                      float beta, float* C, long ldc, long nb)
{
    return syrk(handle, kSplitF16, n, k, alpha, A, lda, false, beta, C, ldc, nb);
}
int tc_ozimmu_syrk(tb_handle_t handle, long n, long k, double alpha, double* A, long lda,
                   double beta, double* C, long ldc, long nb)
{
    return syrk(handle, kOzakiI8, n, k, alpha, A, lda, false, beta, C, ldc, nb);
}
int tc_trsm(tb_handle_t handle, long m, long n, float* A, long lda, float* B, long ldb, __half*,
            long nb)
{
    return trsm(handle, kPlainF16, m, n, A, lda, B, ldb, nb);
}
int tc_cumpsgemm_trsm(tb_handle_t handle, tb_handle_t, long m, long n, float* A, long lda,
                      float* B, long ldb, long nb)
{
    return trsm(handle, kSplitF16, m, n, A, lda, B, ldb, nb);
}
int tc_ozimmu_trsm(tb_handle_t handle, long m, long n, double* A, long lda, double* B, long ldb,
                   long nb)
{
    return trsm(handle, kOzakiI8, m, n, A, lda, B, ldb, nb);
}
int tc_trmm(tb_handle_t handle, long m, long n, float alpha, float* A, long lda, float* B,
            long ldb, float* C, long ldc, __half*, long nb)
{
    return trmm(handle, kPlainF16, m, n, alpha, A, lda, B, ldb, C, ldc, nb);
}
int tc_cumpsgemm_trmm(tb_handle_t handle, long m, long n, float alpha, float* A, long lda,
                      float* B, long ldb, float* C, long ldc, long nb)
{
    return trmm(handle, kSplitF16, m, n, alpha, A, lda, B, ldb, C, ldc, nb);
}
int tc_ozimmu_trmm(tb_handle_t handle, long m, long n, double alpha, double* A, long lda,
                   double* B, long ldb, double* C, long ldc, long nb)
{
    return trmm(handle, kOzakiI8, m, n, alpha, A, lda, B, ldb, C, ldc, nb);
}
int tc_syr2k(tb_handle_t handle, long n, long k, float alpha, float* A, long lda, float* B,
             long ldb, float beta, float* C, long ldc, __half*, long nb)
{
    return syr2k(handle, kPlainF16, n, k, alpha, A, lda, B, ldb, beta, C, ldc, nb);
}
int tc_cumpsgemm_syr2k(tb_handle_t handle, long n, long k, float alpha, float* A, long lda,
                       float* B, long ldb, float beta, float* C, long ldc, long nb)
{
    return syr2k(handle, kSplitF16, n, k, alpha, A, lda, B, ldb, beta, C, ldc, nb);
}
int tc_ozimmu_syr2k(tb_handle_t handle, long n, long k, double alpha, double* A, long lda,
                    double* B, long ldb, double beta, double* C, long ldc, long nb)
{
    return syr2k(handle, kOzakiI8, n, k, alpha, A, lda, B, ldb, beta, C, ldc, nb);
}
int tc_symm(tb_handle_t handle, long m, long n, float alpha, float* A, long lda, float* B, int ldb,
            float beta, float* C, long ldc, __half*)
{
    return symm(handle, kPlainF16, m, n, alpha, A, lda, B, ldb, beta, C, ldc);
}
int tc_cumpsgemm_symm(tb_handle_t handle, long m, long n, float alpha, float* A, long lda,
                      float* B, int ldb, float beta, float* C, long ldc)
{
    return symm(handle, kSplitF16, m, n, alpha, A, lda, B, ldb, beta, C, ldc);
}
int tc_ozimmu_symm(tb_handle_t handle, long m, long n, double alpha, double* A, long lda,
                   double* B, long ldb, double beta, double* C, long ldc)
{
    return symm(handle, kOzakiI8, m, n, alpha, A, lda, B, ldb, beta, C, ldc);
}
