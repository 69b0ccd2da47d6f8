// A C++ caller of the reference's 15 C++ entry points (include/TensorBLAS.h:139-167 of
// the reference, plus tc_ozimmu_symm), written as the reference's testing/ drivers call
// them (e.g. testing/testing_syrk.cu:97).  Each tc_* call is compared bit for bit with
// the tb_* C-ABI call of the same routine and tier on the same inputs, so that an
// argument-order slip in a forwarder (alpha/beta, lda/ldb, m/n) shows up.
//
// The problem sizes, leading dimensions and scalars are all distinct, and every buffer
// carries padding rows, so that exchanging any two of them changes the result.
//
// Built into tests/cpp/build/libtc_boundary.so (tests/cpp/Makefile); the GPU test
// tests/test_gpu_boundary.py loads it with ctypes and calls tc_boundary_check.
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/TensorBLAS.h"

namespace {

// deterministic values in [-1, 1)
struct Lcg {
    unsigned long long s;
    double next()
    {
        s = s * 6364136223846793005ull + 1442695040888963407ull;
        return (double)(s >> 11) * (1.0 / 9007199254740992.0) * 2.0 - 1.0;
    }
};

template <class T>
struct Buf {
    std::vector<T> host;
    T* dev = nullptr;
    explicit Buf(size_t n, unsigned long long seed, double scale = 1.0) : host(n)
    {
        Lcg g{seed};
        for (auto& x : host) x = (T)(g.next() * scale);
    }
    bool upload()
    {
        if (hipMalloc(&dev, host.size() * sizeof(T)) != hipSuccess) return false;
        return hipMemcpy(dev, host.data(), host.size() * sizeof(T), hipMemcpyHostToDevice) == hipSuccess;
    }
    bool reset() { return hipMemcpy(dev, host.data(), host.size() * sizeof(T), hipMemcpyHostToDevice) == hipSuccess; }
    std::vector<T> download() const
    {
        std::vector<T> out(host.size());
        if (hipMemcpy(out.data(), dev, out.size() * sizeof(T), hipMemcpyDeviceToHost) != hipSuccess)
            out.clear();
        return out;
    }
    ~Buf()
    {
        if (dev) (void)hipFree(dev);
    }
};

// diagonally dominant lower triangle (diagonal n + [0, 1)) in the n x n block of A
template <class T>
void make_dominant(Buf<T>& A, long n, long lda)
{
    for (long j = 0; j < n; ++j) A.host[j + j * lda] = (T)(n + 0.5 + 0.25 * (double)A.host[j + j * lda]);
}

struct Checker {
    std::string report;
    int failures = 0, checked = 0;
    void note(const char* name, int st_ref, int st_tb, bool same)
    {
        ++checked;
        char line[160];
        const bool ok = st_ref == 0 && st_tb == 0 && same;
        std::snprintf(line, sizeof(line), "%s: tc status %d, tb status %d, %s\n", name, st_ref, st_tb,
                      same ? "bitwise equal" : "DIFFERENT");
        report += line;
        if (!ok) ++failures;
    }
};

template <class T>
bool equal(const std::vector<T>& a, const std::vector<T>& b)
{
    return !a.empty() && a.size() == b.size() && std::memcmp(a.data(), b.data(), a.size() * sizeof(T)) == 0;
}

// run `ref` and `tb` on the same initial contents of `out` and compare the results
template <class T, class F, class G>
void compare(Checker& ck, tb_handle_t h, const char* name, Buf<T>& out, F&& ref, G&& tb)
{
    if (!out.reset()) { ck.note(name, -1, -1, false); return; }
    const int s1 = ref();
    (void)tb_synchronize(h);
    const std::vector<T> r1 = out.download();
    if (!out.reset()) { ck.note(name, -1, -1, false); return; }
    const int s2 = tb();
    (void)tb_synchronize(h);
    const std::vector<T> r2 = out.download();
    ck.note(name, s1, s2, equal(r1, r2));
}

template <class T>
void run_family(Checker& ck, tb_handle_t h, tb_handle_t h2, int tier, unsigned long long seed);

// one routine family per precision: plain and split take float*, Ozaki double*
template <>
void run_family<float>(Checker& ck, tb_handle_t h, tb_handle_t h2, int tier, unsigned long long seed)
{
    (void)h2;
    const bool split = tier == TB_TIER_SPLIT;
    const long n = 67, k = 45, nb = 256;
    const long lda = n + 3, ldb = n + 5, ldc = n + 7;
    const float alpha = 1.5f, beta = -0.75f;
    Buf<float> A(lda * k, seed + 1), B(ldb * k, seed + 2), C(ldc * n, seed + 3);
    if (!A.upload() || !B.upload() || !C.upload()) { ck.note("upload", -1, -1, false); return; }

    // SYRK (include/TensorBLAS.h:139-143)
    compare(ck, h, split ? "tc_cumpsgemm_syrk" : "tc_syrk", C,
            [&] {
                return split ? tc_cumpsgemm_syrk(h, n, k, alpha, A.dev, lda, beta, C.dev, ldc, nb)
                             : tc_syrk(h, n, k, alpha, A.dev, lda, beta, C.dev, ldc, nullptr, nb);
            },
            [&] { return tb_syrk(h, tier, n, k, alpha, A.dev, lda, beta, C.dev, ldc, nb); });
    // SYR2K (include/TensorBLAS.h:159-163)
    compare(ck, h, split ? "tc_cumpsgemm_syr2k" : "tc_syr2k", C,
            [&] {
                return split ? tc_cumpsgemm_syr2k(h, n, k, alpha, A.dev, lda, B.dev, ldb, beta, C.dev, ldc, nb)
                             : tc_syr2k(h, n, k, alpha, A.dev, lda, B.dev, ldb, beta, C.dev, ldc, nullptr, nb);
            },
            [&] { return tb_syr2k(h, tier, n, k, alpha, A.dev, lda, B.dev, ldb, beta, C.dev, ldc, nb); });

    // TRMM (include/TensorBLAS.h:153-157): L m x m, B m x nn, C m x nn
    const long m = 59, nn = 41, ldl = m + 2, ldbt = m + 6, ldct = m + 9;
    Buf<float> L(ldl * m, seed + 4), Bt(ldbt * nn, seed + 5), Ct(ldct * nn, seed + 6);
    make_dominant(L, m, ldl);
    if (!L.upload() || !Bt.upload() || !Ct.upload()) { ck.note("upload", -1, -1, false); return; }
    compare(ck, h, split ? "tc_cumpsgemm_trmm" : "tc_trmm", Ct,
            [&] {
                return split ? tc_cumpsgemm_trmm(h, m, nn, alpha, L.dev, ldl, Bt.dev, ldbt, Ct.dev, ldct, nb)
                             : tc_trmm(h, m, nn, alpha, L.dev, ldl, Bt.dev, ldbt, Ct.dev, ldct, nullptr, nb);
            },
            [&] { return tb_trmm(h, tier, m, nn, alpha, L.dev, ldl, Bt.dev, ldbt, Ct.dev, ldct, nb); });
    // SYMM (include/TensorBLAS.h:165-167): A m x m symmetric (lower), B m x nn (int ldb)
    compare(ck, h, split ? "tc_cumpsgemm_symm" : "tc_symm", Ct,
            [&] {
                return split ? tc_cumpsgemm_symm(h, m, nn, alpha, L.dev, ldl, Bt.dev, (int)ldbt, beta, Ct.dev, ldct)
                             : tc_symm(h, m, nn, alpha, L.dev, ldl, Bt.dev, (int)ldbt, beta, Ct.dev, ldct, nullptr);
            },
            [&] { return tb_symm(h, tier, m, nn, alpha, L.dev, ldl, Bt.dev, ldbt, beta, Ct.dev, ldct); });
    // TRSM (include/TensorBLAS.h:147-151): X L^T = B, B mr x m
    const long mr = 53, ldx = mr + 4;
    Buf<float> X(ldx * m, seed + 7);
    if (!X.upload()) { ck.note("upload", -1, -1, false); return; }
    compare(ck, h, split ? "tc_cumpsgemm_trsm" : "tc_trsm", X,
            [&] {
                return split ? tc_cumpsgemm_trsm(h, h2, mr, m, L.dev, ldl, X.dev, ldx, nb)
                             : tc_trsm(h, mr, m, L.dev, ldl, X.dev, ldx, nullptr, nb);
            },
            [&] { return tb_trsm(h, tier, mr, m, L.dev, ldl, X.dev, ldx, nb); });
}

template <>
void run_family<double>(Checker& ck, tb_handle_t h, tb_handle_t, int tier, unsigned long long seed)
{
    const long n = 61, k = 38, nb = 256;
    const long lda = n + 4, ldb = n + 2, ldc = n + 5;
    const double alpha = -1.25, beta = 0.625;
    Buf<double> A(lda * k, seed + 1), B(ldb * k, seed + 2), C(ldc * n, seed + 3);
    if (!A.upload() || !B.upload() || !C.upload()) { ck.note("upload", -1, -1, false); return; }
    compare(ck, h, "tc_ozimmu_syrk", C,
            [&] { return tc_ozimmu_syrk(h, n, k, alpha, A.dev, lda, beta, C.dev, ldc, nb); },
            [&] { return tb_syrk(h, tier, n, k, alpha, A.dev, lda, beta, C.dev, ldc, nb); });
    compare(ck, h, "tc_ozimmu_syr2k", C,
            [&] { return tc_ozimmu_syr2k(h, n, k, alpha, A.dev, lda, B.dev, ldb, beta, C.dev, ldc, nb); },
            [&] { return tb_syr2k(h, tier, n, k, alpha, A.dev, lda, B.dev, ldb, beta, C.dev, ldc, nb); });
    const long m = 57, nn = 43, ldl = m + 3, ldbt = m + 7, ldct = m + 8;
    Buf<double> L(ldl * m, seed + 4), Bt(ldbt * nn, seed + 5), Ct(ldct * nn, seed + 6);
    make_dominant(L, m, ldl);
    if (!L.upload() || !Bt.upload() || !Ct.upload()) { ck.note("upload", -1, -1, false); return; }
    compare(ck, h, "tc_ozimmu_trmm", Ct,
            [&] { return tc_ozimmu_trmm(h, m, nn, alpha, L.dev, ldl, Bt.dev, ldbt, Ct.dev, ldct, nb); },
            [&] { return tb_trmm(h, tier, m, nn, alpha, L.dev, ldl, Bt.dev, ldbt, Ct.dev, ldct, nb); });
    compare(ck, h, "tc_ozimmu_symm", Ct,
            [&] { return tc_ozimmu_symm(h, m, nn, alpha, L.dev, ldl, Bt.dev, ldbt, beta, Ct.dev, ldct); },
            [&] { return tb_symm(h, tier, m, nn, alpha, L.dev, ldl, Bt.dev, ldbt, beta, Ct.dev, ldct); });
    const long mr = 49, ldx = mr + 6;
    Buf<double> X(ldx * m, seed + 7);
    if (!X.upload()) { ck.note("upload", -1, -1, false); return; }
    compare(ck, h, "tc_ozimmu_trsm", X,
            [&] { return tc_ozimmu_trsm(h, mr, m, L.dev, ldl, X.dev, ldx, nb); },
            [&] { return tb_trsm(h, tier, mr, m, L.dev, ldl, X.dev, ldx, nb); });
}

}  // namespace

// Runs every tc_* entry point against its tb_* call; writes one line per entry point
// into report (cap bytes, NUL-terminated) and returns the number of entry points whose
// results differ or whose calls failed (0: all 15 equal), or -1 when no handle could be
// created.
extern "C" __attribute__((visibility("default"))) int tc_boundary_check(char* report, int cap, int* checked)
{
    tb_handle_t h = nullptr, h2 = nullptr;
    if (tb_create(&h) != TB_SUCCESS || tb_create(&h2) != TB_SUCCESS) return -1;
    Checker ck;
    run_family<float>(ck, h, h2, TB_TIER_PLAIN, 100);
    run_family<float>(ck, h, h2, TB_TIER_SPLIT, 200);
    run_family<double>(ck, h, h2, TB_TIER_OZAKI, 300);
    tb_destroy(h2);
    tb_destroy(h);
    if (report && cap > 0) {
        std::strncpy(report, ck.report.c_str(), (size_t)cap - 1);
        report[cap - 1] = '\0';
    }
    if (checked) *checked = ck.checked;
    return ck.failures;
}
