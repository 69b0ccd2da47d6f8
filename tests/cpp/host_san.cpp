// Host-code driver for the AddressSanitizer / UndefinedBehaviorSanitizer build of the
// library's host sources (tests/cpp/Makefile, target `san`).  It runs on a machine
// without a GPU and exercises what needs none: the schedule planners
// (util/mat_size.cu:5-55, BLAS/tc_trsm.cu:31-137), the status strings, argument
// checks of every entry point family, and the error paths of handle creation and of
// the reference's wrappers when no device is present.  It prints the planners'
// results for the test to compare with the oracle and with the unsanitized library.
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "../../include/TensorBLAS.h"

namespace {

int failures = 0;

void expect(bool ok, const char* what)
{
    if (!ok) {
        std::printf("FAIL %s\n", what);
        ++failures;
    }
}

void print_chunks(int trsm, long n)
{
    std::vector<long> out(64);
    const int c = tb_plan_chunks(trsm, n, out.data(), (int)out.size());
    std::printf("chunks %d %ld:", trsm, n);
    for (int i = 0; i < c; ++i) std::printf(" %ld", out[i]);
    std::printf("\n");
}

void print_trsm(long n, long nb)
{
    const int cap = 1 << 16;
    std::vector<long> out(5 * (size_t)cap);
    const int c = tb_plan_trsm(n, nb, out.data(), cap);
    unsigned long long sum = 0;
    for (long i = 0; i < 5l * (c > 0 ? c : 0); ++i) sum = sum * 31 + (unsigned long long)out[i];
    std::printf("trsm %ld %ld: %d %llu\n", n, nb, c, sum);
}

}  // namespace

int main()
{
    // planners, including sizes past the reference's power table and tiny ones
    const long ns[] = {0, 1, 7, 2047, 2048, 2049, 8192, 12345, 16384, 16385, 50000, 188416,
                       188416 + 8192 + 5, 250001};
    for (long n : ns) {
        print_chunks(0, n);
        print_chunks(1, n);
    }
    const long tn[][2] = {{1, 256}, {255, 256}, {256, 256}, {1000, 64}, {4096, 256},
                          {16384, 256}, {20000, 512}, {65536, 256}};
    for (auto& p : tn) print_trsm(p[0], p[1]);
    {
        long small[2];
        expect(tb_plan_chunks(0, 100000, small, 2) == -TB_ERR_INVALID, "chunks cap");
        expect(tb_plan_chunks(0, -1, small, 2) == -TB_ERR_INVALID, "chunks n < 0");
        expect(tb_plan_trsm(4096, 0, small, 2) == -TB_ERR_INVALID, "trsm nb 0");
    }
    // status strings of every code and of codes past the table
    for (int s = -2; s < 8; ++s) expect(tb_status_string(s) != nullptr, "status string");
    expect(tb_version() > 0, "version");

    // argument checks: a null handle is refused before anything is launched
    double d = 0.0;
    float f = 0.0f;
    expect(tb_syrk(nullptr, TB_TIER_SPLIT, 4, 4, 1.0, &f, 4, 0.0, &f, 4, 256) == TB_ERR_INVALID, "syrk");
    expect(tb_syr2k(nullptr, TB_TIER_PLAIN, 4, 4, 1.0, &f, 4, &f, 4, 0.0, &f, 4, 256) == TB_ERR_INVALID,
           "syr2k");
    expect(tb_trmm(nullptr, TB_TIER_FP32, 4, 4, 1.0, &f, 4, &f, 4, &f, 4, 256) == TB_ERR_INVALID, "trmm");
    expect(tb_trsm(nullptr, TB_TIER_FP64, 4, 4, &d, 4, &d, 4, 256) == TB_ERR_INVALID, "trsm");
    expect(tb_symm(nullptr, TB_TIER_OZAKI, 4, 4, 1.0, &d, 4, &d, 4, 0.0, &d, 4) == TB_ERR_INVALID, "symm");
    expect(tb_gemm(nullptr, TB_TIER_SPLIT, 0, 1, 4, 4, 4, 1.0, &f, 4, &f, 4, 0.0, &f, 4) == TB_ERR_INVALID,
           "gemm");
    expect(tb_syrk_noalloc(nullptr, TB_TIER_SPLIT, 4, 4, 1.0, &f, 4, 0.0, &f, 4, 256) == TB_ERR_INVALID,
           "syrk_noalloc");
    expect(tb_trsm_ex(nullptr, TB_TIER_SPLIT, 'C', 'L', 'L', 'N', 'N', 4, 4, 1.0, &f, 4, &f, 4, 256) ==
               TB_ERR_INVALID,
           "trsm_ex");
    expect(tb_set_ozaki_slices(nullptr, 7) == TB_ERR_INVALID, "ozaki slices");
    expect(tb_set_stream(nullptr, nullptr) == TB_ERR_INVALID, "set stream");
    tb_batch_t batch = nullptr;
    expect(tb_batch_create(nullptr, TB_TIER_SPLIT, nullptr, 0, 256, 0, &batch) == TB_ERR_INVALID, "batch");
    expect(tb_s2h(nullptr, 4, 4, &f, 4, &f, 4) == TB_ERR_INVALID, "s2h");

    // no device: handle creation and the reference's wrappers report an error
    tb_handle_t h = nullptr;
    const int c = tb_create(&h);
    std::printf("create %d\n", c);
    if (c == TB_SUCCESS) tb_destroy(h);
    expect(tb_create(nullptr) == TB_ERR_INVALID, "create null");

    std::printf("done %d\n", failures);
    return failures ? 1 : 0;
}
