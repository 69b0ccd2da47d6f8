// Host check of the Ozaki kernels' tile orders (tensor-blas_amd/csrc/ozaki_tiles.hpp):
// over many grid shapes, every flat index maps to a tile of the grid and every tile is
// visited exactly once.  Prints "FAIL ..." per failure and "done <failures>" at the end.
#include <cstdio>
#include <vector>

#include "ozaki_tiles.hpp"

int main()
{
    int fails = 0;
    // 64 x 64 kernel, lower-triangular C: T x T tiles, tile (ti, tj) with tj <= ti
    for (int T = 1; T <= 300; T += (T < 40 ? 1 : 37)) {
        const int total = T * (T + 1) / 2;
        std::vector<int> seen((size_t)T * T, 0);
        for (int g = 0; g < total; ++g) {
            int ti = -1, tj = -1;
            tb::lower_tile_banded(g, T, ti, tj);
            if (ti < 0 || ti >= T || tj < 0 || tj > ti || seen[(size_t)ti * T + tj]++) {
                std::printf("FAIL lower_tile_banded T=%d g=%d -> (%d, %d)\n", T, g, ti, tj);
                ++fails;
                break;
            }
        }
    }
    // wide kernel, lower part of C: T tile rows of 128, tn tile columns of 64; row ti
    // holds columns tj < min(2 ti + 2, tn)
    for (int T = 1; T <= 300; T += (T < 40 ? 1 : 37))
        for (int tn : {2 * T - 1, 2 * T, T, 1}) {
            if (tn < 1) continue;
            int total = 0;
            for (int t = 0; t < T; ++t) total += tb::oz_imin(2 * t + 2, tn);
            std::vector<int> seen((size_t)T * tn, 0);
            for (int g = 0; g < total; ++g) {
                int ti = -1, tj = -1;
                tb::lower_wide_banded(g, T, tn, ti, tj);
                if (ti < 0 || ti >= T || tj < 0 || tj >= tb::oz_imin(2 * ti + 2, tn) ||
                    seen[(size_t)ti * tn + tj]++) {
                    std::printf("FAIL lower_wide_banded T=%d tn=%d g=%d -> (%d, %d)\n", T, tn, g, ti, tj);
                    ++fails;
                    break;
                }
            }
        }
    std::printf("done %d\n", fails);
    return fails != 0;
}
