// Tile orders of the Ozaki product kernels.  Host and device code: the CPU tests check
// that each order visits every tile of its grid exactly once.
#pragma once
#include <hip/hip_runtime.h>

namespace tb {

__host__ __device__ constexpr int oz_imin(int a, int b) { return a < b ? a : b; }

// Lower-triangular tile g of a T x T tile grid, walked in bands of kBand tile rows:
// inside a band column by column (rows fastest), first the columns left of the band,
// then its triangle.  Consecutive tiles -- those an XCD runs together -- then share
// kBand row panels and a few column panels instead of one row panel and many columns.
constexpr int kBand = 8;

__host__ __device__ __forceinline__ void lower_tile_banded(int g, int T, int& ti, int& tj)
{
    int r0 = 0;
    for (;;) {
        const int gb = oz_imin(kBand, T - r0);                     // rows of this band
        const int cnt = r0 * gb + gb * (gb + 1) / 2;
        if (g < cnt || r0 + gb >= T) break;
        g -= cnt;
        r0 += gb;
    }
    const int gb = oz_imin(kBand, T - r0);
    if (g < r0 * gb) {                                            // columns left of the band
        tj = g / gb;
        ti = r0 + g % gb;
        return;
    }
    g -= r0 * gb;
    tj = r0;                                                      // triangle: column c has gb - c rows
    for (int c = 0; c < gb; ++c) {
        const int rows = gb - c;
        if (g < rows) {
            tj = r0 + c;
            ti = r0 + c + g;
            return;
        }
        g -= rows;
    }
    ti = tj = T - 1;                                              // unreachable for g < total
}

// tile rows per band of the wide kernel's tile order: the 32 workgroups of an XCD then
// cover 4 x 8 tiles (512 rows of A, 512 columns of B per k block); bands of 8 and 16
// rows measured 1-2 % and 4-7 % slower on SYR2K / SYMM 32768 x 8192
constexpr int kBandW = 4;

// Wide tile g (128 x 64) of the lower part of C, walked in bands of kBandW tile rows:
// inside a band column by column (rows fastest), first the columns left of the band's
// diagonal (every row of the band), then those it crosses, column tj needing rows
// ti >= tj / 2 (128 ti + 127 >= 64 tj).
__host__ __device__ __forceinline__ void lower_wide_banded(int g, int T, int tn, int& ti, int& tj)
{
    int r0 = 0;
    for (;;) {
        const int gb = oz_imin(kBandW, T - r0);
        const int left = oz_imin(2 * r0, tn);
        int cnt = left * gb;
        for (int c = left; c < oz_imin(2 * (r0 + gb), tn); ++c) cnt += r0 + gb - c / 2;
        if (g < cnt || r0 + gb >= T) break;
        g -= cnt;
        r0 += gb;
    }
    const int gb = oz_imin(kBandW, T - r0);
    const int left = oz_imin(2 * r0, tn);
    if (g < left * gb) {
        tj = g / gb;
        ti = r0 + g % gb;
        return;
    }
    g -= left * gb;
    for (int c = left; c < oz_imin(2 * (r0 + gb), tn); ++c) {
        const int first = c / 2 > r0 ? c / 2 : r0;
        const int rows = r0 + gb - first;
        if (g < rows) {
            tj = c;
            ti = first + g;
            return;
        }
        g -= rows;
    }
    ti = T - 1;                                                   // unreachable for g < total
    tj = tn - 1;
}

}  // namespace tb
