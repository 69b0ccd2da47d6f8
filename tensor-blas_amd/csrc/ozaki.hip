// Ozaki-scheme FP64 GEMM on INT8 MFMA (the tc_ozimmu_* tier).
//
// Published method: Ozaki, Ogita, Oishi, Rump (2012) error-free splitting, in the
// INT8 form of Ootomo, Ozaki, Yokota (2024).  Each row of op(A) (column of op(B)) is
// scaled by 2^-e (e = frexp exponent of its max |x|) and cut into S signed 7-bit
// digits:  x = 2^e * sum_s d_s 2^(-7(s+1)) + tail,  d_s = trunc(r_s * 2^7),
// r_{s+1} = r_s * 2^7 - d_s.  Then
//   C(i,j) = 2^(ea_i+eb_j) * sum_g 2^(-7(g+2)) * D_g(i,j),
//   D_g = sum_{s+t=g} A_s * B_t^T     (exact int32 sums, g = 0..S-1)
// The S(S+1)/2 slice-pair products run on v_mfma_i32_32x32x32_i8; all pairs of one g
// share an int32 accumulator (exact while S*kchunk*127^2 < 2^31, K is processed in
// chunks), and each chunk's D_g are added to an FP64 accumulator in increasing g.
// The operation order is exactly oracle/oracle.c orc_ozaki_gemm's, so the result
// is reproducible bit for bit.
#include <hip/hip_runtime.h>

#include <cstdint>

#include "gemm_kernel.hpp"
#include "tb_kernels.hpp"

#pragma clang fp contract(off)

namespace tb {

namespace {

typedef int intx4_t __attribute__((ext_vector_type(4)));
typedef int intx16_t __attribute__((ext_vector_type(16)));

constexpr int kSliceTile = 64;

// Slices of a 64 x 64 block of op(X) rows: loaded through LDS so that both the
// source read and the int8 write are coalesced.
__global__ void __launch_bounds__(256)
ozaki_slice_kernel(const double* __restrict__ X, long long ld, int mode, int rows, int K, int kpad,
                   int S, long long sx, const int* __restrict__ exps, int8_t* __restrict__ out)
{
    __shared__ double t[kSliceTile][kSliceTile + 1];   // t[row][k]
    const int q = blockIdx.z;
    const double* Xq = X + (long long)q * sx;
    const int r0 = blockIdx.y * kSliceTile, k0 = blockIdx.x * kSliceTile;
    const int tid = threadIdx.x;
    // load: mode N / TriL / SymL-lower read rows contiguous, mode T reads k contiguous
    for (int e = tid; e < kSliceTile * kSliceTile; e += 256) {
        int rl, kl;
        if (mode == kOpT) { kl = e % kSliceTile; rl = e / kSliceTile; }
        else { rl = e % kSliceTile; kl = e / kSliceTile; }
        const int r = r0 + rl, k = k0 + kl;
        double v = 0.0;
        if (r < rows && k < K) {
            if (mode == kOpN) v = Xq[(long long)r + (long long)k * ld];
            else if (mode == kOpT) v = Xq[(long long)k + (long long)r * ld];
            else if (r >= k) v = Xq[(long long)r + (long long)k * ld];
            else if (mode == kOpSymL) v = Xq[(long long)k + (long long)r * ld];
        }
        t[rl][kl] = v;
    }
    __syncthreads();
    // each thread: one row, 16 consecutive k
    const int rl = tid >> 2, kl0 = (tid & 3) * 16;
    const int r = r0 + rl;
    if (r >= rows || k0 + kl0 >= kpad) return;
    const int e = exps[(long long)q * rows + r];
    const int sh = (e < -5000) ? 0 : e;   // all-zero row: digits are zero anyway
    double rem[16];
#pragma unroll
    for (int j = 0; j < 16; ++j) rem[j] = __builtin_ldexp(t[rl][kl0 + j], -sh);
    for (int s = 0; s < S; ++s) {
        intx4_t pk;
#pragma unroll
        for (int w = 0; w < 4; ++w) {
            uint32_t word = 0;
#pragma unroll
            for (int b = 0; b < 4; ++b) {
                const int j = w * 4 + b;
                const double tt = rem[j] * 128.0;
                const double d = __builtin_trunc(tt);
                rem[j] = tt - d;
                word |= (uint32_t)(uint8_t)(int8_t)(int)d << (8 * b);
            }
            pk[w] = (int)word;
        }
        int8_t* dst = out + (((long long)q * S + s) * rows + r) * (long long)kpad + k0 + kl0;
        *reinterpret_cast<intx4_t*>(dst) = pk;
    }
}

constexpr int OBM = 64, OBN = 64, OBK = 64, OPAD = 16;
constexpr int OLDK = OBK + OPAD;   // bytes per LDS row
constexpr int kMinSlices = 3, kMaxSlices = 10;

// One 64 x 64 tile of C per workgroup; each wave owns a 32 x 32 block.
template <int S>
__global__ void __launch_bounds__(256)
ozaki_gemm_kernel(OzakiProb P, int tilesM, int totalTiles)
{
    __shared__ __attribute__((aligned(16))) int8_t sA[S * OBM * OLDK];   // [S][OBM][OLDK]
    __shared__ __attribute__((aligned(16))) int8_t sB[S * OBN * OLDK];

    const int g = xcd_remap(blockIdx.x, totalTiles);
    int ti, tj;
    if (P.lowerC) {
        ti = (int)((__builtin_sqrtf(8.0f * (float)g + 1.0f) - 1.0f) * 0.5f);
        while ((ti + 1) * (ti + 2) / 2 <= g) ++ti;
        while (ti * (ti + 1) / 2 > g) --ti;
        tj = g - ti * (ti + 1) / 2;
    } else {
        ti = g % tilesM;
        tj = g / tilesM;
    }
    const int i0 = ti * OBM, j0 = tj * OBN;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wm = wave & 1, wn = wave >> 1;

    intx16_t acc[S];
    double facc[16];
#pragma unroll
    for (int r = 0; r < 16; ++r) facc[r] = 0.0;

    const int npairs = P.sa2 ? 2 : 1;
    const int kblocks = P.kpad / OBK;
    const int cblocks = P.kchunk / OBK;       // k blocks per exactness chunk
    const int lrow = tid >> 2, lseg = (tid & 3) * 16;

    for (int pair = 0; pair < npairs; ++pair) {
        const int8_t* GA = pair ? P.sa2 : P.sa;
        const int8_t* GB = pair ? P.sb2 : P.sb;
        for (int c0 = 0; c0 < kblocks; c0 += cblocks) {
            const int c1 = min(kblocks, c0 + cblocks);
#pragma unroll
            for (int q = 0; q < S; ++q) acc[q] = intx16_t{};
            for (int kb = c0; kb < c1; ++kb) {
                __syncthreads();
                // stage all slices of both panels for this k block
#pragma unroll
                for (int s = 0; s < S; ++s) {
                    const int ra = i0 + lrow, rb = j0 + lrow;
                    intx4_t va = intx4_t{}, vb = intx4_t{};
                    if (ra < P.m)
                        va = *reinterpret_cast<const intx4_t*>(
                            GA + ((long long)s * P.m + ra) * P.kpad + kb * OBK + lseg);
                    if (rb < P.n)
                        vb = *reinterpret_cast<const intx4_t*>(
                            GB + ((long long)s * P.n + rb) * P.kpad + kb * OBK + lseg);
                    *reinterpret_cast<intx4_t*>(sA + (s * OBM + lrow) * OLDK + lseg) = va;
                    *reinterpret_cast<intx4_t*>(sB + (s * OBN + lrow) * OLDK + lseg) = vb;
                }
                __syncthreads();
                const int fr = lane & 31, fk = (lane >> 5) * 16;
#pragma unroll
                for (int kk = 0; kk < OBK; kk += 32) {
                    intx4_t af[S], bf[S];
#pragma unroll
                    for (int s = 0; s < S; ++s) {
                        af[s] = *reinterpret_cast<const intx4_t*>(sA + (s * OBM + wm * 32 + fr) * OLDK + kk + fk);
                        bf[s] = *reinterpret_cast<const intx4_t*>(sB + (s * OBN + wn * 32 + fr) * OLDK + kk + fk);
                    }
#pragma unroll
                    for (int gg = 0; gg < S; ++gg) {
#pragma unroll
                        for (int s = 0; s <= gg; ++s)
                            acc[gg] = __builtin_amdgcn_mfma_i32_32x32x32_i8(bf[gg - s], af[s], acc[gg], 0, 0, 0);
                    }
                }
            }
            // chunk done: D_g into FP64, increasing g
#pragma unroll
            for (int gg = 0; gg < S; ++gg) {
#pragma unroll
                for (int r = 0; r < 16; ++r)
                    facc[r] = facc[r] + __builtin_ldexp((double)acc[gg][r], -7 * (gg + 2));
            }
        }
    }

    // epilogue: D[jj][ii], ii = lane%32 runs along C's rows
    const int i = i0 + wm * 32 + (lane & 31);
    if (i >= P.m) return;
    const int ea = P.ea[i];
    const int sea = ea < -5000 ? 0 : ea;
#pragma unroll
    for (int r = 0; r < 16; ++r) {
        const int j = j0 + wn * 32 + (r >> 2) * 8 + (lane >> 5) * 4 + (r & 3);
        if (j >= P.n || (P.lowerC && i < j)) continue;
        const int eb = P.eb[j];
        const int seb = eb < -5000 ? 0 : eb;
        const double v = __builtin_ldexp(facc[r], sea + seb);
        double* c = P.C + (long long)i + (long long)j * P.ldc;
        *c = (P.beta == 0.0) ? P.alpha * v : __builtin_fma(P.alpha, v, P.beta * *c);
    }
}

}  // namespace

hipError_t ozaki_slice(const double* X, long long ld, int mode, int rows, int K, int kpad,
                       int nslices, int nb, long long sx, const int* exps, int8_t* slices,
                       hipStream_t s)
{
    if (rows <= 0 || kpad <= 0 || nb <= 0) return hipSuccess;
    if (kpad % kSliceTile || nslices < kMinSlices || nslices > kMaxSlices) return hipErrorInvalidValue;
    dim3 g(kpad / kSliceTile, (rows + kSliceTile - 1) / kSliceTile, nb);
    hipLaunchKernelGGL(ozaki_slice_kernel, g, dim3(256), 0, s, X, ld, mode, rows, K, kpad, nslices,
                       sx, exps, slices);
    return hipGetLastError();
}

hipError_t ozaki_gemm(const OzakiProb& p, hipStream_t s)
{
    if (p.m <= 0 || p.n <= 0) return hipSuccess;
    if (p.nslices < kMinSlices || p.nslices > kMaxSlices || p.kpad % OBK || p.kchunk < OBK ||
        p.kchunk % OBK)
        return hipErrorInvalidValue;
    const int tm = (p.m + OBM - 1) / OBM, tn = (p.n + OBN - 1) / OBN;
    if (p.lowerC && (p.m != p.n)) return hipErrorInvalidValue;
    const int total = p.lowerC ? tm * (tm + 1) / 2 : tm * tn;
    switch (p.nslices) {
#define TB_OZ_CASE(S_)                                                                            \
    case S_:                                                                                      \
        hipLaunchKernelGGL(ozaki_gemm_kernel<S_>, dim3(total), dim3(256), 0, s, p, tm, total);    \
        break;
        TB_OZ_CASE(3) TB_OZ_CASE(4) TB_OZ_CASE(5) TB_OZ_CASE(6) TB_OZ_CASE(7) TB_OZ_CASE(8)
        TB_OZ_CASE(9) TB_OZ_CASE(10)
#undef TB_OZ_CASE
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

}  // namespace tb
