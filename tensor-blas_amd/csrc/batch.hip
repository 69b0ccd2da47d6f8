// Persistent kernel of a batch plan (tb_batch.hpp) and its batched prologue kernels.
//
// One workgroup per CU loops: take the next work item from an atomic counter (items
// are sorted by decreasing work), run it, until the list is exhausted.  A TRSM item
// walks its instance's recursion for 64 rows: each leaf X = (B + B W^T) D^-1 is one
// 64 x 256 tile, each update B2 -= X1 L21^T is ceil(w / 256) such tiles.  Every tile
// converts its FP32 operands while staging them (the plane values the single-call
// path pre-converts, with the same exponents) and uses the same MFMA sequence per
// output element, so a batched instance equals its single call bit for bit.
#include <hip/hip_runtime.h>

#include "aux_device.hpp"
#include "gemm16_kernel.hpp"
#include "tb_batch.hpp"
#include "tb_kernels.hpp"

namespace tb {

namespace {

// tile configurations: TRSM panel tiles and the 64 x 256 SYRK tiles; 128 x 128 SYRK tiles
template <int NPL>
using CfgPanel = Gemm16Cfg<NPL, kSrcF32, kSrcF32, kPanelRows, kPanelCols, 1, 4, 2>;
template <int NPL>
using CfgSq = Gemm16Cfg<NPL, kSrcF32, kSrcF32, 128, 128, 2, 2, 2>;
template <int NPL>
using LdsPanel = Gemm16Lds<CfgPanel<NPL>, kPanelRows, kPanelCols>;
template <int NPL>
using LdsSq = Gemm16Lds<CfgSq<NPL>, 128, 128>;

constexpr size_t cmax(size_t a, size_t b) { return a > b ? a : b; }

template <int NPL>
__device__ __forceinline__ void panel_tile(const GemmProb& P, int i0, int j0, char* raw)
{
    auto* L = reinterpret_cast<LdsPanel<NPL>*>(raw);
    gemm16_tile<NPL, kSrcF32, kSrcF32, kPanelRows, kPanelCols, 1, 4, 2>(P, 0, i0, j0, L->ring,
                                                                        L->scA, L->scB, L->emax);
}

// running exponents are raised by atomics at the L2: read them past the L1
__device__ __forceinline__ int load_exp(const int* p)
{
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// rows i0 .. i0+63 of one TRSM instance through its whole schedule
template <int UNPL, int ANPL>
__device__ void trsm_panel(const BatchInst& I, const BatchStep* __restrict__ steps, int i0,
                           char* raw)
{
    const int rows = min(kPanelRows, I.m - i0);
    for (int q = 0; q < I.nsteps; ++q) {
        const BatchStep st = steps[I.step0 + q];
        // the exponents this step's conversions use: the single-call path converts the
        // solved block once per step, before the step's epilogues raise eRow
        for (int r = threadIdx.x; r < rows; r += blockDim.x) I.eSnap[i0 + r] = load_exp(I.eRow + i0 + r);
        __syncthreads();
        GemmProb P = {};
        P.m = I.m;
        P.batch = 1;
        P.modeC = kCFull;
        P.srcA = P.srcB = kSrcF32;
        P.eA = I.eSnap;
        P.eOut = I.eRow;
        if (st.type == 0) {
            // leaf: B(:, off:off+s) := (B + B W^T) D^-1 in place (one tile spans all s columns)
            const int off = st.a, s = st.b;
            P.n = s;
            P.k = s;
            P.A = I.C + (long long)off * I.ldc;
            P.lda = I.ldc;
            P.modeA = kOpN;
            P.B = I.Z + st.zoff;
            P.ldb = s;
            P.modeB = kOpN;
            P.C = I.C + (long long)off * I.ldc;
            P.ldc = I.ldc;
            P.alpha = 1.0;
            P.beta = 1.0;
            P.cscale = I.dinv + off;
            P.eB = I.eZ + st.eoff;
            panel_tile<ANPL>(P, i0, 0, raw);
        } else {
            // update: B(:, c0:c0+w) -= B(:, k0:k0+kw) * L(c0:c0+w, k0:k0+kw)^T
            const int c0 = st.a, w = st.b, k0 = st.c, kw = st.d;
            P.n = w;
            P.k = kw;
            P.A = I.C + (long long)k0 * I.ldc;
            P.lda = I.ldc;
            P.modeA = kOpN;
            P.B = I.A + c0 + (long long)k0 * I.lda;
            P.ldb = I.lda;
            P.modeB = kOpT;
            P.C = I.C + (long long)c0 * I.ldc;
            P.ldc = I.ldc;
            P.alpha = -1.0;
            P.beta = 1.0;
            P.eB = I.eL + c0;
            if (st.type == 1)
                for (int j0 = 0; j0 < w; j0 += kPanelCols) panel_tile<UNPL>(P, i0, j0, raw);
            else
                for (int j0 = 0; j0 < w; j0 += kPanelCols) panel_tile<ANPL>(P, i0, j0, raw);
        }
    }
}

template <int UNPL, int ANPL, int SQ>
__global__ void __launch_bounds__(256)
batch_kernel(const BatchItem* __restrict__ items, int nitems, const BatchInst* __restrict__ insts,
             const BatchStep* __restrict__ steps, int* counter)
{
    constexpr size_t kLds = cmax(cmax(sizeof(LdsPanel<UNPL>), sizeof(LdsPanel<ANPL>)),
                                 SQ ? sizeof(LdsSq<UNPL>) : 0);
    __shared__ __attribute__((aligned(1024))) char raw[kLds];
    __shared__ int next_s;
    for (;;) {
        if (threadIdx.x == 0) next_s = atomicAdd(counter, 1);
        __syncthreads();
        const int it = __builtin_amdgcn_readfirstlane(next_s);
        __syncthreads();           // every wave has read next_s before it is replaced
        if (it >= nitems) return;  // the counter only grows: every workgroup gets here
        const BatchItem w = items[it];
        const BatchInst& I = insts[w.inst];
        if (I.op == kBatchTrsm) {
            trsm_panel<UNPL, ANPL>(I, steps, w.i0, raw);
        } else {
            GemmProb P = {};
            P.m = P.n = I.n;
            P.k = I.k;
            P.A = P.B = I.A;
            P.lda = P.ldb = I.lda;
            P.modeA = kOpN;
            P.modeB = kOpT;
            P.srcA = P.srcB = kSrcF32;
            P.C = I.C;
            P.ldc = I.ldc;
            P.alpha = I.alpha;
            P.beta = I.beta;
            P.modeC = kCLower;
            P.eA = P.eB = I.eRow;
            P.batch = 1;
            if constexpr (SQ) {
                auto* L = reinterpret_cast<LdsSq<UNPL>*>(raw);
                gemm16_tile<UNPL, kSrcF32, kSrcF32, 128, 128, 2, 2, 2>(P, 0, w.i0, w.j0, L->ring,
                                                                       L->scA, L->scB, L->emax);
            } else {
                panel_tile<UNPL>(P, w.i0, w.j0, raw);
            }
        }
    }
}

template <typename J>
__device__ __forceinline__ const J& find_job(const J* jobs, int njobs, int b)
{
    int lo = 0, hi = njobs - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (jobs[mid].blockBegin <= b) lo = mid;
        else hi = mid - 1;
    }
    return jobs[lo];
}

__global__ void __launch_bounds__(256)
batch_rowexp_kernel(const RowexpJob* __restrict__ jobs, int njobs)
{
    const RowexpJob& J = find_job(jobs, njobs, blockIdx.x);
    const int local = blockIdx.x - J.blockBegin;
    const int rb = local % J.rowBlocks, kb = local / J.rowBlocks;
    if (J.kcontig) rowexp_k_block(J.X, J.ld, J.rows, J.K, false, J.out, rb, kb);
    else rowexp_mn_block(J.X, J.ld, J.rows, J.K, J.lower != 0, J.out, rb, kb);
}

__global__ void __launch_bounds__(64)
batch_trtri_kernel(const TrtriJob* __restrict__ jobs, int njobs)
{
    __shared__ double y[256 * 64];
    __shared__ double lrow[256];
    const TrtriJob& J = find_job(jobs, njobs, blockIdx.x);
    const int local = blockIdx.x - J.blockBegin;
    const int q = local / J.colBlocks, cb = (local % J.colBlocks) * 64;
    trtri_block(J.L + (long long)q * J.bstride, J.lda, J.s, J.Z + (long long)q * J.s * J.s,
                J.dinv + (long long)q * J.s, cb, y, lrow);
}

template <int UNPL, int ANPL>
hipError_t run_sq(int sq, const BatchItem* items, int nitems, const BatchInst* insts,
                  const BatchStep* steps, int* counter, int grid, hipStream_t s)
{
    if (sq)
        hipLaunchKernelGGL((batch_kernel<UNPL, ANPL, 1>), dim3(grid), dim3(256), 0, s, items, nitems,
                           insts, steps, counter);
    else
        hipLaunchKernelGGL((batch_kernel<UNPL, ANPL, 0>), dim3(grid), dim3(256), 0, s, items, nitems,
                           insts, steps, counter);
    return hipGetLastError();
}

}  // namespace

int batch_grid(int device)
{
    // one workgroup per CU: the LDS ring and registers of the tiles leave room for one
    int cus = 0;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device) != hipSuccess ||
        cus <= 0)
        cus = 256;
    return cus;
}

hipError_t batch_rowexp(const RowexpJob* jobs, int njobs, int nblocks, hipStream_t s)
{
    if (njobs <= 0 || nblocks <= 0) return hipSuccess;
    hipLaunchKernelGGL(batch_rowexp_kernel, dim3(nblocks), dim3(256), 0, s, jobs, njobs);
    return hipGetLastError();
}

hipError_t batch_trtri(const TrtriJob* jobs, int njobs, int nblocks, hipStream_t s)
{
    if (njobs <= 0 || nblocks <= 0) return hipSuccess;
    hipLaunchKernelGGL(batch_trtri_kernel, dim3(nblocks), dim3(64), 0, s, jobs, njobs);
    return hipGetLastError();
}

hipError_t batch_run(int unpl, int anpl, int sq, const BatchItem* items, int nitems,
                     const BatchInst* insts, const BatchStep* steps, int* counter, int grid,
                     hipStream_t s)
{
    if (nitems <= 0) return hipSuccess;
    if (grid <= 0) return hipErrorInvalidValue;
    if (unpl == 1 && anpl == 2) return run_sq<1, 2>(sq, items, nitems, insts, steps, counter, grid, s);
    if (unpl == 1 && anpl == 3) return run_sq<1, 3>(sq, items, nitems, insts, steps, counter, grid, s);
    if (unpl == 2 && anpl == 2) return run_sq<2, 2>(sq, items, nitems, insts, steps, counter, grid, s);
    if (unpl == 3 && anpl == 3) return run_sq<3, 3>(sq, items, nitems, insts, steps, counter, grid, s);
    return hipErrorInvalidValue;
}

}  // namespace tb
