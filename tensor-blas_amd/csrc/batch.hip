// Persistent kernel of a batch plan (tb_batch.hpp) and its batched prologue kernels.
//
// One workgroup per CU loops: take the next work item from an atomic counter (items
// are sorted by decreasing work), run it, until the list is exhausted.  A SYRK or TRMM
// item is one output tile, 64 x 256 or 128 x 128 as the plan chose for its instance,
// on operands the prologue put into FP16 planes.  A TRSM item
// walks its instance's recursion for 64 rows: at each step it converts its rows of the
// solved block (the leaf's B block, or the update's X1) into FP16 planes in its own
// scratch, with the exponents the single call uses; then each leaf X = (B + B W^T) D^-1
// is one 64 x 256 tile and each update B2 -= X1 L21^T is ceil(w / 256) tiles, reading
// L and W^T from planes the prologue made.  All these tiles stage their operands with
// direct-to-LDS loads.  The plane values and the MFMA sequence per output element are
// the single call's, so a batched instance equals its single call bit for bit.
#include <hip/hip_runtime.h>

#include "aux_device.hpp"
#include "gemm16_kernel.hpp"
#include "tb_batch.hpp"
#include "tb_kernels.hpp"

namespace tb {

namespace {

constexpr int kPP = kSrcPlanes;

// stages of the plane-sourced panel / square tiles: as many as fit in the LDS
template <int NPL>
constexpr int panel_stages() { return NPL == 3 ? 2 : 3; }
template <int NPL>
constexpr int sq_stages() { return NPL == 3 ? 3 : 4; }
template <int NPL, int SRC>
constexpr int panel_nst() { return SRC == kSrcPlanes ? panel_stages<NPL>() : 2; }

// waves per workgroup: four, or eight (two per SIMD, as the single calls' split tiles)
// when every tile is 2-plane; 64-row panels of 1 or 3 planes do not split into whole
// direct-to-LDS instructions over eight waves.  Both give the same bits: each element
// sums the same MFMA chain in the same order.
template <int UNPL, int ANPL>
constexpr bool waves_ok(int nw) { return nw == 4 || (nw == 8 && UNPL == 2 && ANPL == 2); }

template <int NPL, int SRC, int NW>
using CfgPanel = Gemm16Cfg<NPL, SRC, SRC, kPanelRows, kPanelCols, 1, NW, panel_nst<NPL, SRC>()>;
template <int NPL, int NW>
using CfgSq = Gemm16Cfg<NPL, kSrcPlanes, kSrcPlanes, 128, 128, 2, NW / 2, sq_stages<NPL>()>;
template <int NPL, int SRC, int NW>
using LdsPanel = Gemm16Lds<CfgPanel<NPL, SRC, NW>, kPanelRows, kPanelCols>;
template <int NPL, int NW>
using LdsSq = Gemm16Lds<CfgSq<NPL, NW>, 128, 128>;

constexpr size_t cmax(size_t a, size_t b) { return a > b ? a : b; }

template <int NPL, int SRC, int NW>
__device__ __forceinline__ void panel_tile(const GemmProb& P, int i0, int j0, char* raw)
{
    auto* L = reinterpret_cast<LdsPanel<NPL, SRC, NW>*>(raw);
    gemm16_tile<NPL, SRC, SRC, kPanelRows, kPanelCols, 1, NW, panel_nst<NPL, SRC>()>(
        P, 0, i0, j0, L->ring, L->scA, L->scB, L->emax);
}

// running exponents are raised by atomics at the L2: read them past the L1
__device__ __forceinline__ int load_exp(const int* p)
{
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Rows i0 .. i0+63 (those < m) of the column-major X (ld), columns 0 .. K-1, into NPL
// planes at dst (plane p, row r, column k at dst[plane_off(NPL, ldp, sp, p, r, k)]) with columns
// K .. round_up(K, 64) and rows >= m zero.  Split planes are scaled per row by
// 2^(15 - e[row]) exactly as split_planes_block converts them.  t: a 64 x 65 LDS tile.
// NT threads (256 or 512): four per row, each converting 16 (NT = 256) or 8 columns of
// every 64-column block.
template <int NPL, int NT>
__device__ __forceinline__ void panel_planes(const float* __restrict__ X, long long ld, int m, int i0, int K,
                             const int* __restrict__ e, _Float16* __restrict__ dst, long long ldp,
                             long long sp, float (*t)[65])
{
    typedef _Float16 half8 __attribute__((ext_vector_type(8)));
    static_assert(NT == 256 || NT == 512, "threads");
    constexpr int kHalves = NT == 256 ? 2 : 1;       // 8-column halves per thread
    const int tid = threadIdx.x;
    const int rl = (tid >> 2) & 63, cl0 = (tid & 3) * 16, h0 = NT == 256 ? 0 : tid >> 8;
    int sc = 0;
    if (NPL >= 2 && i0 + rl < m) {
        const int ex = e[i0 + rl];
        sc = 15 - (ex < -1000 ? 0 : (ex > 1000 ? 1000 : ex));
    }
    const int kend = (K + 63) / 64 * 64;
    for (int c0 = 0; c0 < kend; c0 += 64) {
        __syncthreads();                                 // t is free
        // every load of the block in flight before the LDS stores (clamped, masked)
        constexpr int kPer = 64 * 64 / NT;
        float v[kPer];
#pragma unroll
        for (int it = 0; it < kPer; ++it) {
            const int q = tid + it * NT;
            const int r = i0 + (q & 63), k = c0 + (q >> 6);
            const bool keep = r < m && k < K;
            v[it] = (float)masked_load(X + (keep ? (long long)r + (long long)k * ld : 0ll), keep);
        }
#pragma unroll
        for (int it = 0; it < kPer; ++it) {
            const int q = tid + it * NT;
            t[q & 63][q >> 6] = v[it];
        }
        __syncthreads();
        for (int hq = 0; hq < kHalves; ++hq) {
            const int half = h0 + hq;
            half8 hh, mm, ll;
#pragma unroll
            for (int j = 0; j < 8; ++j) {
                const float x = t[rl][cl0 + half * 8 + j];
                if constexpr (NPL == 1) {
                    hh[j] = (_Float16)x;
                } else {
                    const float xs = __builtin_ldexpf(x, sc);
                    const _Float16 h = (_Float16)xs;
                    hh[j] = h;
                    if constexpr (NPL == 2) {
                        mm[j] = (_Float16)(xs - (float)h);      // the 2-plane form: xs ~ h + m
                    } else {
                        const float r1 = (xs - (float)h) * 2048.0f;
                        const _Float16 mq = (_Float16)r1;
                        mm[j] = mq;
                        ll[j] = (_Float16)((r1 - (float)mq) * 2048.0f);
                    }
                }
            }
            const int c = c0 + cl0 + half * 8;
            *reinterpret_cast<half8*>(dst + plane_off(NPL, ldp, sp, 0, rl, c)) = hh;
            if constexpr (NPL >= 2) *reinterpret_cast<half8*>(dst + plane_off(NPL, ldp, sp, 1, rl, c)) = mm;
            if constexpr (NPL == 3) *reinterpret_cast<half8*>(dst + plane_off(NPL, ldp, sp, 2, rl, c)) = ll;
        }
    }
    // the tiles read these planes with direct-to-LDS loads: stores complete, then the
    // workgroup's vector L1 drops whatever it held of the scratch
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
    __syncthreads();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
}

// rows i0 .. i0+63 of one TRSM instance through its whole schedule
template <int UNPL, int ANPL, int NW>
__device__ __forceinline__ void trsm_panel(const BatchInst& I, const BatchStep* __restrict__ steps, int i0,
                           const BatchScratch& S, char* raw)
{
    const int rows = min(kPanelRows, I.m - i0);
    _Float16* scr = reinterpret_cast<_Float16*>(static_cast<char*>(S.base) +
                                                (long long)blockIdx.x * S.bytes_per_wg);
    // the scratch holds rows i0..i0+63: row i of the operand sits (i - i0) plane rows
    // after scr, in the layout of the step's plane count
    auto scr_at = [&](int npl) -> const _Float16* {
        return scr - (long long)i0 * plane_row_stride(npl, S.ldp);
    };
    auto* t = reinterpret_cast<float (*)[65]>(raw);
    for (int q = 0; q < I.nsteps; ++q) {
        const BatchStep st = steps[I.step0 + q];
        // the exponents this step's conversions use: the single-call path converts the
        // solved block once per step, before the step's epilogues raise eRow
        __syncthreads();
        for (int r = threadIdx.x; r < rows; r += blockDim.x) I.eSnap[i0 + r] = load_exp(I.eRow + i0 + r);
        __syncthreads();
        GemmProb P = {};
        P.m = I.m;
        P.batch = 1;
        P.modeA = P.modeB = kOpN;
        P.modeC = kCFull;
        P.eA = I.eSnap;
        P.eOut = I.eRow;
        P.ldc = I.ldc;
        P.beta = 1.0;
        P.beta32 = 1.0f;
        if (st.type == 0) {
            // leaf: B(:, off:off+s) := (B + B W^T) D^-1 in place (one tile spans all s columns)
            const int off = st.a, s = st.b;
            panel_planes<ANPL, 64 * NW>(I.C + (long long)off * I.ldc, I.ldc, I.m, i0, s, I.eSnap, scr,
                                        S.ldp, S.sp, t);
            P.n = s;
            P.k = s;
            P.A = scr_at(ANPL);
            P.lda = S.ldp;
            P.spA = S.sp;
            P.B = static_cast<const _Float16*>(I.Zp) + (long long)st.eoff * plane_row_stride(ANPL, kZpLd);
            P.ldb = kZpLd;
            P.spB = I.spZ;
            P.srcA = P.srcB = kPP;
            P.C = I.C + (long long)off * I.ldc;
            P.alpha = 1.0;
            P.alpha32 = 1.0f;
            P.cdiv = I.ldiag + off;
            P.eB = I.eZ + st.eoff;
            panel_tile<ANPL, kPP, NW>(P, i0, 0, raw);
            continue;
        }
        // update: B(:, c0:c0+w) -= B(:, k0:k0+kw) * L(c0:c0+w, k0:k0+kw)^T, in the update
        // tier (type 1) or, inside a wide leaf block, in the leaf tier (type 2)
        const int c0 = st.a, w = st.b, k0 = st.c, kw = st.d;
        P.n = w;
        P.k = kw;
        P.C = I.C + (long long)c0 * I.ldc;
        P.alpha = -1.0;
        P.alpha32 = -1.0f;
        P.eB = I.eL + c0;
        {
            // L's plane rows give 16-byte aligned 32-k stages from column k0 (in the copy
            // shifted by kshift1 when k0 is not a multiple of 8); past kw they meet the
            // zero padding of X1's planes
            const bool sh = (k0 & 7) != 0;
            const long long kc = (long long)k0 + (sh ? I.kshift1 : 0);
            P.lda = S.ldp;
            P.spA = S.sp;
            P.ldb = I.ldpL;
            P.srcA = P.srcB = kPP;
            if (st.type == 1) {
                panel_planes<UNPL, 64 * NW>(I.C + (long long)k0 * I.ldc, I.ldc, I.m, i0, kw, I.eSnap,
                                            scr, S.ldp, S.sp, t);
                const PlaneRef lr = plane_ref(UNPL, static_cast<const _Float16*>(sh ? I.LpU1 : I.LpU),
                                              I.ldpL, c0, kc);
                P.A = scr_at(UNPL);
                P.B = lr.p;
                P.kofsB = lr.kofs;
                P.spB = I.spLU;
                for (int j0 = 0; j0 < w; j0 += kPanelCols) panel_tile<UNPL, kPP, NW>(P, i0, j0, raw);
            } else {
                panel_planes<ANPL, 64 * NW>(I.C + (long long)k0 * I.ldc, I.ldc, I.m, i0, kw, I.eSnap,
                                            scr, S.ldp, S.sp, t);
                const PlaneRef lr = plane_ref(ANPL, static_cast<const _Float16*>(sh ? I.LpA1 : I.LpA),
                                              I.ldpL, c0, kc);
                P.A = scr_at(ANPL);
                P.B = lr.p;
                P.kofsB = lr.kofs;
                P.spB = I.spLA;
                for (int j0 = 0; j0 < w; j0 += kPanelCols) panel_tile<ANPL, kPP, NW>(P, i0, j0, raw);
            }
        }
    }
}

template <int UNPL, int ANPL, int NW>
__global__ void __launch_bounds__(64 * NW)
batch_kernel(const BatchItem* __restrict__ items, int nitems, const BatchInst* __restrict__ insts,
             const BatchStep* __restrict__ steps, int* counter, BatchScratch S)
{
    static_assert(waves_ok<UNPL, ANPL>(NW), "waves");
    constexpr size_t kPanels = cmax(sizeof(LdsPanel<UNPL, kPP, NW>), sizeof(LdsPanel<ANPL, kPP, NW>));
    constexpr size_t kLds = cmax(cmax(kPanels, sizeof(LdsSq<UNPL, NW>)), 64 * 65 * sizeof(float));
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
            trsm_panel<UNPL, ANPL, NW>(I, steps, w.i0, S, raw);
        } else {
            // SYRK: lower tile of A A^T (one operand); TRMM: full tile of L B with L's rows
            // zero above the diagonal (the tile's k range stops at its last row)
            const bool trmm = I.op == kBatchTrmm;
            GemmProb P = {};
            P.m = trmm ? I.m : I.n;
            P.n = I.n;
            P.k = trmm ? I.m : I.k;
            P.A = I.Ap;
            P.lda = I.ldpA;
            P.spA = I.spA;
            P.B = trmm ? I.Bp : I.Ap;
            P.ldb = trmm ? I.ldpB : I.ldpA;
            P.spB = trmm ? I.spB : I.spA;
            P.modeA = trmm ? kOpTriL : kOpN;
            P.modeB = kOpN;
            P.srcA = P.srcB = kPP;
            P.C = I.C;
            P.ldc = I.ldc;
            P.alpha = I.alpha;
            P.beta = I.beta;
            P.alpha32 = I.alpha;
            P.beta32 = I.beta;
            P.modeC = trmm ? kCFull : kCLower;
            P.eA = I.eRow;
            P.eB = trmm ? I.eCol : I.eRow;
            P.batch = 1;
            if (w.shape == kShapeSq) {
                auto* L = reinterpret_cast<LdsSq<UNPL, NW>*>(raw);
                gemm16_tile<UNPL, kPP, kPP, 128, 128, 2, NW / 2, sq_stages<UNPL>()>(
                    P, 0, w.i0, w.j0, L->ring, L->scA, L->scB, L->emax);
            } else {
                panel_tile<UNPL, kPP, NW>(P, w.i0, w.j0, raw);
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
    const TrtriJob& J = find_job(jobs, njobs, blockIdx.x);
    const int local = blockIdx.x - J.blockBegin;
    const int q = local / J.colBlocks, cb = (local % J.colBlocks) * kTrtriCols;
    trtri_block<float>(J.L + (long long)q * J.bstride, J.lda, J.s, J.Z + (long long)q * J.s * J.s,
                       J.ldiag + (long long)q * J.s, cb);
}

__global__ void __launch_bounds__(256)
batch_planes_kernel(const PlanesJob* __restrict__ jobs, int njobs)
{
    __shared__ float t[64][65];
    const PlanesJob& J = find_job(jobs, njobs, blockIdx.x);
    const int local = blockIdx.x - J.blockBegin;
    const int bx = local % J.gx, by = local / J.gx;
    _Float16* out = static_cast<_Float16*>(J.out);
    if (J.npl == 1)
        split_planes_block<1>(J.X, J.ld, J.mode, J.rows, J.K, J.e, nullptr, out, J.ldp, J.sp, J.kshift, bx, by, t);
    else if (J.npl == 2)
        split_planes_block<2>(J.X, J.ld, J.mode, J.rows, J.K, J.e, nullptr, out, J.ldp, J.sp, J.kshift, bx, by, t);
    else
        split_planes_block<3>(J.X, J.ld, J.mode, J.rows, J.K, J.e, nullptr, out, J.ldp, J.sp, J.kshift, bx, by, t);
}

template <int UNPL, int ANPL, int NW>
hipError_t run_nw(const BatchItem* items, int nitems, const BatchInst* insts, const BatchStep* steps,
                  int* counter, const BatchScratch& S, int grid, hipStream_t s)
{
    hipLaunchKernelGGL((batch_kernel<UNPL, ANPL, NW>), dim3(grid), dim3(64 * NW), 0, s, items, nitems,
                       insts, steps, counter, S);
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

hipError_t batch_planes(const PlanesJob* jobs, int njobs, int nblocks, hipStream_t s)
{
    if (njobs <= 0 || nblocks <= 0) return hipSuccess;
    hipLaunchKernelGGL(batch_planes_kernel, dim3(nblocks), dim3(256), 0, s, jobs, njobs);
    return hipGetLastError();
}

hipError_t batch_run(int unpl, int anpl, int waves, const BatchItem* items, int nitems,
                     const BatchInst* insts, const BatchStep* steps, int* counter,
                     const BatchScratch& S, int grid, hipStream_t s)
{
    if (nitems <= 0) return hipSuccess;
    if (grid <= 0) return hipErrorInvalidValue;
    if (unpl == 2 && anpl == 2 && waves == 8)
        return run_nw<2, 2, 8>(items, nitems, insts, steps, counter, S, grid, s);
    if (waves != 4) return hipErrorInvalidValue;
    if (unpl == 1 && anpl == 2) return run_nw<1, 2, 4>(items, nitems, insts, steps, counter, S, grid, s);
    if (unpl == 1 && anpl == 3) return run_nw<1, 3, 4>(items, nitems, insts, steps, counter, S, grid, s);
    if (unpl == 2 && anpl == 2) return run_nw<2, 2, 4>(items, nitems, insts, steps, counter, S, grid, s);
    if (unpl == 3 && anpl == 3) return run_nw<3, 3, 4>(items, nitems, insts, steps, counter, S, grid, s);
    return hipErrorInvalidValue;
}

}  // namespace tb
