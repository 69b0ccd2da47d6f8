// Batch plans (tb_batch_*): many independent SYRK / TRMM / TRSM instances of mixed sizes
// in one HIP graph -- batched row exponents, diagonal-block inverses and plane
// conversions, then the persistent batch kernel of batch.hip over a work list sorted by
// decreasing work.
//
// The plan reproduces the single calls exactly: the TRSM schedule comes from the same
// planner (trsm_schedule, BLAS/tc_trsm.cu:31-137; the reference's batched TRMM levels,
// BLAS/tc_trmm.cu:4-27, become output tiles of one triangular product), the exponents
// from the same row scans, the inverses from the same trtri body, and the tiles run the
// same MFMA sequence per element.  tests/test_gpu_batch.py checks batch == single calls
// bitwise.
//
// Tile choices: each SYRK / TRMM instance runs on 64 x 256 or 128 x 128 tiles, and the
// plan's workgroups have 4 or 8 waves (2-plane plans).  They come from the tuning table
// (batch_tuning.inc, measured on MI355X by python/tune_batch.py), keyed by (tier, op,
// size bucket), or, with TB_BATCH_AUTOTUNE, are timed on the plan's own instances.
#include <algorithm>
#include <cstring>
#include <new>
#include <vector>

#include "../tb_batch.hpp"
#include "tb_internal.hpp"

using namespace tb;

namespace {

constexpr int kBuckets = 6;            // size buckets: n < 128, < 256, < 512, < 1024, < 2048, more
constexpr int kTunedOps = 2;           // table rows: SYRK, TRMM
#include "../batch_tuning.inc"          // kTunedShape[tier][op][bucket], kTunedWaves[tier][bucket]

int bucket_of(long n)
{
    int b = 0;
    for (long v = 128; b < kBuckets - 1 && n >= v; v *= 2) ++b;
    return b;
}

// the tile choices of one plan
struct Choice {
    int shape[kTunedOps][kBuckets];    // BatchShape of SYRK (0) / TRMM (1) instances per bucket
    int waves;                         // waves per workgroup of the persistent kernel
};

}  // namespace

struct tb_batch_s {
    tb_handle_s* h = nullptr;
    int tier = 0, unpl = 0, anpl = 0, waves = 4, flags = 0;
    int count = 0, ninst_trsm = 0, ninst_syrk = 0, ninst_trmm = 0, ninst_sq = 0;
    Choice choice = {};
    bool tuned = false;
    char* dev = nullptr;
    size_t dev_bytes = 0;
    BatchItem* items = nullptr;
    int nitems = 0;
    BatchInst* insts = nullptr;
    BatchStep* steps = nullptr;
    RowexpJob* rj[2] = {nullptr, nullptr};
    int nrj[2] = {0, 0}, rjblocks[2] = {0, 0};
    TrtriJob* tj = nullptr;
    int ntj = 0, tjblocks = 0;
    PlanesJob* pj = nullptr;
    int npj = 0, pjblocks = 0;
    BatchScratch scratch = {nullptr, 0, 0, 0};
    int* exps = nullptr;
    long long nexps = 0;
    int* counter = nullptr;
    int grid = 0;
    hipStream_t cap = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    double tune_us = 0.0;        // autotuning time (0: the table's choices)
    double flops = 0.0;          // algorithmic flops of one run
};

namespace {

bool dim_ok(long v) { return v >= 0 && v < (1l << 31); }

int check_entry(const tb_batch_entry& e)
{
    if (e.op == TB_OP_SYRK) {
        if (!dim_ok(e.n) || !dim_ok(e.k)) return TB_ERR_INVALID;
        if (e.lda < std::max(1l, e.n) || e.ldc < std::max(1l, e.n)) return TB_ERR_INVALID;
        if (e.n > 0 && (!e.C || (e.k > 0 && !e.A))) return TB_ERR_INVALID;
        return TB_SUCCESS;
    }
    if (e.op == TB_OP_TRSM) {
        if (!dim_ok(e.m) || !dim_ok(e.n)) return TB_ERR_INVALID;
        if (e.lda < std::max(1l, e.n) || e.ldc < std::max(1l, e.m)) return TB_ERR_INVALID;
        if (e.m > 0 && e.n > 0 && (!e.A || !e.C)) return TB_ERR_INVALID;
        return TB_SUCCESS;
    }
    if (e.op == TB_OP_TRMM) {
        if (!dim_ok(e.m) || !dim_ok(e.n)) return TB_ERR_INVALID;
        if (e.lda < std::max(1l, e.m) || e.ldb < std::max(1l, e.m) || e.ldc < std::max(1l, e.m))
            return TB_ERR_INVALID;
        if (e.m > 0 && e.n > 0 && (!e.A || !e.B || !e.C)) return TB_ERR_INVALID;
        return TB_SUCCESS;
    }
    return TB_ERR_INVALID;
}

// size bucket of an instance: SYRK n, TRMM m (the triangle), TRSM n
int inst_bucket(const tb_batch_entry& e) { return bucket_of(e.op == TB_OP_TRMM ? e.m : e.n); }

Choice table_choice(int tier, const tb_batch_entry* E, int count)
{
    Choice c;
    const int t = tier == kSplitF16 ? 1 : 0;
    for (int o = 0; o < kTunedOps; ++o)
        for (int q = 0; q < kBuckets; ++q) c.shape[o][q] = kTunedShape[t][o][q];
    // waves: the table's choice for the bucket that holds most of the plan's TRSM work
    double work[kBuckets] = {};
    for (int i = 0; i < count; ++i)
        if (E[i].op == TB_OP_TRSM) work[inst_bucket(E[i])] += (double)E[i].m * E[i].n * E[i].n;
    int best = kBuckets - 1;
    for (int q = 0; q < kBuckets; ++q)
        if (work[q] > work[best]) best = q;
    c.waves = kTunedWaves[t][best];
    return c;
}

struct Costed {
    BatchItem it;
    double cost;
};

void sort_items(std::vector<Costed>& v)
{
    std::stable_sort(v.begin(), v.end(), [](const Costed& a, const Costed& b) { return a.cost > b.cost; });
}

// the launch sequence of one run on stream s (captured once into the plan's graph)
hipError_t enqueue(const tb_batch_s* b, hipStream_t s)
{
    hipError_t e;
    if (b->nexps > 0 && (e = fill_i32(b->exps, b->nexps, kExpZero, s)) != hipSuccess) return e;
    if ((e = fill_i32(b->counter, 1, 0, s)) != hipSuccess) return e;
    if ((e = batch_rowexp(b->rj[0], b->nrj[0], b->rjblocks[0], s)) != hipSuccess) return e;
    if ((e = batch_trtri(b->tj, b->ntj, b->tjblocks, s)) != hipSuccess) return e;
    if ((e = batch_rowexp(b->rj[1], b->nrj[1], b->rjblocks[1], s)) != hipSuccess) return e;
    if ((e = batch_planes(b->pj, b->npj, b->pjblocks, s)) != hipSuccess) return e;
    return batch_run(b->unpl, b->anpl, b->waves, b->items, b->nitems, b->insts, b->steps,
                     b->counter, b->scratch, b->grid, s);
}

void release(tb_batch_s* b)
{
    // the plan may still run on its handle's stream; the handle itself may be gone
    (void)hipDeviceSynchronize();
    if (b->exec) (void)hipGraphExecDestroy(b->exec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    if (b->cap) (void)hipStreamDestroy(b->cap);
    if (b->dev) (void)hipFree(b->dev);
    delete b;
}

// Builds the device tables and workspace of a plan with the tile choices ch (no graph,
// no tuning).
int build(tb_handle_s* h, int tier, const tb_batch_entry* E, int count, long nb, const Choice& ch,
          tb_batch_s* b)
{
    const bool split = tier == kSplitF16;
    b->h = h;
    b->tier = tier;
    b->count = count;
    b->unpl = planes_of(h, tier);
    b->anpl = planes_of(h, kSplitF16);      // leaves run in the split tier (as single calls do)
    b->grid = batch_grid(h->device);
    b->choice = ch;
    b->waves = (b->unpl == 2 && b->anpl == 2 && ch.waves == 8) ? 8 : 4;

    std::vector<BatchInst> insts(count);
    std::vector<BatchStep> steps;
    std::vector<Costed> items;
    std::vector<RowexpJob> rj[2];
    std::vector<TrtriJob> tj;
    // workspace offsets, fixed up to pointers after the allocation
    struct Off {
        long long eRow = -1, eSnap = -1, eL = -1, eZ = -1, Z = -1, ldiag = -1;
        long long LpU = -1, LpA = -1, Zp = -1, Ap = -1, ldp = 0, esum = 0;   // planes: halves
        long long LpU1 = -1, LpA1 = -1;
        long long Bp = -1, ldpB = 0, eCol = -1;     // TRMM: columns of B in planes
        int kshift1 = 0;
    };
    std::vector<Off> off(count);
    long long nexps = 0, nsnap = 0, nz = 0, ndiag = 0, nhalves = 0, maxk = 64;
    auto rup = [](long long v, long long q) { return (v + q - 1) / q * q; };
    auto reserve_halves = [&](long long h) {
        const long long o = nhalves;
        nhalves += rup(h, 128);        // 256-byte aligned
        return o;
    };

    std::vector<TrsmStep> sched;
    std::vector<TrsmLeaf> leaves;
    for (int i = 0; i < count; ++i) {
        const tb_batch_entry& e = E[i];
        BatchInst& I = insts[i];
        std::memset(&I, 0, sizeof(I));
        I.op = e.op == TB_OP_SYRK ? kBatchSyrk : (e.op == TB_OP_TRMM ? kBatchTrmm : kBatchTrsm);
        I.m = (int)e.m;
        I.n = (int)e.n;
        I.k = (int)e.k;
        I.alpha = (float)e.alpha;
        I.beta = (float)e.beta;
        I.A = e.A;
        I.lda = e.lda;
        I.C = e.C;
        I.ldc = e.ldc;
        b->flops += e.op == TB_OP_SYRK ? (double)e.n * e.n * e.k
                  : e.op == TB_OP_TRMM ? (double)e.m * e.m * e.n : (double)e.m * e.n * e.n;
        if (e.op == TB_OP_SYRK || e.op == TB_OP_TRMM) {
            const bool trmm = e.op == TB_OP_TRMM;
            ++(trmm ? b->ninst_trmm : b->ninst_syrk);
            const long rows = trmm ? e.m : e.n, cols = e.n, K = trmm ? e.m : e.k;
            if (rows == 0 || cols == 0) continue;
            if (split) {
                off[i].eRow = nexps;
                nexps += rows;
                if (trmm) {
                    off[i].eCol = nexps;
                    nexps += cols;
                }
            }
            off[i].ldp = rup(std::max(1l, K), 64);
            off[i].Ap = reserve_halves((long long)b->unpl * rows * off[i].ldp);
            if (trmm) {
                off[i].ldpB = off[i].ldp;
                off[i].Bp = reserve_halves((long long)b->unpl * cols * off[i].ldp);
            }
            // tiles: 128 x 128 (kShapeSq) or 64 x 256 (kShapePanel), lower ones only for
            // SYRK; a TRMM tile's k range ends at its last row
            const int shape = ch.shape[trmm ? 1 : 0][bucket_of(rows)];
            if (shape == kShapeSq) ++b->ninst_sq;
            const long bm = shape == kShapeSq ? 128 : 64, bn = shape == kShapeSq ? 128 : 256;
            for (long i0 = 0; i0 < rows; i0 += bm) {
                const long kk = trmm ? std::min(K, i0 + bm) : K;
                const long jend = trmm ? cols : std::min(cols, i0 + bm);
                for (long j0 = 0; j0 < jend; j0 += bn)
                    items.push_back({{i, (int)i0, (int)j0, shape}, (double)bm * bn * kk});
            }
            continue;
        }
        ++b->ninst_trsm;
        if (e.m == 0 || e.n == 0) continue;
        trsm_schedule(e.n, nb, &sched, &leaves);
        I.step0 = (int)steps.size();
        I.nsteps = (int)sched.size();
        // leaf blocks: W^T offsets and exponent offsets in schedule order
        std::vector<long long> zoff(leaves.size()), eoff(leaves.size());
        long long z = 0, ez = 0;
        for (size_t q = 0; q < leaves.size(); ++q) {
            zoff[q] = z;
            eoff[q] = ez;
            z += leaves[q].s * leaves[q].s;
            ez += leaves[q].s;
        }
        size_t lq = 0;
        int resid = -1;        // the one residue mod 8 of the unaligned update offsets
        for (const TrsmStep& s : sched) {
            maxk = std::max<long long>(maxk, s.type == 0 ? s.b : s.d);
            if (s.type != 0 && s.c % 8 != 0) {
                if (resid >= 0 && resid != (int)(s.c % 8)) return TB_ERR_INVALID;   // not from this planner
                resid = (int)(s.c % 8);
            }
            BatchStep st;
            std::memset(&st, 0, sizeof(st));
            st.type = s.type == 0 ? 0 : (s.acc ? 2 : 1);
            st.a = (int)s.a;
            st.b = (int)s.b;
            st.c = (int)s.c;
            st.d = (int)s.d;
            if (s.type == 0) {
                st.zoff = zoff[lq];
                st.eoff = (int)eoff[lq];
                ++lq;
            }
            steps.push_back(st);
        }
        off[i].eRow = nexps;
        nexps += e.m;
        off[i].eL = nexps;
        nexps += e.n;
        off[i].eZ = nexps;
        nexps += ez;
        off[i].eSnap = nsnap;
        nsnap += e.m;
        off[i].Z = nz;
        nz += z;
        off[i].ldiag = ndiag;
        ndiag += e.n;
        // L's rows in planes: 32 columns of slack past n (and past the shift), so that a
        // 32-k stage starting inside the row never leaves it
        off[i].kshift1 = resid > 0 ? 8 - resid : 0;
        off[i].ldp = rup(e.n + off[i].kshift1 + 32, 64);
        off[i].esum = ez;
        off[i].LpU = reserve_halves((long long)b->unpl * e.n * off[i].ldp);
        off[i].LpA = b->unpl == 1 ? reserve_halves((long long)b->anpl * e.n * off[i].ldp) : off[i].LpU;
        if (resid > 0) {
            off[i].LpU1 = reserve_halves((long long)b->unpl * e.n * off[i].ldp);
            off[i].LpA1 = b->unpl == 1 ? reserve_halves((long long)b->anpl * e.n * off[i].ldp) : off[i].LpU1;
        }
        off[i].Zp = reserve_halves((long long)b->anpl * ez * kZpLd);
        for (long r0 = 0; r0 < e.m; r0 += kPanelRows) {
            const double rows = (double)std::min<long>(kPanelRows, e.m - r0);
            items.push_back({{i, (int)r0, 0, 0}, rows * (double)e.n * (double)e.n * 0.5});
        }
    }

    // device layout
    Arena ar;
    const size_t oInst = ar.reserve(sizeof(BatchInst) * std::max(1, count));
    const size_t oStep = ar.reserve(sizeof(BatchStep) * std::max<size_t>(1, steps.size()));
    const size_t oIt = ar.reserve(sizeof(BatchItem) * std::max<size_t>(1, items.size()));
    const size_t oCnt = ar.reserve(sizeof(int));
    const size_t oExp = ar.reserve(sizeof(int) * std::max<long long>(1, nexps));
    const size_t oSnap = ar.reserve(sizeof(int) * std::max<long long>(1, nsnap));
    const size_t oZ = ar.reserve(sizeof(float) * std::max<long long>(1, nz));
    const size_t oD = ar.reserve(sizeof(float) * std::max<long long>(1, ndiag));
    // job tables are sized after the pointers are known; reserve their upper bounds
    const size_t oRj0 = ar.reserve(sizeof(RowexpJob) * (2 * (size_t)count + 1));
    const size_t oRj1 = ar.reserve(sizeof(RowexpJob) * (size_t)std::max<size_t>(1, steps.size() + 1));
    const size_t oTj = ar.reserve(sizeof(TrtriJob) * (size_t)std::max<size_t>(1, steps.size() + 1));
    const size_t oPj = ar.reserve(sizeof(PlanesJob) * (4 * (size_t)count + steps.size() + 1));
    const size_t oPl = ar.reserve(sizeof(_Float16) * (size_t)std::max<long long>(1, nhalves));
    // per-workgroup scratch: a panel's solved block in the leaf tier's planes
    const long long sldp = rup(maxk, 64);
    const long long sper = rup((long long)std::max(b->unpl, b->anpl) * kPanelRows * sldp * 2, 256);
    const size_t oScr = ar.reserve((size_t)b->grid * sper);
    if (hipMalloc(&b->dev, ar.total()) != hipSuccess) {
        (void)hipGetLastError();
        b->dev = nullptr;
        return TB_ERR_NOMEM;
    }
    b->dev_bytes = ar.total();
    char* d = b->dev;
    b->insts = reinterpret_cast<BatchInst*>(d + oInst);
    b->steps = reinterpret_cast<BatchStep*>(d + oStep);
    b->items = reinterpret_cast<BatchItem*>(d + oIt);
    b->counter = reinterpret_cast<int*>(d + oCnt);
    b->exps = reinterpret_cast<int*>(d + oExp);
    b->nexps = nexps;
    int* snap = reinterpret_cast<int*>(d + oSnap);
    float* Z = reinterpret_cast<float*>(d + oZ);
    float* ldiag = reinterpret_cast<float*>(d + oD);
    b->rj[0] = reinterpret_cast<RowexpJob*>(d + oRj0);
    b->rj[1] = reinterpret_cast<RowexpJob*>(d + oRj1);
    b->tj = reinterpret_cast<TrtriJob*>(d + oTj);
    b->pj = reinterpret_cast<PlanesJob*>(d + oPj);
    _Float16* planes = reinterpret_cast<_Float16*>(d + oPl);
    b->scratch = {d + oScr, sper, sldp, kPanelRows * sldp};
    std::vector<PlanesJob> pj;
    auto add_planes = [&](const float* X, long long ld, int mode, long rows, long K, const int* ex,
                          int npl, _Float16* out, long long ldp, long long sp, int kshift = 0) {
        if (rows <= 0) return;
        PlanesJob j;
        std::memset(&j, 0, sizeof(j));
        j.X = X;
        j.ld = ld;
        j.mode = mode;
        j.rows = (int)rows;
        j.K = (int)K;
        j.npl = npl;
        j.kshift = kshift;
        j.e = ex;
        j.out = out;
        j.ldp = ldp;
        j.sp = sp;
        j.gx = (int)(ldp / 64);
        j.gy = (int)((rows + 63) / 64);
        j.blockBegin = b->pjblocks;
        b->pjblocks += j.gx * j.gy;
        pj.push_back(j);
    };

    auto add_rowexp = [](std::vector<RowexpJob>& v, int& blocks, const float* X, long long ld,
                         int mode, long rows, long K, int* out) {
        if (rows <= 0 || K <= 0) return;
        RowexpJob j;
        std::memset(&j, 0, sizeof(j));
        j.X = X;
        j.ld = ld;
        j.rows = (int)rows;
        j.K = (int)K;
        j.kcontig = mode == kOpT;
        j.lower = mode == kOpTriL;
        j.out = out;
        j.rowBlocks = j.kcontig ? (int)((rows + 3) / 4) : (int)((rows + 255) / 256);
        j.kBlocks = j.kcontig ? (int)((K + 1023) / 1024) : (int)((K + 255) / 256);
        j.blockBegin = blocks;
        blocks += j.rowBlocks * j.kBlocks;
        v.push_back(j);
    };
    for (int i = 0; i < count; ++i) {
        const tb_batch_entry& e = E[i];
        BatchInst& I = insts[i];
        const Off& o = off[i];
        if (o.eRow >= 0) I.eRow = b->exps + o.eRow;
        if (e.op == TB_OP_TRMM) {
            if (e.m == 0 || e.n == 0) continue;
            // rows of L (lower, zero above the diagonal) and columns of B, exponents and
            // planes as the single call's fp16_product makes them
            if (o.eCol >= 0) I.eCol = b->exps + o.eCol;
            if (o.eRow >= 0) {
                add_rowexp(rj[0], b->rjblocks[0], e.A, e.lda, kOpTriL, e.m, e.m, I.eRow);
                add_rowexp(rj[0], b->rjblocks[0], e.B, e.ldb, kOpT, e.n, e.m, (int*)I.eCol);
            }
            I.Ap = planes + o.Ap;
            I.ldpA = o.ldp;
            I.spA = (long long)e.m * o.ldp;
            add_planes(e.A, e.lda, kOpTriL, e.m, e.m, split ? I.eRow : nullptr, b->unpl,
                       planes + o.Ap, o.ldp, I.spA);
            I.Bp = planes + o.Bp;
            I.ldpB = o.ldpB;
            I.spB = (long long)e.n * o.ldpB;
            add_planes(e.B, e.ldb, kOpT, e.n, e.m, split ? I.eCol : nullptr, b->unpl,
                       planes + o.Bp, o.ldpB, I.spB);
            continue;
        }
        if (e.op == TB_OP_SYRK) {
            if (o.eRow >= 0) add_rowexp(rj[0], b->rjblocks[0], e.A, e.lda, kOpN, e.n, e.k, I.eRow);
            if (e.n == 0) continue;
            // the rows of A in planes, as the single call converts them
            I.Ap = planes + o.Ap;
            I.ldpA = o.ldp;
            I.spA = (long long)e.n * o.ldp;
            add_planes(e.A, e.lda, kOpN, e.n, e.k, split ? I.eRow : nullptr, b->unpl, planes + o.Ap,
                       o.ldp, I.spA);
            continue;
        }
        if (e.m == 0 || e.n == 0) continue;
        I.eSnap = snap + o.eSnap;
        I.eL = b->exps + o.eL;
        I.eZ = b->exps + o.eZ;
        I.Z = Z + o.Z;
        I.ldiag = ldiag + o.ldiag;
        // exponents of B's rows and L's rows (split updates), as trsm_run scans them
        add_rowexp(rj[0], b->rjblocks[0], e.C, e.ldc, kOpN, e.m, e.n, I.eRow);
        add_rowexp(rj[0], b->rjblocks[0], e.A, e.lda, kOpTriL, e.n, e.n, (int*)I.eL);
        // L's rows in the update tier's planes and, if it differs, the leaf tier's
        I.ldpL = o.ldp;
        I.spLU = (long long)e.n * o.ldp;
        I.spLA = I.spLU;
        I.LpU = planes + o.LpU;
        I.LpA = planes + o.LpA;
        add_planes(e.A, e.lda, kOpTriL, e.n, e.n, split ? I.eL : nullptr, b->unpl, planes + o.LpU,
                   o.ldp, I.spLU);
        if (o.LpA != o.LpU)
            add_planes(e.A, e.lda, kOpTriL, e.n, e.n, I.eL, b->anpl, planes + o.LpA, o.ldp, I.spLA);
        I.kshift1 = o.kshift1;
        I.LpU1 = o.LpU1 >= 0 ? planes + o.LpU1 : I.LpU;
        I.LpA1 = o.LpA1 >= 0 ? planes + o.LpA1 : I.LpA;
        if (o.LpU1 >= 0) {
            add_planes(e.A, e.lda, kOpTriL, e.n, e.n, split ? I.eL : nullptr, b->unpl,
                       planes + o.LpU1, o.ldp, I.spLU, o.kshift1);
            if (o.LpA1 != o.LpU1)
                add_planes(e.A, e.lda, kOpTriL, e.n, e.n, I.eL, b->anpl, planes + o.LpA1, o.ldp,
                           I.spLA, o.kshift1);
        }
        I.Zp = planes + o.Zp;
        I.spZ = o.esum * kZpLd;
        // inverses: one job per run of equal adjacent leaves, as trsm_run launches them
        trsm_schedule(e.n, nb, &sched, &leaves);
        long long z = 0, ez = 0;
        for (size_t q = 0; q < leaves.size();) {
            size_t r = q + 1;
            while (r < leaves.size() && leaves[r].s == leaves[q].s &&
                   leaves[r].off == leaves[r - 1].off + leaves[q].s)
                ++r;
            const long s = leaves[q].s, lo = leaves[q].off;
            TrtriJob t;
            std::memset(&t, 0, sizeof(t));
            t.L = e.A + lo * (e.lda + 1);
            t.lda = e.lda;
            t.bstride = s * (e.lda + 1);
            t.s = (int)s;
            t.count = (int)(r - q);
            t.Z = Z + o.Z + z;
            t.ldiag = ldiag + o.ldiag + lo;
            t.colBlocks = (int)((s + kTrtriCols - 1) / kTrtriCols);
            t.blockBegin = b->tjblocks;
            b->tjblocks += t.colBlocks * t.count;
            tj.push_back(t);
            // exponents of the columns of the run's W^T blocks (stacked rows of Z^T), and
            // those rows in the leaf tier's planes
            add_rowexp(rj[1], b->rjblocks[1], Z + o.Z + z, s, kOpT, s * (long)(r - q), s,
                       (int*)I.eZ + ez);
            add_planes(Z + o.Z + z, s, kOpT, s * (long)(r - q), s, I.eZ + ez, b->anpl,
                       planes + o.Zp + ez * plane_row_stride(b->anpl, kZpLd), kZpLd, I.spZ);
            z += s * s * (long long)(r - q);
            ez += s * (long long)(r - q);
            q = r;
        }
    }
    b->nrj[0] = (int)rj[0].size();
    b->nrj[1] = (int)rj[1].size();
    b->ntj = (int)tj.size();
    b->npj = (int)pj.size();
    if (rj[0].size() > 2 * (size_t)count + 1 || rj[1].size() > steps.size() + 1 ||
        tj.size() > steps.size() + 1 || pj.size() > 4 * (size_t)count + steps.size() + 1)
        return TB_ERR_INVALID;   // cannot happen: each leaf run is one step at least
    if (!pj.empty())
        TB_HIP(hipMemcpy(b->pj, pj.data(), sizeof(PlanesJob) * pj.size(), hipMemcpyHostToDevice));

    sort_items(items);
    std::vector<BatchItem> flat(items.size());
    for (size_t q = 0; q < flat.size(); ++q) flat[q] = items[q].it;
    b->nitems = (int)flat.size();
    if (!flat.empty())
        TB_HIP(hipMemcpy(b->items, flat.data(), sizeof(BatchItem) * flat.size(), hipMemcpyHostToDevice));
    if (count > 0) TB_HIP(hipMemcpy(b->insts, insts.data(), sizeof(BatchInst) * count, hipMemcpyHostToDevice));
    if (!steps.empty())
        TB_HIP(hipMemcpy(b->steps, steps.data(), sizeof(BatchStep) * steps.size(), hipMemcpyHostToDevice));
    for (int v = 0; v < 2; ++v)
        if (!rj[v].empty())
            TB_HIP(hipMemcpy(b->rj[v], rj[v].data(), sizeof(RowexpJob) * rj[v].size(), hipMemcpyHostToDevice));
    if (!tj.empty()) TB_HIP(hipMemcpy(b->tj, tj.data(), sizeof(TrtriJob) * tj.size(), hipMemcpyHostToDevice));
    return TB_SUCCESS;
}

// Run time of a plan over entries E (on its handle's stream, HIP events; the faster of
// two runs after a warm-up), in microseconds, or a negative status.
double time_plan(tb_handle_s* h, int tier, const std::vector<tb_batch_entry>& E, long nb, const Choice& ch)
{
    tb_batch_s* t = new (std::nothrow) tb_batch_s();
    if (!t) return -TB_ERR_NOMEM;
    int st = build(h, tier, E.data(), (int)E.size(), nb, ch, t);
    double best = -1.0;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (st == TB_SUCCESS && (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess))
        st = TB_ERR_DEVICE;
    if (st == TB_SUCCESS && enqueue(t, h->stream) != hipSuccess) st = TB_ERR_DEVICE;
    for (int r = 0; r < 2 && st == TB_SUCCESS; ++r) {
        float ms = 0.0f;
        if (hipEventRecord(e0, h->stream) != hipSuccess || enqueue(t, h->stream) != hipSuccess ||
            hipEventRecord(e1, h->stream) != hipSuccess || hipEventSynchronize(e1) != hipSuccess ||
            hipEventElapsedTime(&ms, e0, e1) != hipSuccess)
            st = TB_ERR_DEVICE;
        else if (best < 0.0 || 1e3 * ms < best)
            best = 1e3 * ms;
    }
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    release(t);
    return st == TB_SUCCESS ? best : -(double)st;
}

// Times the tile choices on the plan's own instances, with scratch copies of their
// outputs (the caller's data is not touched): for each (op, size bucket) of SYRK and
// TRMM instances, their sub-plan on 64 x 256 against 128 x 128 tiles; for 2-plane plans
// with TRSM instances, the TRSM sub-plan on four against eight waves.  Choices without
// instances keep the table's values.  The choices never change results.
int autotune(tb_handle_s* h, int tier, const tb_batch_entry* E, int count, long nb, Choice* ch,
             double* us)
{
    std::vector<tb_batch_entry> T(E, E + count);
    std::vector<void*> scratch;
    int st = TB_SUCCESS;
    hipEvent_t a0 = nullptr, a1 = nullptr;
    if (hipEventCreate(&a0) != hipSuccess || hipEventCreate(&a1) != hipSuccess) st = TB_ERR_DEVICE;
    if (st == TB_SUCCESS && hipEventRecord(a0, h->stream) != hipSuccess) st = TB_ERR_DEVICE;
    for (int i = 0; i < count && st == TB_SUCCESS; ++i) {
        // the output is rows x cols with leading dimension ldc: only (cols-1)*ldc + rows
        // elements of it are the caller's, so it is copied as a 2D region
        const long rows = E[i].op == TB_OP_SYRK ? E[i].n : E[i].m, cols = E[i].n;
        if (!E[i].C || rows <= 0 || cols <= 0) continue;
        void* p = nullptr;
        if (hipMalloc(&p, sizeof(float) * (size_t)E[i].ldc * (size_t)cols) != hipSuccess) {
            (void)hipGetLastError();
            st = TB_ERR_NOMEM;
            break;
        }
        scratch.push_back(p);
        const size_t pitch = sizeof(float) * (size_t)E[i].ldc;
        if (hipMemcpy2D(p, pitch, E[i].C, pitch, sizeof(float) * (size_t)rows, (size_t)cols,
                        hipMemcpyDeviceToDevice) != hipSuccess)
            st = TB_ERR_DEVICE;
        T[i].C = static_cast<float*>(p);
    }
    // SYRK / TRMM tile shapes per (op, bucket)
    for (int o = 0; o < kTunedOps && st == TB_SUCCESS; ++o) {
        const int op = o == 0 ? TB_OP_SYRK : TB_OP_TRMM;
        for (int q = 0; q < kBuckets && st == TB_SUCCESS; ++q) {
            std::vector<tb_batch_entry> sub;
            for (const tb_batch_entry& e : T)
                if (e.op == op && inst_bucket(e) == q) sub.push_back(e);
            if (sub.empty()) continue;
            double t[2];
            for (int v = 0; v < 2 && st == TB_SUCCESS; ++v) {
                Choice c = *ch;
                c.shape[o][q] = v;
                t[v] = time_plan(h, tier, sub, nb, c);
                if (t[v] < 0.0) st = (int)-t[v];
            }
            if (st == TB_SUCCESS) ch->shape[o][q] = t[1] < t[0] ? kShapeSq : kShapePanel;
        }
    }
    // waves per workgroup (2-plane plans): on the TRSM instances
    if (st == TB_SUCCESS && planes_of(h, tier) == 2) {
        std::vector<tb_batch_entry> sub;
        for (const tb_batch_entry& e : T)
            if (e.op == TB_OP_TRSM) sub.push_back(e);
        if (!sub.empty()) {
            double t[2];
            for (int v = 0; v < 2 && st == TB_SUCCESS; ++v) {
                Choice c = *ch;
                c.waves = v == 0 ? 4 : 8;
                t[v] = time_plan(h, tier, sub, nb, c);
                if (t[v] < 0.0) st = (int)-t[v];
            }
            if (st == TB_SUCCESS) ch->waves = t[1] < t[0] ? 8 : 4;
        }
    }
    float ms = 0.0f;
    if (st == TB_SUCCESS && (hipEventRecord(a1, h->stream) != hipSuccess ||
                             hipEventSynchronize(a1) != hipSuccess ||
                             hipEventElapsedTime(&ms, a0, a1) != hipSuccess))
        st = TB_ERR_DEVICE;
    *us = 1e3 * ms;
    if (a0) (void)hipEventDestroy(a0);
    if (a1) (void)hipEventDestroy(a1);
    (void)hipStreamSynchronize(h->stream);
    for (void* p : scratch) (void)hipFree(p);
    return st;
}

}  // namespace

int tb_batch_create(tb_handle_t h, int tier, const tb_batch_entry* entries, int count, long nb,
                    int flags, tb_batch_t* batch)
{
    if (!h || !batch || count < 0 || (count > 0 && !entries) || nb <= 0) return TB_ERR_INVALID;
    if (tier != TB_TIER_PLAIN && tier != TB_TIER_SPLIT) return TB_ERR_INVALID;
    if (flags & ~(TB_BATCH_AUTOTUNE | TB_BATCH_NO_GRAPH | TB_BATCH_TILES_PANEL | TB_BATCH_TILES_SQUARE |
                  TB_BATCH_WAVES_4 | TB_BATCH_WAVES_8))
        return TB_ERR_INVALID;
    if ((flags & TB_BATCH_TILES_PANEL) && (flags & TB_BATCH_TILES_SQUARE)) return TB_ERR_INVALID;
    if ((flags & TB_BATCH_WAVES_4) && (flags & TB_BATCH_WAVES_8)) return TB_ERR_INVALID;
    if ((flags & TB_BATCH_WAVES_8) && planes_of(h, tier) != 2) return TB_ERR_INVALID;
    for (int i = 0; i < count; ++i) TB_TRY(check_entry(entries[i]));
    *batch = nullptr;
    tb_batch_s* b = new (std::nothrow) tb_batch_s();
    if (!b) return TB_ERR_NOMEM;
    b->flags = flags;
    Choice ch = table_choice(tier, entries, count);
    double tune_us = 0.0;
    int st = TB_SUCCESS;
    if (flags & TB_BATCH_AUTOTUNE) st = autotune(h, tier, entries, count, nb, &ch, &tune_us);
    if (flags & (TB_BATCH_TILES_PANEL | TB_BATCH_TILES_SQUARE))
        for (int o = 0; o < kTunedOps; ++o)
            for (int q = 0; q < kBuckets; ++q)
                ch.shape[o][q] = (flags & TB_BATCH_TILES_SQUARE) ? kShapeSq : kShapePanel;
    if (flags & TB_BATCH_WAVES_4) ch.waves = 4;
    if (flags & TB_BATCH_WAVES_8) ch.waves = 8;
    if (st == TB_SUCCESS) st = build(h, tier, entries, count, nb, ch, b);
    b->tune_us = tune_us;
    b->tuned = (flags & TB_BATCH_AUTOTUNE) != 0;
    if (st == TB_SUCCESS && !(flags & TB_BATCH_NO_GRAPH)) {
        if (hipStreamCreateWithFlags(&b->cap, hipStreamNonBlocking) != hipSuccess ||
            hipStreamBeginCapture(b->cap, hipStreamCaptureModeThreadLocal) != hipSuccess) {
            st = TB_ERR_DEVICE;
        } else {
            const hipError_t eq = enqueue(b, b->cap);
            const hipError_t ec = hipStreamEndCapture(b->cap, &b->graph);
            if (eq != hipSuccess || ec != hipSuccess ||
                hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0) != hipSuccess)
                st = TB_ERR_DEVICE;
        }
    }
    if (st != TB_SUCCESS) {
        (void)hipGetLastError();
        release(b);
        return st;
    }
    *batch = b;
    return TB_SUCCESS;
}

int tb_batch_run(tb_batch_t b)
{
    if (!b) return TB_ERR_INVALID;
    hipEvent_t e0;
    TB_TRY(prof_begin(b->h, &e0));
    if (b->exec) TB_HIP(hipGraphLaunch(b->exec, b->h->stream));
    else TB_HIP(enqueue(b, b->h->stream));
    return prof_end(b->h, kProfBatch, e0, b->flops);
}

int tb_batch_info(tb_batch_t b, long* out, int n)
{
    if (!b || n < 0 || (n > 0 && !out)) return TB_ERR_INVALID;
    const long v[6] = {b->waves, b->nitems, b->count, b->ninst_sq, (long)b->tune_us,
                       (long)b->dev_bytes};
    for (int i = 0; i < n && i < 6; ++i) out[i] = v[i];
    return TB_SUCCESS;
}

int tb_batch_tiles(tb_batch_t b, int* out, int n)
{
    if (!b || n < 0 || (n > 0 && !out)) return TB_ERR_INVALID;
    int v[2 * kBuckets + 1];
    for (int o = 0; o < kTunedOps; ++o)
        for (int q = 0; q < kBuckets; ++q) v[o * kBuckets + q] = b->choice.shape[o][q];
    v[2 * kBuckets] = b->waves;
    for (int i = 0; i < n && i < 2 * kBuckets + 1; ++i) out[i] = v[i];
    return TB_SUCCESS;
}

int tb_batch_destroy(tb_batch_t b)
{
    if (!b) return TB_ERR_INVALID;
    release(b);
    return TB_SUCCESS;
}
