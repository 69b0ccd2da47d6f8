// Request loop (tb_serve): a stream of mixed BLAS-3 requests of every tier, admitted
// as they arrive onto a set of slots -- each slot a handle with its own HIP stream
// and workspace -- within an HBM budget.  Requests run concurrently on the device;
// completions are detected by polling events, and each request's times come from HIP
// events against one start event, so they are device times.
//
// Admission: two first-come first-served lanes.  A request whose estimated time alone
// (flops / per-tier rate, below) exceeds cfg->large_us is large; at most
// cfg->max_large large requests run at once, so the longest requests do not share the
// device with each other, and small requests never queue behind a large one.  Within
// a lane the oldest waiting request is admitted as soon as a slot is free and the
// bytes it holds fit the budget; later arrivals in its lane never overtake it.  Work
// submitted on one slot stream never waits for another slot.
#include <algorithm>
#include <chrono>
#include <thread>
#include <vector>

#include "tb_internal.hpp"

namespace {

using Clock = std::chrono::steady_clock;

size_t elem_bytes(int tier) { return (tier == TB_TIER_OZAKI || tier == TB_TIER_FP64) ? 8 : 4; }

// Operand elements of a request (what it keeps resident while admitted).
double operand_elems(const tb_request& r)
{
    const double m = (double)r.m, n = (double)r.n, k = (double)r.k;
    switch (r.op) {
    case TB_REQ_SYRK: return n * k + n * n;
    case TB_REQ_SYR2K: return 2 * n * k + n * n;
    case TB_REQ_TRMM: return m * m + 2 * m * n;
    case TB_REQ_TRSM: return n * n + m * n;
    case TB_REQ_SYMM: return m * m + 2 * m * n;
    default: return 0;
    }
}

// Bytes a request holds while admitted when the caller gives none: its operands plus
// the library's workspace for them (FP16 planes: at most 3 halves per FP32 element;
// Ozaki: up to 10 INT8 slices per FP64 element and FP64 copies in TRSM refinement),
// bounded above by twice the operands.
size_t default_bytes(const tb_request& r) { return (size_t)(2.0 * operand_elems(r) * elem_bytes(r.tier)); }

// Algorithmic flops of a request and its estimated time alone: a fixed launch cost
// plus flops at a per-tier rate (MI355X, large sizes; small sizes run below these
// rates, which the fixed term covers in part).
double req_flops(const tb_request& r)
{
    const double m = (double)r.m, n = (double)r.n, k = (double)r.k;
    switch (r.op) {
    case TB_REQ_SYRK: return n * n * k;
    case TB_REQ_SYR2K: return 2 * n * n * k;
    case TB_REQ_TRMM: return m * m * n;
    case TB_REQ_TRSM: return m * n * n;
    case TB_REQ_SYMM: return 2 * m * m * n;
    default: return 0;
    }
}

double est_us(const tb_request& r)
{
    static const double tflops[5] = {250.0, 150.0, 28.0, 70.0, 40.0};   // by tier
    const int t = r.tier >= 0 && r.tier < 5 ? r.tier : 0;
    return 50.0 + req_flops(r) / (tflops[t] * 1e6);
}

int launch(tb_handle_t h, const tb_request& r)
{
    switch (r.op) {
    case TB_REQ_SYRK:
        return tb_syrk(h, r.tier, r.n, r.k, r.alpha, r.A, r.lda, r.beta, r.C, r.ldc, r.nb);
    case TB_REQ_SYR2K:
        return tb_syr2k(h, r.tier, r.n, r.k, r.alpha, r.A, r.lda, r.B, r.ldb, r.beta, r.C, r.ldc, r.nb);
    case TB_REQ_TRMM:
        return tb_trmm(h, r.tier, r.m, r.n, r.alpha, r.A, r.lda, r.B, r.ldb, r.C, r.ldc, r.nb);
    case TB_REQ_TRSM:
        return tb_trsm(h, r.tier, r.m, r.n, r.A, r.lda, r.C, r.ldc, r.nb);
    case TB_REQ_SYMM:
        return tb_symm(h, r.tier, r.m, r.n, r.alpha, r.A, r.lda, r.B, r.ldb, r.beta, r.C, r.ldc);
    default:
        return TB_ERR_INVALID;
    }
}

struct Running {
    int req;
    int slot;
    size_t bytes;
    bool large;
};

}  // namespace

int tb_reserve_workspace(tb_handle_t h, long bytes, long ozaki_bytes)
{
    if (!h || bytes < 0 || ozaki_bytes < 0) return TB_ERR_INVALID;
    char* p;
    if (bytes > 0) TB_TRY(tb::workspace_get(h, (size_t)bytes, &p, 0));
    if (ozaki_bytes > 0) TB_TRY(tb::workspace_get(h, (size_t)ozaki_bytes, &p, 1));
    return TB_SUCCESS;
}

int tb_workspace_sizes(tb_handle_t h, long* out)
{
    if (!h || !out) return TB_ERR_INVALID;
    out[0] = (long)h->ws_bytes[0];
    out[1] = (long)h->ws_bytes[1];
    return TB_SUCCESS;
}

int tb_serve(tb_handle_t* slots, int nslots, const tb_serve_config* cfg, const tb_request* reqs,
             int count, tb_request_result* out)
{
    if (!slots || nslots <= 0 || !cfg || cfg->hbm_budget <= 0 || cfg->large_us < 0 ||
        cfg->max_large < 1 || count < 0 || (count > 0 && (!reqs || !out)))
        return TB_ERR_INVALID;
    const size_t budget = (size_t)cfg->hbm_budget;
    for (int s = 0; s < nslots; ++s)
        if (!slots[s]) return TB_ERR_INVALID;
    for (int i = 0; i < count; ++i) {
        if (reqs[i].arrival_us < 0 || reqs[i].op < TB_REQ_SYRK || reqs[i].op > TB_REQ_SYMM)
            return TB_ERR_INVALID;
        out[i].admit_us = out[i].done_us = -1.0;
        out[i].slot = -1;
        out[i].status = TB_SUCCESS;
        out[i].est_us = est_us(reqs[i]);
    }
    if (count == 0) return TB_SUCCESS;

    // the two lanes, each in arrival order
    std::vector<int> lane[2];
    std::vector<size_t> need(count);
    for (int i = 0; i < count; ++i) {
        need[i] = reqs[i].bytes > 0 ? (size_t)reqs[i].bytes : default_bytes(reqs[i]);
        lane[cfg->large_us > 0 && out[i].est_us > cfg->large_us ? 1 : 0].push_back(i);
    }
    for (auto& l : lane)
        std::stable_sort(l.begin(), l.end(),
                         [&](int a, int b) { return reqs[a].arrival_us < reqs[b].arrival_us; });

    // one completion event per slot, one start event
    std::vector<hipEvent_t> done(nslots, nullptr);
    hipEvent_t t0 = nullptr;
    int st = TB_SUCCESS;
    for (int s = 0; s < nslots && st == TB_SUCCESS; ++s)
        if (hipEventCreate(&done[s]) != hipSuccess) st = TB_ERR_DEVICE;
    if (st == TB_SUCCESS && hipEventCreate(&t0) != hipSuccess) st = TB_ERR_DEVICE;
    // every slot is idle at t0: its stream has finished all earlier work
    for (int s = 0; s < nslots && st == TB_SUCCESS; ++s)
        if (hipStreamSynchronize(slots[s]->stream) != hipSuccess) st = TB_ERR_DEVICE;
    if (st == TB_SUCCESS && (hipEventRecord(t0, slots[0]->stream) != hipSuccess ||
                             hipEventSynchronize(t0) != hipSuccess))
        st = TB_ERR_DEVICE;
    const Clock::time_point host0 = Clock::now();

    std::vector<int> idle;
    for (int s = nslots - 1; s >= 0; --s) idle.push_back(s);
    std::vector<Running> run;
    size_t used = 0;
    int large_running = 0, finished = 0;
    size_t next[2] = {0, 0};
    while (st == TB_SUCCESS && finished < count) {
        bool progress = false;
        // completions
        for (size_t q = 0; q < run.size();) {
            const hipError_t e = hipEventQuery(done[run[q].slot]);
            if (e == hipErrorNotReady) {
                ++q;
                continue;
            }
            float ms = 0.0f;
            if (e != hipSuccess || hipEventElapsedTime(&ms, t0, done[run[q].slot]) != hipSuccess) {
                st = TB_ERR_DEVICE;
                break;
            }
            out[run[q].req].done_us = 1e3 * (double)ms;
            used -= run[q].bytes;
            large_running -= run[q].large ? 1 : 0;
            idle.push_back(run[q].slot);
            run[q] = run.back();
            run.pop_back();
            ++finished;
            progress = true;
        }
        // admissions, oldest first within each lane; the small lane first
        const double now_us =
            std::chrono::duration<double, std::micro>(Clock::now() - host0).count();
        for (int L = 0; L < 2 && st == TB_SUCCESS; ++L) {
            while (next[L] < lane[L].size()) {
                const int r = lane[L][next[L]];
                if (reqs[r].arrival_us > now_us) break;
                if (need[r] > budget) {                       // can never fit: refuse it
                    out[r].status = TB_ERR_NOMEM;
                    out[r].admit_us = now_us;
                    ++next[L];
                    ++finished;
                    progress = true;
                    continue;
                }
                if (idle.empty() || used + need[r] > budget) break;
                if (L == 1 && large_running >= cfg->max_large) break;
                const int s = idle.back();
                idle.pop_back();
                out[r].admit_us = now_us;
                out[r].slot = s;
                out[r].status = launch(slots[s], reqs[r]);
                if (hipEventRecord(done[s], slots[s]->stream) != hipSuccess) st = TB_ERR_DEVICE;
                used += need[r];
                large_running += L;
                run.push_back({r, s, need[r], L == 1});
                ++next[L];
                progress = true;
            }
        }
        if (!progress) std::this_thread::yield();
    }
    // drain whatever is still running after an error
    for (int s = 0; s < nslots; ++s) (void)hipStreamSynchronize(slots[s]->stream);
    for (hipEvent_t e : done)
        if (e) (void)hipEventDestroy(e);
    if (t0) (void)hipEventDestroy(t0);
    return st;
}
