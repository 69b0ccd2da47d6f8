// Request loop (tb_serve): a stream of mixed BLAS-3 requests of every tier, admitted
// as they arrive onto a set of slots -- each slot a handle with its own HIP stream
// and workspace -- within an HBM budget.  Requests run concurrently on the device;
// completions are detected by polling events, and each request's times come from HIP
// events against one start event, so they are device times.
//
// Admission: three lanes by the estimated time alone (a per-routine, per-tier model,
// below).  A request above cfg->large_us is large; at most cfg->max_large large
// requests run at once, so the longest requests do not share the device with each
// other.  A request below cfg->small_us is small: cfg->small_slots slots are kept for
// small requests, so at most nslots - small_slots of the others run at once and a
// small request finds a slot without waiting for a long one to end.  The lanes are
// served small first, then the rest, then large; no request waits behind one of a
// lane served after its own.  Within a lane the next request -- the oldest waiting one
// (TB_SERVE_FIFO), the one with the shortest estimate (TB_SERVE_SHORTEST), or the
// shortest in the small and middle lanes and the oldest in the large lane
// (TB_SERVE_SHORTEST_SMALL: a long request cannot be passed over indefinitely by
// shorter ones that keep arriving) -- is admitted as soon as a slot is free and the
// bytes it holds fit the budget; no other request of its lane overtakes it.  Work submitted on one slot stream never waits for
// another slot.
//
// A kernel whose workgroups cannot all be resident keeps the CUs it is given as they
// free up, so kernels from other streams start only once it has dispatched its last
// workgroup: a small request admitted next to a large one would wait for most of it.
// With cfg->reserve_cus > 0 the requests of the upper lanes -- the large lane; with
// slots kept for a small lane, every request outside it -- leave that many CUs to the
// lower ones:
//   partition 0: their GEMMs are capped (tb_set_max_cus) at (CUs - reserve_cus) / (how
//     many of them may run at once) CUs; the workgroups then all fit and loop over the
//     tiles;
//   partition 1: each runs on a stream created with a CU mask without the reserved
//     CUs (hipExtStreamCreateWithCUMask), so none of its kernels -- GEMMs, conversions,
//     slicing, inversions -- ever occupies them.  One such stream per request that may
//     run at once, swapped in as the slot's stream for that request; each needs a
//     hardware queue of its own (GPU_MAX_HW_QUEUES >= slots + that many).
// partition 2 turns the reservation around: the lower lanes run on masked streams (one
// per slot; GPU_MAX_HW_QUEUES >= 2 slots) without the reserved CUs, which stay free for
// the upper ones.  Latency targets are absolute, so the longest requests have the
// least slack: this keeps small requests from slowing them down.
#include <algorithm>
#include <chrono>
#include <cstdint>
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

// Algorithmic flops of a request and its estimated time alone.
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

// Time alone on MI355X, t = a + b n + flops / c, per routine and tier (plain, split,
// Ozaki): a covers launches and conversions, b the dependent chain of a recursion
// (TRSM's n/nb leaves), c the rate of its products.  Fitted (non-negative least
// squares on relative error) to the isolated times of a stream of 500 requests, every
// routine and tier, n log-uniform in 128 .. 8192 (k = n, m = n); median error 2-7 %,
// TRSM 11-16 %.
// The FP32 and FP64 comparator tiers use a fixed cost and a flat rate.
struct TimeModel {
    double a_us, b_us_per_n, c_tflops;
};
double est_us(const tb_request& r)
{
    static const TimeModel model[5][3] = {
        /* SYRK  */ {{27.8, 0.0062, 716.0}, {39.5, 0.0141, 296.5}, {45.8, 0.0279, 72.2}},
        /* SYR2K */ {{30.2, 0.0130, 909.3}, {51.8, 0.0243, 276.3}, {83.2, 0.0422, 71.3}},
        /* TRMM  */ {{29.5, 0.0099, 406.5}, {45.9, 0.0181, 172.8}, {48.4, 0.0267, 41.6}},
        /* TRSM  */ {{70.4, 0.3406, 1e9}, {61.0, 0.3890, 390.2}, {8.1, 0.7155, 158.0}},
        /* SYMM  */ {{32.3, 0.0086, 702.5}, {50.5, 0.0153, 299.8}, {55.7, 0.0152, 78.8}},
    };
    const double f = req_flops(r);
    if (r.tier == TB_TIER_FP32) return 50.0 + f / (70.0 * 1e6);
    if (r.tier == TB_TIER_FP64) return 50.0 + f / (40.0 * 1e6);
    const int t = r.tier == TB_TIER_PLAIN ? 0 : (r.tier == TB_TIER_SPLIT ? 1 : 2);
    const TimeModel& m = model[r.op][t];
    const double n = (double)(r.op == TB_REQ_TRMM || r.op == TB_REQ_SYMM ? r.m : r.n);
    return m.a_us + m.b_us_per_n * n + f / (m.c_tflops * 1e6);
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
    int lane;         // 0 small, 1 the others, 2 large
    int mstream;      // index of the masked stream the request runs on, or -1
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
        cfg->max_large < 1 || cfg->order < TB_SERVE_FIFO || cfg->order > TB_SERVE_SHORTEST_SMALL ||
        cfg->reserve_cus < 0 || cfg->partition < 0 || cfg->partition > 2 || cfg->small_us < 0 ||
        cfg->small_slots < 0 || cfg->small_slots >= nslots || count < 0 ||
        (count > 0 && (!reqs || !out)))
        return TB_ERR_INVALID;
    // lanes 0 small, 1 the others, 2 large; requests outside the small lane that may
    // run at once; the first upper lane (the lanes that leave reserved CUs free) and
    // how many upper-lane requests may run at once
    const int max_regular = nslots - (cfg->small_us > 0 ? cfg->small_slots : 0);
    const int upper = cfg->small_us > 0 && cfg->small_slots > 0 ? 1 : 2;
    const int max_upper = upper == 1 ? max_regular : cfg->max_large;
    int upper_cus = 0;                  // CU cap of an upper-lane request's GEMMs (0: none)
    int cus = 0;
    if (cfg->reserve_cus > 0) {
        if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, slots[0]->device) !=
            hipSuccess)
            return TB_ERR_DEVICE;
        if (cfg->reserve_cus >= cus) return TB_ERR_INVALID;
        if (cfg->partition == 0) upper_cus = std::max(1, (cus - cfg->reserve_cus) / max_upper);
    }
    // which lanes run on masked streams
    const bool mask_on = cfg->reserve_cus > 0 && cfg->partition > 0;
    auto masked = [&](int L) { return mask_on && (cfg->partition == 1 ? L >= upper : L < upper); };
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

    // the three lanes, each in arrival order
    std::vector<int> lane[3];
    std::vector<size_t> need(count);
    for (int i = 0; i < count; ++i) {
        need[i] = reqs[i].bytes > 0 ? (size_t)reqs[i].bytes : default_bytes(reqs[i]);
        const double e = out[i].est_us;
        lane[cfg->large_us > 0 && e > cfg->large_us ? 2 : (cfg->small_us > 0 && e < cfg->small_us ? 0 : 1)]
            .push_back(i);
    }
    for (auto& l : lane)
        std::stable_sort(l.begin(), l.end(),
                         [&](int a, int b) { return reqs[a].arrival_us < reqs[b].arrival_us; });

    // one completion event per slot, one start event
    std::vector<hipEvent_t> done(nslots, nullptr);
    hipEvent_t t0 = nullptr;
    int st = TB_SUCCESS;
    // masked streams (CUs reserve_cus .. cus-1): one per upper-lane request that may
    // run at once (partition 1), or one per slot (partition 2)
    std::vector<hipStream_t> mstream;
    std::vector<int> mfree;
    std::vector<hipStream_t> own(nslots);
    for (int s = 0; s < nslots; ++s) own[s] = slots[s]->stream;
    if (mask_on) {
        std::vector<uint32_t> mask((size_t)(cus + 31) / 32, 0u);
        for (int c = cfg->reserve_cus; c < cus; ++c) mask[(size_t)c / 32] |= 1u << (c % 32);
        const int nm = cfg->partition == 1 ? max_upper : nslots;
        for (int q = 0; q < nm && st == TB_SUCCESS; ++q) {
            hipStream_t m = nullptr;
            if (hipExtStreamCreateWithCUMask(&m, (uint32_t)mask.size(), mask.data()) != hipSuccess) {
                (void)hipGetLastError();
                st = TB_ERR_DEVICE;
                break;
            }
            mstream.push_back(m);
            mfree.push_back(q);
        }
    }
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
    int large_running = 0, regular_running = 0, finished = 0;
    size_t next[3] = {0, 0, 0};        // next request of each lane still to arrive
    std::vector<int> waiting[3];       // arrived, not admitted, in arrival order
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
            large_running -= run[q].lane == 2 ? 1 : 0;
            regular_running -= run[q].lane > 0 ? 1 : 0;
            if (run[q].mstream >= 0) {               // the slot's own stream again
                slots[run[q].slot]->stream = own[run[q].slot];
                mfree.push_back(run[q].mstream);
            }
            idle.push_back(run[q].slot);
            run[q] = run.back();
            run.pop_back();
            ++finished;
            progress = true;
        }
        // admissions, lane by lane: small, the others, large
        const double now_us =
            std::chrono::duration<double, std::micro>(Clock::now() - host0).count();
        for (int L = 0; L < 3 && st == TB_SUCCESS; ++L) {
            while (next[L] < lane[L].size()) {
                const int r = lane[L][next[L]];
                if (reqs[r].arrival_us > now_us) break;
                ++next[L];
                progress = true;
                if (need[r] > budget) {                       // can never fit: refuse it
                    out[r].status = TB_ERR_NOMEM;
                    out[r].admit_us = now_us;
                    ++finished;
                    continue;
                }
                waiting[L].push_back(r);
            }
            while (!waiting[L].empty()) {
                size_t w = 0;
                if (cfg->order == TB_SERVE_SHORTEST || (cfg->order == TB_SERVE_SHORTEST_SMALL && L < 2))
                    for (size_t q = 1; q < waiting[L].size(); ++q)
                        if (out[waiting[L][q]].est_us < out[waiting[L][w]].est_us) w = q;
                const int r = waiting[L][w];
                if (idle.empty() || used + need[r] > budget) break;
                if (L > 0 && regular_running >= max_regular) break;
                if (L == 2 && large_running >= cfg->max_large) break;
                const int s = idle.back();
                idle.pop_back();
                out[r].admit_us = now_us;
                out[r].slot = s;
                const int keep_cus = slots[s]->max_cus;
                if (L >= upper && upper_cus > 0) slots[s]->max_cus = upper_cus;
                int ms = -1;
                if (masked(L)) {               // a stream per request the lane may run
                    ms = mfree.back();
                    mfree.pop_back();
                    slots[s]->stream = mstream[(size_t)ms];
                }
                out[r].status = launch(slots[s], reqs[r]);
                slots[s]->max_cus = keep_cus;
                if (hipEventRecord(done[s], slots[s]->stream) != hipSuccess) st = TB_ERR_DEVICE;
                used += need[r];
                large_running += L == 2 ? 1 : 0;
                regular_running += L > 0 ? 1 : 0;
                run.push_back({r, s, need[r], L, ms});
                waiting[L].erase(waiting[L].begin() + (long)w);
                progress = true;
            }
        }
        if (!progress) std::this_thread::yield();
    }
    // drain whatever is still running after an error; every slot gets its own stream back
    for (int s = 0; s < nslots; ++s) (void)hipStreamSynchronize(slots[s]->stream);
    for (int s = 0; s < nslots; ++s) slots[s]->stream = own[s];
    for (hipStream_t m : mstream) (void)hipStreamDestroy(m);
    for (hipEvent_t e : done)
        if (e) (void)hipEventDestroy(e);
    if (t0) (void)hipEventDestroy(t0);
    return st;
}
