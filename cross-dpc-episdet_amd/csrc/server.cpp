// server.cpp -- continuous-batching request loop behind the C ABI.
//
// No reference counterpart: the reference scans one generated dataset per
// process (GPU/v4.cpp:198-624).  Here a worker thread owns the context's
// stream and runs scheduling rounds:
//   1. one round: every active request gets a max-min fair share of a work
//      budget (sized from the measured throughput to take about round_ms of
//      device time) as a contiguous range of its remaining tile rows; the ranges
//      of all requests become the segments of one persistent scan launch,
//      replayed from a captured HIP graph (queue reset, scan, merge),
//      followed on the stream by the rescoring of the requests whose last
//      rows it scans and the copy of their results to pinned memory;
//   2. admission, while that round runs on the device: queued requests are
//      packed into HBM (cohort planes and pair margins on a stream of the
//      server's own, then the resident row table) in FIFO order while the
//      device footprint stays within the budget and fewer than max_active
//      run; they join the next round;
//   3. after the round (one wait), the finished requests complete and
//      their device memory goes back to the context's pool.
// Every triplet of a request is scored exactly once whatever the round
// split, and each round folds the new candidates into the request's exact
// running top-64, so a result is identical to epi_scan of the same data.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <memory>
#include <mutex>
#include <thread>
#include <unordered_map>
#include <vector>

#include "../../include/episdet.h"
#include "episdet_internal.h"
#include "episdet_lib.h"
#include "server_poll.h"

namespace epi {

using Clock = std::chrono::steady_clock;

static double ms_between(Clock::time_point a, Clock::time_point b)
{
    return std::chrono::duration<double, std::milli>(b - a).count();
}

struct Request {
    uint64_t id = 0;
    std::vector<uint8_t> geno, pheno;
    uint32_t N = 0, M = 0;
    uint64_t footprint = 0;
    Clock::time_point submitted, admitted;
    bool cancel = false;
    bool finishing = false;  // its last rows are in the launched round
    bool done = false;
    bool polling = false;  // a thread waits in epi_poll for it
    bool charged = false;  // footprint counted in bytes_in_use
    int status = EPI_OK;
    // resident: `co` is the caller's cohort (epi_submit_cohort), counted in
    // its server_refs; otherwise the server creates and frees it
    bool resident = false;
    epi_cohort *co = nullptr;
    // resident: the row table the rounds read, a copy of the request's own
    // (the cohort's cached one is rebuilt when the caller scans it with
    // another tile shape)
    DevBuf rows;
    Job job;
    std::vector<double> cum;  // work units of the job's rows [0, r), r = 0..rows
    epi_result result;
};

// Work units of a row: (items of the row) x (64-bit words of both classes);
// the scan kernel's time per item is proportional to the words it loops over.
static double row_units(const Row &r, uint32_t nw, const epi_cohort *co)
{
    const double items = (double)((r.icount + nw - 1) / nw);
    return items * (double)(co->W[0] + co->W[1] + 2);
}

}  // namespace epi

using namespace epi;

struct epi_server {
    epi_ctx *ctx = nullptr;
    epi_server_config cfg;
    ScanConfig scfg;
    LaunchPlan plan;
    std::thread worker;
    std::mutex m;
    std::condition_variable cv_work, cv_done;
    bool stop = false;
    uint64_t next_id = 1;
    std::deque<Request *> pending;
    std::vector<Request *> active;
    std::unordered_map<uint64_t, std::unique_ptr<Request>> all;
    uint32_t pollers = 0;  // threads inside epi_poll
    epi_server_stats stats;
    double units_per_ms = 1e5;  // throughput estimate, refined by every round
    hipGraphExec_t graph = nullptr;
    std::vector<uint64_t> graph_sig;
    hipStream_t upload = nullptr;  // cohort creation, overlapping the rounds
    // requests whose last rows the launched round scans: their rescoring
    // and the copy of its results into `fin_out` follow the round on the
    // stream, and `fin_ev` marks the end of both
    std::vector<Request *> fin;
    HostBuf fin_out;
    hipEvent_t fin_ev = nullptr;
    // the launched round's overflow flag (a candidate that found no slot),
    // copied to pinned memory after the round
    HostBuf ovf;

    void run();
    int round_launch(std::vector<Request *> &act, double &units);
    int round_wait(double units);
    int admit_locked_ctx(Request *r);
    void release_locked_ctx(Request *r);
    void complete(Request *r, int status);
    void fail_all_locked_ctx(std::vector<Request *> &admit, int status,
                             std::vector<Request *> &failed);
};

void epi_server::complete(Request *r, int status)
{
    std::lock_guard<std::mutex> lk(m);
    r->status = status;
    r->done = true;
    r->geno.clear();
    r->geno.shrink_to_fit();
    r->pheno.clear();
    r->pheno.shrink_to_fit();
    if (status == EPI_OK)
        stats.completed++;
    else if (status == EPI_ECANCELED)
        stats.cancelled++;
    else
        stats.failed++;
    if (r->charged) {
        stats.bytes_in_use -= std::min(stats.bytes_in_use, r->footprint);
        r->charged = false;
    }
    cv_done.notify_all();
}

// ctx->mu held.
int epi_server::admit_locked_ctx(Request *r)
{
    int st = EPI_OK;
    if (!r->resident)
        // no seed: its sample's scoring waits for the round in flight (the
        // persistent scan kernel holds every CU), and the next round waits
        // for the admissions (serving capacity 1.30e14 without, 1.18e14
        // with); resident cohorts (epi_submit_cohort) keep theirs
        st = cohort_create_locked(ctx, r->geno.data(), nullptr, r->pheno.data(), r->N, r->M,
                                  &r->co, upload, true, false);
    if (st == EPI_OK)
        st = job_init(ctx, r->job, r->co, scfg, nullptr);
    if (st == EPI_OK && r->resident) {
        const size_t bytes = sizeof(Row) * std::max<size_t>(r->job.rows.size(), 1);
        r->rows.pool = &ctx->pool;
        st = r->rows.reserve(bytes);
        if (st == EPI_OK)
            st = stage_upload(ctx, r->rows.ptr, r->job.rows.data(),
                              sizeof(Row) * r->job.rows.size());
        r->job.drows = (const Row *)r->rows.ptr;
    }
    if (st == EPI_OK) {
        const Job &J = r->job;
        r->cum.assign(J.rows.size() + 1, 0.0);
        for (size_t q = 0; q < J.rows.size(); ++q)
            r->cum[q + 1] = r->cum[q] + row_units(J.rows[q], J.nw, J.co);
    }
    r->admitted = Clock::now();
    return st;
}

// ctx->mu held; the stream has no work left that reads the cohort.
void epi_server::release_locked_ctx(Request *r)
{
    if (r->co) {
        if (r->resident)
            r->co->server_refs--;
        else
            delete r->co;
        r->co = nullptr;
    }
    r->rows.release();
    r->job.drows = nullptr;
    r->job.list = nullptr;
    r->job.own_list.release();
    r->job.rows.clear();
    r->job.rows.shrink_to_fit();
    r->cum.clear();
    r->cum.shrink_to_fit();
}

// Launches one scheduling round over `act` (ctx->mu held), followed by the
// rescoring of the requests it finishes; `units` is its work estimate.
// Fraction of a round's budget reserved for the oldest active request
// (measured on the benchmark's stream: 0.9 gave a p99 a few ms above the
// one-at-a-time FIFO queue's, 1.0 about 5 ms below it, at a higher median).
constexpr double kOldestShare = 1.0;

int epi_server::round_launch(std::vector<Request *> &act, double &units)
{
    const uint32_t cap = cfg.max_active;
    const uint32_t n = (uint32_t)act.size();
    // Shares of the round budget.  The oldest request (act is in admission
    // order) first gets up to kOldestShare of it: a large request then runs
    // at the speed it would have alone, which bounds the tail latency
    // against a one-at-a-time FIFO queue.  What it leaves (all of a round
    // once it needs less) is split max-min fair among the others: a request
    // that needs less than an equal share gets what it needs and what it
    // leaves goes to the others, so rounds stay full while small requests
    // finish in their first round.  At least one row each.
    const double budget = std::max(units_per_ms, 1.0) * cfg.round_ms;
    std::vector<double> need(n, 0.0), share(n, 0.0);
    for (uint32_t b = 0; b < n; ++b)
        need[b] = act[b]->cum.back() - act[b]->cum[act[b]->job.next_row];
    double left = budget;
    if (n > 0) {
        share[0] = std::min(need[0], kOldestShare * budget);
        need[0] -= share[0];
        left -= share[0];
    }
    std::vector<uint32_t> by_need(n);
    for (uint32_t b = 0; b < n; ++b)
        by_need[b] = b;
    std::sort(by_need.begin(), by_need.end(),
              [&](uint32_t a, uint32_t c) { return need[a] < need[c]; });
    for (uint32_t q = 0; q < n; ++q) {
        const uint32_t b = by_need[q];
        const double s = std::min(need[b], left / (double)(n - q));
        share[b] += s;
        left -= s;
    }
    std::vector<Job *> jp(n);
    std::vector<std::pair<size_t, size_t>> ranges(n);
    units = 0.0;
    for (uint32_t b = 0; b < n; ++b) {
        Job &J = act[b]->job;
        const std::vector<double> &cum = act[b]->cum;
        jp[b] = &J;
        // the most rows whose units fit the share, at least one
        const size_t r0 = J.next_row;
        size_t r = r0;
        if (r0 < J.rows.size()) {
            r = (size_t)(std::upper_bound(cum.begin() + (r0 + 1), cum.end(), cum[r0] + share[b]) -
                         cum.begin()) - 1;
            r = std::max(r, r0 + 1);
        }
        ranges[b] = std::make_pair(r0, r);
        units += cum[r] - cum[r0];
    }
    std::vector<unsigned char> host;
    RoundLayout L;
    uint64_t items = 0;
    EPI_TRY(build_round(jp.data(), ranges.data(), n, cap, host, L, items));
    EPI_TRY(reserve_round(ctx, plan, cap, L.total));
    EPI_TRY(stage_upload(ctx, ctx->round.ptr, host.data(), L.total));
    std::vector<uint64_t> sig = {(uint64_t)cap, (uint64_t)plan.grid, (uint64_t)plan.cap, L.total,
                                 (uint64_t)(uintptr_t)ctx->round.ptr,
                                 (uint64_t)(uintptr_t)ctx->counters.ptr,
                                 (uint64_t)(uintptr_t)ctx->slots.ptr,
                                 (uint64_t)(uintptr_t)ctx->t32.ptr};
    if (!graph || sig != graph_sig) {
        if (graph) {
            (void)hipGraphExecDestroy(graph);
            graph = nullptr;
        }
        hipGraph_t g = nullptr;
        EPI_HIP(hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
        int st = enqueue_round_kernels(ctx, plan, cap, L, ctx_round_bufs(ctx), true, false,
                                       nullptr, nullptr);
        hipError_t ce = hipStreamEndCapture(ctx->stream, &g);
        if (st != EPI_OK) {
            if (g)
                (void)hipGraphDestroy(g);
            return st;
        }
        EPI_HIP(ce);
        hipError_t ie = hipGraphInstantiate(&graph, g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        EPI_HIP(ie);
        graph_sig = sig;
    }
    EPI_HIP(hipEventRecord(ctx->ev0, ctx->stream));
    EPI_HIP(hipGraphLaunch(graph, ctx->stream));
    EPI_HIP(hipEventRecord(ctx->ev1, ctx->stream));
    EPI_TRY(ovf.reserve(sizeof(uint32_t)));
    EPI_HIP(hipMemcpyAsync(ovf.ptr, (const uint32_t *)ctx->counters.ptr + 1 + cap,
                           sizeof(uint32_t), hipMemcpyDeviceToHost, ctx->stream));
    fin.clear();
    std::vector<Job *> fj;
    for (uint32_t b = 0; b < n; ++b)
        if (act[b]->job.next_row >= act[b]->job.rows.size()) {
            act[b]->finishing = true;
            fin.push_back(act[b]);
            fj.push_back(&act[b]->job);
        }
    if (!fj.empty())
        EPI_TRY(enqueue_finish(ctx, fj.data(), (uint32_t)fj.size(), fin_out));
    EPI_HIP(hipEventRecord(fin_ev, ctx->stream));
    return EPI_OK;
}

// Waits for the launched round and its rescoring (ctx->mu held) and refines
// the throughput estimate from the round's device time.  A round in which a
// candidate found no slot fails (its lists may miss a triplet).
int epi_server::round_wait(double units)
{
    EPI_HIP(hipEventSynchronize(fin_ev));
    float dms = 0.0f;
    EPI_HIP(hipEventElapsedTime(&dms, ctx->ev0, ctx->ev1));
    if (dms > 0.05f && units > 0.0) {
        const double rate = units / dms;
        units_per_ms = 0.7 * units_per_ms + 0.3 * rate;
    }
    {
        std::lock_guard<std::mutex> lk(m);
        stats.rounds++;
        stats.device_ms += dms;
    }
    return *(const volatile uint32_t *)ovf.ptr != 0 ? EPI_EINTERNAL : EPI_OK;
}

// The device cannot be selected: every active request and the ones being
// admitted fail with `status` (ctx->mu held), so that no poller waits on
// them; `failed` collects them for completion.
void epi_server::fail_all_locked_ctx(std::vector<Request *> &admit, int status,
                                     std::vector<Request *> &failed)
{
    std::vector<Request *> act;
    {
        std::lock_guard<std::mutex> lk(m);
        act.swap(active);
        stats.active = 0;
    }
    act.insert(act.end(), admit.begin(), admit.end());
    admit.clear();
    for (Request *r : act) {
        release_locked_ctx(r);
        r->status = status;
        failed.push_back(r);
    }
}

void epi_server::run()
{
    for (;;) {
        // ---- admission (FIFO, by device footprint) ----
        std::vector<Request *> admit, cancelled;
        {
            std::unique_lock<std::mutex> lk(m);
            cv_work.wait(lk, [&] { return stop || !pending.empty() || !active.empty(); });
            if (stop && pending.empty() && active.empty())
                return;
            while (!pending.empty()) {
                Request *r = pending.front();
                if (r->cancel || stop) {
                    pending.pop_front();
                    cancelled.push_back(r);
                    continue;
                }
                if (active.size() + admit.size() >= cfg.max_active ||
                    stats.bytes_in_use + r->footprint > cfg.memory_budget)
                    break;
                pending.pop_front();
                stats.bytes_in_use += r->footprint;
                r->charged = true;
                admit.push_back(r);
            }
            stats.pending = (uint32_t)pending.size();
        }
        if (!cancelled.empty()) {
            std::lock_guard<epi::FairMutex> g(ctx->mu);
            for (Request *r : cancelled)
                release_locked_ctx(r);  // resident cohorts: drop the reference
        }
        for (Request *r : cancelled)
            complete(r, EPI_ECANCELED);

        std::vector<Request *> finished, failed;
        {
            std::lock_guard<epi::FairMutex> g(ctx->mu);
            const hipError_t de = hipSetDevice(ctx->device);
            if (de != hipSuccess) {
                fail_all_locked_ctx(admit, hip_status(de), failed);
                std::lock_guard<std::mutex> lk(m);
                // requests still queued fail at the next pass, which admits
                // them: none is left waiting
                stats.pending = (uint32_t)pending.size();
            }
            // the round of the requests already active
            std::vector<Request *> act;
            if (de == hipSuccess) {
                std::lock_guard<std::mutex> lk(m);
                for (Request *r : active)
                    if (!(r->cancel || stop))
                        act.push_back(r);
            }
            double units = 0.0;
            int st = act.empty() ? EPI_OK : round_launch(act, units);
            const bool launched = !act.empty() && st == EPI_OK;
            // admission while it runs (the new requests join the next round)
            for (Request *r : admit) {
                const int as = admit_locked_ctx(r);
                if (as != EPI_OK) {
                    (void)hipStreamSynchronize(upload);
                    (void)hipStreamSynchronize(ctx->stream);
                    stage_synced(ctx);
                    release_locked_ctx(r);
                    r->status = as;
                    failed.push_back(r);
                    continue;
                }
                std::lock_guard<std::mutex> lk(m);
                active.push_back(r);
                stats.active = (uint32_t)active.size();
            }
            if (launched)
                st = round_wait(units);
            // requests that are cancelled, finished, or hit by a failed round
            std::vector<Request *> keep, done_ok, drop;
            {
                std::lock_guard<std::mutex> lk(m);
                for (Request *r : active) {
                    if (r->cancel || stop)
                        drop.push_back(r);
                    else if (st != EPI_OK)
                        drop.push_back(r);
                    else if (r->finishing)
                        done_ok.push_back(r);
                    else
                        keep.push_back(r);
                }
                active = keep;
                stats.active = (uint32_t)active.size();
            }
            if (!done_ok.empty()) {
                // the rescored lists of `fin`, in its order (a request of
                // it that was cancelled meanwhile is dropped instead)
                std::vector<Job *> jp;
                for (Request *r : fin)
                    jp.push_back(&r->job);
                std::vector<epi_result> res(fin.size());
                std::vector<int> fs(fin.size(), EPI_OK);
                assemble_finished(fin_out.ptr, jp.data(), (uint32_t)jp.size(), res.data(),
                                  fs.data());
                const auto now = Clock::now();
                for (size_t b = 0; b < fin.size(); ++b) {
                    Request *r = fin[b];
                    if (std::find(done_ok.begin(), done_ok.end(), r) == done_ok.end())
                        continue;
                    r->result = res[b];
                    r->result.latency_ms = ms_between(r->submitted, now);
                    r->result.queue_ms = ms_between(r->submitted, r->admitted);
                    r->result.scan_ms = ms_between(r->admitted, now);
                    r->status = fs[b];
                    finished.push_back(r);
                }
            }
            fin.clear();
            (void)hipStreamSynchronize(ctx->stream);
            stage_synced(ctx);
            for (Request *r : finished)
                release_locked_ctx(r);
            for (Request *r : drop) {
                release_locked_ctx(r);
                r->status = (r->cancel || stop) ? EPI_ECANCELED : st;
                failed.push_back(r);
            }
        }
        for (Request *r : finished)
            complete(r, r->status);
        for (Request *r : failed)
            complete(r, r->status);
    }
}

extern "C" int epi_server_create(epi_ctx *ctx, const epi_server_config *config,
                                 epi_server **out)
{
    if (!ctx || !out)
        return EPI_EINVAL;
    *out = nullptr;
    epi_server_config cfg;
    std::memset(&cfg, 0, sizeof(cfg));
    if (config)
        cfg = *config;
    if (cfg.max_active == 0)
        cfg.max_active = 32;
    if (cfg.max_active > 64)
        return EPI_EINVAL;
    if (cfg.max_samples == 0)
        cfg.max_samples = 32768;
    if (cfg.round_ms <= 0.0)
        cfg.round_ms = 20.0;
    std::unique_ptr<epi_server> s(new (std::nothrow) epi_server());
    if (!s)
        return EPI_ENOMEM;
    {
        std::lock_guard<epi::FairMutex> lk(ctx->mu);
        EPI_HIP(hipSetDevice(ctx->device));
        if (cfg.memory_budget == 0) {
            size_t free_b = 0, total_b = 0;
            EPI_HIP(hipMemGetInfo(&free_b, &total_b));
            cfg.memory_budget = total_b / 4;
        }
        EPI_TRY(ctx_ensure_tables(ctx, cfg.max_samples));
        s->scfg = resolve_config(ctx, nullptr, cfg.max_samples);
        EPI_TRY(plan_launch(ctx, s->scfg, cfg.max_samples + 2, false, s->plan));
        EPI_HIP(hipStreamCreateWithFlags(&s->upload, hipStreamNonBlocking));
        EPI_HIP(hipEventCreateWithFlags(&s->fin_ev, hipEventDisableTiming));
    }
    s->ctx = ctx;
    s->cfg = cfg;
    std::memset(&s->stats, 0, sizeof(s->stats));
    epi_server *raw = s.get();
    try {
        s->worker = std::thread([raw] { raw->run(); });
    } catch (...) {
        return EPI_ENOMEM;
    }
    *out = s.release();
    return EPI_OK;
}

extern "C" int epi_server_destroy(epi_server *s)
{
    if (!s)
        return EPI_EINVAL;
    {
        std::lock_guard<std::mutex> lk(s->m);
        s->stop = true;
    }
    s->cv_work.notify_all();
    if (s->worker.joinable())
        s->worker.join();
    {
        // every request has completed; let the pollers return
        std::unique_lock<std::mutex> lk(s->m);
        s->cv_done.notify_all();
        s->cv_done.wait(lk, [s] { return s->pollers == 0; });
    }
    {
        std::lock_guard<epi::FairMutex> g(s->ctx->mu);
        if (s->graph)
            (void)hipGraphExecDestroy(s->graph);
        if (s->upload)
            (void)hipStreamDestroy(s->upload);
        if (s->fin_ev)
            (void)hipEventDestroy(s->fin_ev);
    }
    delete s;
    return EPI_OK;
}

extern "C" int epi_submit(epi_server *s, const uint8_t *geno, const uint8_t *pheno,
                          uint32_t n_samples, uint32_t n_snps, uint64_t *request_id)
{
    if (!s || !geno || !pheno || !request_id || !cohort_shape_ok(n_samples, n_snps) ||
        n_samples > s->cfg.max_samples)
        return EPI_EINVAL;
    const uint64_t fp = cohort_footprint(n_samples, n_snps, s->scfg.variant);
    if (fp > s->cfg.memory_budget)
        return EPI_ENOMEM;
    std::unique_ptr<Request> r(new (std::nothrow) Request());
    if (!r)
        return EPI_ENOMEM;
    try {
        r->geno.assign(geno, geno + (size_t)n_samples * n_snps);
        r->pheno.assign(pheno, pheno + n_samples);
    } catch (...) {
        return EPI_ENOMEM;
    }
    r->N = n_samples;
    r->M = n_snps;
    r->footprint = fp;
    r->submitted = Clock::now();
    {
        std::lock_guard<std::mutex> lk(s->m);
        if (s->stop)
            return EPI_EBUSY;
        r->id = s->next_id++;
        *request_id = r->id;
        s->pending.push_back(r.get());
        s->stats.submitted++;
        s->stats.pending = (uint32_t)s->pending.size();
        s->all.emplace(r->id, std::move(r));
    }
    s->cv_work.notify_all();
    return EPI_OK;
}

extern "C" int epi_submit_cohort(epi_server *s, epi_cohort *co, uint64_t *request_id)
{
    // a cohort past its pair budget is scanned in strip rounds by epi_scan only
    if (!s || !co || !request_id || co->ctx != s->ctx || co->N > s->cfg.max_samples ||
        co->blocked)
        return EPI_EINVAL;
    std::unique_ptr<Request> r(new (std::nothrow) Request());
    if (!r)
        return EPI_ENOMEM;
    r->N = co->N;
    r->M = co->M;
    r->resident = true;
    r->co = co;
    // device bytes the request adds: its row table copy and running list
    const uint64_t blocks = (uint64_t)(co->Mp / 64 + 1);
    r->footprint = blocks * blocks * 8 * sizeof(Row) + sizeof(Cand) * kTop;
    r->submitted = Clock::now();
    {
        std::lock_guard<epi::FairMutex> g(s->ctx->mu);
        co->server_refs++;
    }
    {
        std::lock_guard<std::mutex> lk(s->m);
        if (!s->stop) {
            r->id = s->next_id++;
            *request_id = r->id;
            s->pending.push_back(r.get());
            s->stats.submitted++;
            s->stats.pending = (uint32_t)s->pending.size();
            s->all.emplace(r->id, std::move(r));
        }
    }
    if (r) {  // not queued: the server is stopping
        std::lock_guard<epi::FairMutex> g(s->ctx->mu);
        co->server_refs--;
        return EPI_EBUSY;
    }
    s->cv_work.notify_all();
    return EPI_OK;
}

extern "C" int epi_poll(epi_server *s, uint64_t id, int32_t timeout_ms, epi_result *result)
{
    if (!s)
        return EPI_EINVAL;
    std::unique_lock<std::mutex> lk(s->m);
    return poll_request(s->all, id, lk, s->cv_done, timeout_ms, s->pollers,
                        [result](const Request &r) {
                            if (r.status == EPI_OK && result)
                                *result = r.result;
                        });
}

extern "C" int epi_cancel(epi_server *s, uint64_t id)
{
    if (!s)
        return EPI_EINVAL;
    {
        std::lock_guard<std::mutex> lk(s->m);
        auto it = s->all.find(id);
        if (it == s->all.end())
            return EPI_ENOENT;
        it->second->cancel = true;
    }
    s->cv_work.notify_all();
    return EPI_OK;
}

extern "C" int epi_server_get_stats(epi_server *s, epi_server_stats *out)
{
    if (!s || !out)
        return EPI_EINVAL;
    std::lock_guard<std::mutex> lk(s->m);
    *out = s->stats;
    return EPI_OK;
}
