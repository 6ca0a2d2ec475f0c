// scan.cpp -- scans behind the C ABI: triplet-tile rows, launch planning,
// rounds (one persistent scan launch over the rows of several cohorts),
// graph-captured batches, rescoring and result assembly.
//
// Replaces the reference's host launch loop and argmin (GPU/v4.cpp:283-608):
// there, one kernel per SNP-block triple with a queue.wait() and a parameter
// copy before each launch, then a host scan over bs^3 slots.  Here, one
// launch pulls every tile from a device queue and the result is a merged
// top-64 list per cohort.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <vector>

#include "../../include/episdet.h"
#include "episdet_internal.h"
#include "episdet_lib.h"

namespace epi {

static uint64_t choose3(uint64_t m) { return m < 3 ? 0 : m * (m - 1) * (m - 2) / 6; }

static double ms_since(std::chrono::steady_clock::time_point t0)
{
    return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
}

HostBuf::~HostBuf()
{
    if (ptr)
        (void)hipHostFree(ptr);
}

int HostBuf::reserve(size_t want)
{
    if (want <= bytes && ptr)
        return EPI_OK;
    if (ptr)
        (void)hipHostFree(ptr);
    ptr = nullptr;
    bytes = 0;
    hipError_t e = hipHostMalloc(&ptr, want, hipHostMallocDefault);
    if (e != hipSuccess) {
        ptr = nullptr;
        return hip_status(e);
    }
    bytes = want;
    return EPI_OK;
}

// Triplet-tile rows of one cohort: k-block kb (64 k), j-group jg (snps_j
// j), i in [0, icount).  Only rows holding at least one i < j < k < M.
void build_rows(uint32_t cohort, uint32_t M, uint32_t Mp, int jp, int nw,
                std::vector<Row> &rows, uint64_t &items)
{
    for (uint32_t kb = 0; kb * 64 < Mp; ++kb) {
        const uint32_t k_lo = kb * 64;
        if (k_lo >= M)
            break;
        const uint32_t k_hi = std::min(kb * 64 + 63, M - 1);
        for (uint32_t jg = 0;; ++jg) {
            const uint32_t j_lo = jg * (uint32_t)jp;
            if (j_lo >= k_hi || j_lo > M - 2)
                break;
            const uint32_t j_max = std::min(std::min(j_lo + (uint32_t)jp - 1, k_hi - 1), M - 2);
            const uint32_t icount = j_max;
            if (icount == 0)
                continue;
            Row r;
            r.cohort = cohort;
            r.kb = kb;
            r.jg = jg;
            r.icount = icount;
            r.item_start = (uint32_t)items;
            r.pad = 0;
            rows.push_back(r);
            items += (icount + (uint32_t)nw - 1) / (uint32_t)nw;
        }
    }
}

// With the SNPs in allele-frequency order, rows in k-block order would
// visit the triplets by allele frequency of their last SNP; where the best
// scores sit at one end, the wavefronts' thresholds would trail the scan
// and far more lanes would pass them (and be summed twice).  A fixed
// pseudo-random row order (xorshift64 Fisher-Yates) decorrelates them; each
// row still runs its items together, so its j and k blocks stay in cache.
// Rows of one i per item (nw = 1).
void shuffle_rows(std::vector<Row> &rows)
{
    uint64_t h = 0x9E3779B97F4A7C15ull;
    for (size_t t = rows.size(); t > 1; --t) {
        h ^= h << 13;
        h ^= h >> 7;
        h ^= h << 17;
        std::swap(rows[t - 1], rows[h % t]);
    }
    uint32_t it = 0;
    for (Row &r : rows) {
        r.item_start = it;
        it += r.icount;
    }
}

// Kernel and tile shape: variant 2 or 3 in `tile` selects that MFMA kernel
// (int8 or FP4); a tile naming the VALU variant or a VALU shape selects the
// VALU kernel; otherwise the context's tile table (per sample-count bucket)
// decides.
// A VALU shape not given by the tile or the table follows the ln(n!) table
// placement: in LDS, 4 j-SNPs per wavefront fit in 128 VGPRs; a large table
// in LDS allows one workgroup per CU, so it gets 16 wavefronts; with the
// table in global memory, 2 j-SNPs per wavefront.
ScanConfig resolve_config(const epi_ctx *ctx, const epi_tile *tile, uint32_t n_samples)
{
    ScanConfig mfma;
    mfma.variant = 2;  // MFMA: 64-SNP j- and k-blocks, kMfmaWaves wavefronts
    mfma.snps_j = 64;
    mfma.waves = kMfmaWaves;
    mfma.words = 1;
    if (tile && (tile->variant == 2 || tile->variant == 3)) {
        mfma.variant = (int)tile->variant;
        return mfma;
    }
    // a tile that names a VALU shape (or the VALU variant) asks for VALU
    const bool valu = tile && (tile->variant == 1 || tile->snps_j || tile->waves);
    ScanConfig cfg;
    cfg.variant = 1;
    cfg.words = 1;
    cfg.snps_j = 4;
    cfg.waves = 8;
    ScanConfig tuned;
    const bool have = tune_lookup(ctx, n_samples, tuned);
    if (have && tuned.variant >= 2 && !valu) {
        mfma.variant = tuned.variant;
        return mfma;
    }
    if (have && tuned.variant == 1) {
        cfg = tuned;
    } else {
        // no tuned VALU shape: the larger workgroup when the ln(n!) table
        // takes more than 64 KiB of LDS, fewer j per wavefront when even
        // that fails
        ScanConfig probe = cfg;
        if (scan_lds_bytes(probe, true, n_samples + 2) > 64 * 1024) {
            probe.waves = 16;
            if (scan_lds_bytes(probe, true, n_samples + 2) <= ctx->lds_per_block)
                cfg.waves = 16;
            else
                cfg.snps_j = 2;
        }
    }
    cfg.variant = 1;
    if (tile) {
        if (tile->snps_j)
            cfg.snps_j = (int)tile->snps_j;
        if (tile->waves)
            cfg.waves = (int)tile->waves;
        if (tile->words)
            cfg.words = (int)tile->words;
    }
    return cfg;
}

int plan_launch(epi_ctx *ctx, const ScanConfig &cfg, uint32_t table_len, bool dump,
                LaunchPlan &plan)
{
    if (!scan_config_supported(cfg))
        return EPI_EINVAL;
    plan.cfg = cfg;
    plan.dump = dump;
    plan.table_len = table_len;
    plan.table_lds = scan_lds_bytes(cfg, true, table_len) <= ctx->lds_per_block;
    plan.lds = scan_lds_bytes(cfg, plan.table_lds, table_len);
    if (plan.lds > 64 * 1024)
        EPI_HIP(scan_allow_lds(cfg, plan.table_lds, dump, plan.lds));
    int per_cu = 0;
    EPI_HIP(scan_occupancy(cfg, plan.table_lds, dump, plan.lds, &per_cu));
    if (per_cu <= 0)
        return EPI_EDEVICE;
    plan.grid = (uint32_t)per_cu * (uint32_t)ctx->cu_count;
    // every wavefront flushes each cohort at most once per launch, at most
    // kTop candidates each
    plan.cap = plan.grid * (uint32_t)cfg.waves * kTop;
    return EPI_OK;
}

// The cohort's triplet-tile rows for tile width jp, built and uploaded once
// (one i per queue item: a wavefront of the VALU kernel or a workgroup of
// the MFMA kernel takes one item at a time).
static int cohort_rows(epi_ctx *ctx, epi_cohort *co, int jp)
{
    if (co->rows_jp == jp)
        return EPI_OK;
    std::vector<Row> rows;
    uint64_t items = 0;
    build_rows(0, co->M, co->Mp, jp, 1, rows, items);
    if (items >= 0x7FFFFFFFull)
        return EPI_EINVAL;
    if (co->ordered)
        shuffle_rows(rows);
    EPI_TRY(co->drows.reserve(sizeof(Row) * std::max<size_t>(rows.size(), 1)));
    EPI_TRY(stage_upload(ctx, co->drows.ptr, rows.data(), sizeof(Row) * rows.size()));
    co->rows = std::move(rows);
    co->rows_items = items;
    co->rows_jp = jp;
    return EPI_OK;
}

int job_init(epi_ctx *ctx, Job &job, epi_cohort *co, const ScanConfig &cfg, Cand *list)
{
    job.co = co;
    job.next_row = 0;
    job.nw = 1;
    EPI_TRY(cohort_rows(ctx, co, cfg.snps_j));
    job.rows = co->rows;
    job.items = co->rows_items;
    job.drows = (const Row *)co->drows.ptr;
    if (cfg.variant >= 2)
        EPI_TRY(ensure_expanded(ctx, co, cfg.variant));
    if (!list) {
        job.own_list.pool = &ctx->pool;
        EPI_TRY(job.own_list.reserve(sizeof(Cand) * kTop));
        list = (Cand *)job.own_list.ptr;
        EPI_HIP(launch_fill_empty(list, kTop, ctx->stream));
    }
    job.list = list;
    return EPI_OK;
}

int reserve_lists(epi_ctx *ctx, uint32_t n)
{
    EPI_TRY(ctx->lists.reserve(sizeof(Cand) * kTop * std::max<size_t>(n, 1)));
    EPI_HIP(launch_fill_empty((Cand *)ctx->lists.ptr, (size_t)kTop * n, ctx->stream));
    return EPI_OK;
}

int ensure_expanded(epi_ctx *ctx, epi_cohort *co, int variant)
{
    const bool fp4 = variant == 3;
    // rows of whole 128-byte stages (128 samples as bytes, 256 as nibbles),
    // one pitch for both classes so that the kernel's copy offsets serve both
    const uint32_t Wm = std::max(co->W[0], co->W[1]);
    const uint32_t L = fp4 ? (Wm + 3) / 4 * 128 : (Wm + 1) / 2 * 128;
    DevBuf &buf = fp4 ? co->F : co->E;
    if (buf.ptr)
        return EPI_OK;
    const size_t per_class = (size_t)co->Mp * 2 * L;
    EPI_TRY(buf.reserve(2 * per_class));
    (fp4 ? co->FL : co->EL) = L;
    for (int c = 0; c < 2; ++c) {
        uint8_t *dst = (uint8_t *)buf.ptr + c * per_class;
        if (fp4)
            EPI_HIP(launch_expand_fp4((const uint4 *)co->sm[c].ptr, co->Mp, co->W[c], L, dst,
                                      ctx->stream));
        else
            EPI_HIP(launch_expand((const uint4 *)co->sm[c].ptr, co->Mp, co->W[c], L, dst,
                                  ctx->stream));
    }
    co->bytes += 2 * per_class;
    return EPI_OK;
}

// Copies `bytes` from host memory to `dst` through the next pinned staging
// buffer of the ring, after the upload that last used it has left it.
int stage_upload(epi_ctx *ctx, void *dst, const void *src, size_t bytes)
{
    if (bytes == 0)
        return EPI_OK;
    const uint32_t b = ctx->stage_next++ % epi_ctx::kStageRing;
    if (ctx->stage_pending[b]) {
        EPI_HIP(hipEventSynchronize(ctx->stage_ev[b]));
        ctx->stage_pending[b] = false;
    }
    EPI_TRY(ctx->stage[b].reserve(bytes));
    std::memcpy(ctx->stage[b].ptr, src, bytes);
    EPI_HIP(hipMemcpyAsync(dst, ctx->stage[b].ptr, bytes, hipMemcpyHostToDevice, ctx->stream));
    EPI_HIP(hipEventRecord(ctx->stage_ev[b], ctx->stream));
    ctx->stage_pending[b] = true;
    return EPI_OK;
}

void stage_synced(epi_ctx *ctx)
{
    for (bool &p : ctx->stage_pending)
        p = false;
}

size_t align16(size_t x) { return (x + 15) & ~(size_t)15; }

RoundLayout round_layout(uint32_t cap)
{
    RoundLayout L;
    L.segs = align16(sizeof(RoundHdr));
    L.descs = L.segs + align16(sizeof(Segment) * cap);
    L.lists = L.descs + align16(sizeof(CohortDev) * cap);
    L.total = L.lists + align16(sizeof(Cand *) * cap);
    return L;
}

// Host image of the round tables: one segment per job with rows
// [ranges[b].first, .second) of jobs[b]; advances each job's next_row.
int build_round(Job *const *jobs, const std::pair<size_t, size_t> *ranges, uint32_t n,
                       uint32_t cap, std::vector<unsigned char> &host, RoundLayout &L,
                       uint64_t &items)
{
    L = round_layout(cap);
    host.assign(L.total, 0);
    RoundHdr *hdr = (RoundHdr *)host.data();
    Segment *segs = (Segment *)(host.data() + L.segs);
    CohortDev *descs = (CohortDev *)(host.data() + L.descs);
    Cand **lists = (Cand **)(host.data() + L.lists);
    items = 0;
    uint32_t ns = 0;
    for (uint32_t b = 0; b < n; ++b) {
        const Job &J = *jobs[b];
        descs[b] = J.co->host_desc();
        lists[b] = J.list;
        const size_t r0 = ranges[b].first, r1 = ranges[b].second;
        if (r1 <= r0)
            continue;
        const Row &last = J.rows[r1 - 1];
        const uint32_t first = J.rows[r0].item_start;
        const uint32_t end = last.item_start + (last.icount + J.nw - 1) / J.nw;
        Segment sg;
        sg.rows = J.drows;
        sg.cohort = b;
        sg.row_begin = (uint32_t)r0;
        sg.row_end = (uint32_t)r1;
        sg.item_base = (uint32_t)items;
        sg.item_first = first;
        sg.n_items = end - first;
        segs[ns++] = sg;
        items += sg.n_items;
    }
    if (items >= 0x7FFFFFFFull)
        return EPI_EINVAL;
    hdr->n_segments = ns;
    hdr->n_items = (uint32_t)items;
    hdr->n_cohorts = n;
    for (uint32_t b = 0; b < n; ++b)
        jobs[b]->next_row = ranges[b].second;
    return EPI_OK;
}

int reserve_round(epi_ctx *ctx, const LaunchPlan &plan, uint32_t n, size_t bytes)
{
    EPI_TRY(ctx->round.reserve(bytes));
    // [0] queue head, [1..n] appended counts, [1+n] overflow, [2+n..2+2n) thresholds
    EPI_TRY(ctx->counters.reserve(sizeof(uint32_t) * (2 + 2 * (size_t)n)));
    EPI_TRY(ctx->slots.reserve(sizeof(Cand) * (size_t)n * plan.cap));
    return EPI_OK;
}

RoundBufs ctx_round_bufs(epi_ctx *ctx)
{
    RoundBufs B;
    B.round = ctx->round.ptr;
    B.counters = (uint32_t *)ctx->counters.ptr;
    B.slots = (Cand *)ctx->slots.ptr;
    B.t32 = (const float *)ctx->t32.ptr;
    return B;
}

// Queue reset, scan kernel and merge over the round tables in B.round.
// Only device pointers that stay fixed for a given buffer set are passed,
// so the sequence can be captured into a graph.
int enqueue_round_kernels(epi_ctx *ctx, const LaunchPlan &plan, uint32_t n, const RoundLayout &L,
                          const RoundBufs &B, bool any_items, bool timed, float *dump_scores,
                          uint32_t *dump_counts)
{
    unsigned char *base = (unsigned char *)B.round;
    uint32_t *counters = B.counters;
    EPI_HIP(launch_fill_u32(counters, 0u, 2 + (size_t)n, ctx->stream));
    ScanParams P;
    std::memset(&P, 0, sizeof(P));
    P.hdr = (const RoundHdr *)base;
    P.cohorts = (const CohortDev *)(base + L.descs);
    P.segs = (const Segment *)(base + L.segs);
    P.counters = counters;
    P.thr = counters + 2 + n;
    P.table = B.t32;
    P.table_len = plan.table_len;
    P.slot_cap = plan.cap;
    P.slots = B.slots;
    P.overflow = counters + 1 + n;
    P.dump_scores = dump_scores;
    P.dump_counts = dump_counts;
    EPI_HIP(launch_init_thr(P.hdr, P.cohorts, (Cand *const *)(base + L.lists), P.thr, n,
                            ctx->stream));
    if (timed)
        EPI_HIP(hipEventRecord(ctx->ev0, ctx->stream));
    if (any_items)
        EPI_HIP(launch_scan(plan.cfg, plan.table_lds, plan.dump, P, plan.grid, ctx->stream));
    if (timed)
        EPI_HIP(hipEventRecord(ctx->ev1, ctx->stream));
    if (!plan.dump)
        EPI_HIP(launch_merge(P.hdr, P.slots, counters, P.thr, plan.cap,
                             (Cand *const *)(base + L.lists), n, ctx->stream));
    return EPI_OK;
}

int enqueue_round(epi_ctx *ctx, const LaunchPlan &plan, Job *const *jobs,
                  const std::pair<size_t, size_t> *ranges, uint32_t n, bool timed,
                  float *dump_scores, uint32_t *dump_counts, RoundLayout *layout)
{
    std::vector<unsigned char> host;
    RoundLayout L;
    uint64_t items = 0;
    EPI_TRY(build_round(jobs, ranges, n, n, host, L, items));
    EPI_TRY(reserve_round(ctx, plan, n, L.total));
    EPI_TRY(stage_upload(ctx, ctx->round.ptr, host.data(), L.total));
    if (layout)
        *layout = L;
    return enqueue_round_kernels(ctx, plan, n, L, ctx_round_bufs(ctx), items > 0, timed,
                                 dump_scores, dump_counts);
}

// Rescoring tables [descs][lists] of the jobs, in ctx->fin.
static int upload_rescore(epi_ctx *ctx, Job *const *jobs, uint32_t n)
{
    const size_t ld = align16(sizeof(CohortDev) * n);
    const size_t total = ld + sizeof(Cand *) * n;
    std::vector<unsigned char> host(total, 0);
    for (uint32_t b = 0; b < n; ++b) {
        ((CohortDev *)host.data())[b] = jobs[b]->co->host_desc();
        ((Cand **)(host.data() + ld))[b] = jobs[b]->list;
    }
    EPI_TRY(ctx->fin.reserve(total));
    EPI_TRY(ctx->rescored.reserve(sizeof(Rescored) * (size_t)n * kTop));
    return stage_upload(ctx, ctx->fin.ptr, host.data(), total);
}

static int enqueue_rescore_kernel(epi_ctx *ctx, uint32_t n)
{
    const size_t ld = align16(sizeof(CohortDev) * n);
    const unsigned char *base = (const unsigned char *)ctx->fin.ptr;
    EPI_HIP(launch_rescore((const CohortDev *)base, nullptr, (Cand *const *)(base + ld), kTop, n,
                           (const float *)ctx->t32.ptr, (const double *)ctx->t64.ptr,
                           (Rescored *)ctx->rescored.ptr, ctx->stream));
    return EPI_OK;
}

static int enqueue_rescore(epi_ctx *ctx, Job *const *jobs, uint32_t n)
{
    EPI_TRY(upload_rescore(ctx, jobs, n));
    return enqueue_rescore_kernel(ctx, n);
}

int assemble(const Rescored *rs, const epi_cohort *co, epi_result &R, double scan_ms,
                    double kernel_ms)
{
    std::memset(&R, 0, sizeof(R));
    R.triplets = choose3(co->M);
    R.n_samples = co->N;
    R.n_snps = co->M;
    R.n_controls = co->n[0];
    R.n_cases = co->n[1];
    R.scan_ms = scan_ms;
    R.kernel_ms = kernel_ms;
    uint32_t nt = 0;
    double maxerr = 0.0;
    for (int e = 0; e < kTop; ++e) {
        if (!rs[e].valid)
            break;
        // the scan's FP32 score must equal the explicit 27-cell recount
        if (std::memcmp(&rs[e].s32, &rs[e].s32_scan, sizeof(float)) != 0)
            return EPI_EINTERNAL;
        epi_candidate &c = R.top[nt++];
        c.i = rs[e].i;
        c.j = rs[e].j;
        c.k = rs[e].k;
        c.score = rs[e].s32;
        c.score_fp64 = rs[e].s64;
        std::memcpy(c.counts, rs[e].counts, sizeof(c.counts));
        maxerr = std::max(maxerr, std::fabs((double)c.score - c.score_fp64));
    }
    R.n_top = nt;
    R.max_abs_err_fp32 = maxerr;
    if (nt > 0) {
        R.i = R.top[0].i;
        R.j = R.top[0].j;
        R.k = R.top[0].k;
        R.score = R.top[0].score;
        R.score_fp64 = R.top[0].score_fp64;
        // near-tie flag: does FP64 order the FP32-ranked candidates differently?
        for (uint32_t e = 0; e + 1 < nt; ++e)
            if (R.top[e + 1].score_fp64 < R.top[e].score_fp64)
                R.fp64_order_differs = 1;
    }
    return EPI_OK;
}

static int copy_back_and_assemble(epi_ctx *ctx, Job *const *jobs, uint32_t n, epi_result *results,
                                  std::chrono::steady_clock::time_point t0, bool timed)
{
    // into pinned memory: both copies are plain DMA and one wait covers them
    // (a copy into pageable memory waits on its own and then copies again)
    const size_t rb = sizeof(Rescored) * (size_t)n * kTop;
    EPI_TRY(ctx->readback.reserve(rb + sizeof(uint32_t)));
    unsigned char *pin = (unsigned char *)ctx->readback.ptr;
    EPI_HIP(hipMemcpyAsync(pin, ctx->rescored.ptr, rb, hipMemcpyDeviceToHost, ctx->stream));
    EPI_HIP(hipMemcpyAsync(pin + rb, (uint32_t *)ctx->counters.ptr + 1 + n, sizeof(uint32_t),
                           hipMemcpyDeviceToHost, ctx->stream));
    EPI_HIP(hipStreamSynchronize(ctx->stream));
    uint32_t overflow = 0;
    std::memcpy(&overflow, pin + rb, sizeof(uint32_t));
    const Rescored *host = (const Rescored *)pin;
    stage_synced(ctx);
    const double wall = ms_since(t0);
    float kms = 0.0f;
    if (timed)
        EPI_HIP(hipEventElapsedTime(&kms, ctx->ev0, ctx->ev1));
    if (overflow != 0)
        return EPI_EINTERNAL;  // a workgroup slot was lost
    for (uint32_t b = 0; b < n; ++b)
        EPI_TRY(assemble(host + (size_t)b * kTop, jobs[b]->co, results[b], wall, kms));
    return EPI_OK;
}

int enqueue_finish(epi_ctx *ctx, Job *const *jobs, uint32_t n, HostBuf &out)
{
    EPI_TRY(enqueue_rescore(ctx, jobs, n));
    const size_t bytes = sizeof(Rescored) * (size_t)n * kTop;
    EPI_TRY(out.reserve(bytes));
    EPI_HIP(hipMemcpyAsync(out.ptr, ctx->rescored.ptr, bytes, hipMemcpyDeviceToHost, ctx->stream));
    return EPI_OK;
}

void assemble_finished(const void *in, Job *const *jobs, uint32_t n, epi_result *results,
                       int *status)
{
    const Rescored *rs = (const Rescored *)in;
    for (uint32_t b = 0; b < n; ++b)
        status[b] = assemble(rs + (size_t)b * kTop, jobs[b]->co, results[b], 0.0, 0.0);
}

// Scan of a cohort that keeps only some pair strips (blocked): its column
// blocks are cut into groups of s = nslots / 2; for each pair of groups
// A <= B the strips of A go to slots [0, s) and those of B to slots
// [s, 2 s) (none more when B = A), each computed only when its slot does not
// hold it yet, the block -> slot map is uploaded, and one round scans the
// rows whose j block is in A and whose k block is in B into the job's
// running list.  Each row belongs to exactly one group pair, so every
// triplet is scored once; every (j, k), (i, j) and (i, k) margin a round
// reads is in the strip of the k block or of the j block.  kernel_ms (may
// be null): summed device time of the rounds' scan kernels.
static int scan_strip_rounds(epi_ctx *ctx, const LaunchPlan &plan, Job &job, float *dump_scores,
                             uint32_t *dump_counts, double *kernel_ms)
{
    epi_cohort *co = job.co;
    const uint32_t nb = co->Mp / 64;
    const uint32_t s = std::max<uint32_t>(co->nslots / 2, 1);
    const uint32_t G = (nb + s - 1) / s;
    const uint32_t jp = (uint32_t)plan.cfg.snps_j;  // j per row group (divides 64)
    std::vector<std::vector<Row>> bucket((size_t)G * G);
    size_t most = 1;
    for (const Row &r : job.rows) {
        std::vector<Row> &b = bucket[(size_t)(r.jg * jp / 64 / s) * G + r.kb / s];
        b.push_back(r);
        most = std::max(most, b.size());
    }
    EPI_TRY(ctx->brows.reserve(sizeof(Row) * most));
    // what the slots hold is known only for this scan's own uploads
    std::fill(co->slot_block.begin(), co->slot_block.end(), 0xFFFFFFFFu);
    std::vector<uint32_t> map(nb, 0);
    std::vector<uint2> todo;
    if (kernel_ms)
        *kernel_ms = 0.0;
    for (uint32_t ga = 0; ga < G; ++ga)
        for (uint32_t gb = ga; gb < G; ++gb) {
            std::vector<Row> &rows = bucket[(size_t)ga * G + gb];
            if (rows.empty())
                continue;
            todo.clear();
            auto want = [&](uint32_t g, uint32_t slot) {
                if (co->slot_block[slot] != g) {
                    co->slot_block[slot] = g;
                    todo.push_back(make_uint2(g, slot));
                }
            };
            for (uint32_t g = ga * s; g < std::min(nb, ga * s + s); ++g)
                want(g, g - ga * s);
            if (gb != ga)
                for (uint32_t g = gb * s; g < std::min(nb, gb * s + s); ++g)
                    want(g, s + g - gb * s);
            std::fill(map.begin(), map.end(), 0u);
            for (uint32_t t = 0; t < co->nslots; ++t)
                if (co->slot_block[t] < nb)
                    map[co->slot_block[t]] = t;
            EPI_TRY(stage_upload(ctx, co->pslot.ptr, map.data(), sizeof(uint32_t) * nb));
            if (!todo.empty()) {
                EPI_TRY(stage_upload(ctx, co->todo.ptr, todo.data(), sizeof(uint2) * todo.size()));
                EPI_HIP(launch_strips((const CohortDev *)co->desc.ptr, (const uint2 *)co->todo.ptr,
                                      (uint32_t)todo.size(), co->Mp, ctx->stream));
            }
            // the group pair's rows, numbered from item 0
            uint32_t item = 0;
            for (Row &r : rows) {
                r.item_start = item;
                item += (r.icount + job.nw - 1) / job.nw;
            }
            EPI_TRY(stage_upload(ctx, ctx->brows.ptr, rows.data(), sizeof(Row) * rows.size()));
            Job sub;
            sub.co = co;
            sub.rows = rows;
            sub.items = item;
            sub.nw = job.nw;
            sub.list = job.list;
            sub.drows = (const Row *)ctx->brows.ptr;
            Job *sp[1] = {&sub};
            const std::pair<size_t, size_t> all(0, rows.size());
            EPI_TRY(enqueue_round(ctx, plan, sp, &all, 1, kernel_ms != nullptr, dump_scores,
                                  dump_counts));
            if (kernel_ms) {
                float ms = 0.0f;
                EPI_HIP(hipEventSynchronize(ctx->ev1));
                EPI_HIP(hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
                *kernel_ms += ms;
            }
        }
    return EPI_OK;
}

static int check_cohorts(epi_ctx *ctx, epi_cohort *const *cohorts, uint32_t n)
{
    if (!ctx || !cohorts || n == 0)
        return EPI_EINVAL;
    for (uint32_t b = 0; b < n; ++b)
        if (!cohorts[b] || cohorts[b]->ctx != ctx)
            return EPI_EINVAL;
    return EPI_OK;
}

}  // namespace epi

using namespace epi;

extern "C" int epi_scan(epi_ctx *ctx, epi_cohort *cohort, const epi_tile *tile,
                        epi_result *result)
{
    if (!result)
        return EPI_EINVAL;
    EPI_TRY(check_cohorts(ctx, &cohort, 1));
    if (tile && tile->variant > 3)
        return EPI_EINVAL;
    std::lock_guard<epi::FairMutex> lk(ctx->mu);
    EPI_HIP(hipSetDevice(ctx->device));
    EPI_TRY(ctx_ensure_tables(ctx, cohort->N));
    const ScanConfig cfg = resolve_config(ctx, tile, cohort->N);
    LaunchPlan plan;
    EPI_TRY(plan_launch(ctx, cfg, cohort->N + 2, false, plan));
    const auto t0 = std::chrono::steady_clock::now();
    Job job;
    EPI_TRY(reserve_lists(ctx, 1));
    EPI_TRY(job_init(ctx, job, cohort, cfg, (Cand *)ctx->lists.ptr));
    Job *jp[1] = {&job};
    if (cohort->blocked) {
        double kms = 0.0;
        EPI_TRY(scan_strip_rounds(ctx, plan, job, nullptr, nullptr, &kms));
        EPI_TRY(enqueue_rescore(ctx, jp, 1));
        EPI_TRY(copy_back_and_assemble(ctx, jp, 1, result, t0, false));
        result->kernel_ms = kms;
        return EPI_OK;
    }
    const std::pair<size_t, size_t> all(0, job.rows.size());
    RoundLayout L;
    EPI_TRY(enqueue_round(ctx, plan, jp, &all, 1, true, nullptr, nullptr, &L));
    // the rescoring reads the round's own tables (the cohort's descriptor
    // and its list pointer), which the scan left on the device
    EPI_TRY(ctx->rescored.reserve(sizeof(Rescored) * kTop));
    const unsigned char *rb = (const unsigned char *)ctx->round.ptr;
    EPI_HIP(launch_rescore((const CohortDev *)(rb + L.descs), nullptr,
                           (Cand *const *)(rb + L.lists), kTop, 1, (const float *)ctx->t32.ptr,
                           (const double *)ctx->t64.ptr, (Rescored *)ctx->rescored.ptr,
                           ctx->stream));
    return copy_back_and_assemble(ctx, jp, 1, result, t0, true);
}

// All cohorts in one persistent launch over one queue.  The kernel sequence
// (queue reset, scan, merge, rescoring) is captured into a HIP graph once
// and replayed while the batch keeps the same shape and buffers; the round
// tables it reads are uploaded before each replay.
extern "C" int epi_scan_batch(epi_ctx *ctx, epi_cohort *const *cohorts, uint32_t n,
                              const epi_tile *tile, epi_result *results)
{
    if (!results)
        return EPI_EINVAL;
    EPI_TRY(check_cohorts(ctx, cohorts, n));
    if (tile && tile->variant > 3)
        return EPI_EINVAL;
    std::lock_guard<epi::FairMutex> lk(ctx->mu);
    EPI_HIP(hipSetDevice(ctx->device));
    uint32_t n_max = 0;
    for (uint32_t b = 0; b < n; ++b) {
        if (cohorts[b]->blocked)
            return EPI_EINVAL;  // one launch reads every strip: scan it alone
        n_max = std::max(n_max, cohorts[b]->N);
    }
    EPI_TRY(ctx_ensure_tables(ctx, n_max));
    const ScanConfig cfg = resolve_config(ctx, tile, n_max);
    LaunchPlan plan;
    EPI_TRY(plan_launch(ctx, cfg, n_max + 2, false, plan));
    const auto t0 = std::chrono::steady_clock::now();
    std::vector<Job> jobs(n);
    std::vector<Job *> jp(n);
    std::vector<std::pair<size_t, size_t>> ranges(n);
    EPI_TRY(reserve_lists(ctx, n));
    for (uint32_t b = 0; b < n; ++b) {
        EPI_TRY(job_init(ctx, jobs[b], cohorts[b], cfg, (Cand *)ctx->lists.ptr + (size_t)kTop * b));
        jp[b] = &jobs[b];
        ranges[b] = std::make_pair((size_t)0, jobs[b].rows.size());
    }
    std::vector<unsigned char> host;
    RoundLayout L;
    uint64_t items = 0;
    EPI_TRY(build_round(jp.data(), ranges.data(), n, n, host, L, items));
    EPI_TRY(reserve_round(ctx, plan, n, L.total));
    EPI_TRY(stage_upload(ctx, ctx->round.ptr, host.data(), L.total));
    EPI_TRY(upload_rescore(ctx, jp.data(), n));

    std::vector<uint64_t> sig = {(uint64_t)n, (uint64_t)(items > 0), (uint64_t)plan.cfg.snps_j,
                                 (uint64_t)plan.cfg.waves, (uint64_t)plan.grid, plan.cap,
                                 (uint64_t)plan.cfg.variant, plan.table_len,
                                 (uint64_t)plan.table_lds, L.total,
                                 (uint64_t)(uintptr_t)ctx->round.ptr,
                                 (uint64_t)(uintptr_t)ctx->counters.ptr,
                                 (uint64_t)(uintptr_t)ctx->slots.ptr,
                                 (uint64_t)(uintptr_t)ctx->fin.ptr,
                                 (uint64_t)(uintptr_t)ctx->rescored.ptr,
                                 (uint64_t)(uintptr_t)ctx->t32.ptr,
                                 (uint64_t)(uintptr_t)ctx->t64.ptr};
    if (!ctx->batch_graph || sig != ctx->batch_sig) {
        if (ctx->batch_graph) {
            (void)hipGraphExecDestroy(ctx->batch_graph);
            ctx->batch_graph = nullptr;
        }
        hipGraph_t g = nullptr;
        EPI_HIP(hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
        int st = enqueue_round_kernels(ctx, plan, n, L, ctx_round_bufs(ctx), items > 0, false,
                                       nullptr, nullptr);
        if (st == EPI_OK)
            st = enqueue_rescore_kernel(ctx, n);
        hipError_t ce = hipStreamEndCapture(ctx->stream, &g);
        if (st != EPI_OK) {
            if (g)
                (void)hipGraphDestroy(g);
            return st;
        }
        EPI_HIP(ce);
        hipError_t ie = hipGraphInstantiate(&ctx->batch_graph, g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        EPI_HIP(ie);
        ctx->batch_sig = sig;
    }
    EPI_HIP(hipEventRecord(ctx->ev0, ctx->stream));
    EPI_HIP(hipGraphLaunch(ctx->batch_graph, ctx->stream));
    EPI_HIP(hipEventRecord(ctx->ev1, ctx->stream));
    return copy_back_and_assemble(ctx, jp.data(), n, results, t0, true);
}

extern "C" int epi_scan_dump(epi_ctx *ctx, epi_cohort *cohort, const epi_tile *tile,
                             float *scores, uint32_t *counts)
{
    if (!scores)
        return EPI_EINVAL;
    EPI_TRY(check_cohorts(ctx, &cohort, 1));
    if (cohort->M > 512 || (tile && tile->variant > 3))
        return EPI_EINVAL;
    std::lock_guard<epi::FairMutex> lk(ctx->mu);
    EPI_HIP(hipSetDevice(ctx->device));
    EPI_TRY(ctx_ensure_tables(ctx, cohort->N));
    const ScanConfig cfg = resolve_config(ctx, tile, cohort->N);
    LaunchPlan plan;
    EPI_TRY(plan_launch(ctx, cfg, cohort->N + 2, true, plan));
    const uint64_t ntrip = choose3(cohort->M);
    DevBuf ds, dc;
    EPI_TRY(ds.reserve(std::max<uint64_t>(ntrip, 1) * sizeof(float)));
    if (counts)
        EPI_TRY(dc.reserve(std::max<uint64_t>(ntrip, 1) * 54 * sizeof(uint32_t)));
    EPI_HIP(launch_fill_u32((uint32_t *)ds.ptr, 0x7F800000u, ntrip, ctx->stream));  // +inf
    if (counts)
        EPI_HIP(launch_fill_u32((uint32_t *)dc.ptr, 0xFFFFFFFFu, ntrip * 54, ctx->stream));
    Job job;
    EPI_TRY(job_init(ctx, job, cohort, cfg, nullptr));
    Job *jp[1] = {&job};
    const std::pair<size_t, size_t> all(0, job.rows.size());
    if (cohort->blocked)
        EPI_TRY(scan_strip_rounds(ctx, plan, job, (float *)ds.ptr, (uint32_t *)dc.ptr, nullptr));
    else
        EPI_TRY(enqueue_round(ctx, plan, jp, &all, 1, false, (float *)ds.ptr,
                              (uint32_t *)dc.ptr));
    EPI_HIP(hipMemcpyAsync(scores, ds.ptr, ntrip * sizeof(float), hipMemcpyDeviceToHost,
                           ctx->stream));
    if (counts)
        EPI_HIP(hipMemcpyAsync(counts, dc.ptr, ntrip * 54 * sizeof(uint32_t),
                               hipMemcpyDeviceToHost, ctx->stream));
    EPI_HIP(hipStreamSynchronize(ctx->stream));
    stage_synced(ctx);
    return EPI_OK;
}

extern "C" int epi_count(epi_ctx *ctx, epi_cohort *co, const uint32_t *triplets, uint32_t n,
                         uint32_t *counts, float *scores, double *scores_fp64)
{
    if (!ctx || !co || co->ctx != ctx || (!triplets && n))
        return EPI_EINVAL;
    if (n == 0)
        return EPI_OK;
    for (uint32_t t = 0; t < n; ++t) {
        const uint32_t i = triplets[3 * t], j = triplets[3 * t + 1], k = triplets[3 * t + 2];
        if (!(i < j && j < k && k < co->M))
            return EPI_EINVAL;
    }
    std::lock_guard<epi::FairMutex> lk(ctx->mu);
    EPI_HIP(hipSetDevice(ctx->device));
    EPI_TRY(ctx_ensure_tables(ctx, co->N));
    std::vector<Cand> cands(n);
    for (uint32_t t = 0; t < n; ++t) {
        cands[t].s = 0.0f;
        cands[t].pad = 0;
        cands[t].key = make_key(triplets[3 * t], triplets[3 * t + 1], triplets[3 * t + 2]);
    }
    DevBuf dc, dr;
    EPI_TRY(dc.reserve(sizeof(Cand) * n));
    EPI_TRY(dr.reserve(sizeof(Rescored) * n));
    EPI_HIP(hipMemcpyAsync(dc.ptr, cands.data(), sizeof(Cand) * n, hipMemcpyHostToDevice,
                           ctx->stream));
    EPI_HIP(launch_rescore((const CohortDev *)co->desc.ptr, (const Cand *)dc.ptr, nullptr, n, 1,
                           (const float *)ctx->t32.ptr, (const double *)ctx->t64.ptr,
                           (Rescored *)dr.ptr, ctx->stream));
    std::vector<Rescored> host(n);
    EPI_HIP(hipMemcpyAsync(host.data(), dr.ptr, sizeof(Rescored) * n, hipMemcpyDeviceToHost,
                           ctx->stream));
    EPI_HIP(hipStreamSynchronize(ctx->stream));
    stage_synced(ctx);
    for (uint32_t t = 0; t < n; ++t) {
        if (counts)
            std::memcpy(counts + 54 * (size_t)t, host[t].counts, sizeof(host[t].counts));
        if (scores)
            scores[t] = host[t].s32;
        if (scores_fp64)
            scores_fp64[t] = host[t].s64;
    }
    return EPI_OK;
}
