// episdet_lib.h -- host-side objects behind the opaque C ABI handles.
#pragma once
#include <hip/hip_runtime.h>

#include <map>
#include <cmath>
#include <condition_variable>
#include <mutex>
#include <string>
#include <utility>
#include <vector>

#include "../../include/episdet.h"
#include "episdet_internal.h"

#define EPI_HIP(call)                   \
    do {                                \
        hipError_t e_ = (call);         \
        if (e_ != hipSuccess)           \
            return epi::hip_status(e_); \
    } while (0)

#define EPI_TRY(call)          \
    do {                       \
        int s_ = (call);       \
        if (s_ != EPI_OK)      \
            return s_;         \
    } while (0)

namespace epi {

// Device blocks that destroyed cohorts gave back, kept for the cohorts
// created after them.  A hipMalloc / hipFree pair of a large cohort buffer
// costs up to milliseconds of host time, and a free waits for the whole
// device, which a server running scans of other cohorts cannot afford.  A
// block is given back only when no queued work reads it (the callers have
// synchronized the stream), and every call is made with ctx->mu held.
struct DevPool {
    std::multimap<size_t, void *> blocks;  // by size
    size_t held = 0;
    size_t limit = 0;  // most bytes kept; a block beyond it is freed
    DevPool() = default;
    DevPool(const DevPool &) = delete;
    DevPool &operator=(const DevPool &) = delete;
    ~DevPool() { trim(); }
    // A kept block of at least `want` bytes and at most about a quarter
    // more (its size in `got`), or null.
    void *take(size_t want, size_t &got);
    void give(void *p, size_t bytes);
    void trim();  // frees every kept block
};

// Owned device allocation that only grows.
struct DevBuf {
    void *ptr = nullptr;
    size_t bytes = 0;
    DevPool *pool = nullptr;  // where release() gives the block (null: hipFree)
    DevBuf() = default;
    DevBuf(const DevBuf &) = delete;
    DevBuf &operator=(const DevBuf &) = delete;
    ~DevBuf();
    int reserve(size_t want);
    void release();
};

// A lock taken in arrival order (a ticket lock): the server's worker, which
// takes the context lock again right after each round, then queues behind
// calls that were already waiting instead of starving them.
struct FairMutex {
    std::mutex m;
    std::condition_variable cv;
    uint64_t next = 0, serving = 0;
    void lock()
    {
        std::unique_lock<std::mutex> lk(m);
        const uint64_t mine = next++;
        cv.wait(lk, [&] { return serving == mine; });
    }
    void unlock()
    {
        {
            std::lock_guard<std::mutex> lk(m);
            ++serving;
        }
        cv.notify_all();
    }
};

// Pinned host staging buffer that only grows (async uploads of per-launch
// tables).
struct HostBuf {
    void *ptr = nullptr;
    size_t bytes = 0;
    HostBuf() = default;
    HostBuf(const HostBuf &) = delete;
    HostBuf &operator=(const HostBuf &) = delete;
    ~HostBuf();
    int reserve(size_t want);
};

// One cohort's scan in progress: its triplet-tile rows (cohort-local item
// numbering; the device copy is the cohort's cached one), how many rows
// have been launched, and its running top-kTop candidate list on the device
// (a slice of the context's list pool, or `own_list` for server jobs).
struct Job {
    epi_cohort *co = nullptr;
    std::vector<Row> rows;        // host copy (ranges are chosen on the host)
    const Row *drows = nullptr;   // resident copy read by the scan kernel
    uint64_t items = 0;
    uint32_t nw = 1;              // wavefronts per item (i values per item)
    size_t next_row = 0;
    Cand *list = nullptr;
    DevBuf own_list;
};

// Tile table entry: applies to n_samples <= max_samples.
struct TuneEntry {
    uint32_t max_samples;
    ScanConfig cfg;
};

// Launch geometry of the scan kernel for one configuration.
struct LaunchPlan {
    ScanConfig cfg;
    bool table_lds = false;
    bool dump = false;
    size_t lds = 0;
    uint32_t grid = 0;
    uint32_t cap = 0;  // appended-candidate capacity per cohort and launch
    uint32_t table_len = 0;
};

int hip_status(hipError_t e);
int tune_load_file(epi_ctx *ctx, const std::string &path);
void tune_load_default(epi_ctx *ctx);
bool tune_lookup(const epi_ctx *ctx, uint32_t n_samples, ScanConfig &cfg);
bool cohort_shape_ok(uint32_t n_samples, uint32_t n_snps);
uint64_t cohort_footprint(uint32_t n_samples, uint32_t n_snps, int variant);
// Packs a cohort into device memory on stream `s` (synchronized before
// returning).  Caller holds ctx->mu.
// full_pairs: every pair strip resident whatever the context's pair budget
// (cohorts of the server, whose rounds read any strip).  seed: sample the
// cohort threshold's seed (cohort_seed).
int cohort_create_locked(epi_ctx *ctx, const uint8_t *geno, const uint8_t *bed,
                         const uint8_t *pheno, uint32_t n_samples, uint32_t n_snps,
                         epi_cohort **out, hipStream_t s, bool full_pairs = false,
                         bool seed = true);
void build_log_tables(uint32_t n_max, std::vector<float> &t32, std::vector<double> &t64);
int ctx_ensure_tables(epi_ctx *ctx, uint32_t n_samples);
ScanConfig resolve_config(const epi_ctx *ctx, const epi_tile *tile, uint32_t n_samples);
void build_rows(uint32_t cohort, uint32_t M, uint32_t Mp, int jp, int nw,
                std::vector<Row> &rows, uint64_t &items);
// Puts the rows (one i per item) of a cohort in allele-frequency SNP order
// into a fixed pseudo-random order and renumbers their items.
void shuffle_rows(std::vector<Row> &rows);
// Round tables in one device buffer: [hdr][segments][descs][lists], sized
// for `cap` cohorts.
struct RoundLayout {
    size_t segs, descs, lists, total;
};

size_t align16(size_t x);
RoundLayout round_layout(uint32_t cap);
// Copies host memory to the device through the pinned staging buffer.
int stage_upload(epi_ctx *ctx, void *dst, const void *src, size_t bytes);
// Host image of the round tables: one segment per job with rows
// [ranges[b].first, .second) of jobs[b]; advances each job's next_row.
int build_round(Job *const *jobs, const std::pair<size_t, size_t> *ranges, uint32_t n,
                uint32_t cap, std::vector<unsigned char> &host, RoundLayout &L, uint64_t &items);
int reserve_round(epi_ctx *ctx, const LaunchPlan &plan, uint32_t n, size_t bytes);
// Device buffers of one round's kernels: the context's own (ctx_round_bufs)
// or a caller's workspace.
struct RoundBufs {
    void *round = nullptr;         // [hdr][segments][descs][lists]
    uint32_t *counters = nullptr;  // 2 + 2n words
    Cand *slots = nullptr;         // n x plan.cap candidates
    const float *t32 = nullptr;    // ln(n!) table, plan.table_len entries
};
RoundBufs ctx_round_bufs(epi_ctx *ctx);
// Queue reset, scan kernel and merge over the round tables in B.round
// (fixed device pointers only, so the sequence can be captured into a graph).
int enqueue_round_kernels(epi_ctx *ctx, const LaunchPlan &plan, uint32_t n, const RoundLayout &L,
                          const RoundBufs &B, bool any_items, bool timed, float *dump_scores,
                          uint32_t *dump_counts);
int plan_launch(epi_ctx *ctx, const ScanConfig &cfg, uint32_t table_len, bool dump,
                LaunchPlan &plan);
// Prepares a job: the cohort's cached rows for the tile width, its expanded
// planes for an MFMA variant, and the running list: `list` (kTop entries,
// cleared by the caller) or, when null, a list of the job's own, cleared here.
int job_init(epi_ctx *ctx, Job &job, epi_cohort *co, const ScanConfig &cfg, Cand *list);
// Reserves the context's list pool for n jobs and clears it (one launch).
int reserve_lists(epi_ctx *ctx, uint32_t n);
// Builds the cohort's expanded planes for an MFMA variant (2: 0/1 bytes,
// 3: e2m1 nibbles) if they do not exist.
int ensure_expanded(epi_ctx *ctx, epi_cohort *co, int variant);
// Uploads the round tables (rows [ranges[b].first, ranges[b].second) of
// jobs[b]) and enqueues queue reset, scan kernel and merge.
// layout (may be null): where the round's tables sit in ctx->round.
int enqueue_round(epi_ctx *ctx, const LaunchPlan &plan, Job *const *jobs,
                  const std::pair<size_t, size_t> *ranges, uint32_t n, bool timed,
                  float *dump_scores, uint32_t *dump_counts, RoundLayout *layout = nullptr);
// Enqueues the rescoring of the jobs' running lists and the copy of the
// rescored candidates into `out` (pinned, grown as needed); once the stream
// has passed that copy, assemble_finished turns them into epi_results, with
// one status per job (a failed consistency check fails only its own job).
int enqueue_finish(epi_ctx *ctx, Job *const *jobs, uint32_t n, HostBuf &out);
void assemble_finished(const void *in, Job *const *jobs, uint32_t n, epi_result *results,
                       int *status);
// One cohort's epi_result from its kTop rescored candidates (EPI_EINTERNAL
// if a scan score differs from the explicit recount).
int assemble(const Rescored *rs, const epi_cohort *co, epi_result &R, double scan_ms,
             double kernel_ms);
// After the context's stream has been synchronized: no staged upload is
// still in flight.
void stage_synced(epi_ctx *ctx);
// With `device` current: takes the device address of the static workspace
// of epi_scan_noalloc (once per device; called by epi_open).
int noalloc_init(int device);

}  // namespace epi

struct epi_ctx {
    int device = 0;
    epi::DevPool pool;  // blocks of destroyed cohorts
    int cu_count = 0;
    size_t lds_per_block = 65536;
    std::string name, arch;
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    epi::FairMutex mu;
    // device bytes of pair-margin strips a cohort keeps resident (both
    // classes); a cohort whose full set would exceed it keeps a few strips
    // that its scans refill (epi_set_pair_budget)
    uint64_t pair_budget = 0, pair_budget_default = 0;
    // stored SNP order of cohorts created from now on (epi_set_snp_order):
    // 0 original, 1 allele frequency, 2 allele frequency from
    // kOrderMinSnps SNPs on
    int snp_order = 2;
    // ln(n!) tables (FP32 reference semantics, FP64 twin), table_len entries
    uint32_t table_len = 0;
    epi::DevBuf t32, t64;
    std::vector<epi::TuneEntry> tune;  // autotuned tile table
    // per-launch scratch, grown on demand
    epi::DevBuf round, fin, counters, slots, rescored;
    epi::DevBuf lists;  // running top-kTop lists of the jobs of one call
    epi::DevBuf brows;  // row table of one round of a scan over strip slots
    // cohort creation inputs on the device (genotypes, sample order, .bed
    // bytes), kept so that creating a cohort frees nothing (a free waits for
    // the device, which may be running a scan of other cohorts)
    epi::DevBuf cgeno, corder, cbed;
    // pinned staging buffers of host -> device uploads, used in turn, so
    // that an upload queued behind running work does not hold up the next
    // one (each buffer's event marks the copy out of it)
    static constexpr int kStageRing = 8;
    epi::HostBuf stage[kStageRing];
    hipEvent_t stage_ev[kStageRing] = {};
    bool stage_pending[kStageRing] = {};
    uint32_t stage_next = 0;
    epi::HostBuf readback;  // pinned destination of a call's results
    // a cohort creation's seed sample: triplets and scores on the device,
    // the triplets and the selected score in pinned host memory
    epi::DevBuf seedc, seedr;
    epi::HostBuf seedh;
    // captured batch graph, reused while the batch signature is unchanged
    hipGraphExec_t batch_graph = nullptr;
    std::vector<uint64_t> batch_sig;
    ~epi_ctx();
};

struct epi_cohort {
    epi_ctx *ctx = nullptr;
    uint32_t N = 0, M = 0, Mp = 0;
    uint32_t n[2] = {0, 0};
    uint32_t W[2] = {0, 0};
    epi::DevBuf sm[2], wm[2], pair[2], single[2], desc;
    // pair-margin strips: nslots slots of [Mp][64] per class; pslot (device,
    // [Mp / 64]) maps a column block to its slot.  Not blocked: every block
    // in its own slot.  Blocked: slot_block[t] is the block slot t holds
    // (UINT32_MAX: none), set by the scan that fills it; todo is that scan's
    // list of strips to compute.
    uint32_t nslots = 0;
    bool blocked = false;
    std::vector<uint32_t> slot_block;
    epi::DevBuf pslot, todo;
    // expanded planes of both classes, class 1 right after class 0, built on
    // first use: 0/1 bytes for the int8 MFMA variant (E), e2m1 nibbles for
    // the FP4 one (F); rows of one pitch (EL, FL bytes) in both classes
    epi::DevBuf E, F;
    uint32_t EL = 0, FL = 0;
    // stored SNP order (ordered): perm holds orig[Mp] (row -> original SNP)
    // then pos[Mp] (original SNP -> row)
    epi::DevBuf perm;
    bool ordered = false;
    // the 64th best FP32 score of a sample of distinct triplets (an upper
    // bound of the cohort's 64th best; +inf when none was taken)
    float seed = INFINITY;
    // triplet-tile rows for tile width rows_jp (0: none yet), built on first use
    std::vector<epi::Row> rows;
    uint64_t rows_items = 0;
    int rows_jp = 0;
    epi::DevBuf drows;
    uint64_t bytes = 0;
    // server requests queued or running on this cohort (epi_submit_cohort);
    // changed with ctx->mu held, and the cohort is not destroyed while nonzero
    uint32_t server_refs = 0;
    epi::CohortDev host_desc() const;
    void use_pool(epi::DevPool *p);
    ~epi_cohort();
};
