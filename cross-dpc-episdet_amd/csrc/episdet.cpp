// episdet.cpp -- host side of the C ABI (include/episdet.h).
//
// Replaces the host part of the reference's GPU program (GPU/v4.cpp:198-624):
// device selection, USM uploads, the ln(n!) table, the block-triple launch
// loop with its per-launch queue.wait(), and the host argmin.  Here a scan
// is: pair margins -> one persistent scan launch over a queue of triplet
// tiles -> per-cohort top-64 merge -> FP32/FP64 rescoring of the top-64,
// all on one HIP stream, with a single copy back at the end.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <new>
#include <string>
#include <vector>

#include "../../include/episdet.h"
#include "episdet_internal.h"
#include "episdet_lib.h"

using namespace epi;

// ---------------------------------------------------------------------------
// errors
// ---------------------------------------------------------------------------
extern "C" const char *epi_version(void) { return "episdet-mi355x 0.1.0 (gfx950)"; }

extern "C" const char *epi_strerror(int code)
{
    switch (code) {
    case EPI_OK: return "ok";
    case EPI_EINVAL: return "invalid argument";
    case EPI_ENOMEM: return "out of memory";
    case EPI_EDEVICE: return "device error";
    case EPI_EBUSY: return "busy";
    case EPI_ECANCELED: return "cancelled";
    case EPI_EAGAIN: return "not ready";
    case EPI_ENOENT: return "no such request";
    case EPI_EINTERNAL: return "internal consistency check failed";
    default: return "unknown error";
    }
}

namespace epi {

int hip_status(hipError_t e)
{
    if (e == hipSuccess)
        return EPI_OK;
    if (e == hipErrorOutOfMemory)
        return EPI_ENOMEM;
    std::fprintf(stderr, "episdet: HIP error %d: %s\n", (int)e, hipGetErrorString(e));
    return EPI_EDEVICE;
}

#define EPI_HIP(call)                              \
    do {                                           \
        hipError_t e_ = (call);                    \
        if (e_ != hipSuccess)                      \
            return epi::hip_status(e_);            \
    } while (0)

#define EPI_TRY(call)                              \
    do {                                           \
        int s_ = (call);                           \
        if (s_ != EPI_OK)                          \
            return s_;                             \
    } while (0)

// ---------------------------------------------------------------------------
// ln(n!) with the reference's CPU semantics (CPU/v2.cpp:167-195, table built
// at :580-583): below 748 a float recurrence T[n] = (float)(T[n-1] + log n);
// from 748 on the Stirling form (n + 0.5) ln n - (n - 1) ln e - 0.0810
// evaluated in double and rounded to float.  The FP64 twin skips the float
// roundings.
// ---------------------------------------------------------------------------
void build_log_tables(uint32_t n_max, std::vector<float> &t32, std::vector<double> &t64)
{
    t32.assign((size_t)n_max + 1, 0.0f);
    t64.assign((size_t)n_max + 1, 0.0);
    volatile double one = 1.0;
    const double ln_e = std::log(std::exp(one));
    for (uint32_t n = 0; n <= n_max; ++n) {
        if (n < 748) {
            if (n >= 2) {
                t32[n] = (float)((double)t32[n - 1] + std::log((double)n));
                t64[n] = t64[n - 1] + std::log((double)n);
            }
        } else {
            const double nd = (double)n;
            const double x = (nd + 0.5) * std::log(nd) - (double)(n - 1) * ln_e + (-0.0810);
            t32[n] = (float)x;
            t64[n] = x;
        }
    }
}

DevBuf::~DevBuf() { release(); }

void DevBuf::release()
{
    if (ptr)
        (void)hipFree(ptr);
    ptr = nullptr;
    bytes = 0;
}

int DevBuf::reserve(size_t want)
{
    if (want <= bytes && ptr)
        return EPI_OK;
    release();
    if (want == 0)
        return EPI_OK;
    hipError_t e = hipMalloc(&ptr, want);
    if (e != hipSuccess) {
        ptr = nullptr;
        return hip_status(e);
    }
    bytes = want;
    return EPI_OK;
}

int ctx_ensure_tables(epi_ctx *ctx, uint32_t n_samples)
{
    const uint32_t need = n_samples + 2;  // indices up to n0 + n1 + 1 = N + 1
    if (ctx->table_len >= need)
        return EPI_OK;
    uint32_t len = std::max<uint32_t>(need, 1024);
    std::vector<float> t32;
    std::vector<double> t64;
    build_log_tables(len - 1, t32, t64);
    EPI_TRY(ctx->t32.reserve((size_t)len * sizeof(float)));
    EPI_TRY(ctx->t64.reserve((size_t)len * sizeof(double)));
    EPI_HIP(hipMemcpy(ctx->t32.ptr, t32.data(), (size_t)len * sizeof(float), hipMemcpyHostToDevice));
    EPI_HIP(hipMemcpy(ctx->t64.ptr, t64.data(), (size_t)len * sizeof(double), hipMemcpyHostToDevice));
    ctx->table_len = len;
    return EPI_OK;
}

}  // namespace epi

// ---------------------------------------------------------------------------
// context
// ---------------------------------------------------------------------------
extern "C" int epi_open(int device, epi_ctx **out)
{
    if (!out)
        return EPI_EINVAL;
    *out = nullptr;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n <= 0)
        return EPI_EDEVICE;
    if (device < 0 || device >= n)
        return EPI_EINVAL;
    EPI_HIP(hipSetDevice(device));
    epi_ctx *ctx = new (std::nothrow) epi_ctx();
    if (!ctx)
        return EPI_ENOMEM;
    ctx->device = device;
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, device) != hipSuccess) {
        delete ctx;
        return EPI_EDEVICE;
    }
    ctx->cu_count = prop.multiProcessorCount;
    ctx->lds_per_block = prop.sharedMemPerBlock;
    ctx->name = prop.name;
    ctx->arch = prop.gcnArchName;
    if (hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreate(&ctx->ev0) != hipSuccess || hipEventCreate(&ctx->ev1) != hipSuccess) {
        delete ctx;
        return EPI_EDEVICE;
    }
    *out = ctx;
    return EPI_OK;
}

epi_ctx::~epi_ctx()
{
    if (ev0)
        (void)hipEventDestroy(ev0);
    if (ev1)
        (void)hipEventDestroy(ev1);
    if (stream)
        (void)hipStreamDestroy(stream);
}

extern "C" int epi_close(epi_ctx *ctx)
{
    if (!ctx)
        return EPI_EINVAL;
    (void)hipSetDevice(ctx->device);
    (void)hipStreamSynchronize(ctx->stream);
    delete ctx;
    return EPI_OK;
}

extern "C" int epi_device_name(epi_ctx *ctx, char *buf, size_t len)
{
    if (!ctx || !buf || len == 0)
        return EPI_EINVAL;
    std::snprintf(buf, len, "%s", ctx->name.c_str());
    return EPI_OK;
}

// ---------------------------------------------------------------------------
// cohorts
// ---------------------------------------------------------------------------
epi_cohort::~epi_cohort()
{
    for (int c = 0; c < 2; ++c) {
        sm[c].release();
        wm[c].release();
        pair[c].release();
        single[c].release();
    }
    desc.release();
}

extern "C" int epi_cohort_create(epi_ctx *ctx, const uint8_t *geno, const uint8_t *pheno,
                                 uint32_t n_samples, uint32_t n_snps, epi_cohort **out)
{
    if (!ctx || !geno || !pheno || !out || n_samples == 0 || n_snps < 3 ||
        n_snps > (1u << 21) - 1)
        return EPI_EINVAL;
    *out = nullptr;
    std::lock_guard<std::mutex> lk(ctx->mu);
    EPI_HIP(hipSetDevice(ctx->device));
    epi_cohort *co = new (std::nothrow) epi_cohort();
    if (!co)
        return EPI_ENOMEM;
    co->ctx = ctx;
    co->N = n_samples;
    co->M = n_snps;
    co->Mp = (n_snps + 63) / 64 * 64;
    // class split in sample order (GPU/v4.cpp:135-171): phenotype 1 = case
    std::vector<uint32_t> order;
    order.reserve(n_samples);
    for (uint32_t s = 0; s < n_samples; ++s)
        if (pheno[s] != 1)
            order.push_back(s);
    co->n[0] = (uint32_t)order.size();
    for (uint32_t s = 0; s < n_samples; ++s)
        if (pheno[s] == 1)
            order.push_back(s);
    co->n[1] = n_samples - co->n[0];
    for (int c = 0; c < 2; ++c)
        co->W[c] = (co->n[c] + 63) / 64;

    int st = EPI_OK;
    DevBuf dgeno, dorder;
    const size_t gbytes = (size_t)n_samples * n_snps;
    st = dgeno.reserve(gbytes);
    if (st == EPI_OK)
        st = dorder.reserve((size_t)n_samples * 4);
    for (int c = 0; c < 2 && st == EPI_OK; ++c) {
        const size_t pb = (size_t)std::max<uint32_t>(co->W[c], 1) * co->Mp * sizeof(uint4);
        st = co->sm[c].reserve(pb);
        if (st == EPI_OK)
            st = co->wm[c].reserve(pb);
        if (st == EPI_OK)
            st = co->pair[c].reserve((size_t)co->Mp * co->Mp * sizeof(uint4));
        if (st == EPI_OK)
            st = co->single[c].reserve((size_t)co->Mp * sizeof(uint2));
    }
    if (st == EPI_OK)
        st = co->desc.reserve(sizeof(CohortDev));
    hipError_t e = hipSuccess;
    if (st == EPI_OK)
        e = hipMemcpyAsync(dgeno.ptr, geno, gbytes, hipMemcpyHostToDevice, ctx->stream);
    if (st == EPI_OK && e == hipSuccess)
        e = hipMemcpyAsync(dorder.ptr, order.data(), (size_t)n_samples * 4,
                           hipMemcpyHostToDevice, ctx->stream);
    if (st == EPI_OK && e == hipSuccess)
        e = launch_pack((const uint8_t *)dgeno.ptr, n_samples, n_snps, co->Mp,
                        (const uint32_t *)dorder.ptr, co->n[0], co->n[1],
                        (uint4 *)co->sm[0].ptr, (uint4 *)co->wm[0].ptr,
                        (uint4 *)co->sm[1].ptr, (uint4 *)co->wm[1].ptr, co->W[0], co->W[1],
                        ctx->stream);
    CohortDev h = co->host_desc();
    if (st == EPI_OK && e == hipSuccess)
        e = hipMemcpyAsync(co->desc.ptr, &h, sizeof(h), hipMemcpyHostToDevice, ctx->stream);
    if (st == EPI_OK && e == hipSuccess)
        e = hipStreamSynchronize(ctx->stream);
    if (st == EPI_OK && e != hipSuccess)
        st = hip_status(e);
    if (st != EPI_OK) {
        delete co;
        return st;
    }
    co->bytes = 0;
    for (int c = 0; c < 2; ++c)
        co->bytes += co->sm[c].bytes + co->wm[c].bytes + co->pair[c].bytes + co->single[c].bytes;
    co->bytes += co->desc.bytes;
    *out = co;
    return EPI_OK;
}

CohortDev epi_cohort::host_desc() const
{
    CohortDev h;
    std::memset(&h, 0, sizeof(h));
    for (int c = 0; c < 2; ++c) {
        h.sm[c] = (const uint4 *)sm[c].ptr;
        h.wm[c] = (const uint4 *)wm[c].ptr;
        h.pair[c] = (uint4 *)pair[c].ptr;
        h.single[c] = (uint2 *)single[c].ptr;
        h.W[c] = W[c];
        h.n[c] = n[c];
    }
    h.M = M;
    h.Mp = Mp;
    return h;
}

extern "C" int epi_cohort_destroy(epi_cohort *co)
{
    if (!co)
        return EPI_EINVAL;
    epi_ctx *ctx = co->ctx;
    std::lock_guard<std::mutex> lk(ctx->mu);
    (void)hipSetDevice(ctx->device);
    (void)hipStreamSynchronize(ctx->stream);
    delete co;
    return EPI_OK;
}

extern "C" int epi_cohort_bytes(const epi_cohort *co, uint64_t *bytes)
{
    if (!co || !bytes)
        return EPI_EINVAL;
    *bytes = co->bytes;
    return EPI_OK;
}

// ---------------------------------------------------------------------------
// scans
// ---------------------------------------------------------------------------
namespace epi {

static uint64_t choose3(uint64_t m) { return m < 3 ? 0 : m * (m - 1) * (m - 2) / 6; }

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

// Tile shape: explicit fields of `tile` win; the rest follows the table
// placement.  With the ln(n!) table in LDS, 4 j-SNPs per wavefront fit in
// 128 VGPRs; a large table in LDS allows one workgroup per CU, so it gets
// 16 wavefronts.  With the table in global memory, 2 j-SNPs per wavefront.
ScanConfig resolve_config(const epi_ctx *ctx, const epi_tile *tile, uint32_t n_samples)
{
    ScanConfig cfg;
    cfg.words = 1;
    cfg.snps_j = 4;
    cfg.waves = 8;
    const size_t table = ((size_t)n_samples + 2) * 4;
    ScanConfig probe = cfg;
    if (scan_lds_bytes(probe, true, (uint32_t)(n_samples + 2)) > 64 * 1024) {
        probe.waves = 16;
        if (scan_lds_bytes(probe, true, (uint32_t)(n_samples + 2)) <= ctx->lds_per_block)
            cfg.waves = 16;
        else
            cfg.snps_j = 2;
    }
    (void)table;
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

static double ms_since(std::chrono::steady_clock::time_point t0)
{
    return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
}

// One persistent launch over the tiles of all `n` cohorts.
int run_scans(epi_ctx *ctx, epi_cohort *const *cohorts, uint32_t n, const epi_tile *tile,
              epi_result *results, float *dump_scores, uint32_t *dump_counts)
{
    if (!ctx || !cohorts || n == 0 || (!results && !dump_scores))
        return EPI_EINVAL;
    for (uint32_t b = 0; b < n; ++b)
        if (!cohorts[b] || cohorts[b]->ctx != ctx)
            return EPI_EINVAL;
    if (tile && tile->variant > 1)
        return EPI_EINVAL;  // only the VALU popcount path is built
    const bool dump = dump_scores != nullptr;
    if (dump && (n != 1 || cohorts[0]->M > 512))
        return EPI_EINVAL;

    std::lock_guard<std::mutex> lk(ctx->mu);
    EPI_HIP(hipSetDevice(ctx->device));
    const auto t_start = std::chrono::steady_clock::now();

    uint32_t n_max = 0;
    for (uint32_t b = 0; b < n; ++b)
        n_max = std::max(n_max, cohorts[b]->N);
    EPI_TRY(ctx_ensure_tables(ctx, n_max));

    const uint32_t table_len = n_max + 2;  // indices up to n0 + n1 + 1
    const ScanConfig cfg = resolve_config(ctx, tile, n_max);
    if (!scan_config_supported(cfg))
        return EPI_EINVAL;

    std::vector<Row> rows;
    uint64_t items = 0;
    for (uint32_t b = 0; b < n; ++b)
        build_rows(b, cohorts[b]->M, cohorts[b]->Mp, cfg.snps_j, cfg.waves, rows, items);
    if (items >= 0x7FFFFFFFull)
        return EPI_EINVAL;

    // LDS-resident table when it fits next to the candidate buffers
    const bool table_lds = scan_lds_bytes(cfg, true, table_len) <= ctx->lds_per_block;
    const size_t lds = scan_lds_bytes(cfg, table_lds, table_len);
    if (lds > 64 * 1024)
        EPI_HIP(scan_allow_lds(cfg, table_lds, dump, lds));
    int per_cu = 0;
    EPI_HIP(scan_occupancy(cfg, table_lds, dump, lds, &per_cu));
    if (per_cu <= 0)
        return EPI_EDEVICE;
    uint64_t grid64 = (uint64_t)per_cu * (uint64_t)ctx->cu_count;
    if (grid64 > items)
        grid64 = std::max<uint64_t>(items, 1);
    const uint32_t grid = (uint32_t)grid64;
    const uint32_t slot_cap = grid;

    // device scratch
    std::vector<CohortDev> descs(n);
    for (uint32_t b = 0; b < n; ++b)
        descs[b] = cohorts[b]->host_desc();
    EPI_TRY(ctx->descs.reserve(sizeof(CohortDev) * n));
    EPI_TRY(ctx->rows.reserve(sizeof(Row) * std::max<size_t>(rows.size(), 1)));
    EPI_TRY(ctx->counters.reserve(sizeof(uint32_t) * (2 + (size_t)n)));
    EPI_TRY(ctx->slots.reserve(sizeof(Cand) * (size_t)n * slot_cap * kTop));
    EPI_TRY(ctx->merged.reserve(sizeof(Cand) * (size_t)n * kTop));
    EPI_TRY(ctx->rescored.reserve(sizeof(Rescored) * (size_t)n * kTop));
    DevBuf ddump_s, ddump_c;
    uint64_t ntrip = dump ? choose3(cohorts[0]->M) : 0;
    if (dump) {
        EPI_TRY(ddump_s.reserve(std::max<uint64_t>(ntrip, 1) * sizeof(float)));
        if (dump_counts)
            EPI_TRY(ddump_c.reserve(std::max<uint64_t>(ntrip, 1) * 54 * sizeof(uint32_t)));
    }

    hipStream_t st = ctx->stream;
    uint32_t *counters = (uint32_t *)ctx->counters.ptr;
    EPI_HIP(hipMemcpyAsync(ctx->descs.ptr, descs.data(), sizeof(CohortDev) * n,
                           hipMemcpyHostToDevice, st));
    if (!rows.empty())
        EPI_HIP(hipMemcpyAsync(ctx->rows.ptr, rows.data(), sizeof(Row) * rows.size(),
                               hipMemcpyHostToDevice, st));
    EPI_HIP(launch_fill_u32(counters, 0u, 2 + (size_t)n, st));

    // pair / single margins (part of the scan: they depend only on the cohort)
    for (uint32_t b = 0; b < n; ++b)
        EPI_HIP(launch_pairs((const CohortDev *)cohorts[b]->desc.ptr, descs[b], st));

    ScanParams P;
    std::memset(&P, 0, sizeof(P));
    P.cohorts = (const CohortDev *)ctx->descs.ptr;
    P.rows = (const Row *)ctx->rows.ptr;
    P.n_rows = (uint32_t)rows.size();
    P.n_items = (uint32_t)items;
    P.counters = counters;
    P.table = (const float *)ctx->t32.ptr;
    P.table_len = table_len;
    P.slot_cap = slot_cap;
    P.slots = (Cand *)ctx->slots.ptr;
    P.overflow = counters + 1 + n;
    P.dump_scores = (float *)ddump_s.ptr;
    P.dump_counts = (uint32_t *)ddump_c.ptr;
    if (dump) {
        EPI_HIP(launch_fill_u32((uint32_t *)ddump_s.ptr, 0x7F800000u, ntrip, st));  // +inf
        if (dump_counts)
            EPI_HIP(launch_fill_u32((uint32_t *)ddump_c.ptr, 0xFFFFFFFFu, ntrip * 54, st));
    }

    EPI_HIP(hipEventRecord(ctx->ev0, st));
    if (items > 0)
        EPI_HIP(launch_scan(cfg, table_lds, dump, P, grid, st));
    EPI_HIP(hipEventRecord(ctx->ev1, st));

    if (dump) {
        EPI_HIP(hipMemcpyAsync(dump_scores, ddump_s.ptr, ntrip * sizeof(float),
                               hipMemcpyDeviceToHost, st));
        if (dump_counts)
            EPI_HIP(hipMemcpyAsync(dump_counts, ddump_c.ptr, ntrip * 54 * sizeof(uint32_t),
                                   hipMemcpyDeviceToHost, st));
        EPI_HIP(hipStreamSynchronize(st));
        return EPI_OK;
    }

    EPI_HIP(launch_merge((const Cand *)ctx->slots.ptr, counters, slot_cap, n,
                         (Cand *)ctx->merged.ptr, st));
    EPI_HIP(launch_rescore((const CohortDev *)ctx->descs.ptr, (const Cand *)ctx->merged.ptr,
                           kTop, n, (const float *)ctx->t32.ptr, (const double *)ctx->t64.ptr,
                           (Rescored *)ctx->rescored.ptr, st));
    std::vector<Rescored> host((size_t)n * kTop);
    uint32_t host_counters[2];
    EPI_HIP(hipMemcpyAsync(host.data(), ctx->rescored.ptr, sizeof(Rescored) * host.size(),
                           hipMemcpyDeviceToHost, st));
    EPI_HIP(hipMemcpyAsync(host_counters, counters + 1 + n, sizeof(uint32_t),
                           hipMemcpyDeviceToHost, st));
    EPI_HIP(hipStreamSynchronize(st));
    float kms = 0.0f;
    EPI_HIP(hipEventElapsedTime(&kms, ctx->ev0, ctx->ev1));
    const double wall = ms_since(t_start);
    if (host_counters[0] != 0)
        return EPI_EINTERNAL;  // a workgroup slot was lost

    for (uint32_t b = 0; b < n; ++b) {
        epi_result &R = results[b];
        std::memset(&R, 0, sizeof(R));
        const epi_cohort *co = cohorts[b];
        R.triplets = choose3(co->M);
        R.n_samples = co->N;
        R.n_snps = co->M;
        R.n_controls = co->n[0];
        R.n_cases = co->n[1];
        R.scan_ms = wall;
        R.kernel_ms = kms;
        const Rescored *rs = host.data() + (size_t)b * kTop;
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
    }
    return EPI_OK;
}

}  // namespace epi

extern "C" int epi_scan(epi_ctx *ctx, epi_cohort *cohort, const epi_tile *tile,
                        epi_result *result)
{
    if (!result)
        return EPI_EINVAL;
    epi_cohort *list[1] = {cohort};
    return run_scans(ctx, list, 1, tile, result, nullptr, nullptr);
}

extern "C" int epi_scan_batch(epi_ctx *ctx, epi_cohort *const *cohorts, uint32_t n,
                              const epi_tile *tile, epi_result *results)
{
    if (!results)
        return EPI_EINVAL;
    return run_scans(ctx, cohorts, n, tile, results, nullptr, nullptr);
}

extern "C" int epi_scan_dump(epi_ctx *ctx, epi_cohort *cohort, const epi_tile *tile,
                             float *scores, uint32_t *counts)
{
    if (!scores)
        return EPI_EINVAL;
    epi_cohort *list[1] = {cohort};
    return run_scans(ctx, list, 1, tile, nullptr, scores, counts);
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
    std::lock_guard<std::mutex> lk(ctx->mu);
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
    hipStream_t st = ctx->stream;
    EPI_HIP(hipMemcpyAsync(dc.ptr, cands.data(), sizeof(Cand) * n, hipMemcpyHostToDevice, st));
    EPI_HIP(launch_rescore((const CohortDev *)co->desc.ptr, (const Cand *)dc.ptr, n, 1,
                           (const float *)ctx->t32.ptr, (const double *)ctx->t64.ptr,
                           (Rescored *)dr.ptr, st));
    std::vector<Rescored> host(n);
    EPI_HIP(hipMemcpyAsync(host.data(), dr.ptr, sizeof(Rescored) * n, hipMemcpyDeviceToHost, st));
    EPI_HIP(hipStreamSynchronize(st));
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
