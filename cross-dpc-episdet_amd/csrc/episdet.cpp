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
#include <numeric>
#include <string>
#include <vector>

#include "../../include/episdet.h"
#include "episdet_internal.h"
#include "episdet_lib.h"
#include "episdet_sample.h"

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
    case EPI_EIO: return "file cannot be opened, read or written";
    case EPI_EFORMAT: return "malformed input file";
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

void *DevPool::take(size_t want, size_t &got)
{
    auto it = blocks.lower_bound(want);
    if (it == blocks.end() || it->first > want + want / 4 + ((size_t)1 << 20))
        return nullptr;
    got = it->first;
    void *p = it->second;
    held -= got;
    blocks.erase(it);
    return p;
}

void DevPool::give(void *p, size_t bytes)
{
    if (held + bytes > limit) {
        (void)hipFree(p);
        return;
    }
    blocks.emplace(bytes, p);
    held += bytes;
}

void DevPool::trim()
{
    for (auto &b : blocks)
        (void)hipFree(b.second);
    blocks.clear();
    held = 0;
}

DevBuf::~DevBuf() { release(); }

void DevBuf::release()
{
    if (ptr) {
        if (pool)
            pool->give(ptr, bytes);
        else
            (void)hipFree(ptr);
    }
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
    if (pool) {
        size_t got = 0;
        if (void *p = pool->take(want, got)) {
            ptr = p;
            bytes = got;
            return EPI_OK;
        }
    }
    hipError_t e = hipMalloc(&ptr, want);
    if (e != hipSuccess && pool && pool->held) {
        // the kept blocks may be what is missing
        (void)hipGetLastError();
        pool->trim();
        e = hipMalloc(&ptr, want);
    }
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
    // destroyed cohorts' blocks are kept up to a quarter of the device
    ctx->pool.limit = prop.totalGlobalMem / 4;
    ctx->pair_budget = prop.totalGlobalMem / 8;
    ctx->pair_budget_default = ctx->pair_budget;
    ctx->lds_per_block = prop.sharedMemPerBlock;
    char name[256] = {0};
    if (hipDeviceGetName(name, (int)sizeof(name) - 1, device) == hipSuccess && name[0])
        ctx->name = name;
    else if (prop.name[0])
        ctx->name = prop.name;
    else
        ctx->name = std::string("AMD GPU ") + prop.gcnArchName;
    ctx->arch = prop.gcnArchName;
    if (hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreate(&ctx->ev0) != hipSuccess || hipEventCreate(&ctx->ev1) != hipSuccess) {
        delete ctx;
        return EPI_EDEVICE;
    }
    for (hipEvent_t &e : ctx->stage_ev)
        if (hipEventCreateWithFlags(&e, hipEventDisableTiming) != hipSuccess) {
            delete ctx;
            return EPI_EDEVICE;
        }
    const int ns = noalloc_init(device);
    if (ns != EPI_OK) {
        delete ctx;
        return ns;
    }
    tune_load_default(ctx);
    *out = ctx;
    return EPI_OK;
}

epi_ctx::~epi_ctx()
{
    if (batch_graph)
        (void)hipGraphExecDestroy(batch_graph);
    for (hipEvent_t e : stage_ev)
        if (e)
            (void)hipEventDestroy(e);
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
    E.release();
    F.release();
    perm.release();
    desc.release();
}

void epi_cohort::use_pool(epi::DevPool *p)
{
    for (int c = 0; c < 2; ++c) {
        sm[c].pool = wm[c].pool = pair[c].pool = single[c].pool = p;
    }
    E.pool = F.pool = perm.pool = desc.pool = drows.pool = pslot.pool = todo.pool = p;
}

namespace epi {

// Shapes every kernel can launch and count exactly: SNP indices fit the 21
// bits of a candidate key; the FP4 matrix sums are 4 x a class count in
// FP32, exact below 2^24, so fewer than 2^22 samples; and the genotype
// bytes (staged on the device) stay within 2^38, which keeps every
// one-dimensional launch grid below 2^31 blocks.
bool cohort_shape_ok(uint32_t n_samples, uint32_t n_snps)
{
    return n_samples > 0 && n_samples < kMaxSamples && n_snps >= 3 && n_snps <= kMaxSnps &&
           (uint64_t)n_samples * n_snps <= (1ull << 38);
}

// Device bytes a cohort of this shape may hold when scanned by `variant`
// (planes in two layouts, pair and single margins, the stored SNP order,
// descriptor, and the
// expansion an MFMA scan adds: 2 classes x 2 planes x Mp rows of one pitch,
// the larger class's samples rounded up to a 128-byte stage, at most
// n_samples / 2 + 128 bytes as nibbles and n_samples + 128 as bytes), the
// unit of server admission.
uint64_t cohort_footprint(uint32_t n_samples, uint32_t n_snps, int variant)
{
    const uint64_t Mp = (n_snps + 63) / 64 * 64;
    const uint64_t words = (n_samples + 63) / 64 + 2;  // both classes, rounded up
    const uint64_t N = n_samples;
    const uint64_t expanded = variant == 3 ? 4 * Mp * (N / 2 + 128)
                            : variant == 2 ? 4 * Mp * (N + 128)
                                           : 0;
    return 2 * words * Mp * 16 + 2 * Mp * Mp * 16 + 2 * Mp * 8 + 2 * Mp * 4 + expanded +
           sizeof(CohortDev) + 4096;
}

// The cohort's seed (episdet_sample.h): the 64th best FP32 score (the
// reference's, from the rescoring kernel, as epi_count gives it) of the
// sample triplets.  Every scan may start its cohort threshold there: the
// wavefronts of a short scan then offer and compact only the lanes below it
// instead of everything they meet before their own lists fill.  The
// triplets are unranked on the host, scored and the 64th best selected on
// the device, and the selected score is copied back on `strm`, which the
// caller synchronizes before cohort_seed_take.  Fewer than 64 triplets: no
// seed.
static int cohort_seed_enqueue(epi_ctx *ctx, epi_cohort *co, hipStream_t strm, bool &pending)
{
    pending = false;
    const uint64_t total = choose3(co->M);
    if (total < (uint64_t)kTop)
        return EPI_OK;
    const uint32_t n = (uint32_t)std::min<uint64_t>(total, kSeedSample);
    EPI_TRY(ctx_ensure_tables(ctx, co->N));
    EPI_TRY(ctx->seedc.reserve(sizeof(Cand) * kSeedSample));
    EPI_TRY(ctx->seedr.reserve(sizeof(Rescored) * kSeedSample + 256));
    EPI_TRY(ctx->seedh.reserve(sizeof(Cand) * kSeedSample + 256));
    Cand *hc = (Cand *)ctx->seedh.ptr;
    const SeedSampler smp(total, co->N);
    for (uint32_t t = 0; t < n; ++t) {
        uint32_t x[3];
        colex_triplet(smp.rank(t), co->M, x);
        hc[t].s = 0.0f;
        hc[t].pad = 0;
        hc[t].key = make_key(x[0], x[1], x[2]);
    }
    float *dsel = (float *)((unsigned char *)ctx->seedr.ptr + sizeof(Rescored) * kSeedSample);
    float *hsel = (float *)((unsigned char *)ctx->seedh.ptr + sizeof(Cand) * kSeedSample);
    EPI_HIP(hipMemcpyAsync(ctx->seedc.ptr, hc, sizeof(Cand) * n, hipMemcpyHostToDevice, strm));
    EPI_HIP(launch_rescore((const CohortDev *)co->desc.ptr, (const Cand *)ctx->seedc.ptr, nullptr,
                           n, 1, (const float *)ctx->t32.ptr, (const double *)ctx->t64.ptr,
                           (Rescored *)ctx->seedr.ptr, strm));
    EPI_HIP(launch_seed_select((const Rescored *)ctx->seedr.ptr, n, dsel, strm));
    EPI_HIP(hipMemcpyAsync(hsel, dsel, sizeof(float), hipMemcpyDeviceToHost, strm));
    pending = true;
    return EPI_OK;
}

// After `strm` has been synchronized: the selected score.
static void cohort_seed_take(const epi_ctx *ctx, epi_cohort *co)
{
    co->seed = *(const float *)((const unsigned char *)ctx->seedh.ptr + sizeof(Cand) * kSeedSample);
}

// Caller holds ctx->mu.  Exactly one of geno (sample-major bytes) and bed
// (PLINK 1 .bed body, unpacked on the device) is given.  With geno, any
// phenotype other than 1 is a control (GPU/v4.cpp:141); with bed, samples
// whose phenotype is neither 0 nor 1 are left out, as PLINK leaves out
// samples with a missing phenotype.
int cohort_create_locked(epi_ctx *ctx, const uint8_t *geno, const uint8_t *bed,
                         const uint8_t *pheno, uint32_t n_samples, uint32_t n_snps,
                         epi_cohort **out, hipStream_t strm, bool full_pairs, bool seed)
{
    // class split in sample order (GPU/v4.cpp:135-171): phenotype 1 = case
    std::vector<uint32_t> order;
    order.reserve(n_samples);
    for (uint32_t s = 0; s < n_samples; ++s)
        if (bed ? pheno[s] == 0 : pheno[s] != 1)
            order.push_back(s);
    const uint32_t n0 = (uint32_t)order.size();
    for (uint32_t s = 0; s < n_samples; ++s)
        if (pheno[s] == 1)
            order.push_back(s);
    const uint32_t kept = (uint32_t)order.size();
    if (kept == 0)
        return EPI_EINVAL;

    EPI_HIP(hipSetDevice(ctx->device));
    epi_cohort *co = new (std::nothrow) epi_cohort();
    if (!co)
        return EPI_ENOMEM;
    co->ctx = ctx;
    co->use_pool(&ctx->pool);
    co->N = kept;
    co->M = n_snps;
    co->Mp = (n_snps + 63) / 64 * 64;
    co->n[0] = n0;
    co->n[1] = kept - n0;
    for (int c = 0; c < 2; ++c)
        co->W[c] = (co->n[c] + 63) / 64;

    int st = EPI_OK;
    DevBuf &dgeno = ctx->cgeno, &dorder = ctx->corder, &dbed = ctx->cbed;
    const size_t gbytes = (size_t)n_samples * n_snps;
    const size_t bbytes = bed ? (size_t)((n_samples + 3) / 4) * n_snps : 0;
    st = dgeno.reserve(gbytes);
    if (st == EPI_OK)
        st = dorder.reserve((size_t)kept * 4);
    if (st == EPI_OK && bed)
        st = dbed.reserve(bbytes);
    // pair-margin strips: all of them within the context's budget, else as
    // many slots as the budget holds (at least two: a scan round needs the
    // strips of one j block and one k block)
    const uint32_t nb = co->Mp / 64;
    const uint64_t strip_bytes = 2ull * co->Mp * 64 * sizeof(uint4);  // both classes
    co->nslots = nb;
    if (!full_pairs && strip_bytes * nb > ctx->pair_budget) {
        const uint64_t fit = ctx->pair_budget / strip_bytes;
        co->nslots = (uint32_t)std::min<uint64_t>(std::max<uint64_t>(fit, 2), nb);
    }
    co->blocked = co->nslots < nb;
    co->slot_block.assign(co->nslots, co->blocked ? 0xFFFFFFFFu : 0u);
    std::vector<uint32_t> map(nb, 0);  // blocked: no block resident yet (slot 0 keeps reads in bounds)
    for (uint32_t g = 0; g < nb && !co->blocked; ++g) {
        map[g] = g;
        co->slot_block[g] = g;
    }
    for (int c = 0; c < 2 && st == EPI_OK; ++c) {
        const size_t pb = (size_t)std::max<uint32_t>(co->W[c], 1) * co->Mp * sizeof(uint4);
        st = co->sm[c].reserve(pb);
        if (st == EPI_OK)
            st = co->wm[c].reserve(pb);
        if (st == EPI_OK)
            st = co->pair[c].reserve((size_t)co->nslots * co->Mp * 64 * sizeof(uint4));
        if (st == EPI_OK)
            st = co->single[c].reserve((size_t)co->Mp * sizeof(uint2));
    }
    if (st == EPI_OK)
        st = co->desc.reserve(sizeof(CohortDev));
    if (st == EPI_OK)
        st = co->pslot.reserve((size_t)nb * sizeof(uint32_t));
    if (st == EPI_OK && co->blocked)
        st = co->todo.reserve((size_t)co->nslots * sizeof(uint2));
    hipError_t e = hipSuccess;
    if (st == EPI_OK && bed) {
        e = hipMemcpyAsync(dbed.ptr, bed, bbytes, hipMemcpyHostToDevice, strm);
        if (e == hipSuccess)
            e = launch_bed_unpack((const uint8_t *)dbed.ptr, n_samples, n_snps,
                                  (uint8_t *)dgeno.ptr, strm);
    } else if (st == EPI_OK) {
        e = hipMemcpyAsync(dgeno.ptr, geno, gbytes, hipMemcpyHostToDevice, strm);
    }
    if (st == EPI_OK && e == hipSuccess)
        e = hipMemcpyAsync(dorder.ptr, order.data(), (size_t)kept * 4, hipMemcpyHostToDevice,
                           strm);
    auto pack = [&](const uint32_t *src) {
        return launch_pack((const uint8_t *)dgeno.ptr, n_samples, n_snps, co->Mp, src,
                           (const uint32_t *)dorder.ptr, co->n[0], co->n[1],
                           (uint4 *)co->sm[0].ptr, (uint4 *)co->wm[0].ptr,
                           (uint4 *)co->sm[1].ptr, (uint4 *)co->wm[1].ptr, co->W[0], co->W[1],
                           strm);
    };
    if (st == EPI_OK && e == hipSuccess)
        e = pack(nullptr);
    CohortDev h = co->host_desc();
    if (st == EPI_OK && e == hipSuccess)
        e = hipMemcpyAsync(co->desc.ptr, &h, sizeof(h), hipMemcpyHostToDevice, strm);
    if (st == EPI_OK && e == hipSuccess)
        e = hipMemcpyAsync(co->pslot.ptr, map.data(), (size_t)nb * sizeof(uint32_t),
                           hipMemcpyHostToDevice, strm);
    // the single margins and (within the budget) every pair strip are part
    // of the resident cohort: every scan reads them, none recomputes them
    if (st == EPI_OK && e == hipSuccess)
        e = launch_singles((const CohortDev *)co->desc.ptr, co->Mp, strm);
    // Stored SNP order: by allele frequency (count of allele 0 over both
    // classes, 2 n(g = 0) + n(g = 1); ties by original index), so that the
    // 64 SNPs of a k-block have similar genotype counts.  The planes are
    // packed again in that order; results map back to original indices.
    // Lanes that may reach the top list and whose SNPs are out of original
    // order are summed twice, which costs more than the order saves until
    // each wavefront's threshold has settled (the cohort's seed settles it
    // early): below kOrderMinSnps SNPs the scan is too short for the order
    // to pay, and below kOrderMinSamples samples the counts are small and
    // the lookups conflict little in either order (by default the order is
    // used from there).
    std::vector<uint32_t> perm;
    const bool ordered =
        ctx->snp_order == 1 ||
        (ctx->snp_order == 2 && n_snps >= kOrderMinSnps && kept >= kOrderMinSamples);
    if (st == EPI_OK && e == hipSuccess && ordered)
        st = co->perm.reserve((size_t)2 * co->Mp * sizeof(uint32_t));
    if (st == EPI_OK && e == hipSuccess && ordered) {
        std::vector<uint2> sg((size_t)2 * co->Mp);
        for (int c = 0; c < 2 && e == hipSuccess; ++c)
            e = hipMemcpyAsync(sg.data() + (size_t)c * co->Mp, co->single[c].ptr,
                               (size_t)co->Mp * sizeof(uint2), hipMemcpyDeviceToHost, strm);
        if (e == hipSuccess)
            e = hipStreamSynchronize(strm);
        if (e == hipSuccess) {
            const uint32_t Mp = co->Mp;
            std::vector<uint64_t> key(n_snps);
            for (uint32_t p = 0; p < n_snps; ++p) {
                const uint2 a = sg[p], b = sg[Mp + p];
                key[p] = (uint64_t)(2 * (a.x + b.x) + (a.y + b.y)) << 32 | p;
            }
            std::sort(key.begin(), key.end());
            perm.resize((size_t)2 * Mp);
            for (uint32_t p = 0; p < Mp; ++p)
                perm[p] = p < n_snps ? (uint32_t)key[p] : p;
            for (uint32_t p = 0; p < Mp; ++p)
                perm[Mp + perm[p]] = p;
            co->ordered = true;
            h = co->host_desc();
            e = hipMemcpyAsync(co->perm.ptr, perm.data(), perm.size() * sizeof(uint32_t),
                               hipMemcpyHostToDevice, strm);
        }
        if (e == hipSuccess)
            e = pack((const uint32_t *)co->perm.ptr);
        if (e == hipSuccess)
            e = hipMemcpyAsync(co->desc.ptr, &h, sizeof(h), hipMemcpyHostToDevice, strm);
        if (e == hipSuccess)
            e = launch_singles((const CohortDev *)co->desc.ptr, co->Mp, strm);
    }
    if (st == EPI_OK && e == hipSuccess && !co->blocked)
        e = launch_strips((const CohortDev *)co->desc.ptr, nullptr, nb, co->Mp, strm);
    bool seeding = false;
    if (st == EPI_OK && e == hipSuccess && seed)
        st = cohort_seed_enqueue(ctx, co, strm, seeding);
    if (st == EPI_OK && e == hipSuccess)
        e = hipStreamSynchronize(strm);
    if (st == EPI_OK && e != hipSuccess)
        st = hip_status(e);
    if (st == EPI_OK && seeding)
        cohort_seed_take(ctx, co);
    if (st != EPI_OK) {
        (void)hipStreamSynchronize(strm);  // nothing in flight may use the blocks
        delete co;
        return st;
    }
    co->bytes = 0;
    for (int c = 0; c < 2; ++c)
        co->bytes += co->sm[c].bytes + co->wm[c].bytes + co->pair[c].bytes + co->single[c].bytes;
    co->bytes += co->desc.bytes + co->pslot.bytes + co->todo.bytes + co->perm.bytes;
    *out = co;
    return EPI_OK;
}

}  // namespace epi

extern "C" int epi_cohort_create(epi_ctx *ctx, const uint8_t *geno, const uint8_t *pheno,
                                 uint32_t n_samples, uint32_t n_snps, epi_cohort **out)
{
    if (!ctx || !geno || !pheno || !out || !epi::cohort_shape_ok(n_samples, n_snps))
        return EPI_EINVAL;
    *out = nullptr;
    std::lock_guard<epi::FairMutex> lk(ctx->mu);
    return epi::cohort_create_locked(ctx, geno, nullptr, pheno, n_samples, n_snps, out,
                                     ctx->stream);
}

extern "C" int epi_cohort_create_bed(epi_ctx *ctx, const uint8_t *bed, const uint8_t *pheno,
                                     uint32_t n_samples, uint32_t n_snps, epi_cohort **out)
{
    if (out)
        *out = nullptr;
    if (!ctx || !bed || !pheno || !out || !epi::cohort_shape_ok(n_samples, n_snps))
        return EPI_EINVAL;
    std::lock_guard<epi::FairMutex> lk(ctx->mu);
    return epi::cohort_create_locked(ctx, nullptr, bed, pheno, n_samples, n_snps, out,
                                     ctx->stream);
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
        h.E[c] = E.ptr ? (const uint8_t *)E.ptr + c * (size_t)Mp * 2 * EL : nullptr;
        h.EL[c] = EL;
        h.F[c] = F.ptr ? (const uint8_t *)F.ptr + c * (size_t)Mp * 2 * FL : nullptr;
        h.FL[c] = FL;
        h.W[c] = W[c];
        h.n[c] = n[c];
    }
    h.M = M;
    h.Mp = Mp;
    h.pslot = (const uint32_t *)pslot.ptr;
    h.orig = ordered ? (const uint32_t *)perm.ptr : nullptr;
    h.pos = ordered ? (const uint32_t *)perm.ptr + Mp : nullptr;
    h.seed = seed;
    // class 1's planes as a byte offset from class 0's, when it fits 32 bits
    const size_t e1 = (size_t)Mp * 2 * EL, f1 = (size_t)Mp * 2 * FL;
    h.E1off = e1 < ((size_t)1 << 32) ? (uint32_t)e1 : 0;
    h.F1off = f1 < ((size_t)1 << 32) ? (uint32_t)f1 : 0;
    return h;
}

extern "C" int epi_cohort_destroy(epi_cohort *co)
{
    if (!co)
        return EPI_EINVAL;
    epi_ctx *ctx = co->ctx;
    std::lock_guard<epi::FairMutex> lk(ctx->mu);
    if (co->server_refs != 0)
        return EPI_EBUSY;  // a server request still scans it
    (void)hipSetDevice(ctx->device);
    (void)hipStreamSynchronize(ctx->stream);
    delete co;
    return EPI_OK;
}

extern "C" int epi_cohort_footprint(uint32_t n_samples, uint32_t n_snps, uint64_t *bytes)
{
    if (!bytes || !epi::cohort_shape_ok(n_samples, n_snps))
        return EPI_EINVAL;
    *bytes = epi::cohort_footprint(n_samples, n_snps, 3);
    return EPI_OK;
}

extern "C" int epi_set_pair_budget(epi_ctx *ctx, uint64_t bytes)
{
    if (!ctx)
        return EPI_EINVAL;
    std::lock_guard<epi::FairMutex> lk(ctx->mu);
    ctx->pair_budget = bytes ? bytes : ctx->pair_budget_default;
    return EPI_OK;
}

extern "C" int epi_set_snp_order(epi_ctx *ctx, int mode)
{
    if (!ctx || mode < 0 || mode > 2)
        return EPI_EINVAL;
    std::lock_guard<epi::FairMutex> lk(ctx->mu);
    ctx->snp_order = mode;
    return EPI_OK;
}

extern "C" int epi_cohort_bytes(const epi_cohort *co, uint64_t *bytes)
{
    if (!co || !bytes)
        return EPI_EINVAL;
    *bytes = co->bytes;
    return EPI_OK;
}

