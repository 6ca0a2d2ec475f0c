// noalloc.hip -- epi_scan_noalloc: the scan of epi_scan with every device
// buffer the scan needs beyond its inputs and outputs taken from one
// statically allocated workspace, for callers that cannot allocate device
// memory.  No reference counterpart.
//
// The workspace is `scratch`, a fixed array in device memory that every GPU
// the library runs on holds until the process ends.  Its device address on
// each GPU comes from hipGetSymbolAddress, taken once per device in
// epi_open (noalloc_init) with that device current; host code uses only
// that address, never the array's name.  One call carves these regions out
// of it, each 256-byte aligned:
//   running top-64 list            kTop candidates
//   triplet-tile row table         one Row per row of the cohort
//   round tables                   header, one segment, one cohort, one list
//   queue counters                 4 words (cleared by the round's own fill)
//   workgroup candidate slots      the launch plan's capacity
//   ln(n!) tables                  FP32 and FP64, n_samples + 2 entries each
//   rescoring tables and output    one cohort, kTop results
// Calls on one GPU share the array and race when they overlap in time;
// each call leaves its contents behind, and every region a call reads is
// written (or cleared by a kernel) earlier in the same call.
#include <hip/hip_runtime.h>

#include <chrono>
#include <cstring>
#include <mutex>
#include <vector>

#include "../../include/episdet.h"
#include "episdet_internal.h"
#include "episdet_lib.h"

namespace {

constexpr size_t kScratchBytes = (size_t)16 << 20;  // per device
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

constexpr int kMaxDevices = 64;
std::mutex g_mu;
unsigned char *g_scratch[kMaxDevices];  // device address of `scratch`, per device

// Consecutive 256-byte aligned regions of the workspace.
struct Carve {
    unsigned char *base;
    size_t off = 0;
    bool over = false;
    explicit Carve(unsigned char *b) : base(b) {}
    template <class T>
    T *take(size_t bytes)
    {
        const size_t at = (off + 255) & ~(size_t)255;
        if (at + bytes > kScratchBytes) {
            over = true;
            return nullptr;
        }
        off = at + bytes;
        return (T *)(base + at);
    }
};

}  // namespace

namespace epi {

// With `device` current: its address of the workspace (once per device).
int noalloc_init(int device)
{
    if (device < 0 || device >= kMaxDevices)
        return EPI_EINVAL;
    std::lock_guard<std::mutex> lk(g_mu);
    if (g_scratch[device])
        return EPI_OK;
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
        return e != hipSuccess ? hip_status(e) : EPI_EDEVICE;
    g_scratch[device] = (unsigned char *)p;
    return EPI_OK;
}

}  // namespace epi

using namespace epi;

extern "C" int epi_scan_noalloc(epi_ctx *ctx, epi_cohort *co, const epi_tile *tile,
                                epi_result *result)
{
    // a cohort past its pair budget needs its strips recomputed per round
    if (!ctx || !co || !result || co->ctx != ctx || (tile && tile->variant > 3) || co->blocked)
        return EPI_EINVAL;
    unsigned char *base = nullptr;
    {
        std::lock_guard<std::mutex> lk(g_mu);
        if (ctx->device >= 0 && ctx->device < kMaxDevices)
            base = g_scratch[ctx->device];
    }
    if (!base)
        return EPI_EINTERNAL;  // epi_open takes the address
    std::lock_guard<epi::FairMutex> lk(ctx->mu);
    EPI_HIP(hipSetDevice(ctx->device));
    // an MFMA tile needs the cohort's expanded planes, which only an earlier
    // scan with that variant builds; without them the VALU kernel runs (the
    // results are the same bit for bit)
    ScanConfig cfg = resolve_config(ctx, tile, co->N);
    if (cfg.variant >= 2 && !(cfg.variant == 3 ? co->F.ptr : co->E.ptr)) {
        epi_tile valu;
        std::memset(&valu, 0, sizeof(valu));
        valu.variant = 1;
        cfg = resolve_config(ctx, &valu, co->N);
    }
    const uint32_t table_len = co->N + 2;  // indices up to n0 + n1 + 1 = N + 1
    LaunchPlan plan;
    EPI_TRY(plan_launch(ctx, cfg, table_len, false, plan));
    const auto t0 = std::chrono::steady_clock::now();

    // host images (host memory only)
    Job job;
    job.co = co;
    job.nw = 1;
    build_rows(0, co->M, co->Mp, cfg.snps_j, 1, job.rows, job.items);
    if (job.items >= 0x7FFFFFFFull)
        return EPI_EINVAL;
    if (co->ordered)
        shuffle_rows(job.rows);
    std::vector<float> t32;
    std::vector<double> t64;
    build_log_tables(table_len - 1, t32, t64);

    // the workspace
    Carve cv(base);
    Cand *list = cv.take<Cand>(sizeof(Cand) * kTop);
    Row *drows = cv.take<Row>(sizeof(Row) * std::max<size_t>(job.rows.size(), 1));
    job.list = list;
    job.drows = drows;
    Job *jp[1] = {&job};
    const std::pair<size_t, size_t> all(0, job.rows.size());
    std::vector<unsigned char> round_img;
    RoundLayout L;
    uint64_t items = 0;
    EPI_TRY(build_round(jp, &all, 1, 1, round_img, L, items));
    RoundBufs B;
    B.round = cv.take<unsigned char>(L.total);
    B.counters = cv.take<uint32_t>(sizeof(uint32_t) * 4);
    B.slots = cv.take<Cand>(sizeof(Cand) * plan.cap);
    float *dt32 = cv.take<float>(sizeof(float) * t32.size());
    double *dt64 = cv.take<double>(sizeof(double) * t64.size());
    B.t32 = dt32;
    CohortDev *rdesc = cv.take<CohortDev>(sizeof(CohortDev));
    Cand **rlists = cv.take<Cand *>(sizeof(Cand *));
    Rescored *rescored = cv.take<Rescored>(sizeof(Rescored) * kTop);
    if (cv.over)
        return EPI_EINVAL;  // the scan needs more than the workspace holds
    const CohortDev desc = co->host_desc();

    // uploads from host memory, then the kernels, all on the context's stream
    hipStream_t s = ctx->stream;
    EPI_HIP(hipMemcpyAsync(drows, job.rows.data(), sizeof(Row) * job.rows.size(),
                           hipMemcpyHostToDevice, s));
    EPI_HIP(hipMemcpyAsync(B.round, round_img.data(), L.total, hipMemcpyHostToDevice, s));
    EPI_HIP(hipMemcpyAsync(dt32, t32.data(), sizeof(float) * t32.size(), hipMemcpyHostToDevice, s));
    EPI_HIP(hipMemcpyAsync(dt64, t64.data(), sizeof(double) * t64.size(), hipMemcpyHostToDevice,
                           s));
    EPI_HIP(hipMemcpyAsync(rdesc, &desc, sizeof(desc), hipMemcpyHostToDevice, s));
    EPI_HIP(hipMemcpyAsync(rlists, &list, sizeof(list), hipMemcpyHostToDevice, s));
    EPI_HIP(launch_fill_empty(list, kTop, s));
    EPI_TRY(enqueue_round_kernels(ctx, plan, 1, L, B, items > 0, true, nullptr, nullptr));
    EPI_HIP(launch_rescore(rdesc, nullptr, rlists, kTop, 1, dt32, dt64, rescored, s));
    std::vector<Rescored> host(kTop);
    uint32_t overflow = 0;
    EPI_HIP(hipMemcpyAsync(host.data(), rescored, sizeof(Rescored) * kTop, hipMemcpyDeviceToHost,
                           s));
    EPI_HIP(hipMemcpyAsync(&overflow, B.counters + 2, sizeof(uint32_t), hipMemcpyDeviceToHost, s));
    EPI_HIP(hipStreamSynchronize(s));  // the host images above are read until here
    const double wall =
        std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    float kms = 0.0f;
    EPI_HIP(hipEventElapsedTime(&kms, ctx->ev0, ctx->ev1));
    if (overflow != 0)
        return EPI_EINTERNAL;  // a workgroup slot was lost
    return assemble(host.data(), co, *result, wall, kms);
}
