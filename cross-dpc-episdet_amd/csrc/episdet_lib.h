// episdet_lib.h -- host-side objects behind the opaque C ABI handles.
#pragma once
#include <hip/hip_runtime.h>

#include <mutex>
#include <string>
#include <vector>

#include "../../include/episdet.h"
#include "episdet_internal.h"

namespace epi {

// Owned device allocation that only grows.
struct DevBuf {
    void *ptr = nullptr;
    size_t bytes = 0;
    DevBuf() = default;
    DevBuf(const DevBuf &) = delete;
    DevBuf &operator=(const DevBuf &) = delete;
    ~DevBuf();
    int reserve(size_t want);
    void release();
};

int hip_status(hipError_t e);
void build_log_tables(uint32_t n_max, std::vector<float> &t32, std::vector<double> &t64);
int ctx_ensure_tables(epi_ctx *ctx, uint32_t n_samples);
ScanConfig resolve_config(const epi_ctx *ctx, const epi_tile *tile, uint32_t n_samples);
void build_rows(uint32_t cohort, uint32_t M, uint32_t Mp, int jp, int nw,
                std::vector<Row> &rows, uint64_t &items);
int run_scans(epi_ctx *ctx, epi_cohort *const *cohorts, uint32_t n, const epi_tile *tile,
              epi_result *results, float *dump_scores, uint32_t *dump_counts);

}  // namespace epi

struct epi_ctx {
    int device = 0;
    int cu_count = 0;
    size_t lds_per_block = 65536;
    std::string name, arch;
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    std::mutex mu;
    // ln(n!) tables (FP32 reference semantics, FP64 twin), table_len entries
    uint32_t table_len = 0;
    epi::DevBuf t32, t64;
    // per-scan scratch, grown on demand
    epi::DevBuf descs, rows, counters, slots, merged, rescored;
    ~epi_ctx();
};

struct epi_cohort {
    epi_ctx *ctx = nullptr;
    uint32_t N = 0, M = 0, Mp = 0;
    uint32_t n[2] = {0, 0};
    uint32_t W[2] = {0, 0};
    epi::DevBuf sm[2], wm[2], pair[2], single[2], desc;
    uint64_t bytes = 0;
    epi::CohortDev host_desc() const;
    ~epi_cohort();
};
