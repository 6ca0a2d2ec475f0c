// episdet_internal.h -- structures shared by the HIP kernels and the host
// library.  Not part of the public ABI (include/episdet.h is).
#pragma once
#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

namespace epi {

constexpr int kLanes = 64;     // CDNA wavefront
constexpr int kTop = 64;       // candidates kept per cohort
constexpr int kWaveCap = 128;  // per-wavefront LDS candidate buffer
constexpr int kMfmaWaves = 8;  // wavefronts per workgroup of the MFMA scan kernel (2 per SIMD)
constexpr uint32_t kMaxSamples = 1u << 22;    // n_samples below this (exact FP4 sums)
constexpr uint32_t kMaxSnps = (1u << 21) - 1;  // n_snps at most this (21-bit key fields)
constexpr uint32_t kOrderMinSnps = 512;      // allele-frequency SNP order by default from this many
constexpr uint32_t kOrderMinSamples = 2048;  // SNPs and samples on

// Candidate: FP32 score and a lexicographic key (i << 42 | j << 21 | k).
struct Cand {
    float s;
    uint32_t pad;
    uint64_t key;
};

__host__ __device__ inline uint64_t make_key(uint32_t i, uint32_t j, uint32_t k)
{
    return ((uint64_t)i << 42) | ((uint64_t)j << 21) | (uint64_t)k;
}
__host__ __device__ inline uint32_t key_i(uint64_t key) { return (uint32_t)(key >> 42); }
__host__ __device__ inline uint32_t key_j(uint64_t key) { return (uint32_t)((key >> 21) & 0x1FFFFF); }
__host__ __device__ inline uint32_t key_k(uint64_t key) { return (uint32_t)(key & 0x1FFFFF); }
constexpr uint64_t kEmptyKey = ~0ull;

// Resident device layout of one cohort.  Class 0 = controls, 1 = cases.
// A 64-bit sample word of one SNP is stored as uint4 {g0.lo, g0.hi, g1.lo,
// g1.hi}; sample t of a class sits at bit t % 64 of word t / 64, padding
// bits are zero in both planes.
struct CohortDev {
    const uint4 *sm[2];   // SNP-major  [Mp][W[c]]   (uniform loads)
    const uint4 *wm[2];   // word-major [W[c]][Mp]   (per-lane loads)
    // pair margins {n00, n01, n10, n11} of (p, q), every p < Mp, in 64-column
    // strips: column block q / 64 lives in strip slot pslot[q / 64] as
    // [Mp rows][64], so (p, q) is at (pslot[q / 64] Mp + p) 64 + q % 64.
    // A cohort holds every strip (pslot[g] = g) or, past its pair budget, a
    // few slots that a scan refills per round (pair_at()).
    uint4 *pair[2];
    uint2 *single[2];     // [Mp]: {count g0, count g1}
    const uint8_t *E[2];  // int8 MFMA variant: 0/1 bytes [Mp][2 planes][EL[c]], or null
    uint32_t EL[2];       // bytes per E row, both classes: 64 max(W) rounded up to 128, zero padded
    const uint8_t *F[2];  // FP4 MFMA variant: e2m1 nibbles 0 / 1.0 [Mp][2 planes][FL[c]], or null
    uint32_t FL[2];       // bytes per F row, both classes: 32 max(W) rounded up to 128, zero padded
    uint32_t E1off, F1off;  // E[1] - E[0], F[1] - F[0] in bytes (0: too far for 32 bits)
    uint32_t W[2];        // 64-bit words per class
    uint32_t n[2];        // class sizes
    uint32_t M, Mp;       // SNPs, padded to a multiple of 64
    const uint32_t *pslot;  // [Mp / 64] strip slot of each column block
    // Stored SNP order: row p of every array above holds original SNP
    // orig[p] (p < M; orig[p] = p for the padding rows), and original SNP s
    // is stored at row pos[s].  Both null: rows are in original order.
    const uint32_t *orig;   // [Mp]
    const uint32_t *pos;    // [Mp]
    // An upper bound of the cohort's 64th best score, set at creation (the
    // 64th best FP32 score of a sample of distinct triplets; +inf: none).
    float seed;
};

// Offset of pair (p, q) in a class's strips, given the slot of q's block.
__host__ __device__ inline size_t pair_at(uint32_t slot, uint32_t Mp, uint32_t p, uint32_t q)
{
    return ((size_t)slot * Mp + p) * 64 + (q & 63u);
}

// One (k-block, j-group) row of a cohort's triplet tiles.  Items of the
// row are consecutive chunks of `waves` i values, i in [0, icount); item
// numbers are local to the cohort's row table.
struct Row {
    uint32_t cohort;      // unused by the kernel (segments carry the cohort)
    uint32_t kb;          // k in [64 kb, 64 kb + 64)
    uint32_t jg;          // j in [snps_j * jg, snps_j * jg + snps_j)
    uint32_t icount;
    uint32_t item_start;  // first item of this row in the cohort's numbering
    uint32_t pad;
};

// A contiguous range of one cohort's rows inside a launch.  Launch items
// [item_base, item_base + n_items) map to the cohort's items starting at
// item_first = rows[row_begin].item_start.
struct Segment {
    const Row *rows;      // the cohort's resident row table
    uint32_t cohort;      // index into the launch's cohort descriptors
    uint32_t row_begin, row_end;
    uint32_t item_base;
    uint32_t item_first;
    uint32_t n_items;
};

// Work of one launch, read by the kernels from device memory so that a
// captured graph can be replayed with new work.
struct RoundHdr {
    uint32_t n_segments;
    uint32_t n_items;
    uint32_t n_cohorts;
    uint32_t pad;
};

struct ScanParams {
    const RoundHdr *hdr;
    const CohortDev *cohorts;  // [n_cohorts]
    const Segment *segs;       // [n_segments]
    uint32_t *counters;        // [0] item queue head, [1 + c] candidates appended for cohort c
    uint32_t *thr;             // [c] score threshold of cohort c (float bits)
    const float *table;        // ln(n!) FP32, table_len entries
    uint32_t table_len;
    uint32_t slot_cap;         // candidate capacity per cohort
    Cand *slots;               // [cohort][slot_cap] appended candidates
    uint32_t *overflow;        // set to 1 if a candidate could not be stored
    float *dump_scores;        // dump mode: score per triplet rank (cohort 0)
    uint32_t *dump_counts;     // dump mode: 54 counts per triplet rank
};

// Rescoring output for one candidate.
struct Rescored {
    uint32_t i, j, k, valid;
    float s32;
    float s32_scan;
    double s64;
    uint32_t counts[54];
    uint32_t pad[2];
};

// Scan kernel configuration (template instance selector).
struct ScanConfig {
    int variant = 1;  // 1 VALU AND+popcount, 2 int8 MFMA, 3 FP4 MFMA
    int snps_j = 4;   // VALU: 1, 2, 4, 8 j SNPs per wavefront; MFMA: 64 (j-block)
    int waves = 8;    // wavefronts per workgroup
    int words = 1;    // 64-bit sample words per loop step
};

// Launch wrappers (episdet_kernels.hip).
hipError_t launch_fill_u32(uint32_t *p, uint32_t v, size_t n, hipStream_t s);
// Row p of the planes holds genotype column src[p] (src null: column p).
hipError_t launch_pack(const uint8_t *geno, uint32_t N, uint32_t M, uint32_t Mp,
                       const uint32_t *src, const uint32_t *order, uint32_t n0, uint32_t n1,
                       uint4 *sm0, uint4 *wm0, uint4 *sm1, uint4 *wm1,
                       uint32_t W0, uint32_t W1, hipStream_t s);
hipError_t launch_bed_unpack(const uint8_t *bed, uint32_t N, uint32_t M, uint8_t *geno,
                             hipStream_t s);
// Pair-margin strips todo[t] = (column block, slot), t < n_todo (todo
// null: strip t = block t in slot t).
hipError_t launch_strips(const CohortDev *dev_cohort, const uint2 *todo, uint32_t n_todo,
                         uint32_t Mp, hipStream_t s);
hipError_t launch_singles(const CohortDev *dev_cohort, uint32_t Mp, hipStream_t s);
bool scan_config_supported(const ScanConfig &cfg);
// MFMA variants (episdet_mfma.hip); fp4 selects the FP4 kernel (variant 3)
hipError_t launch_expand(const uint4 *sm, uint32_t Mp, uint32_t W, uint32_t EL, uint8_t *E,
                         hipStream_t s);
hipError_t launch_expand_fp4(const uint4 *sm, uint32_t Mp, uint32_t W, uint32_t FL, uint8_t *F,
                             hipStream_t s);
size_t mfma_lds_bytes(bool table_lds, uint32_t table_len);
hipError_t mfma_occupancy(bool fp4, bool table_lds, bool dump, size_t lds_bytes,
                          int *blocks_per_cu);
hipError_t mfma_allow_lds(bool fp4, bool table_lds, bool dump, size_t lds_bytes);
hipError_t launch_mfma_scan(bool fp4, bool table_lds, bool dump, const ScanParams &p,
                            uint32_t grid, hipStream_t s);
hipError_t scan_occupancy(const ScanConfig &cfg, bool table_lds, bool dump,
                          size_t lds_bytes, int *blocks_per_cu);
size_t scan_lds_bytes(const ScanConfig &cfg, bool table_lds, uint32_t table_len);
hipError_t scan_allow_lds(const ScanConfig &cfg, bool table_lds, bool dump, size_t lds_bytes);
hipError_t launch_scan(const ScanConfig &cfg, bool table_lds, bool dump,
                       const ScanParams &p, uint32_t grid, hipStream_t s);
hipError_t launch_fill_empty(Cand *p, size_t n, hipStream_t s);
// Folds the workgroup slots of each cohort of the round into its running
// top-kTop list lists[c] (read and written).
hipError_t launch_merge(const RoundHdr *hdr, const Cand *slots, const uint32_t *counters,
                        const uint32_t *thr, uint32_t slot_cap, Cand *const *lists,
                        uint32_t max_cohorts, hipStream_t s);
hipError_t launch_init_thr(const RoundHdr *hdr, const CohortDev *cohorts, Cand *const *lists,
                           uint32_t *thr,
                           uint32_t max_cohorts, hipStream_t s);
// Rescores per_cohort candidates of each cohort: cands given either as one
// array [n_cohorts][per_cohort] or (lists != nullptr) as per-cohort lists.
// The kTop-th smallest s32 of rs[0..n) (kTop <= n <= 8192) into *out.
hipError_t launch_seed_select(const Rescored *rs, uint32_t n, float *out, hipStream_t s);
hipError_t launch_rescore(const CohortDev *cohorts, const Cand *cands,
                          Cand *const *lists, uint32_t per_cohort, uint32_t n_cohorts,
                          const float *t32, const double *t64, Rescored *out,
                          hipStream_t s);

}  // namespace epi
