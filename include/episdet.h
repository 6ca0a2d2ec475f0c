/*
 * episdet.h -- C ABI of the MI355X-native exhaustive third-order epistasis
 * scan with Bayesian K2 scoring.
 *
 * The reference (hiperbio/cross-dpc-episdet) has no library interface: its
 * GPU path is one program, `./v4 <num_samples> <num_snps> <block_size>`
 * (GPU/README.md:29-30, parsed at GPU/v4.cpp:210-212), whose kernel is a
 * lambda inside main (GPU/v4.cpp:297-593).  Each entry point below names
 * the piece of that program it replaces.  The `v4` command-line program
 * shipped with this library is built on these entry points only.
 *
 * Conventions
 *  - Every function returns 0 (EPI_OK) or a negative EPI_E* code; no C++
 *    exception crosses this boundary.  epi_strerror() names the code.
 *  - The caller owns every host buffer it passes; inputs may be freed as
 *    soon as the call returns.  The library owns device memory until the
 *    matching destroy/close call.
 *  - A context is used by one host thread at a time, except the server
 *    calls (epi_submit / epi_poll / epi_cancel), which are thread-safe.
 *  - Genotypes are uint8 values 0/1/2, sample-major [n_samples][n_snps]
 *    like the reference generator's matrix (GPU/v4.cpp:88-95); any value
 *    other than 0 or 1 counts as genotype 2 (GPU/v4.cpp:156,170).
 *    Phenotype value 1 marks a case; anything else is a control
 *    (GPU/v4.cpp:141).
 *  - Controls are class 0 and cases class 1, so counts[c] and
 *    counts[27 + c] for cell c = 9*g_i + 3*g_j + g_k follow the
 *    reference's frequency table ft[54] (GPU/v4.cpp:317-319).
 */
#ifndef EPISDET_H
#define EPISDET_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define EPI_OK 0
#define EPI_EINVAL (-1)     /* bad argument or shape */
#define EPI_ENOMEM (-2)     /* host or device allocation failed */
#define EPI_EDEVICE (-3)    /* HIP runtime or kernel error */
#define EPI_EBUSY (-4)      /* server stopping / request or cohort in use */
#define EPI_ECANCELED (-5)  /* request was cancelled */
#define EPI_EAGAIN (-6)     /* result not ready yet (epi_poll) */
#define EPI_ENOENT (-7)     /* unknown request id */
#define EPI_EINTERNAL (-8)  /* internal consistency check failed */
#define EPI_EIO (-9)        /* a file cannot be opened, read or written */
#define EPI_EFORMAT (-10)   /* malformed input file */

#define EPI_TOPK 64

typedef struct epi_ctx epi_ctx;
typedef struct epi_cohort epi_cohort;
typedef struct epi_server epi_server;

/* One scored triplet. */
typedef struct {
    uint32_t i, j, k;   /* SNP indices, i < j < k */
    float score;        /* K2 score in FP32, summed in the reference order */
    double score_fp64;  /* the same function evaluated in FP64 */
    uint32_t counts[54];/* controls ft[0..26], cases ft[27..53] */
} epi_candidate;

/* Result of one scan.  Replaces the reference's final host argmin and
 * print (GPU/v4.cpp:598-612); the reported triplet is the FP32 argmin,
 * ties going to the lexicographically first (i,j,k). */
typedef struct {
    uint32_t i, j, k;
    float score;
    double score_fp64;
    uint64_t triplets;        /* C(n_snps, 3) */
    uint32_t n_samples, n_snps;
    uint32_t n_controls, n_cases;
    double scan_ms;           /* host wall time, launch to result on host */
    double kernel_ms;         /* device time of the scan kernel(s) */
    double latency_ms;        /* server requests: submission to result ready */
    double queue_ms;          /* server requests: submission to admission */
    double max_abs_err_fp32;  /* max |score - score_fp64| over top[] */
    int32_t fp64_order_differs; /* 1 if FP64 ranks top[] differently */
    uint32_t n_top;           /* entries of top[] that are filled */
    epi_candidate top[EPI_TOPK]; /* best candidates by (score, i, j, k) */
} epi_result;

/* Tile shape of the scan kernel.  Zero fields are chosen by the library
 * (autotuned table).  Replaces the reference's block_size argument
 * (GPU/v4.cpp:212, nd_range at :277-278). */
typedef struct {
    uint32_t variant;   /* 0 auto, 1 VALU AND+popcount, 2 int8 MFMA, 3 FP4 MFMA */
    uint32_t snps_j;    /* j SNPs per wavefront (1, 2, 4 or 8) */
    uint32_t waves;     /* wavefronts per workgroup (1..16) */
    uint32_t words;     /* 64-bit sample words per loop step (1 or 2) */
} epi_tile;

/* Library / device ---------------------------------------------------- */
const char *epi_version(void);
const char *epi_strerror(int code);

/* Opens device `device` (replaces the SYCL queue creation,
 * GPU/v4.cpp:243-251). */
int epi_open(int device, epi_ctx **ctx);
int epi_close(epi_ctx *ctx);
/* Device name, as the reference prints it ("Selected device:",
 * GPU/v4.cpp:251). */
int epi_device_name(epi_ctx *ctx, char *buf, size_t len);

/* Cohorts -------------------------------------------------------------- */
/* Class split and bit-packing into resident device planes.  Replaces
 * data_to_binary + binary_to_tiled + the USM uploads (GPU/v4.cpp:112-196,
 * :220-230, :256-274).  n_snps >= 3, n_samples >= 1.  Also scores 8192
 * distinct triplets (all of them, when there are fewer) and keeps the 64th
 * best score as the starting threshold of every scan of the cohort; scans
 * give the same results either way. */
int epi_cohort_create(epi_ctx *ctx, const uint8_t *geno, const uint8_t *pheno,
                      uint32_t n_samples, uint32_t n_snps, epi_cohort **out);
/* Waits for the context's queued work, then gives the cohort's device
 * blocks to the context, which keeps them (up to a quarter of device
 * memory) for later cohorts and frees them at epi_close, or earlier when
 * an allocation would otherwise fail.  EPI_EBUSY (nothing freed): a server
 * request on the cohort has not completed (epi_submit_cohort). */
int epi_cohort_destroy(epi_cohort *cohort);
/* Device bytes held by a cohort (its admission footprint). */
int epi_cohort_bytes(const epi_cohort *cohort, uint64_t *bytes);
/* Device bytes a cohort of this shape holds with every pair strip resident
 * and FP4 expanded planes (the server's admission unit), or EPI_EINVAL for
 * a shape no cohort may have: n_samples in [1, 2^22) (the FP4 matrix sums
 * stay exact), n_snps in [3, 2^21 - 1] (21-bit indices), n_samples x
 * n_snps <= 2^38.  Needs no device.  No reference counterpart. */
int epi_cohort_footprint(uint32_t n_samples, uint32_t n_snps, uint64_t *bytes);
/* Most device bytes of pair margins (both classes) a cohort created after
 * this call keeps resident; 0 restores the default, an eighth of device
 * memory.  A cohort whose pair margins exceed it (32 x padded n_snps^2
 * bytes) keeps room for budget / (2048 x padded n_snps) 64-column strips,
 * at least two, and epi_scan / epi_scan_dump then scan it in rounds that
 * recompute the strips each round needs; such a cohort is refused
 * (EPI_EINVAL) by epi_scan_batch, epi_scan_noalloc and the server.  The
 * reference keeps no pair margins (its memory is O(n_snps x n_samples),
 * GPU/v4.cpp:256-257); no counterpart. */
int epi_set_pair_budget(epi_ctx *ctx, uint64_t bytes);
/* Device SNP order of cohorts created after this call: 1 allele
 * frequency, which makes the scan's ln(n!) lookups cheaper on long scans;
 * 0 the original order; 2 (the default) allele frequency for cohorts of
 * at least 512 SNPs and 2048 samples, else original.  Every result (indices, scores,
 * counts, tie order, dumps) is that of the original order bit for bit
 * either way.  The reference keeps the input order (GPU/v4.cpp:112-196);
 * no counterpart. */
int epi_set_snp_order(epi_ctx *ctx, int mode);

/* Scans ---------------------------------------------------------------- */
/* Exhaustive scan of all C(M,3) triplets.  Replaces the block-triple
 * launch loop, the kernel and the host argmin (GPU/v4.cpp:283-608).
 * tile may be NULL. */
int epi_scan(epi_ctx *ctx, epi_cohort *cohort, const epi_tile *tile,
             epi_result *result);

/* epi_scan that allocates no device memory, for callers that cannot
 * allocate (code under HIP graph capture, tight serving loops, harnesses
 * that forbid hipMalloc).  No counterpart in the reference.  Every device
 * buffer beyond the cohort and the result (row table, round tables,
 * counters, candidate slots, ln(n!) tables, rescoring) comes from one
 * static 16 MiB workspace per GPU; a scan needing more returns EPI_EINVAL
 * (about 4 MiB of candidate slots, 12 bytes per sample and 20 bytes per
 * triplet-tile row: roughly n_snps <= 8192 with the VALU tile, more with
 * MFMA tiles).  An MFMA tile is used only when the cohort already holds
 * that variant's expanded planes (an earlier epi_scan with it built them);
 * otherwise the VALU kernel runs.  The result equals epi_scan's bit for
 * bit.  Calls on one GPU share the workspace: calls that overlap in time
 * race. */
int epi_scan_noalloc(epi_ctx *ctx, epi_cohort *cohort, const epi_tile *tile,
                     epi_result *result);

/* Many cohorts in one persistent launch over a shared (cohort, tile) work
 * queue; results[b] is identical to epi_scan of cohorts[b]. */
int epi_scan_batch(epi_ctx *ctx, epi_cohort *const *cohorts, uint32_t n,
                   const epi_tile *tile, epi_result *results);

/* 27x2 contingency counts of explicit triplets (3*n indices in, 54*n
 * counts out), and their FP32/FP64 scores (either may be NULL). */
int epi_count(epi_ctx *ctx, epi_cohort *cohort, const uint32_t *triplets,
              uint32_t n, uint32_t *counts, float *scores, double *scores_fp64);

/* Test hook: runs the scan kernel in dump mode and writes the FP32 score
 * of every triplet at its lexicographic rank (scores has C(M,3) floats),
 * and, if counts is not NULL, its 54 counts (54*C(M,3) uint32).  Only for
 * n_snps <= 512. */
int epi_scan_dump(epi_ctx *ctx, epi_cohort *cohort, const epi_tile *tile,
                  float *scores, uint32_t *counts);

/* Tile autotuning ---------------------------------------------------------- */
/* Times every supported tile shape on a synthetic Hardy-Weinberg cohort of
 * n_samples x min(n_snps, 768) and returns the fastest (no reference
 * counterpart; the reference takes block_size from the command line,
 * GPU/v4.cpp:212).  best_ms may be NULL. */
int epi_autotune(epi_ctx *ctx, uint32_t n_samples, uint32_t n_snps, epi_tile *best,
                 double *best_ms);
/* Tile table used when epi_scan gets no explicit tile: entry e applies to
 * n_samples <= max_samples[e] (entries in increasing order).  epi_open
 * loads tune/<arch>.txt next to the library's directory when it exists
 * (lines "max_samples snps_j waves"). */
int epi_tune_set(epi_ctx *ctx, const uint32_t *max_samples, const epi_tile *tiles,
                 uint32_t n);
int epi_tune_get(epi_ctx *ctx, uint32_t n_samples, epi_tile *tile);

/* Serving ---------------------------------------------------------------- */
/* A request loop on one context (no reference counterpart; the reference
 * scans one generated dataset per process run, GPU/v4.cpp:198-624).
 * Submitted cohorts are packed into HBM when their device footprint fits
 * the budget (FIFO admission); every scheduling round is one persistent
 * scan launch over a share of the remaining tiles of every admitted
 * request, so requests that arrive while others run join at the next
 * round.  Each result equals epi_scan of the same data bit for bit. */
typedef struct {
    uint64_t memory_budget;  /* device bytes for resident cohorts; 0: 1/4 of the device */
    uint32_t max_active;     /* requests on the device at once (1..64); 0: 32 */
    uint32_t max_samples;    /* largest n_samples accepted; 0: 32768 */
    double round_ms;         /* target device time of one round; 0: 20 */
} epi_server_config;

typedef struct {
    uint64_t submitted, completed, cancelled, failed, rounds;
    uint32_t pending, active;
    uint64_t bytes_in_use;
    double device_ms;        /* summed device time of all rounds */
} epi_server_stats;

int epi_server_create(epi_ctx *ctx, const epi_server_config *config, epi_server **out);
/* Cancels what is still queued or running, waits for the threads inside
 * epi_poll to return, then stops the request loop. */
int epi_server_destroy(epi_server *server);
/* Copies the inputs (they may be freed on return) and queues a scan. */
int epi_submit(epi_server *server, const uint8_t *geno, const uint8_t *pheno,
               uint32_t n_samples, uint32_t n_snps, uint64_t *request_id);
/* Queues a scan of a cohort already resident on the server's context
 * (epi_cohort_create*): nothing is uploaded, packed or recomputed.  The
 * cohort stays the caller's; epi_cohort_destroy returns EPI_EBUSY until
 * every request on it has completed.  The result equals epi_scan of the
 * cohort bit for bit.  Like epi_submit, no reference counterpart. */
int epi_submit_cohort(epi_server *server, epi_cohort *cohort, uint64_t *request_id);
/* Waits up to timeout_ms (< 0: forever) for a request.  On success the
 * result is copied out and the request is forgotten.  EPI_EAGAIN: still
 * running; EPI_EBUSY: another thread is polling the same request;
 * EPI_ENOENT: unknown or already collected; EPI_ECANCELED / EPI_ENOMEM /
 * other codes: the request failed (and is forgotten). */
int epi_poll(epi_server *server, uint64_t request_id, int32_t timeout_ms,
             epi_result *result);
int epi_cancel(epi_server *server, uint64_t request_id);
int epi_server_get_stats(epi_server *server, epi_server_stats *stats);

/* On-disk input ----------------------------------------------------------
 * No counterpart in the reference, which only generates its data
 * (GPU/v4.cpp:81-110, called at :219). */

/* Cohort from a PLINK 1 .bed body: n_snps rows (SNP-major) of
 * (n_samples + 3) / 4 bytes; sample t in bits 2(t%4)..2(t%4)+1 of byte t/4,
 * codes 00 homozygous first allele (genotype 2), 10 heterozygous (1),
 * 11 homozygous second allele (0), 01 missing (counts as 2, like any value
 * other than 0/1 in epi_cohort_create).  The body is unpacked on the
 * device.  pheno: 0 control, 1 case; samples with any other value are left
 * out (EPI_EINVAL if none is left). */
int epi_cohort_create_bed(epi_ctx *ctx, const uint8_t *bed, const uint8_t *pheno,
                          uint32_t n_samples, uint32_t n_snps, epi_cohort **out);
/* Reads <prefix>.fam (samples; 6th column 2 -> case 1, 1 -> control 0,
 * anything else -> 0xFF, missing), <prefix>.bim (SNPs) and <prefix>.bed
 * (SNP-major only, magic 6c 1b 01).  Call first with bed = pheno = NULL to
 * learn n_samples and n_snps; then again with bed of
 * n_snps * ((n_samples + 3) / 4) bytes and pheno of n_samples bytes
 * (n_samples and n_snps must still hold the counts the first call gave). */
int epi_read_plink(const char *prefix, uint32_t *n_samples, uint32_t *n_snps,
                   uint8_t *bed, uint8_t *pheno);
/* Writes a PLINK 1 fileset from sample-major genotypes (0 -> 11, 1 -> 10,
 * 2 -> 00, other -> 01 missing) and phenotypes (1 -> 2 case, other -> 1). */
int epi_write_plink(const char *prefix, const uint8_t *geno, const uint8_t *pheno,
                    uint32_t n_samples, uint32_t n_snps);
/* Raw input: geno_path holds n_samples x n_snps uint8 genotypes,
 * sample-major (the reference's A[N][M], GPU/v4.cpp:86-97), pheno_path
 * n_samples uint8 phenotypes (B[N], :99-108); n_snps = file size /
 * n_samples.  Call with geno = pheno = NULL to learn n_snps, then with
 * buffers of n_samples * n_snps and n_samples bytes. */
int epi_read_raw(const char *geno_path, const char *pheno_path, uint32_t n_samples,
                 uint32_t *n_snps, uint8_t *geno, uint8_t *pheno);

/* Data generators ------------------------------------------------------- */
/* The reference generator: srand(100), rand()%3, N/2 cases
 * (GPU/v4.cpp:81-110).  geno is n_samples*n_snps, pheno n_samples. */
int epi_generate_reference(uint32_t n_samples, uint32_t n_snps, uint8_t *geno,
                           uint8_t *pheno);
/* Seeded Hardy-Weinberg data, per-SNP MAF ~ U[maf_lo, maf_hi], exactly
 * n_cases cases.  No reference counterpart. */
int epi_generate_hwe(uint64_t seed, uint32_t n_samples, uint32_t n_snps,
                     double maf_lo, double maf_hi, uint32_t n_cases,
                     uint8_t *geno, uint8_t *pheno);
/* Same with one planted three-way interaction (indices to planted[3]). */
int epi_generate_planted(uint64_t seed, uint32_t n_samples, uint32_t n_snps,
                         double maf_lo, double maf_hi, uint32_t n_cases,
                         uint8_t *geno, uint8_t *pheno, uint32_t planted[3]);

#ifdef __cplusplus
}
#endif
#endif /* EPISDET_H */
