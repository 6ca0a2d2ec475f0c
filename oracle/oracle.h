/*
 * oracle.h -- CPU restatement of the reference's exhaustive third-order
 * epistasis scan with Bayesian K2 scoring.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product links, loads or calls
 * this library; only tests/, __graft_entry__.smoke() and bench.py's CPU
 * baseline leg use it, and only as the checker.
 *
 * Every function cites the reference file:line it restates (paths are
 * relative to the hiperbio/cross-dpc-episdet tree).
 */
#ifndef EPISDET_ORACLE_H
#define EPISDET_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Reference data generator: srand(100); A[s][j] = rand()%3 sample-major;
 * N/2 distinct cases by rejection (GPU/v4.cpp:81-110, CPU/v2.cpp:56-84).
 * A is N*M bytes (row stride M), B is N bytes. */
void orc_ref_generate(int64_t N, int64_t M, uint8_t *A, uint8_t *B);

/* Seeded Hardy-Weinberg generator (no reference counterpart; used for the
 * benchmark configurations).  Sample-major A[N][M], phenotype B[N] with
 * exactly n_cases cases.  maf_out (M doubles) may be NULL. */
void orc_hwe_generate(uint64_t seed, int64_t N, int64_t M, double maf_lo,
                      double maf_hi, int64_t n_cases, uint8_t *A, uint8_t *B,
                      double *maf_out);

/* Same, with one planted three-way interaction (penetrance 0.5 where the
 * genotype sum of the three planted SNPs is odd, 0.05 otherwise), class
 * balanced by acceptance.  planted_out receives the three SNP indices
 * (ascending). */
void orc_planted_generate(uint64_t seed, int64_t N, int64_t M, double maf_lo,
                          double maf_hi, int64_t n_cases, uint8_t *A,
                          uint8_t *B, int64_t planted_out[3]);

/* ln(n!) table with the reference's CPU semantics, T[0..n_max]:
 * float recurrence below 748 (CPU/v2.cpp:179-195, built at :580-583) and
 * the double-precision Stirling form rounded to float from 748 on
 * (CPU/v2.cpp:167-176). */
void orc_log_table(int64_t n_max, float *T);

/* The same function evaluated without float rounding (FP64 everywhere). */
void orc_log_table_f64(int64_t n_max, double *T);

/* Textbook per-sample counting on the uint8 matrix: out[c] for controls
 * (phenotype != 1) and out[27 + c] for cases, c = 9*gi + 3*gj + gk, any
 * genotype other than 0/1 counted as 2 (CPU/v2.cpp:142-160 semantics). */
void orc_count_brute(const uint8_t *A, const uint8_t *B, int64_t N, int64_t M,
                     int64_t i, int64_t j, int64_t k, uint32_t out[54]);

/* Class split + 32-bit bit planes, [snp][word][plane] per class, first
 * sample of a full word at bit 31, last word right-aligned
 * (CPU/v2.cpp:103-164). */
typedef struct {
    int64_t N, M;
    int64_t n[2];  /* class sizes: [0] controls, [1] cases */
    int64_t W[2];  /* 32-bit words per class */
    uint32_t mask[2];
    uint32_t *planes[2]; /* [M][W[c]][2] */
} orc_packed;

int orc_pack(const uint8_t *A, const uint8_t *B, int64_t N, int64_t M,
             orc_packed *out);
void orc_packed_free(orc_packed *p);

/* 27x2 counts from the bit planes: 27 three-way ANDs with g2 = ~(g0|g1)
 * and the padding mask on the last word (CPU/v2.cpp:255-514). */
void orc_count_packed(const orc_packed *p, int64_t i, int64_t j, int64_t k,
                      uint32_t out[54]);

/* K2 score: sum over c = 0..26 in order of
 * ((T[n0+n1+1] - T[n0]) - T[n1]) in FP32, then fabs (CPU/v2.cpp:516-520). */
float orc_k2(const uint32_t c[54], const float *T);
double orc_k2_f64(const uint32_t c[54], const double *T64);

typedef struct {
    float score;
    uint32_t i, j, k;
} orc_cand;

/* Exhaustive scan over i<j<k (CPU/v2.cpp:244-541).  Ties on the FP32 score
 * go to the lexicographically first (i,j,k).  The topk best candidates,
 * ordered by (score, i, j, k), are written to top (topk <= 256).
 * Returns the number of triplets scored. */
int64_t orc_scan(const orc_packed *p, const float *T, int nthreads, int topk,
                 orc_cand *top);

/* Score of every triplet in lexicographic order (rank order), for small M. */
int64_t orc_scan_all_scores(const orc_packed *p, const float *T, int nthreads,
                            float *scores);

/* Combinatorial rank of i<j<k in lexicographic order over C(M,3). */
int64_t orc_rank3(int64_t M, int64_t i, int64_t j, int64_t k);

/* PLINK 1 .bed body (SNP-major), restated from the published PLINK 1
 * binary format; the reference itself has no file input (it generates its
 * data, GPU/v4.cpp:219).  Row j holds (N+3)/4 bytes; sample s sits in bits
 * 2(s%4)..2(s%4)+1 of byte s/4.  decode: 00 -> 2 (homozygous first
 * allele), 01 -> 3 (missing), 10 -> 1, 11 -> 0 into sample-major A[N][M].
 * encode: the inverse, any value other than 0/1/2 -> 01; unused high bits
 * of a row's last byte are 0. */
void orc_bed_decode(const uint8_t *bed, int64_t N, int64_t M, uint8_t *A);
void orc_bed_encode(const uint8_t *A, int64_t N, int64_t M, uint8_t *bed);

/* x86 time-stamp counter (for converting reference cycle counts). */
uint64_t orc_rdtsc(void);

#ifdef __cplusplus
}
#endif
#endif
