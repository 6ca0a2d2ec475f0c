/*
 * oracle.c -- CPU restatement of the reference's exhaustive third-order
 * epistasis scan with Bayesian K2 scoring (hiperbio/cross-dpc-episdet).
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the HIP product.  The product
 * never links or loads this file.  Each function names the reference
 * file:line it restates.  Parity pins (what checks each function):
 *   orc_ref_generate     -- known-answer runs of the compiled reference
 *                           (best triplet + %f score) on its own generator.
 *   orc_count_brute      -- textbook definition; cross-checked against
 *                           orc_count_packed and count invariants.
 *   orc_count_packed     -- orc_count_brute on random inputs incl. odd
 *                           class sizes (padding masks), empty classes.
 *   orc_log_table        -- lgamma closed form (error bounds), and the
 *                           reference binaries' end-to-end scores on data
 *                           whose cell counts reach the Stirling branch.
 *   orc_k2 / orc_scan    -- reference known-answer scores (bit-exact %f) and
 *                           brute-force rescans on tiny inputs.
 *   orc_hwe_generate / orc_planted_generate -- no reference counterpart;
 *                           pinned by distributional checks (HWE genotype
 *                           frequencies, exact class sizes, planted triplet
 *                           recovery).
 *   orc_bed_decode / orc_bed_encode -- no reference counterpart (published
 *                           PLINK 1 format); pinned by hand-decoded
 *                           known-answer bytes, encode/decode round trips
 *                           and the sample-major scan of the same cohort.
 *
 * Build: see oracle/Makefile (gcc, -ffp-contract=off, OpenMP).
 */
#include "oracle.h"

#include <float.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define ORC_TABLE_MAX 748      /* GPU/v4.cpp:15, CPU/v2.cpp:12 */
#define ORC_TABLE_ERROR -0.0810 /* GPU/v4.cpp:16, CPU/v2.cpp:13 */

/* ------------------------------------------------------------------ */
/* Reference generator (GPU/v4.cpp:81-110; CPU/v2.cpp:56-84).           */
/* ------------------------------------------------------------------ */
void orc_ref_generate(int64_t N, int64_t M, uint8_t *A, uint8_t *B)
{
    srand(100);
    for (int64_t s = 0; s < N; s++)
        for (int64_t j = 0; j < M; j++)
            A[s * M + j] = (uint8_t)(rand() % 3);
    memset(B, 0, (size_t)N);
    /* The reference can draw index == N when rand() returns RAND_MAX and
     * then touches B[N]; that slot is modelled as a phantom sample. */
    uint8_t phantom = 0;
    for (int64_t c = 0; c < N / 2; c++) {
        int index = (int)(N * ((double)rand() / (RAND_MAX)));
        while ((index < N ? B[index] : phantom) == 1)
            index = (int)(N * ((double)rand() / (RAND_MAX)));
        if (index < N)
            B[index] = 1;
        else
            phantom = 1;
    }
}

/* ------------------------------------------------------------------ */
/* Seeded generators (no reference counterpart).                        */
/* SplitMix64 (Steele, Lea, Flood 2014) seeds xoshiro256** (Blackman,   */
/* Vigna 2018).                                                         */
/* ------------------------------------------------------------------ */
static uint64_t sm64(uint64_t *s)
{
    uint64_t z = (*s += 0x9E3779B97F4A7C15ULL);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}

typedef struct {
    uint64_t s[4];
} xo256;

static void xo_seed(xo256 *x, uint64_t seed)
{
    uint64_t s = seed;
    for (int t = 0; t < 4; t++)
        x->s[t] = sm64(&s);
}

static inline uint64_t rotl64(uint64_t v, int k) { return (v << k) | (v >> (64 - k)); }

static uint64_t xo_next(xo256 *x)
{
    uint64_t *s = x->s;
    uint64_t r = rotl64(s[1] * 5, 7) * 9;
    uint64_t t = s[1] << 17;
    s[2] ^= s[0];
    s[3] ^= s[1];
    s[1] ^= s[2];
    s[0] ^= s[3];
    s[2] ^= t;
    s[3] = rotl64(s[3], 45);
    return r;
}

static double xo_u01(xo256 *x) { return (double)(xo_next(x) >> 11) * 0x1.0p-53; }

static uint64_t stream_seed(uint64_t seed, uint64_t salt, uint64_t idx)
{
    uint64_t h = seed ^ salt;
    h ^= (idx + 1) * 0xD1B54A32D192ED03ULL;
    return h;
}

#define SALT_MAF 0x4D41465F53545245ULL
#define SALT_SAMPLE 0x53414D504C455321ULL
#define SALT_PHEN 0x50484E5F53545245ULL
#define SALT_PLANT 0x504C414E54454421ULL

static inline uint8_t hwe_draw(xo256 *x, double p)
{
    double u = xo_u01(x);
    double q = 1.0 - p;
    double c0 = q * q;
    double c1 = c0 + (2.0 * p) * q;
    return (uint8_t)(u < c0 ? 0 : (u < c1 ? 1 : 2));
}

static void draw_mafs(uint64_t seed, int64_t M, double lo, double hi, double *maf)
{
    xo256 x;
    xo_seed(&x, stream_seed(seed, SALT_MAF, 0));
    for (int64_t j = 0; j < M; j++)
        maf[j] = lo + (hi - lo) * xo_u01(&x);
}

void orc_hwe_generate(uint64_t seed, int64_t N, int64_t M, double maf_lo,
                      double maf_hi, int64_t n_cases, uint8_t *A, uint8_t *B,
                      double *maf_out)
{
    double *maf = (double *)malloc(sizeof(double) * (size_t)(M > 0 ? M : 1));
    draw_mafs(seed, M, maf_lo, maf_hi, maf);
#pragma omp parallel for schedule(static)
    for (int64_t s = 0; s < N; s++) {
        xo256 x;
        xo_seed(&x, stream_seed(seed, SALT_SAMPLE, (uint64_t)s));
        for (int64_t j = 0; j < M; j++)
            A[s * M + j] = hwe_draw(&x, maf[j]);
    }
    /* exactly n_cases cases: partial Fisher-Yates over sample indices */
    memset(B, 0, (size_t)N);
    int64_t *perm = (int64_t *)malloc(sizeof(int64_t) * (size_t)(N > 0 ? N : 1));
    for (int64_t s = 0; s < N; s++)
        perm[s] = s;
    xo256 x;
    xo_seed(&x, stream_seed(seed, SALT_PHEN, 0));
    if (n_cases > N)
        n_cases = N;
    for (int64_t t = 0; t < n_cases; t++) {
        int64_t r = t + (int64_t)(xo_next(&x) % (uint64_t)(N - t));
        int64_t tmp = perm[t];
        perm[t] = perm[r];
        perm[r] = tmp;
        B[perm[t]] = 1;
    }
    if (maf_out)
        memcpy(maf_out, maf, sizeof(double) * (size_t)M);
    free(perm);
    free(maf);
}

void orc_planted_generate(uint64_t seed, int64_t N, int64_t M, double maf_lo,
                          double maf_hi, int64_t n_cases, uint8_t *A,
                          uint8_t *B, int64_t planted_out[3])
{
    double *maf = (double *)malloc(sizeof(double) * (size_t)M);
    draw_mafs(seed, M, maf_lo, maf_hi, maf);
    xo256 px;
    xo_seed(&px, stream_seed(seed, SALT_PLANT, 0));
    int64_t pl[3];
    int got = 0;
    while (got < 3) {
        int64_t c = (int64_t)(xo_next(&px) % (uint64_t)M);
        int dup = 0;
        for (int t = 0; t < got; t++)
            dup |= (pl[t] == c);
        if (!dup)
            pl[got++] = c;
    }
    /* ascending */
    for (int a = 0; a < 3; a++)
        for (int b = a + 1; b < 3; b++)
            if (pl[b] < pl[a]) {
                int64_t t = pl[a];
                pl[a] = pl[b];
                pl[b] = t;
            }
    for (int t = 0; t < 3; t++) {
        maf[pl[t]] = 0.3;
        planted_out[t] = pl[t];
    }
    int64_t need[2] = {N - n_cases, n_cases};
    int64_t filled = 0;
    uint8_t *row = (uint8_t *)malloc((size_t)M);
    for (uint64_t cand = 0; filled < N && cand < (uint64_t)N * 4096u; cand++) {
        xo256 x;
        xo_seed(&x, stream_seed(seed, SALT_SAMPLE, cand));
        for (int64_t j = 0; j < M; j++)
            row[j] = hwe_draw(&x, maf[j]);
        int odd = (row[pl[0]] + row[pl[1]] + row[pl[2]]) & 1;
        double pen = odd ? 0.5 : 0.05;
        int cls = xo_u01(&x) < pen ? 1 : 0;
        if (need[cls] > 0) {
            need[cls]--;
            memcpy(A + filled * M, row, (size_t)M);
            B[filled] = (uint8_t)cls;
            filled++;
        }
    }
    free(row);
    free(maf);
}

/* ------------------------------------------------------------------ */
/* ln(n!) (CPU/v2.cpp:167-195, table built at :580-583).                */
/* ------------------------------------------------------------------ */
void orc_log_table(int64_t n_max, float *T)
{
    for (int64_t n = 0; n <= n_max; n++) {
        if (n < ORC_TABLE_MAX) {
            /* my_factorial: z = table[n-1] + log(n), stored as float */
            if (n < 2)
                T[n] = 0.0f;
            else
                T[n] = (float)((double)T[n - 1] + log((double)n));
        } else {
            /* addlog beyond the table: Stirling in double, returned as float */
            double nd = (double)n;
            double x = (nd + 0.5) * log(nd) - (double)(n - 1) * log(exp(1.0)) +
                       ORC_TABLE_ERROR;
            T[n] = (float)x;
        }
    }
}

void orc_log_table_f64(int64_t n_max, double *T)
{
    for (int64_t n = 0; n <= n_max; n++) {
        if (n < ORC_TABLE_MAX) {
            T[n] = n < 2 ? 0.0 : T[n - 1] + log((double)n);
        } else {
            double nd = (double)n;
            T[n] = (nd + 0.5) * log(nd) - (double)(n - 1) * log(exp(1.0)) +
                   ORC_TABLE_ERROR;
        }
    }
}

/* ------------------------------------------------------------------ */
/* Counting.                                                            */
/* ------------------------------------------------------------------ */
static inline int geno3(uint8_t g) { return g == 0 ? 0 : (g == 1 ? 1 : 2); }

void orc_count_brute(const uint8_t *A, const uint8_t *B, int64_t N, int64_t M,
                     int64_t i, int64_t j, int64_t k, uint32_t out[54])
{
    memset(out, 0, 54 * sizeof(uint32_t));
    for (int64_t s = 0; s < N; s++) {
        const uint8_t *r = A + s * M;
        int cell = 9 * geno3(r[i]) + 3 * geno3(r[j]) + geno3(r[k]);
        out[(B[s] == 1 ? 27 : 0) + cell]++;
    }
}

int orc_pack(const uint8_t *A, const uint8_t *B, int64_t N, int64_t M,
             orc_packed *p)
{
    memset(p, 0, sizeof(*p));
    p->N = N;
    p->M = M;
    for (int64_t s = 0; s < N; s++)
        p->n[B[s] == 1 ? 1 : 0]++;
    for (int c = 0; c < 2; c++) {
        p->W[c] = (p->n[c] + 31) / 32;
        uint32_t rem = (uint32_t)(p->W[c] * 32 - p->n[c]);
        p->mask[c] = 0xFFFFFFFFu >> rem;
        size_t words = (size_t)(M * (p->W[c] > 0 ? p->W[c] : 1) * 2);
        p->planes[c] = (uint32_t *)calloc(words, sizeof(uint32_t));
        if (!p->planes[c])
            return -1;
    }
    /* transposed_to_binary: shift-in in sample order within each class */
    for (int64_t snp = 0; snp < M; snp++) {
        int64_t cnt[2] = {0, 0};
        for (int64_t s = 0; s < N; s++) {
            int c = B[s] == 1 ? 1 : 0;
            int64_t w = cnt[c] / 32;
            uint32_t *d = p->planes[c] + (snp * p->W[c] + w) * 2;
            d[0] <<= 1;
            d[1] <<= 1;
            uint8_t g = A[s * M + snp];
            if (g == 0 || g == 1)
                d[g] |= 1u;
            cnt[c]++;
        }
    }
    return 0;
}

void orc_packed_free(orc_packed *p)
{
    free(p->planes[0]);
    free(p->planes[1]);
    p->planes[0] = p->planes[1] = NULL;
}

static inline void count_class(const uint32_t *a, const uint32_t *b,
                               const uint32_t *c, int64_t W, uint32_t mask,
                               uint32_t *ft)
{
    for (int64_t w = 0; w < W; w++) {
        uint32_t m = (w == W - 1) ? mask : 0xFFFFFFFFu;
        uint32_t x[3], y[3], z[3];
        x[0] = a[2 * w];
        x[1] = a[2 * w + 1];
        x[2] = ~(x[0] | x[1]) & m;
        y[0] = b[2 * w];
        y[1] = b[2 * w + 1];
        y[2] = ~(y[0] | y[1]) & m;
        z[0] = c[2 * w];
        z[1] = c[2 * w + 1];
        z[2] = ~(z[0] | z[1]) & m;
        for (int gi = 0; gi < 3; gi++)
            for (int gj = 0; gj < 3; gj++) {
                uint32_t xy = x[gi] & y[gj];
                for (int gk = 0; gk < 3; gk++)
                    ft[9 * gi + 3 * gj + gk] += (uint32_t)__builtin_popcount(xy & z[gk]);
            }
    }
}

void orc_count_packed(const orc_packed *p, int64_t i, int64_t j, int64_t k,
                      uint32_t out[54])
{
    memset(out, 0, 54 * sizeof(uint32_t));
    for (int c = 0; c < 2; c++) {
        const uint32_t *P = p->planes[c];
        int64_t W = p->W[c];
        count_class(P + i * W * 2, P + j * W * 2, P + k * W * 2, W, p->mask[c],
                    out + 27 * c);
    }
}

/* ------------------------------------------------------------------ */
/* K2 (CPU/v2.cpp:516-520; GPU/v4.cpp:579-583).                         */
/* ------------------------------------------------------------------ */
float orc_k2(const uint32_t c[54], const float *T)
{
    float score = 0.0f;
    for (int m = 0; m < 27; m++)
        score += T[c[m] + c[27 + m] + 1] - T[c[m]] - T[c[27 + m]];
    return fabsf(score);
}

double orc_k2_f64(const uint32_t c[54], const double *T64)
{
    double score = 0.0;
    for (int m = 0; m < 27; m++)
        score += T64[c[m] + c[27 + m] + 1] - T64[c[m]] - T64[c[27 + m]];
    return fabs(score);
}

/* ------------------------------------------------------------------ */
/* Exhaustive scan (CPU/v2.cpp:244-541) with a deterministic tie rule.  */
/* ------------------------------------------------------------------ */
static inline int cand_less(const orc_cand *a, const orc_cand *b)
{
    if (a->score != b->score)
        return a->score < b->score;
    if (a->i != b->i)
        return a->i < b->i;
    if (a->j != b->j)
        return a->j < b->j;
    return a->k < b->k;
}

static void topk_insert(orc_cand *list, int *len, int cap, const orc_cand *c)
{
    if (*len == cap && !cand_less(c, &list[cap - 1]))
        return;
    int pos = *len < cap ? (*len)++ : cap - 1;
    while (pos > 0 && cand_less(c, &list[pos - 1])) {
        list[pos] = list[pos - 1];
        pos--;
    }
    list[pos] = *c;
}

int64_t orc_rank3(int64_t M, int64_t i, int64_t j, int64_t k)
{
    /* triplets with first index < i, then second index < j, then k */
    int64_t r = 0;
    for (int64_t a = 0; a < i; a++) {
        int64_t m = M - 1 - a;
        r += m * (m - 1) / 2;
    }
    for (int64_t b = i + 1; b < j; b++)
        r += M - 1 - b;
    r += k - j - 1;
    return r;
}

int64_t orc_scan(const orc_packed *p, const float *T, int nthreads, int topk,
                 orc_cand *top)
{
    int64_t M = p->M;
    if (topk < 1)
        topk = 1;
    if (topk > 256)
        topk = 256;
    if (nthreads < 1)
        nthreads = 1;
    orc_cand *lists = (orc_cand *)malloc(sizeof(orc_cand) * (size_t)topk * (size_t)nthreads);
    int *lens = (int *)calloc((size_t)nthreads, sizeof(int));
    int64_t scored = 0;
#pragma omp parallel num_threads(nthreads) reduction(+ : scored)
    {
        int tid = 0;
#ifdef _OPENMP
        tid = omp_get_thread_num();
#endif
        orc_cand *mine = lists + (size_t)tid * (size_t)topk;
        uint32_t ft[54];
#pragma omp for schedule(dynamic)
        for (int64_t i = 0; i < M - 2; i++)
            for (int64_t j = i + 1; j < M - 1; j++)
                for (int64_t k = j + 1; k < M; k++) {
                    orc_count_packed(p, i, j, k, ft);
                    orc_cand c = {orc_k2(ft, T), (uint32_t)i, (uint32_t)j, (uint32_t)k};
                    topk_insert(mine, &lens[tid], topk, &c);
                    scored++;
                }
    }
    int len = 0;
    for (int t = 0; t < nthreads; t++)
        for (int e = 0; e < lens[t]; e++)
            topk_insert(top, &len, topk, &lists[(size_t)t * (size_t)topk + (size_t)e]);
    for (int e = len; e < topk; e++) {
        top[e].score = FLT_MAX;
        top[e].i = top[e].j = top[e].k = 0xFFFFFFFFu;
    }
    free(lists);
    free(lens);
    return scored;
}

int64_t orc_scan_all_scores(const orc_packed *p, const float *T, int nthreads,
                            float *scores)
{
    int64_t M = p->M;
    if (nthreads < 1)
        nthreads = 1;
#pragma omp parallel for num_threads(nthreads) schedule(dynamic)
    for (int64_t i = 0; i < M - 2; i++) {
        uint32_t ft[54];
        int64_t r = orc_rank3(M, i, i + 1, i + 2);
        for (int64_t j = i + 1; j < M - 1; j++)
            for (int64_t k = j + 1; k < M; k++) {
                orc_count_packed(p, i, j, k, ft);
                scores[r++] = orc_k2(ft, T);
            }
    }
    return M < 3 ? 0 : M * (M - 1) * (M - 2) / 6;
}

/* Time-stamp counter read, used only to convert the reference programs'
 * cycle counts (they print (end - start) / FREQ, GPU/v4.cpp:609-610,
 * CPU/v2.cpp:554) into seconds when they are built with FREQ=1. */
uint64_t orc_rdtsc(void)
{
#if defined(__x86_64__)
    uint32_t lo, hi;
    __asm__ __volatile__("rdtsc" : "=a"(lo), "=d"(hi));
    return ((uint64_t)hi << 32) | lo;
#else
    return 0;
#endif
}

/* ---------------------------------------------------------------------- */
/* PLINK 1 .bed body (published format; no reference counterpart).       */
/* ---------------------------------------------------------------------- */
void orc_bed_decode(const uint8_t *bed, int64_t N, int64_t M, uint8_t *A)
{
    static const uint8_t geno_of_code[4] = {2, 3, 1, 0};
    const int64_t bpr = (N + 3) / 4;
    for (int64_t j = 0; j < M; ++j)
        for (int64_t s = 0; s < N; ++s) {
            const unsigned code = (bed[j * bpr + s / 4] >> (2 * (s % 4))) & 3u;
            A[s * M + j] = geno_of_code[code];
        }
}

void orc_bed_encode(const uint8_t *A, int64_t N, int64_t M, uint8_t *bed)
{
    const int64_t bpr = (N + 3) / 4;
    for (int64_t j = 0; j < M; ++j) {
        for (int64_t b = 0; b < bpr; ++b)
            bed[j * bpr + b] = 0;
        for (int64_t s = 0; s < N; ++s) {
            const uint8_t g = A[s * M + j];
            const unsigned code = g == 0 ? 3u : g == 1 ? 2u : g == 2 ? 0u : 1u;
            bed[j * bpr + s / 4] |= (uint8_t)(code << (2 * (s % 4)));
        }
    }
}
