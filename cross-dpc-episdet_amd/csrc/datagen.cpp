// datagen.cpp -- input generators behind the C ABI.
//
// epi_generate_reference restates the reference's generate_data
// (GPU/v4.cpp:81-110; the same draw order as CPU/v2.cpp:56-84): glibc
// srand(100), then rand() % 3 for every (sample, SNP) in sample-major
// order, then N/2 distinct cases drawn as (int)(N * rand() / RAND_MAX) with
// rejection.  The `v4` program uses it so its output matches the
// reference's for the same arguments.
//
// epi_generate_hwe / epi_generate_planted have no reference counterpart.
// They are counter-based (SplitMix64 seeding xoshiro256**, one stream per
// sample), so any thread count produces the same matrix.
#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "../../include/episdet.h"

namespace {

uint64_t splitmix_next(uint64_t &s)
{
    uint64_t z = (s += 0x9E3779B97F4A7C15ULL);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}

struct Xoshiro256 {
    uint64_t s[4];
    explicit Xoshiro256(uint64_t seed)
    {
        uint64_t sm = seed;
        for (auto &w : s)
            w = splitmix_next(sm);
    }
    static uint64_t rotl(uint64_t v, int k) { return (v << k) | (v >> (64 - k)); }
    uint64_t next()
    {
        const uint64_t out = rotl(s[1] * 5, 7) * 9;
        const uint64_t t = s[1] << 17;
        s[2] ^= s[0];
        s[3] ^= s[1];
        s[1] ^= s[2];
        s[0] ^= s[3];
        s[2] ^= t;
        s[3] = rotl(s[3], 45);
        return out;
    }
    double uniform() { return (double)(next() >> 11) * 0x1.0p-53; }
};

constexpr uint64_t kSaltMaf = 0x4D41465F53545245ULL;
constexpr uint64_t kSaltSample = 0x53414D504C455321ULL;
constexpr uint64_t kSaltPhen = 0x50484E5F53545245ULL;
constexpr uint64_t kSaltPlant = 0x504C414E54454421ULL;

uint64_t stream_seed(uint64_t seed, uint64_t salt, uint64_t idx)
{
    return (seed ^ salt) ^ ((idx + 1) * 0xD1B54A32D192ED03ULL);
}

// Genotype under Hardy-Weinberg with minor-allele frequency p:
// P(0) = (1-p)^2, P(1) = 2p(1-p), P(2) = p^2.
uint8_t hwe_genotype(Xoshiro256 &x, double p)
{
    const double u = x.uniform();
    const double q = 1.0 - p;
    const double c0 = q * q;
    const double c1 = c0 + (2.0 * p) * q;
    return (uint8_t)(u < c0 ? 0 : (u < c1 ? 1 : 2));
}

std::vector<double> minor_allele_freqs(uint64_t seed, uint32_t M, double lo, double hi)
{
    std::vector<double> maf(M);
    Xoshiro256 x(stream_seed(seed, kSaltMaf, 0));
    for (uint32_t j = 0; j < M; ++j)
        maf[j] = lo + (hi - lo) * x.uniform();
    return maf;
}

bool maf_range_ok(double lo, double hi)
{
    return lo >= 0.0 && hi <= 1.0 && lo <= hi;
}

}  // namespace

extern "C" int epi_generate_reference(uint32_t n_samples, uint32_t n_snps, uint8_t *geno,
                                      uint8_t *pheno)
{
    if (!geno || !pheno || n_samples == 0 || n_snps == 0)
        return EPI_EINVAL;
    const long long N = n_samples;
    srand(100);
    for (long long s = 0; s < N; ++s)
        for (uint32_t j = 0; j < n_snps; ++j)
            geno[s * n_snps + j] = (uint8_t)(rand() % 3);
    std::memset(pheno, 0, n_samples);
    // index == N (rand() == RAND_MAX) would touch one byte past the
    // reference's array; it is kept as a phantom slot that holds no sample.
    uint8_t phantom = 0;
    for (long long c = 0; c < N / 2; ++c) {
        int index = (int)(N * ((double)rand() / (RAND_MAX)));
        while ((index < N ? pheno[index] : phantom) == 1)
            index = (int)(N * ((double)rand() / (RAND_MAX)));
        if (index < N)
            pheno[index] = 1;
        else
            phantom = 1;
    }
    return EPI_OK;
}

extern "C" int epi_generate_hwe(uint64_t seed, uint32_t n_samples, uint32_t n_snps,
                                double maf_lo, double maf_hi, uint32_t n_cases, uint8_t *geno,
                                uint8_t *pheno)
{
    if (!geno || !pheno || n_samples == 0 || n_snps == 0 || n_cases > n_samples ||
        !maf_range_ok(maf_lo, maf_hi))
        return EPI_EINVAL;
    const std::vector<double> maf = minor_allele_freqs(seed, n_snps, maf_lo, maf_hi);
#pragma omp parallel for schedule(static)
    for (long long s = 0; s < (long long)n_samples; ++s) {
        Xoshiro256 x(stream_seed(seed, kSaltSample, (uint64_t)s));
        uint8_t *row = geno + (size_t)s * n_snps;
        for (uint32_t j = 0; j < n_snps; ++j)
            row[j] = hwe_genotype(x, maf[j]);
    }
    // exactly n_cases cases: partial Fisher-Yates over the sample indices
    std::memset(pheno, 0, n_samples);
    std::vector<uint32_t> perm(n_samples);
    for (uint32_t s = 0; s < n_samples; ++s)
        perm[s] = s;
    Xoshiro256 x(stream_seed(seed, kSaltPhen, 0));
    for (uint32_t t = 0; t < n_cases; ++t) {
        const uint32_t r = t + (uint32_t)(x.next() % (uint64_t)(n_samples - t));
        std::swap(perm[t], perm[r]);
        pheno[perm[t]] = 1;
    }
    return EPI_OK;
}

extern "C" int epi_generate_planted(uint64_t seed, uint32_t n_samples, uint32_t n_snps,
                                    double maf_lo, double maf_hi, uint32_t n_cases,
                                    uint8_t *geno, uint8_t *pheno, uint32_t planted[3])
{
    if (!geno || !pheno || !planted || n_samples == 0 || n_snps < 3 || n_cases > n_samples ||
        !maf_range_ok(maf_lo, maf_hi))
        return EPI_EINVAL;
    std::vector<double> maf = minor_allele_freqs(seed, n_snps, maf_lo, maf_hi);
    Xoshiro256 px(stream_seed(seed, kSaltPlant, 0));
    uint32_t pl[3];
    int got = 0;
    while (got < 3) {
        const uint32_t c = (uint32_t)(px.next() % (uint64_t)n_snps);
        bool dup = false;
        for (int t = 0; t < got; ++t)
            dup |= pl[t] == c;
        if (!dup)
            pl[got++] = c;
    }
    std::sort(pl, pl + 3);
    for (int t = 0; t < 3; ++t) {
        maf[pl[t]] = 0.3;
        planted[t] = pl[t];
    }
    // Candidate samples are drawn in order; the phenotype follows a 27-entry
    // penetrance table (0.5 where g_a + g_b + g_c is odd, 0.05 otherwise)
    // and a sample is kept while its class still has room.
    uint64_t need[2] = {(uint64_t)(n_samples - n_cases), (uint64_t)n_cases};
    uint64_t filled = 0;
    std::vector<uint8_t> row(n_snps);
    const uint64_t max_cand = (uint64_t)n_samples * 4096u;
    for (uint64_t cand = 0; filled < n_samples && cand < max_cand; ++cand) {
        Xoshiro256 x(stream_seed(seed, kSaltSample, cand));
        for (uint32_t j = 0; j < n_snps; ++j)
            row[j] = hwe_genotype(x, maf[j]);
        const int odd = (row[pl[0]] + row[pl[1]] + row[pl[2]]) & 1;
        const double pen = odd ? 0.5 : 0.05;
        const int cls = x.uniform() < pen ? 1 : 0;
        if (need[cls] > 0) {
            --need[cls];
            std::memcpy(geno + filled * n_snps, row.data(), n_snps);
            pheno[filled] = (uint8_t)cls;
            ++filled;
        }
    }
    return filled == n_samples ? EPI_OK : EPI_EINVAL;
}
