// episdet_sample.h -- the triplet sample behind a cohort's threshold seed.
//
// A cohort scores kSeedSample distinct triplets at creation; the 64th best
// of their scores bounds the 64th best of all triplets from above (any 64
// distinct triplets do).  The sample is the triplets of colex ranks
// a t + b mod C(M, 3), t = 0, 1, ..., with a coprime to C(M, 3), so the
// ranks are distinct without sorting; every triplet when there are no more
// than kSeedSample.
//
// Plain C++ so the same code is compiled into the library and, in the CPU
// test suite, into the host test shim.
#pragma once
#include <stdint.h>

#include <numeric>

namespace epi {

constexpr uint32_t kSeedSample = 8192;

// C(n, 3) and C(n, 2); n < 2^21 keeps them below 2^62.
inline uint64_t choose3(uint64_t n) { return n < 3 ? 0 : n * (n - 1) * (n - 2) / 6; }
inline uint64_t choose2(uint64_t n) { return n < 2 ? 0 : n * (n - 1) / 2; }

// The triplet i < j < k < M of colex rank r < C(M, 3): k the largest with
// C(k, 3) <= r, then j the largest with C(j, 2) <= r - C(k, 3), i the rest.
inline void colex_triplet(uint64_t r, uint32_t M, uint32_t t[3])
{
    uint32_t lo = 2, hi = M - 1;
    while (lo < hi) {
        const uint32_t mid = lo + (hi - lo + 1) / 2;
        if (choose3(mid) <= r)
            lo = mid;
        else
            hi = mid - 1;
    }
    const uint32_t k = lo;
    r -= choose3(k);
    lo = 1, hi = k - 1;
    while (lo < hi) {
        const uint32_t mid = lo + (hi - lo + 1) / 2;
        if (choose2(mid) <= r)
            lo = mid;
        else
            hi = mid - 1;
    }
    t[0] = (uint32_t)(r - choose2(lo));
    t[1] = lo;
    t[2] = k;
}

// Ranks of the sample of a cohort with M SNPs (total = C(M, 3) triplets)
// and N samples: t-th rank, t < min(total, kSeedSample).
struct SeedSampler {
    uint64_t total, a, b;
    SeedSampler(uint64_t total_, uint32_t N) : total(total_), a(1), b(0)
    {
        if (total <= kSeedSample)
            return;
        a = 0x9E3779B97F4A7C15ull % total;
        b = (0xD1B54A32D192ED03ull ^ N) % total;
        if (a == 0)
            a = 1;
        while (std::gcd(a, total) != 1)
            a = a + 1 < total ? a + 1 : 1;
    }
    uint64_t rank(uint32_t t) const
    {
        return (uint64_t)(((unsigned __int128)a * t + b) % total);  // t < total: t itself if a = 1, b = 0
    }
};

}  // namespace epi
