// episdet_topk.h -- per-wavefront top-64 candidate selection, shared by the
// scan kernels (device code).
//
// Replaces the reference's per-slot running minimum and host argmin
// (GPU/v4.cpp:586-608).  Candidates are ordered by (FP32 score, key), key =
// i << 42 | j << 21 | k, so ties go to the lexicographically first triplet.
//   wave_offer   every lane offers one triplet; a ballot keeps those better
//                than the wave's current 64th best and than the cohort's
//                global threshold, appended to the wave's LDS buffer;
//   wave_compact selects the buffer's best 64 (keys are unique); once 64 are
//                kept it lowers the cohort's global threshold to their 64th
//                best score;
//   wave_flush   when a wave leaves a cohort: appends its entries that do not
//                exceed the cohort's global threshold to the cohort's
//                candidate buffer (one atomic per wave).
// The global threshold is only ever the 64th best score of some set of 64
// real candidates, so it bounds the cohort's true 64th best from above and
// nothing of the true top 64 is dropped.
#pragma once
#include <hip/hip_runtime.h>
#include <math.h>

#include "episdet_internal.h"

namespace epi {
namespace topk {

__device__ __forceinline__ bool less(float s, uint64_t k, float ts, uint64_t tk)
{
    return s < ts || (s == ts && k < tk);
}

__device__ __forceinline__ void wave_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
    __builtin_amdgcn_wave_barrier();
}

// Per-wavefront state; cnt, ts, tk are wave-uniform.
struct WaveTop {
    Cand *buf;   // kWaveCap entries
    int cnt;
    float ts;    // the wave's 64th best (score, key) once it has 64
    uint64_t tk;
    float gthr;  // cohort-wide score threshold (may be stale, never too low)
    uint32_t *thr_slot;  // the current cohort's threshold word (null: none)

    __device__ void reset()
    {
        cnt = 0;
        ts = INFINITY;
        tk = kEmptyKey;
        gthr = INFINITY;
        thr_slot = nullptr;
    }
};

static_assert(kWaveCap == 2 * kLanes, "compact_buf holds two buffer entries per lane");

struct Compacted {
    int cnt;
    float ts;
    uint64_t tk;
};

typedef __attribute__((address_space(3))) Cand LdsCand;

__device__ __forceinline__ Cand ld_cand(const LdsCand *p)
{
    Cand c;
    c.s = p->s;
    c.pad = p->pad;
    c.key = p->key;
    return c;
}
__device__ __forceinline__ void st_cand(LdsCand *p, const Cand &c)
{
    p->s = c.s;
    p->pad = c.pad;
    p->key = c.key;
}

// Kept out of line (it runs rarely); state goes in and out by value so the
// caller's registers are not forced into scratch memory.  The buffer is
// passed as an LDS pointer: a generic pointer would make every access a
// FLAT instruction waited on with both counters.
//
// Selection on ballots, without LDS traffic beyond one read and one write
// of each entry: lane L holds entries L and L + 64 (n <= kWaveCap = 128).
// A binary search over the score bits (scores are >= 0, so their bits
// order like the values) finds T, the kTop-th smallest score; if entries
// of score T must be split at the cut, a second one over the key finds K,
// the key of the kTop-th best.  Entries (s, key) <= (T, K) are the best
// kTop; they go to buf[0, kTop) in lane order, and (T, K) becomes the
// wave's threshold.  Every loop bound is fixed (32 and 63 halvings); all
// lanes run every iteration.
__device__ __noinline__ Compacted compact_buf(LdsCand *buf, int n, float ts, uint64_t tk,
                                              uint32_t lane)
{
    Compacted r;
    r.cnt = n;
    r.ts = ts;
    r.tk = tk;
    if (n <= kTop)
        return r;  // all kept, in place
    wave_sync();
    const bool h0 = (int)lane < n, h1 = (int)lane + kLanes < n;
    Cand e0, e1;
    e0.s = e1.s = 0.0f;
    e0.pad = e1.pad = 0;
    e0.key = e1.key = kEmptyKey;
    if (h0)
        e0 = ld_cand(buf + lane);
    if (h1)
        e1 = ld_cand(buf + lane + kLanes);
    const uint32_t s0 = h0 ? __float_as_uint(e0.s) : 0xFFFFFFFFu;
    const uint32_t s1 = h1 ? __float_as_uint(e1.s) : 0xFFFFFFFFu;
    auto count = [](bool a, bool b) {
        return __popcll(__ballot(a)) + __popcll(__ballot(b));
    };
    // smallest t with at least kTop scores <= t (absent entries never count)
    uint32_t lo = 0, hi = 0xFFFFFFFEu;
    for (int it = 0; it < 32; ++it) {
        const uint32_t mid = lo + ((hi - lo) >> 1);
        if (count(s0 <= mid, s1 <= mid) >= kTop)
            hi = mid;
        else
            lo = mid + 1;
    }
    const uint32_t T = lo;
    const bool t0 = s0 == T, t1 = s1 == T;
    const int need = kTop - count(s0 < T, s1 < T);  // >= 1 entries of score T
    // smallest key K with at least `need` entries of score T and key <= K
    uint64_t klo = 0, khi = kEmptyKey - 1;
    for (int it = 0; it < 64 && count(t0, t1) > need; ++it) {
        const uint64_t mid = klo + ((khi - klo) >> 1);
        if (count(t0 && e0.key <= mid, t1 && e1.key <= mid) >= need)
            khi = mid;
        else
            klo = mid + 1;
        if (klo == khi)
            break;
    }
    uint64_t K = klo;
    if (count(t0, t1) <= need) {
        // every entry of score T is kept: K is the largest of their keys
        uint64_t m = 0;
        if (t0)
            m = e0.key;
        if (t1 && e1.key > m)
            m = e1.key;
        for (int off = 32; off > 0; off >>= 1) {
            const uint64_t o = ((uint64_t)(uint32_t)__shfl_xor((int)(m >> 32), off) << 32) |
                               (uint32_t)__shfl_xor((int)(uint32_t)m, off);
            m = o > m ? o : m;
        }
        K = m;
    }
    const bool k0 = s0 < T || (t0 && e0.key <= K);
    const bool k1 = s1 < T || (t1 && e1.key <= K);
    const uint64_t b0 = __ballot(k0), b1 = __ballot(k1);
    wave_sync();  // every entry is in registers before the buffer is rewritten
    if (k0)
        st_cand(buf + __builtin_amdgcn_mbcnt_hi((uint32_t)(b0 >> 32),
                                                __builtin_amdgcn_mbcnt_lo((uint32_t)b0, 0u)),
                e0);
    if (k1)
        st_cand(buf + __popcll(b0) +
                    __builtin_amdgcn_mbcnt_hi((uint32_t)(b1 >> 32),
                                              __builtin_amdgcn_mbcnt_lo((uint32_t)b1, 0u)),
                e1);
    wave_sync();
    r.cnt = kTop;
    r.ts = __uint_as_float(T);
    r.tk = K;
    return r;
}

// Compacts the wave's buffer; once it holds 64 candidates their 64th best
// score is published as the cohort threshold, so that other wavefronts
// filter early.
__device__ __forceinline__ void wave_compact(WaveTop &W, uint32_t lane)
{
    const Compacted r = compact_buf((LdsCand *)W.buf, W.cnt, W.ts, W.tk, lane);
    W.cnt = __builtin_amdgcn_readfirstlane(r.cnt);
    W.ts = r.ts;
    W.tk = r.tk;
    if (W.cnt == kTop && W.thr_slot) {
        if (lane == 0)
            atomicMin(W.thr_slot, __float_as_uint(W.ts));
        W.gthr = fminf(W.gthr, W.ts);
    }
}

__device__ __forceinline__ void wave_offer(WaveTop &W, bool valid, float s, uint64_t key,
                                           uint32_t lane)
{
    // cheap pre-filter on the score alone (a superset of what passes); the
    // key comparison of ties runs only when some lane gets that far
    const bool maybe = valid && s <= W.gthr && s <= W.ts;
    if (__ballot(maybe) == 0)
        return;
    const bool pass = maybe && less(s, key, W.ts, W.tk);
    const uint64_t mask = __ballot(pass);
    if (mask == 0)
        return;
    const int nn = __popcll(mask);
    if (W.cnt + nn > kWaveCap)
        wave_compact(W, lane);
    if (pass) {
        const int below = (int)__builtin_amdgcn_mbcnt_hi(
            (uint32_t)(mask >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)mask, 0u));
        Cand c;
        c.s = s;
        c.pad = 0;
        c.key = key;
        W.buf[W.cnt + below] = c;
    }
    W.cnt += nn;
}

// Reads the cohort's global threshold (scores are >= 0, so their bit
// patterns order like the values and the threshold is kept as uint32).
__device__ __forceinline__ float load_thr(const uint32_t *thr, uint32_t cohort)
{
    return __uint_as_float(__atomic_load_n(thr + cohort, __ATOMIC_RELAXED));
}

// The two halves of a flush.  flush_select compacts the buffer (at most
// kTop entries, publishing the 64th best) and marks the entries that do not
// exceed the cohort's global threshold; it returns their number
// (wave-uniform).  flush_append stores the marked entries at slots
// base .. base + nn of the cohort's candidate buffer, or raises the
// overflow flag if they do not fit.
__device__ __forceinline__ uint32_t flush_select(WaveTop &W, const ScanParams &P, uint32_t cohort,
                                                 uint32_t lane, bool &pass)
{
    pass = false;
    if (W.cnt == 0)
        return 0;
    W.thr_slot = P.thr + cohort;
    wave_compact(W, lane);
    const float g = load_thr(P.thr, cohort);
    pass = (int)lane < W.cnt && W.buf[lane].s <= g;
    return (uint32_t)__popcll(__ballot(pass));
}
__device__ __forceinline__ void flush_append(const WaveTop &W, const ScanParams &P,
                                             uint32_t cohort, uint32_t lane, bool pass,
                                             uint32_t base, uint32_t nn)
{
    const uint64_t mask = __ballot(pass);
    if (base + nn <= P.slot_cap) {
        if (pass) {
            const uint32_t below = __builtin_amdgcn_mbcnt_hi(
                (uint32_t)(mask >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)mask, 0u));
            P.slots[(size_t)cohort * P.slot_cap + base + below] = W.buf[lane];
        }
    } else if (lane == 0 && nn) {
        atomicExch(P.overflow, 1u);
    }
}

// A wavefront's flush with its own append atomic (wavefronts that leave a
// cohort independently).
__device__ void wave_flush(WaveTop &W, const ScanParams &P, uint32_t cohort, uint32_t lane)
{
    bool pass;
    const uint32_t nn = flush_select(W, P, cohort, lane, pass);
    if (nn) {
        uint32_t base = 0;
        if (lane == 0)
            base = atomicAdd(&P.counters[1 + cohort], nn);
        base = __builtin_amdgcn_readfirstlane(base);
        flush_append(W, P, cohort, lane, pass, base, nn);
    }
    W.reset();
}

}  // namespace topk
}  // namespace epi
