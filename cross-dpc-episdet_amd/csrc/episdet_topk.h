// episdet_topk.h -- per-wavefront top-64 candidate selection, shared by the
// scan kernels (device code).
//
// Replaces the reference's per-slot running minimum and host argmin
// (GPU/v4.cpp:586-608).  Candidates are ordered by (FP32 score, key), key =
// i << 42 | j << 21 | k, so ties go to the lexicographically first triplet.
//   wave_offer   every lane offers one triplet; a ballot keeps those better
//                than the wave's current 64th best and than the cohort's
//                global threshold, appended to the wave's LDS buffer;
//   wave_compact ranks the buffer (keys are unique) and keeps the best 64,
//                sorted; once 64 are kept it lowers the cohort's global
//                threshold to their 64th best score;
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
    Cand *keep;  // kTop entries
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
// caller's registers are not forced into scratch memory.  The buffers are
// passed as LDS pointers: a generic pointer would make every access a FLAT
// instruction waited on with both counters.  Each lane ranks its entries
// against all n, reading 8 (broadcast) entries per wait.
__device__ __noinline__ Compacted compact_buf(LdsCand *buf, LdsCand *keep, int n, float ts,
                                              uint64_t tk, uint32_t lane)
{
    constexpr int kRankBatch = 8;
    wave_sync();
    for (int e = (int)lane; e < n; e += kLanes) {
        const Cand me = ld_cand(buf + e);
        int rank = 0, f = 0;
        for (; f + kRankBatch <= n; f += kRankBatch) {
            Cand o[kRankBatch];
#pragma unroll
            for (int q = 0; q < kRankBatch; ++q)
                o[q] = ld_cand(buf + f + q);
#pragma unroll
            for (int q = 0; q < kRankBatch; ++q)
                rank += less(o[q].s, o[q].key, me.s, me.key) ? 1 : 0;
        }
        for (; f < n; ++f) {
            const Cand o = ld_cand(buf + f);
            rank += less(o.s, o.key, me.s, me.key) ? 1 : 0;
        }
        if (rank < kTop)
            st_cand(keep + rank, me);
    }
    wave_sync();
    const int m = n < kTop ? n : kTop;
    if ((int)lane < m)
        st_cand(buf + lane, ld_cand(keep + lane));
    Compacted r;
    r.cnt = m;
    r.ts = ts;
    r.tk = tk;
    if (m == kTop) {
        r.ts = keep[kTop - 1].s;
        r.tk = keep[kTop - 1].key;
    }
    wave_sync();
    return r;
}

// Compacts the wave's buffer; once it holds 64 candidates their 64th best
// score is published as the cohort threshold, so that other wavefronts
// filter early.
__device__ __forceinline__ void wave_compact(WaveTop &W, uint32_t lane)
{
    const Compacted r = compact_buf((LdsCand *)W.buf, (LdsCand *)W.keep, W.cnt, W.ts, W.tk, lane);
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

__device__ void wave_flush(WaveTop &W, const ScanParams &P, uint32_t cohort, uint32_t lane)
{
    if (W.cnt > 0) {
        W.thr_slot = P.thr + cohort;
        wave_compact(W, lane);  // sorted, at most kTop entries; publishes the 64th best
        const float g = load_thr(P.thr, cohort);
        const bool pass = (int)lane < W.cnt && W.buf[lane].s <= g;
        const uint64_t mask = __ballot(pass);
        const uint32_t nn = (uint32_t)__popcll(mask);
        if (nn) {
            uint32_t base = 0;
            if (lane == 0)
                base = atomicAdd(&P.counters[1 + cohort], nn);
            base = __builtin_amdgcn_readfirstlane(base);
            if (base + nn <= P.slot_cap) {
                if (pass) {
                    const uint32_t below = __builtin_amdgcn_mbcnt_hi(
                        (uint32_t)(mask >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)mask, 0u));
                    P.slots[(size_t)cohort * P.slot_cap + base + below] = W.buf[lane];
                }
            } else if (lane == 0) {
                atomicExch(P.overflow, 1u);
            }
        }
    }
    W.reset();
}

}  // namespace topk
}  // namespace epi
