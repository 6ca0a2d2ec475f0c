// episdet_kernels.hip -- gfx950 kernels of the exhaustive third-order
// epistasis scan with K2 scoring.
//
// Reference: the DPC++ kernel lambda of GPU/v4.cpp:297-593 (one work-item
// per triplet, 27 three-way ANDs + 27 popcounts per 32-bit word and class,
// K2 from a log-factorial table, per-slot running minimum).  This file
// redesigns it for CDNA4:
//   * pack_kernel    class split + bit-packing into 64-bit planes, two
//                    layouts (SNP-major for scalar loads, word-major for
//                    coalesced per-lane loads)          (GPU/v4.cpp:112-196)
//   * strips_kernel  pair margins of the {0,1} genotype planes, in 64-column
//                    strips (all of them, or those one scan round needs)
//   * singles_kernel single-SNP margins
//   * scan_kernel    persistent; pulls (cohort, k-block, j-group, i-chunk)
//                    items from a device queue.  A wavefront owns one i, a
//                    group of snps_j consecutive j and 64 consecutive k (one
//                    per lane).  Pair products P_ab = g_a(i) & g_b(j) are
//                    wave-uniform and live in SGPRs (SALU work); each lane
//                    then issues v_and + v_bcnt for the 8 cells with genotypes
//                    in {0,1}.  The 19 cells involving genotype 2 are exact
//                    differences of pair/single margins (episdet_cells.h).
//                    K2 from the ln(n!) table (LDS-resident when it fits),
//                    then a ballot-filtered per-wave top-64 in LDS.
//                                                       (GPU/v4.cpp:283-597)
//   * merge_kernel   per-cohort top-64 over all workgroup slots
//                                                       (GPU/v4.cpp:598-608)
//   * rescore_kernel recounts each candidate with all 27 cells explicitly
//                    (~(g0|g1) under the padding mask) and scores it in FP32
//                    and FP64.
#include <math.h>

#include <algorithm>

#include "episdet_cells.h"
#include "episdet_internal.h"
#include "episdet_order.h"
#include "episdet_topk.h"

#define EPI_CONST __attribute__((address_space(4)))

namespace epi {

template <typename T>
__device__ __forceinline__ const EPI_CONST T *as_const(const T *p)
{
    return (const EPI_CONST T *)p;
}

__device__ __forceinline__ uint32_t rfl(uint32_t v) { return __builtin_amdgcn_readfirstlane(v); }

// Loads through the constant address space: with a wave-uniform address the
// compiler issues them as scalar (SMEM) loads into SGPRs.
__device__ __forceinline__ uint4 ldc4(const uint4 *p)
{
    const EPI_CONST uint32_t *q = (const EPI_CONST uint32_t *)p;
    return make_uint4(q[0], q[1], q[2], q[3]);
}
__device__ __forceinline__ uint2 ldc2(const uint2 *p)
{
    const EPI_CONST uint32_t *q = (const EPI_CONST uint32_t *)p;
    return make_uint2(q[0], q[1]);
}
__device__ __forceinline__ uint32_t ldc1(const uint32_t *p)
{
    return *(const EPI_CONST uint32_t *)p;
}

__device__ __forceinline__ bool cand_less(float s, uint64_t k, float ts, uint64_t tk)
{
    return s < ts || (s == ts && k < tk);
}

__device__ __forceinline__ void wave_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
    __builtin_amdgcn_wave_barrier();
}

__device__ __forceinline__ uint64_t c3(uint64_t n) { return n < 3 ? 0 : n * (n - 1) * (n - 2) / 6; }
__device__ __forceinline__ uint64_t c2(uint64_t n) { return n < 2 ? 0 : n * (n - 1) / 2; }

// lexicographic rank of i < j < k among C(M,3)
__device__ __forceinline__ uint64_t rank3(uint32_t M, uint32_t i, uint32_t j, uint32_t k)
{
    return c3(M) - c3(M - i) + c2(M - i - 1) - c2(M - j) + (uint64_t)(k - j - 1);
}

// ------------------------------------------------------------------------
// small utilities
// ------------------------------------------------------------------------
__global__ void fill_u32_kernel(uint32_t *p, uint32_t v, size_t n)
{
    for (size_t t = blockIdx.x * (size_t)blockDim.x + threadIdx.x; t < n;
         t += (size_t)gridDim.x * blockDim.x)
        p[t] = v;
}

hipError_t launch_fill_u32(uint32_t *p, uint32_t v, size_t n, hipStream_t s)
{
    if (n == 0)
        return hipSuccess;
    size_t blocks = (n + 255) / 256;
    if (blocks > 4096)
        blocks = 4096;
    hipLaunchKernelGGL(fill_u32_kernel, dim3((uint32_t)blocks), dim3(256), 0, s, p, v, n);
    return hipGetLastError();
}

// ------------------------------------------------------------------------
// packing (GPU/v4.cpp:112-196): order[] lists controls then cases, each in
// original sample order; sample t of class c goes to bit t%64 of word t/64;
// row p holds genotype column src[p] (src null: column p).
// ------------------------------------------------------------------------
// One-dimensional grid of (class, word, 256-SNP chunk) blocks, chunks
// fastest (no grid dimension bounded by the word count).
__global__ void __launch_bounds__(256)
pack_kernel(const uint8_t *__restrict__ geno, uint32_t M, uint32_t Mp,
            const uint32_t *__restrict__ src, const uint32_t *__restrict__ order,
            uint32_t n0, uint32_t n1,
            uint4 *sm0, uint4 *wm0, uint4 *sm1, uint4 *wm1, uint32_t W0, uint32_t W1,
            uint32_t chunks, uint32_t wmax)
{
    const uint32_t b = blockIdx.x;
    const uint32_t s = (b % chunks) * 256 + threadIdx.x;
    const uint32_t w = (b / chunks) % wmax;
    const uint32_t c = b / chunks / wmax;
    const uint32_t W = c ? W1 : W0;
    if (s >= Mp || w >= W)
        return;
    const uint32_t n = c ? n1 : n0;
    const uint32_t *ord = order + (c ? n0 : 0);
    uint64_t g0 = 0, g1 = 0;
    if (s < M) {
        const uint32_t col = src ? src[s] : s;
        const uint32_t t0 = w * 64;
        const uint32_t tn = (n - t0) < 64u ? (n - t0) : 64u;
        for (uint32_t t = 0; t < tn; ++t) {
            const uint8_t g = geno[(size_t)ord[t0 + t] * M + col];
            g0 |= (uint64_t)(g == 0) << t;
            g1 |= (uint64_t)(g == 1) << t;
        }
    }
    const uint4 v = make_uint4((uint32_t)g0, (uint32_t)(g0 >> 32), (uint32_t)g1,
                               (uint32_t)(g1 >> 32));
    (c ? sm1 : sm0)[(size_t)s * W + w] = v;
    (c ? wm1 : wm0)[(size_t)w * Mp + s] = v;
}

hipError_t launch_pack(const uint8_t *geno, uint32_t N, uint32_t M, uint32_t Mp,
                       const uint32_t *src, const uint32_t *order, uint32_t n0, uint32_t n1,
                       uint4 *sm0,
                       uint4 *wm0, uint4 *sm1, uint4 *wm1, uint32_t W0, uint32_t W1,
                       hipStream_t s)
{
    (void)N;
    const uint32_t wmax = W0 > W1 ? W0 : W1;
    if (wmax == 0)
        return hipSuccess;
    const uint32_t chunks = (Mp + 255) / 256;
    const uint64_t blocks = (uint64_t)chunks * wmax * 2;
    if (blocks > 0x7FFFFFFFull)
        return hipErrorInvalidConfiguration;  // callers bound the shape first
    hipLaunchKernelGGL(pack_kernel, dim3((uint32_t)blocks), dim3(256), 0, s, geno, M, Mp, src,
                       order, n0, n1, sm0, wm0, sm1, wm1, W0, W1, chunks, wmax);
    return hipGetLastError();
}

// ------------------------------------------------------------------------
// PLINK 1 .bed body -> sample-major genotype bytes.  Row s (SNP s) holds
// bpr bytes; sample t sits in bits 2(t%4)..2(t%4)+1 of byte t/4 with codes
// 00 homozygous first allele -> 2, 10 heterozygous -> 1, 11 homozygous
// second allele -> 0, 01 missing -> 3 (neither plane: counted as 2 by the
// packer, like any value other than 0 or 1).  Threads run over SNPs
// fastest, so the byte writes of a wavefront are contiguous.
// ------------------------------------------------------------------------
__global__ void __launch_bounds__(256)
bed_unpack_kernel(const uint8_t *__restrict__ bed, uint32_t bpr, uint32_t N, uint32_t M,
                  uint8_t *__restrict__ geno)
{
    const size_t t = blockIdx.x * (size_t)256 + threadIdx.x;
    if (t >= (size_t)bpr * M)
        return;
    const uint32_t s = (uint32_t)(t % M);
    const uint32_t b = (uint32_t)(t / M);
    const uint32_t v = bed[(size_t)s * bpr + b];
#pragma unroll
    for (uint32_t q = 0; q < 4; ++q) {
        const uint32_t smp = 4 * b + q;
        if (smp < N) {
            const uint32_t code = (v >> (2 * q)) & 3u;
            // 0 -> 2, 1 -> 3, 2 -> 1, 3 -> 0
            geno[(size_t)smp * M + s] = (uint8_t)((0x1Eu >> (2 * code)) & 3u);
        }
    }
}

hipError_t launch_bed_unpack(const uint8_t *bed, uint32_t N, uint32_t M, uint8_t *geno,
                             hipStream_t s)
{
    const uint32_t bpr = (N + 3) / 4;
    const size_t total = (size_t)bpr * M;
    if (total == 0)
        return hipSuccess;
    hipLaunchKernelGGL(bed_unpack_kernel, dim3((uint32_t)((total + 255) / 256)), dim3(256), 0, s,
                       bed, bpr, N, M, geno);
    return hipGetLastError();
}

// ------------------------------------------------------------------------
// pair and single margins
// ------------------------------------------------------------------------
// Strips of pair margins: todo[t] = (column block g, slot) for t < n_todo
// (todo null: strip t is block t in slot t).  One-dimensional grid of
// (strip, class, 4-row group) blocks; wavefront r of a block computes row
// p = 4 group + r, lane l column q = 64 g + l: X (row p's planes) is
// wave-uniform, Z (column q's word-major planes) one coalesced 1 KiB load
// per word and wavefront.
__global__ void __launch_bounds__(256)
strips_kernel(const CohortDev *__restrict__ cd, const uint2 *__restrict__ todo, uint32_t groups,
              uint32_t t0)
{
    const uint32_t b = blockIdx.x;
    const uint32_t grp = b % groups;
    const uint32_t c = (b / groups) & 1u;
    const uint32_t t = t0 + b / groups / 2;
    const uint32_t g = todo ? todo[t].x : t;
    const uint32_t slot = todo ? todo[t].y : t;
    const uint32_t Mp = as_const(cd)->Mp;
    const uint32_t p = grp * 4 + (threadIdx.x >> 6);
    const uint32_t q = g * 64 + (threadIdx.x & 63u);
    if (p >= Mp)
        return;
    const uint32_t W = as_const(cd)->W[c];
    const uint4 *X = as_const(cd)->sm[c] + (size_t)p * W;
    const uint4 *Z = as_const(cd)->wm[c] + q;
    uint32_t a00 = 0, a01 = 0, a10 = 0, a11 = 0;
    for (uint32_t w = 0; w < W; ++w) {
        const uint4 x = ldc4(X + w);
        const uint4 z = Z[(size_t)w * Mp];
        a00 += __popc(x.x & z.x) + __popc(x.y & z.y);
        a01 += __popc(x.x & z.z) + __popc(x.y & z.w);
        a10 += __popc(x.z & z.x) + __popc(x.w & z.y);
        a11 += __popc(x.z & z.z) + __popc(x.w & z.w);
    }
    // every row p, not only p < q: the MFMA kernel's epilogue scores whole
    // 64 x 64 tiles, and with every row the counts of its lanes outside
    // i < j < k are exact too (so every ln(n!) index stays in the table)
    as_const(cd)->pair[c][pair_at(slot, Mp, p, q)] = make_uint4(a00, a01, a10, a11);
}

hipError_t launch_strips(const CohortDev *dev_cohort, const uint2 *todo, uint32_t n_todo,
                         uint32_t Mp, hipStream_t s)
{
    const uint32_t groups = (Mp + 3) / 4;
    // at most 2^30 blocks per launch (2 x groups <= 2^20 blocks per strip)
    const uint32_t per = (uint32_t)std::max<uint64_t>(1, (1ull << 30) / (2ull * groups));
    for (uint32_t t0 = 0; t0 < n_todo; t0 += per) {
        const uint32_t nt = std::min(per, n_todo - t0);
        hipLaunchKernelGGL(strips_kernel, dim3(nt * 2 * groups), dim3(256), 0, s, dev_cohort, todo,
                           groups, t0);
        const hipError_t e = hipGetLastError();
        if (e != hipSuccess)
            return e;
    }
    return hipSuccess;
}

// Single-SNP margins {count g = 0, count g = 1} of every SNP and class.
__global__ void __launch_bounds__(256) singles_kernel(const CohortDev *__restrict__ cd)
{
    const uint32_t Mp = as_const(cd)->Mp;
    const uint32_t t = blockIdx.x * 256 + threadIdx.x;
    if (t >= 2 * Mp)
        return;
    const uint32_t c = t / Mp, p = t % Mp;
    const uint32_t W = as_const(cd)->W[c];
    const uint4 *X = as_const(cd)->sm[c] + (size_t)p * W;
    uint32_t a0 = 0, a1 = 0;
    for (uint32_t w = 0; w < W; ++w) {
        const uint4 x = X[w];
        a0 += __popc(x.x) + __popc(x.y);
        a1 += __popc(x.z) + __popc(x.w);
    }
    as_const(cd)->single[c][p] = make_uint2(a0, a1);
}

hipError_t launch_singles(const CohortDev *dev_cohort, uint32_t Mp, hipStream_t s)
{
    hipLaunchKernelGGL(singles_kernel, dim3((2 * Mp + 255) / 256), dim3(256), 0, s, dev_cohort);
    return hipGetLastError();
}

// ------------------------------------------------------------------------
// scan kernel
// ------------------------------------------------------------------------
// acc += popcount(a) + popcount(b) as two chained v_bcnt_u32_b32 (popcount
// and accumulate in one instruction each).  The empty asm pins the
// intermediate sum in a VGPR; without it the compiler reassociates into
// bcnt(a,0) + bcnt(b,0) + v_add3, one extra VALU op per cell-word.
__device__ __forceinline__ void bcnt2(uint32_t &acc, uint32_t a, uint32_t b)
{
    uint32_t t = __popc(a) + acc;
    asm("" : "+v"(t));
    acc = __popc(b) + t;
}

// 8 cells (a, b, c in {0,1}) of JP triplets (i, j0 + jj, k) over one class.
// x (SNP i) and y (SNP j) are wave-uniform: their loads are scalar and the
// pair products are SALU work.  z (SNP k = lane) is a per-lane uint4.
template <int JP>
__device__ __forceinline__ void count8(const uint4 *X, const uint4 *Y,
                                       uint32_t W, const uint4 *__restrict__ Z, uint32_t Mp,
                                       uint32_t (&acc)[JP][8])
{
#pragma unroll
    for (int jj = 0; jj < JP; ++jj)
#pragma unroll
        for (int e = 0; e < 8; ++e)
            acc[jj][e] = 0;
    for (uint32_t w = 0; w < W; ++w) {
        const uint4 x = ldc4(X + w);
        const uint4 z = Z[(size_t)w * Mp];
#pragma unroll
        for (int jj = 0; jj < JP; ++jj) {
            const uint4 y = ldc4(Y + (size_t)jj * W + w);
            const uint32_t p00l = x.x & y.x, p00h = x.y & y.y;  // g0(i) & g0(j)
            const uint32_t p01l = x.x & y.z, p01h = x.y & y.w;  // g0(i) & g1(j)
            const uint32_t p10l = x.z & y.x, p10h = x.w & y.y;  // g1(i) & g0(j)
            const uint32_t p11l = x.z & y.z, p11h = x.w & y.w;  // g1(i) & g1(j)
            bcnt2(acc[jj][0], p00l & z.x, p00h & z.y);
            bcnt2(acc[jj][1], p00l & z.z, p00h & z.w);
            bcnt2(acc[jj][2], p01l & z.x, p01h & z.y);
            bcnt2(acc[jj][3], p01l & z.z, p01h & z.w);
            bcnt2(acc[jj][4], p10l & z.x, p10h & z.y);
            bcnt2(acc[jj][5], p10l & z.z, p10h & z.w);
            bcnt2(acc[jj][6], p11l & z.x, p11h & z.y);
            bcnt2(acc[jj][7], p11l & z.z, p11h & z.w);
        }
    }
}

// The 27 cells (9 g_a + 3 g_b + g_c) of one class for SNP rows a, b, c
// (per lane), counted word by word from the planes; genotype 2 is neither
// plane, masked to the class's samples in the last word.
__device__ __forceinline__ void recount27(const uint4 *S, uint32_t W, uint32_t n, uint32_t a,
                                          uint32_t b, uint32_t c, uint32_t (&x)[27])
{
#pragma unroll
    for (int m = 0; m < 27; ++m)
        x[m] = 0;
    for (uint32_t w = 0; w < W; ++w) {
        const uint32_t rem = n - w * 64;
        const uint64_t live = rem >= 64 ? ~0ull : ((1ull << rem) - 1ull);
        const uint4 pa = S[(size_t)a * W + w], pb = S[(size_t)b * W + w], pc = S[(size_t)c * W + w];
        uint64_t ga[3], gb[3], gc[3];
        ga[0] = pa.x | ((uint64_t)pa.y << 32);
        ga[1] = pa.z | ((uint64_t)pa.w << 32);
        ga[2] = ~(ga[0] | ga[1]) & live;
        gb[0] = pb.x | ((uint64_t)pb.y << 32);
        gb[1] = pb.z | ((uint64_t)pb.w << 32);
        gb[2] = ~(gb[0] | gb[1]) & live;
        gc[0] = pc.x | ((uint64_t)pc.y << 32);
        gc[1] = pc.z | ((uint64_t)pc.w << 32);
        gc[2] = ~(gc[0] | gc[1]) & live;
#pragma unroll
        for (int u = 0; u < 3; ++u)
#pragma unroll
            for (int v = 0; v < 3; ++v) {
                const uint64_t uv = ga[u] & gb[v];
#pragma unroll
                for (int t = 0; t < 3; ++t)
                    x[9 * u + 3 * v + t] += __popcll(uv & gc[t]);
            }
    }
}

// LDS: per-wavefront candidate buffers, then (optionally) the ln(n!) table.
template <int NW>
struct ScanLds {
    Cand cand[NW][kWaveCap];
};

template <int NW>
__host__ __device__ constexpr size_t scan_lds_fixed()
{
    return (sizeof(ScanLds<NW>) + 15) & ~(size_t)15;
}

// Wavefronts are independent workers: each pulls its own items (one i of a
// (k-block, j-group) row) from the queue, prefetching the next index while
// it counts.  The workgroup only shares the LDS table.
template <int JP, int NW, bool TLDS, bool DUMP>
__global__ void __launch_bounds__(NW * 64, 4) scan_kernel(const ScanParams P)
{
    extern __shared__ __align__(16) unsigned char lds_raw[];
    ScanLds<NW> &L = *reinterpret_cast<ScanLds<NW> *>(lds_raw);
    float *Tl = reinterpret_cast<float *>(lds_raw + scan_lds_fixed<NW>());

    const uint32_t tid = threadIdx.x;
    const uint32_t lane = tid & 63u;
    const uint32_t wave = rfl(tid >> 6);
    topk::WaveTop W;
    W.buf = L.cand[wave];
    W.reset();

    if (TLDS) {
        for (uint32_t t = tid; t < P.table_len; t += NW * 64)
            Tl[t] = P.table[t];
        __syncthreads();
    }

    uint32_t cur = 0xFFFFFFFFu;  // cohort whose candidates W holds
    uint32_t sg = 0;             // current segment
    uint32_t r = 0;              // current row of that segment
    bool fresh = true;           // r must be reset to the segment's first row
    const EPI_CONST Segment *segs = as_const(P.segs);
    const uint32_t n_segs = as_const(P.hdr)->n_segments;
    const uint32_t n_items = as_const(P.hdr)->n_items;

    uint32_t q = 0;
    if (lane == 0)
        q = atomicAdd(&P.counters[0], 1u);
    uint32_t next = rfl(q);
    for (;;) {
        const uint32_t item = next;
        if (item < n_items) {  // prefetch the following item while this one runs
            uint32_t q2 = 0;
            if (lane == 0)
                q2 = atomicAdd(&P.counters[0], 1u);
            next = rfl(q2);
        }
        // Items are pulled in increasing order, so segment and row only move
        // forward: both are found by short forward walks.
        uint32_t cohort = 0xFFFFFFFFu;
        uint32_t u = 0;
        const EPI_CONST Row *rows = nullptr;
        if (item < n_items) {
            while (sg + 1 < n_segs && segs[sg + 1].item_base <= item) {
                ++sg;
                fresh = true;
            }
            rows = (const EPI_CONST Row *)segs[sg].rows;
            if (fresh) {
                r = segs[sg].row_begin;
                fresh = false;
            }
            u = item - segs[sg].item_base + segs[sg].item_first;
            const uint32_t row_end = segs[sg].row_end;
            while (r + 1 < row_end && rows[r + 1].item_start <= u)
                ++r;
            cohort = segs[sg].cohort;
        }
        if (cohort != cur) {
            if (!DUMP && cur != 0xFFFFFFFFu)
                topk::wave_flush(W, P, cur, lane);
            cur = cohort;
        }
        if (item >= n_items)
            break;
        if (!DUMP) {
            W.gthr = fminf(W.gthr, topk::load_thr(P.thr, cohort));
            W.thr_slot = P.thr + cohort;
        }

        const uint32_t kb = rows[r].kb;
        const uint32_t jg = rows[r].jg;
        const uint32_t i = u - rows[r].item_start;  // one i per item

        const EPI_CONST CohortDev *C = as_const(P.cohorts) + cohort;
        const uint32_t M = C->M, Mp = C->Mp;
        const uint32_t j0 = jg * JP;
        const uint32_t k = kb * 64 + lane;

        uint32_t acc0[JP][8], acc1[JP][8];
        {
            const uint32_t Wd = C->W[0];
            const uint4 *S = C->sm[0];
            count8<JP>(S + (size_t)i * Wd, S + (size_t)j0 * Wd, Wd, C->wm[0] + k, Mp, acc0);
        }
        {
            const uint32_t Wd = C->W[1];
            const uint4 *S = C->sm[1];
            count8<JP>(S + (size_t)i * Wd, S + (size_t)j0 * Wd, Wd, C->wm[1] + k, Mp, acc1);
        }

        const uint32_t n0 = C->n[0], n1 = C->n[1];
        const uint32_t tmax = P.table_len - 1;
        // strip slots of the j group's and the k block's columns (the JP
        // j share one 64-block)
        const uint32_t slot_j = ldc1(C->pslot + (j0 >> 6));
        const uint32_t slot_k = ldc1(C->pslot + kb);
#pragma unroll
        for (int jj = 0; jj < JP; ++jj) {
            const uint32_t j = j0 + jj;
            const bool valid = (i < j) && (j < k) && (k < M);
            float s = INFINITY;
            uint64_t key = make_key(i, j, k);
            if (valid) {
                uint32_t x0[27], x1[27];
#pragma unroll
                for (int c = 0; c < 2; ++c) {
                    const uint4 ij = ldc4(C->pair[c] + pair_at(slot_j, Mp, i, j));
                    const uint4 ik = C->pair[c][pair_at(slot_k, Mp, i, k)];
                    const uint4 jk = C->pair[c][pair_at(slot_k, Mp, j, k)];
                    const uint2 si = ldc2(C->single[c] + i);
                    const uint2 sj = ldc2(C->single[c] + j);
                    const uint2 sk = C->single[c][k];
                    const uint32_t aij[4] = {ij.x, ij.y, ij.z, ij.w};
                    const uint32_t aik[4] = {ik.x, ik.y, ik.z, ik.w};
                    const uint32_t ajk[4] = {jk.x, jk.y, jk.z, jk.w};
                    const uint32_t asi[2] = {si.x, si.y};
                    const uint32_t asj[2] = {sj.x, sj.y};
                    const uint32_t ask[2] = {sk.x, sk.y};
                    const uint32_t(&t8)[8] = c ? acc1[jj] : acc0[jj];
                    derive27(t8, aij, aik, ajk, asi, asj, ask, c ? n1 : n0, c ? x1 : x0);
                }
                // K2 in the reference order (GPU/v4.cpp:579-583, CPU/v2.cpp:516-519)
                float sc = 0.0f;
#pragma unroll
                for (int m = 0; m < 27; ++m) {
                    // indices are <= N + 1 < table_len for exact counts; the
                    // clamp only keeps a corrupted count inside the table
                    const uint32_t ia = min(x0[m] + x1[m] + 1, tmax);
                    const uint32_t ib = min(x0[m], tmax);
                    const uint32_t id = min(x1[m], tmax);
                    const float a = TLDS ? Tl[ia] : P.table[ia];
                    const float b = TLDS ? Tl[ib] : P.table[ib];
                    const float d = TLDS ? Tl[id] : P.table[id];
                    sc += (a - b) - d;
                    // bound the number of table loads in flight (register pressure)
                    if (!TLDS && (m % 3) == 2)
                        __builtin_amdgcn_sched_barrier(0);
                }
                s = fabsf(sc);
                // Rows in another order than the original one (C->orig):
                // the key holds the original indices, and the reference sums
                // the cells in the order of those.  The stored-order sum is
                // within kOrderSlack of that sum, so only lanes that could
                // reach the top list are looked at; those whose axis order
                // is not the identity are counted again with their rows in
                // original index order and summed in order.
                uint32_t t3[3] = {i, j, k};
                bool again = false;
                if (C->orig && (DUMP || s * kOrderSlack <= fminf(W.gthr, W.ts))) {
                    const uint32_t oi = C->orig[i], oj = C->orig[j], ok = C->orig[k];
                    sort3(oi, oj, ok, t3);
                    key = make_key(t3[0], t3[1], t3[2]);
                    again = axis_order(oi, oj, ok) != 0;
                }
                if (again) {
                    auto row = [&](uint32_t t) {
                        return t == C->orig[i] ? i : t == C->orig[j] ? j : k;
                    };
                    const uint32_t ra = row(t3[0]), rb = row(t3[1]), rc = row(t3[2]);
                    uint32_t y0[27], y1[27];
                    recount27(C->sm[0], C->W[0], n0, ra, rb, rc, y0);
                    recount27(C->sm[1], C->W[1], n1, ra, rb, rc, y1);
                    float sr = 0.0f;
#pragma unroll
                    for (int m = 0; m < 27; ++m) {
                        const uint32_t ia = min(y0[m] + y1[m] + 1, tmax);
                        const uint32_t ib = min(y0[m], tmax);
                        const uint32_t id = min(y1[m], tmax);
                        const float a = TLDS ? Tl[ia] : P.table[ia];
                        const float b = TLDS ? Tl[ib] : P.table[ib];
                        const float d = TLDS ? Tl[id] : P.table[id];
                        sr += (a - b) - d;
                    }
                    s = fabsf(sr);
                    if (DUMP && P.dump_counts) {
                        const uint64_t rk = rank3(M, t3[0], t3[1], t3[2]);
#pragma unroll
                        for (int m = 0; m < 27; ++m) {
                            P.dump_counts[rk * 54 + m] = y0[m];
                            P.dump_counts[rk * 54 + 27 + m] = y1[m];
                        }
                    }
                } else if (DUMP && P.dump_counts) {
                    const uint64_t rk = rank3(M, t3[0], t3[1], t3[2]);
#pragma unroll
                    for (int m = 0; m < 27; ++m) {
                        P.dump_counts[rk * 54 + m] = x0[m];
                        P.dump_counts[rk * 54 + 27 + m] = x1[m];
                    }
                }
                if (DUMP)
                    P.dump_scores[rank3(M, t3[0], t3[1], t3[2])] = s;
            }
            if (!DUMP)
                topk::wave_offer(W, valid, s, key, lane);
        }
    }
}

template <int NW>
size_t lds_total(bool table_lds, uint32_t table_len)
{
    return scan_lds_fixed<NW>() + (table_lds ? (size_t)table_len * 4 : 0);
}

size_t scan_lds_bytes(const ScanConfig &cfg, bool table_lds, uint32_t table_len)
{
    if (cfg.variant >= 2)
        return mfma_lds_bytes(table_lds, table_len);
    switch (cfg.waves) {
    case 4: return lds_total<4>(table_lds, table_len);
    case 8: return lds_total<8>(table_lds, table_len);
    case 16: return lds_total<16>(table_lds, table_len);
    default: return 0;
    }
}

typedef void (*ScanFn)(const ScanParams);

template <int JP, int NW>
static ScanFn pick3(bool tl, bool dump)
{
    if (tl)
        return dump ? scan_kernel<JP, NW, true, true> : scan_kernel<JP, NW, true, false>;
    return dump ? scan_kernel<JP, NW, false, true> : scan_kernel<JP, NW, false, false>;
}

template <int JP>
static ScanFn pick2(int nw, bool tl, bool dump)
{
    switch (nw) {
    case 4: return pick3<JP, 4>(tl, dump);
    case 8: return pick3<JP, 8>(tl, dump);
    case 16: return pick3<JP, 16>(tl, dump);
    default: return nullptr;
    }
}

static ScanFn pick(const ScanConfig &cfg, bool tl, bool dump)
{
    if (cfg.words != 1 || cfg.variant != 1)
        return nullptr;
    switch (cfg.snps_j) {
    case 1: return pick2<1>(cfg.waves, tl, dump);
    case 2: return pick2<2>(cfg.waves, tl, dump);
    case 4: return pick2<4>(cfg.waves, tl, dump);
    case 8: return pick2<8>(cfg.waves, tl, dump);
    default: return nullptr;
    }
}

bool scan_config_supported(const ScanConfig &cfg)
{
    if (cfg.variant == 2 || cfg.variant == 3)
        return cfg.snps_j == 64 && cfg.waves == kMfmaWaves && cfg.words == 1;
    return pick(cfg, false, false) != nullptr;
}

hipError_t scan_occupancy(const ScanConfig &cfg, bool table_lds, bool dump, size_t lds_bytes,
                          int *blocks_per_cu)
{
    if (cfg.variant >= 2)
        return mfma_occupancy(cfg.variant == 3, table_lds, dump, lds_bytes, blocks_per_cu);
    ScanFn f = pick(cfg, table_lds, dump);
    if (!f)
        return hipErrorInvalidValue;
    return hipOccupancyMaxActiveBlocksPerMultiprocessor(blocks_per_cu, (const void *)f,
                                                        cfg.waves * 64, lds_bytes);
}

hipError_t scan_allow_lds(const ScanConfig &cfg, bool table_lds, bool dump, size_t lds_bytes)
{
    if (cfg.variant >= 2)
        return mfma_allow_lds(cfg.variant == 3, table_lds, dump, lds_bytes);
    ScanFn f = pick(cfg, table_lds, dump);
    if (!f)
        return hipErrorInvalidValue;
    return hipFuncSetAttribute((const void *)f, hipFuncAttributeMaxDynamicSharedMemorySize,
                               (int)lds_bytes);
}

hipError_t launch_scan(const ScanConfig &cfg, bool table_lds, bool dump, const ScanParams &p,
                       uint32_t grid, hipStream_t s)
{
    if (cfg.variant >= 2)
        return launch_mfma_scan(cfg.variant == 3, table_lds, dump, p, grid, s);
    ScanFn f = pick(cfg, table_lds, dump);
    if (!f)
        return hipErrorInvalidValue;
    const size_t lds = scan_lds_bytes(cfg, table_lds, p.table_len);
    hipLaunchKernelGGL(f, dim3(grid), dim3(cfg.waves * 64), lds, s, p);
    return hipGetLastError();
}

// ------------------------------------------------------------------------
// per-cohort merge of workgroup slots into the final top-kTop
// ------------------------------------------------------------------------
__global__ void fill_empty_kernel(Cand *p, size_t n)
{
    for (size_t t = blockIdx.x * (size_t)blockDim.x + threadIdx.x; t < n;
         t += (size_t)gridDim.x * blockDim.x) {
        Cand e;
        e.s = INFINITY;
        e.pad = 0;
        e.key = kEmptyKey;
        p[t] = e;
    }
}

hipError_t launch_fill_empty(Cand *p, size_t n, hipStream_t s)
{
    if (n == 0)
        return hipSuccess;
    size_t blocks = (n + 255) / 256;
    if (blocks > 1024)
        blocks = 1024;
    hipLaunchKernelGGL(fill_empty_kernel, dim3((uint32_t)blocks), dim3(256), 0, s, p, n);
    return hipGetLastError();
}

// A candidate held by one lane: score and the two halves of the key, each
// moved between lanes on its own.
struct LaneCand {
    float s;
    uint32_t lo, hi;
};
__device__ __forceinline__ bool lc_less(LaneCand a, LaneCand b)
{
    // no short-circuit: plain compares and selects, no branches
    const uint64_t ka = ((uint64_t)a.hi << 32) | a.lo, kb = ((uint64_t)b.hi << 32) | b.lo;
    return (a.s < b.s) | ((a.s == b.s) & (ka < kb));
}
__device__ __forceinline__ LaneCand lc_of(const Cand &c)
{
    return {c.s, (uint32_t)c.key, (uint32_t)(c.key >> 32)};
}
__device__ __forceinline__ Cand lc_cand(const LaneCand &v)
{
    Cand c;
    c.s = v.s;
    c.pad = 0;
    c.key = ((uint64_t)v.hi << 32) | v.lo;
    return c;
}
__device__ __forceinline__ LaneCand lc_empty() { return {INFINITY, ~0u, ~0u}; }
// c ? a : b field by field (a select of the whole struct goes through scratch)
__device__ __forceinline__ LaneCand lc_sel(bool c, LaneCand a, LaneCand b)
{
    return {c ? a.s : b.s, c ? a.lo : b.lo, c ? a.hi : b.hi};
}
__device__ __forceinline__ LaneCand lc_xor(const LaneCand &v, int m)
{
    return {__shfl_xor(v.s, m), (uint32_t)__shfl_xor((int)v.lo, m),
            (uint32_t)__shfl_xor((int)v.hi, m)};
}
__device__ __forceinline__ LaneCand lc_from(const LaneCand &v, int src)
{
    return {__shfl(v.s, src), (uint32_t)__shfl((int)v.lo, src), (uint32_t)__shfl((int)v.hi, src)};
}
// lane b's entry (b wave-uniform)
__device__ __forceinline__ LaneCand lc_read(const LaneCand &v, int b)
{
    return {__uint_as_float((uint32_t)__builtin_amdgcn_readlane((int)__float_as_uint(v.s), b)),
            (uint32_t)__builtin_amdgcn_readlane((int)v.lo, b),
            (uint32_t)__builtin_amdgcn_readlane((int)v.hi, b)};
}

// One compare-exchange step of a bitonic network over the 64 lanes: lanes
// `lane` and `lane ^ j` swap unless the lower one holds the smaller entry
// (asc) or the larger (!asc).
__device__ __forceinline__ void lc_step(LaneCand &v, uint32_t lane, int j, bool asc)
{
    const LaneCand p = lc_xor(v, j);
    const bool want_min = ((lane & (uint32_t)j) == 0) == asc;
    const bool p_less = lc_less(p, v), v_less = lc_less(v, p);
    v = lc_sel(want_min ? p_less : v_less, p, v);
}
// Sorts the wavefront's 64 entries, lane 0 first: ascending, or descending.
__device__ __forceinline__ void lc_sort(LaneCand &v, uint32_t lane, bool desc)
{
#pragma unroll
    for (int k = 2; k <= 64; k <<= 1)
#pragma unroll
        for (int j = k >> 1; j > 0; j >>= 1)
            lc_step(v, lane, j, ((lane & (uint32_t)k) == 0) != desc);
}
// a ascending, d descending: the 64 best of both, ascending.
__device__ __forceinline__ LaneCand lc_merge(const LaneCand &a, const LaneCand &d, uint32_t lane)
{
    LaneCand v = lc_sel(lc_less(d, a), d, a);  // bitonic, holding the 64 best
#pragma unroll
    for (int j = 32; j > 0; j >>= 1)
        lc_step(v, lane, j, true);
    return v;
}

// One workgroup of 16 wavefronts per cohort: folds the entries the scan
// appended into the cohort's running list.  Each wavefront keeps a sorted
// list of its 64 best in registers (lane l holds the l-th; wavefront 0
// starts from the running list) and takes the entries 64 at a time; an
// entry counts only if it beats both the wavefront's 64th best and the
// workgroup's bound `bthr`, the smallest 64th-best score any of them holds
// (or the cohort threshold the scan left: 64 distinct triplets at or below
// it were appended or are in the running list).  A few such entries are
// inserted one by one (rank by a ballot, the rest shift up a lane), more
// are sorted and merged in (bitonic, no barriers).  Then the 16 lists are
// merged pairwise in LDS, four levels.
__global__ void __launch_bounds__(1024)
merge_kernel(const RoundHdr *__restrict__ hdr, const Cand *__restrict__ slots,
             const uint32_t *__restrict__ counters, const uint32_t *__restrict__ thr,
             uint32_t cap, Cand *const *lists)
{
    constexpr int kWaves = 16, kInsertMax = 8;
    __shared__ Cand part[kWaves][kTop];
    __shared__ uint32_t bthr;
    const uint32_t cohort = blockIdx.x;
    if (cohort >= hdr->n_cohorts)
        return;
    const uint32_t t = threadIdx.x;
    const uint32_t lane = t & 63u;
    const uint32_t wave = t >> 6;
    uint32_t n = counters[1 + cohort];
    if (n > cap)
        n = cap;
    const Cand *src = slots + (size_t)cohort * cap;
    Cand *list = lists[cohort];
    if (t == 0)
        bthr = thr[cohort];
    LaneCand L = lc_sel(wave == 0, lc_of(list[lane]), lc_empty());
    LaneCand L63 = lc_read(L, 63);
    __syncthreads();
    const uint32_t chunks = (n + 63) / 64;
    uint32_t c = wave;
    LaneCand e = lc_empty();
    if (c * 64 + lane < n)
        e = lc_of(src[c * 64 + lane]);
    for (; c < chunks; c += kWaves) {
        const LaneCand cur = e;
        const uint32_t x = (c + kWaves) * 64 + lane;
        e = lc_empty();
        if (x < n)
            e = lc_of(src[x]);
        const float tb = __uint_as_float(__atomic_load_n(&bthr, __ATOMIC_RELAXED));
        const bool pass = cur.s <= tb && lc_less(cur, L63);
        uint64_t m = __ballot(pass);
        if (m == 0)
            continue;
        if (__popcll(m) <= kInsertMax) {
            while (m) {
                const int b = __builtin_ctzll(m);
                m &= m - 1;
                const LaneCand E = lc_read(cur, b);
                if (!lc_less(E, L63))
                    continue;
                const uint32_t pos = (uint32_t)__popcll(__ballot(lc_less(L, E)));
                const LaneCand up = lc_from(L, (int)lane - 1);
                L = lc_sel(lane > pos, up, lc_sel(lane == pos, E, L));
                L63 = lc_read(L, 63);
            }
        } else {
            LaneCand d = lc_sel(pass, cur, lc_empty());
            lc_sort(d, lane, true);
            L = lc_merge(L, d, lane);
            L63 = lc_read(L, 63);
        }
        if (lane == 0 && L63.s < tb)
            atomicMin(&bthr, __float_as_uint(L63.s));
    }
    part[wave][lane] = lc_cand(L);
    __syncthreads();
    for (uint32_t s = 1; s < (uint32_t)kWaves; s <<= 1) {
        if ((wave & (2 * s - 1)) == 0) {
            const LaneCand a = lc_of(part[wave][lane]);
            const LaneCand d = lc_of(part[wave + s][kTop - 1 - lane]);
            part[wave][lane] = lc_cand(lc_merge(a, d, lane));
        }
        __syncthreads();
    }
    if (wave == 0)
        list[lane] = part[0][lane];
}

// Per-launch cohort thresholds: the running list's 64th best score (+inf
// while it holds fewer than 64 candidates) or the cohort's seed, whichever
// is smaller (both bound the 64th best from above).
__global__ void init_thr_kernel(const RoundHdr *__restrict__ hdr,
                                const CohortDev *__restrict__ cohorts, Cand *const *lists,
                                uint32_t *thr, uint32_t max_cohorts)
{
    const uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= max_cohorts)
        return;
    thr[c] = c < hdr->n_cohorts ? __float_as_uint(fminf(lists[c][kTop - 1].s, cohorts[c].seed))
                                : 0x7F800000u;
}

hipError_t launch_init_thr(const RoundHdr *hdr, const CohortDev *cohorts, Cand *const *lists,
                           uint32_t *thr, uint32_t max_cohorts, hipStream_t s)
{
    if (max_cohorts == 0)
        return hipSuccess;
    hipLaunchKernelGGL(init_thr_kernel, dim3((max_cohorts + 63) / 64), dim3(64), 0, s, hdr,
                       cohorts, lists, thr, max_cohorts);
    return hipGetLastError();
}

hipError_t launch_merge(const RoundHdr *hdr, const Cand *slots, const uint32_t *counters,
                        const uint32_t *thr, uint32_t slot_cap, Cand *const *lists,
                        uint32_t max_cohorts, hipStream_t s)
{
    if (max_cohorts == 0)
        return hipSuccess;
    hipLaunchKernelGGL(merge_kernel, dim3(max_cohorts), dim3(1024), 0, s, hdr, slots, counters,
                       thr, slot_cap, lists);
    return hipGetLastError();
}

// ------------------------------------------------------------------------
// FP32 + FP64 rescoring of candidates with explicit 27-cell counting
// (the reference's counting form, GPU/v4.cpp:321-577, on 64-bit words).
// One wavefront per candidate.
// ------------------------------------------------------------------------
__global__ void __launch_bounds__(64)
rescore_kernel(const CohortDev *__restrict__ cohorts, const Cand *__restrict__ cands,
               Cand *const *lists, uint32_t per_cohort, const float *__restrict__ t32,
               const double *__restrict__ t64, Rescored *__restrict__ out)
{
    const uint32_t e = blockIdx.x;
    const uint32_t cohort = blockIdx.y;
    const uint32_t lane = threadIdx.x;
    const Cand cd = lists ? lists[cohort][e] : cands[(size_t)cohort * per_cohort + e];
    Rescored *o = out + (size_t)cohort * per_cohort + e;
    if (cd.key == kEmptyKey) {
        if (lane == 0)
            o->valid = 0;
        return;
    }
    const uint32_t i = key_i(cd.key), j = key_j(cd.key), k = key_k(cd.key);
    const CohortDev &C = cohorts[cohort];
    // the rows of SNPs i, j, k (in original index order, so the cells come
    // out in the reference's order 9 g_i + 3 g_j + g_k)
    const uint32_t ri = C.pos ? C.pos[i] : i, rj = C.pos ? C.pos[j] : j,
                   rk = C.pos ? C.pos[k] : k;
    uint32_t cnt[54];
#pragma unroll
    for (int m = 0; m < 54; ++m)
        cnt[m] = 0;
#pragma unroll
    for (int c = 0; c < 2; ++c) {
        const uint32_t W = C.W[c];
        const uint32_t n = C.n[c];
        const uint4 *S = C.sm[c];
        for (uint32_t w = lane; w < W; w += 64) {
            const uint4 a = S[(size_t)ri * W + w];
            const uint4 b = S[(size_t)rj * W + w];
            const uint4 d = S[(size_t)rk * W + w];
            const uint32_t rem = n - w * 64;
            const uint64_t live = rem >= 64 ? ~0ull : ((1ull << rem) - 1ull);
            uint64_t x[3], y[3], z[3];
            x[0] = a.x | ((uint64_t)a.y << 32);
            x[1] = a.z | ((uint64_t)a.w << 32);
            x[2] = ~(x[0] | x[1]) & live;
            y[0] = b.x | ((uint64_t)b.y << 32);
            y[1] = b.z | ((uint64_t)b.w << 32);
            y[2] = ~(y[0] | y[1]) & live;
            z[0] = d.x | ((uint64_t)d.y << 32);
            z[1] = d.z | ((uint64_t)d.w << 32);
            z[2] = ~(z[0] | z[1]) & live;
#pragma unroll
            for (int gi = 0; gi < 3; ++gi)
#pragma unroll
                for (int gj = 0; gj < 3; ++gj)
#pragma unroll
                    for (int gk = 0; gk < 3; ++gk)
                        cnt[27 * c + 9 * gi + 3 * gj + gk] += __popcll(x[gi] & y[gj] & z[gk]);
        }
    }
#pragma unroll
    for (int m = 0; m < 54; ++m) {
        uint32_t v = cnt[m];
#pragma unroll
        for (int off = 32; off > 0; off >>= 1)
            v += __shfl_xor(v, off);
        cnt[m] = v;
    }
    if (lane == 0) {
        float s32 = 0.0f;
        double s64 = 0.0;
#pragma unroll
        for (int m = 0; m < 27; ++m) {
            s32 += (t32[cnt[m] + cnt[27 + m] + 1] - t32[cnt[m]]) - t32[cnt[27 + m]];
            s64 += (t64[cnt[m] + cnt[27 + m] + 1] - t64[cnt[m]]) - t64[cnt[27 + m]];
        }
        o->i = i;
        o->j = j;
        o->k = k;
        o->valid = 1;
        o->s32 = fabsf(s32);
        o->s32_scan = cd.s;
        o->s64 = fabs(s64);
#pragma unroll
        for (int m = 0; m < 54; ++m)
            o->counts[m] = cnt[m];
    }
}

// The kTop-th smallest FP32 score of n rescored triplets (kTop <= n <=
// 8 x 1024) into *out: a binary search over the score bits (scores are >= 0,
// so their bits order like the values) with a block-wide count per step,
// one LDS add per wavefront.  The three counters rotate: thread 0 clears the
// one the previous step read (every wavefront has read it before this
// step's barrier), which is next added to two steps later, after the next
// barrier.
__global__ void __launch_bounds__(1024) seed_select_kernel(const Rescored *__restrict__ rs,
                                                           uint32_t n, float *out)
{
    __shared__ uint32_t cnt[3];
    const uint32_t t = threadIdx.x;
    uint32_t v[8];
#pragma unroll
    for (int q = 0; q < 8; ++q) {
        const uint32_t x = t + 1024u * q;
        v[q] = x < n ? __float_as_uint(rs[x].s32) : 0xFFFFFFFFu;
    }
    if (t < 3)
        cnt[t] = 0;
    __syncthreads();
    uint32_t lo = 0, hi = 0x7F800000u;  // at least kTop scores are <= hi
    for (uint32_t it = 0; lo < hi; ++it) {
        const uint32_t mid = lo + (hi - lo) / 2;
        uint32_t c = 0;
#pragma unroll
        for (int q = 0; q < 8; ++q)
            c += v[q] <= mid;
#pragma unroll
        for (int m = 32; m > 0; m >>= 1)
            c += (uint32_t)__shfl_xor((int)c, m);
        if ((t & 63u) == 0)
            atomicAdd(&cnt[it % 3], c);
        __syncthreads();
        const uint32_t tot = cnt[it % 3];
        if (t == 0)
            cnt[(it + 2) % 3] = 0;
        if (tot >= (uint32_t)kTop)
            hi = mid;
        else
            lo = mid + 1;
    }
    if (t == 0)
        *out = __uint_as_float(lo);
}

hipError_t launch_seed_select(const Rescored *rs, uint32_t n, float *out, hipStream_t s)
{
    if (n < (uint32_t)kTop || n > 8192)
        return hipErrorInvalidValue;
    hipLaunchKernelGGL(seed_select_kernel, dim3(1), dim3(1024), 0, s, rs, n, out);
    return hipGetLastError();
}

hipError_t launch_rescore(const CohortDev *cohorts, const Cand *cands, Cand *const *lists,
                          uint32_t per_cohort, uint32_t n_cohorts, const float *t32,
                          const double *t64, Rescored *out, hipStream_t s)
{
    if (per_cohort == 0 || n_cohorts == 0)
        return hipSuccess;
    hipLaunchKernelGGL(rescore_kernel, dim3(per_cohort, n_cohorts), dim3(64), 0, s, cohorts,
                       cands, lists, per_cohort, t32, t64, out);
    return hipGetLastError();
}

}  // namespace epi
