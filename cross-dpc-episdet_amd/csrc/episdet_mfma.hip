// episdet_mfma.hip -- matrix-core formulations of the 8 popcounted cells
// (the alternative to the VALU AND+popcount scan kernel).
//
// For one class and a fixed SNP i, the counts the VALU kernel popcounts,
//   t[a][b][c](i,j,k) = sum_s [g_i(s)=a] [g_j(s)=b] [g_k(s)=c],  a,b,c in {0,1},
// are a matrix product over samples s:
//   A[(j,a,b)][s] = [g_i(s)=a] & [g_j(s)=b]      (0/1)
//   B[s][(k,c)]   = [g_k(s)=c]                    (0/1)
//   C = A B
// in one of two number formats, both exact:
//   int8 (variant 2): 0/1 bytes, v_mfma_i32_32x32x32_i8, int32 accumulation;
//   FP4  (variant 3): e2m1 nibbles 0 and 1.0 (0x2), two samples per byte,
//        v_mfma_scale_f32_32x32x64_f8f6f4 (both block scales 2^0), FP32
//        accumulation.  Every product is 0 or 1 and every partial sum an
//        integer below 2^24 (class sizes are below 2^21), so the FP32 sums are
//        exact; the instruction does twice the int8 instruction's MACs in the
//        same issue time, and the operands take half the bytes.
// The operand bytes come from a per-cohort expansion of the bit planes
// (E[c] or F[c], [snp][plane][sample]).  Both formats use one LDS layout:
// 128-byte rows per stage, 32 bytes per MFMA step (32 samples as bytes, 64
// as nibbles); A and B map a byte position to the same sample, so the sum
// over the K dimension covers each sample once whatever that mapping is.
//
// Work item = (i, j-block of 64 SNPs, k-block of 64 SNPs), one workgroup of
// 4 wavefronts.  Per stage of 128 samples the workgroup stages the j-block
// and k-block bytes (and SNP i's) in LDS; wavefront w computes the 64 rows
// (16 j x 4 (a,b)) of its quarter against all 128 columns (64 k x 2 c):
// 2 x 4 MFMA tiles of 32 x 32 per class, 256 int32 accumulators.
// Pipelining: two LDS stage buffers, so one barrier per stage; the next
// stage is copied from global memory straight into the other buffer
// (global_load_lds, no registers) while this stage's MFMAs run; within a
// stage the operands of the next 32-sample step are read from LDS while the
// current step's MFMAs run, and a stage's last step runs after the barrier
// together with the next stage's first reads.  The epilogue is the VALU kernel's: exact 27-cell
// reconstruction from the margins, K2 in the reference order, per-wavefront
// top-64.
#include <math.h>

#include "episdet_cells.h"
#include "episdet_internal.h"
#include "episdet_topk.h"

#define EPI_CONST __attribute__((address_space(4)))

namespace epi {

typedef int v4i __attribute__((ext_vector_type(4)));
typedef int v16i __attribute__((ext_vector_type(16)));

namespace mf {

template <typename T>
__device__ __forceinline__ const EPI_CONST T *as_const(const T *p)
{
    return (const EPI_CONST T *)p;
}

__device__ __forceinline__ uint32_t rfl(uint32_t v) { return __builtin_amdgcn_readfirstlane(v); }

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

}  // namespace mf

// ------------------------------------------------------------------------
// bit planes -> 0/1 bytes: E[c] row (snp, plane) holds EL bytes (64 W_c
// rounded up to a whole 128-byte stage), byte t = bit t of the plane
// (sample t of the class); padding bytes are 0.  One thread per (snp,
// 64-bit word slot, plane) writes 64 bytes; slots w >= W write zeros.
// ------------------------------------------------------------------------
__global__ void __launch_bounds__(256)
expand_kernel(const uint4 *__restrict__ sm, uint32_t Mp, uint32_t W, uint32_t EL,
              uint8_t *__restrict__ E)
{
    const size_t t = blockIdx.x * (size_t)256 + threadIdx.x;
    const uint32_t WS = EL / 64;  // word slots per row
    const size_t total = (size_t)Mp * WS * 2;
    if (t >= total)
        return;
    const uint32_t plane = (uint32_t)(t & 1);
    const size_t sw = t >> 1;  // snp * WS + w
    const uint32_t w = (uint32_t)(sw % WS);
    const size_t snp = sw / WS;
    const uint4 v = w < W ? sm[snp * W + w] : make_uint4(0, 0, 0, 0);
    const uint32_t lo = plane ? v.z : v.x;
    const uint32_t hi = plane ? v.w : v.y;
    uint32_t out[16];
#pragma unroll
    for (int q = 0; q < 16; ++q) {
        const uint32_t word = q < 8 ? lo : hi;
        const uint32_t nib = (word >> (4 * (q & 7))) & 0xFu;
        // spread 4 bits to the low bit of 4 bytes
        out[q] = (nib * 0x00204081u) & 0x01010101u;
    }
    uint4 *dst = (uint4 *)(E + (snp * 2 + plane) * (size_t)EL + (size_t)w * 64);
#pragma unroll
    for (int q = 0; q < 4; ++q)
        dst[q] = make_uint4(out[4 * q], out[4 * q + 1], out[4 * q + 2], out[4 * q + 3]);
}

hipError_t launch_expand(const uint4 *sm, uint32_t Mp, uint32_t W, uint32_t EL, uint8_t *E,
                         hipStream_t s)
{
    const size_t total = (size_t)Mp * (EL / 64) * 2;
    if (total == 0)
        return hipSuccess;
    hipLaunchKernelGGL(expand_kernel, dim3((uint32_t)((total + 255) / 256)), dim3(256), 0, s, sm,
                       Mp, W, EL, E);
    return hipGetLastError();
}

// bit planes -> e2m1 nibbles: F[c] row (snp, plane) holds FL bytes (32 W_c
// rounded up to a whole 128-byte stage); sample t is nibble t % 8 of dword
// t / 8, 0x2 (1.0) when bit t of the plane is set, else 0; padding is 0.
// One thread per (snp, 64-bit word slot, plane) writes 32 bytes.
__global__ void __launch_bounds__(256)
expand_fp4_kernel(const uint4 *__restrict__ sm, uint32_t Mp, uint32_t W, uint32_t FL,
                  uint8_t *__restrict__ F)
{
    const size_t t = blockIdx.x * (size_t)256 + threadIdx.x;
    const uint32_t WS = FL / 32;  // word slots per row
    const size_t total = (size_t)Mp * WS * 2;
    if (t >= total)
        return;
    const uint32_t plane = (uint32_t)(t & 1);
    const size_t sw = t >> 1;  // snp * WS + w
    const uint32_t w = (uint32_t)(sw % WS);
    const size_t snp = sw / WS;
    const uint4 v = w < W ? sm[snp * W + w] : make_uint4(0, 0, 0, 0);
    const uint32_t lo = plane ? v.z : v.x;
    const uint32_t hi = plane ? v.w : v.y;
    uint32_t out[8];
#pragma unroll
    for (int q = 0; q < 8; ++q) {
        // spread 8 bits to bit 0 of 8 nibbles, then make each a 1.0 (0x2)
        uint32_t x = ((q < 4 ? lo : hi) >> (8 * (q & 3))) & 0xFFu;
        x = (x | (x << 12)) & 0x000F000Fu;
        x = (x | (x << 6)) & 0x03030303u;
        x = (x | (x << 3)) & 0x11111111u;
        out[q] = x << 1;
    }
    uint4 *dst = (uint4 *)(F + (snp * 2 + plane) * (size_t)FL + (size_t)w * 32);
    dst[0] = make_uint4(out[0], out[1], out[2], out[3]);
    dst[1] = make_uint4(out[4], out[5], out[6], out[7]);
}

hipError_t launch_expand_fp4(const uint4 *sm, uint32_t Mp, uint32_t W, uint32_t FL, uint8_t *F,
                             hipStream_t s)
{
    const size_t total = (size_t)Mp * (FL / 32) * 2;
    if (total == 0)
        return hipSuccess;
    hipLaunchKernelGGL(expand_fp4_kernel, dim3((uint32_t)((total + 255) / 256)), dim3(256), 0, s,
                       sm, Mp, W, FL, F);
    return hipGetLastError();
}

// One MFMA step of a 32 x 32 tile: 16 operand bytes per lane, as 16 int8
// samples (int32 accumulators) or 32 e2m1 samples (FP32 accumulators; the
// instruction reads 4 of the 8 operand registers for e2m1, format code 4,
// and scale exponent 127 = 2^0).
typedef int v8i __attribute__((ext_vector_type(8)));
typedef float v16f __attribute__((ext_vector_type(16)));

__device__ __forceinline__ v16i mf_mma(v4i a, v4i b, v16i c)
{
    return __builtin_amdgcn_mfma_i32_32x32x32_i8(a, b, c, 0, 0, 0);
}
__device__ __forceinline__ v16f mf_mma(v4i a, v4i b, v16f c)
{
    const v8i a8 = {a[0], a[1], a[2], a[3], 0, 0, 0, 0};
    const v8i b8 = {b[0], b[1], b[2], b[3], 0, 0, 0, 0};
    return __builtin_amdgcn_mfma_scale_f32_32x32x64_f8f6f4(a8, b8, c, 4, 4, 0, 127, 0, 127);
}
template <bool FP4>
struct MfAcc {
    typedef v16i type;
};
template <>
struct MfAcc<true> {
    typedef v16f type;
};

// ------------------------------------------------------------------------
// the MFMA scan kernel
// ------------------------------------------------------------------------
constexpr int kMfWaves = 4;
constexpr int kStage = 128;           // bytes per row per LDS stage (four 32-byte MFMA steps)
constexpr int kRowBytes = kStage;     // LDS bytes per staged (snp, plane) row
constexpr int kStageRows = 64 * 2;    // 64 SNPs x 2 planes
constexpr int kStageCopies = 2 * kStageRows * kRowBytes / 1024 / kMfWaves;  // 1 KiB copies per wave

// Staged rows are 128 bytes with no padding; 16-byte slot s of row r holds
// chunk s ^ ((r >> 1) & 7) of the row.  LDS serves 16 lanes of a 16-byte
// read per cycle from 64 banks (256 bytes): even rows fall on one half of
// the banks and odd rows on the other, and the swizzle spreads the 8 even
// (or odd) rows of 16 consecutive rows over the 8 slots of their half, so
// the operand reads (16 lanes: up to 16 different rows at one chunk) are
// free of bank conflicts.
struct MfStage {
    uint8_t jrows[kStageRows * kRowBytes];  // j-block bytes of the stage
    uint8_t krows[kStageRows * kRowBytes];  // k-block bytes of the stage (right after)
    uint8_t irow[2 * kStage];               // SNP i, both planes, unswizzled
};
static_assert(__builtin_offsetof(MfStage, krows) == kStageRows * kRowBytes, "k rows follow j rows");

// The margins an item's epilogue reads, staged in LDS over the (then idle)
// stage buffers.  The j-block is split in two halves of 32 SNPs, one per
// accumulator row tile fa: half fa holds j_local = 16w + 8fa + r (w < 4,
// r < 8) as row 8w + r.
constexpr int kJkPitch = 64 + 1;  // uint4 per JK row (+1: bank offset)
struct MfMargins {
    uint4 jk[2][32][kJkPitch];  // pair margins of (j, k), one j half, both classes
    uint4 ij[2][64];            // pair margins of (i, j)
    uint4 ik[2][64];            // pair margins of (i, k)
    uint2 sj[2][64];            // single-SNP counts of the j-block
    uint2 sk[2][64];            // and of the k-block
};

struct MfLds {
    Cand cand[kMfWaves][kWaveCap + kTop];
    uint32_t misc[4];
    union {
        MfStage stage[2];  // double buffer: one is read while the other is filled
        MfMargins mg;      // the epilogue's margins (after the stage loop)
    };
};

__host__ __device__ constexpr size_t mf_lds_fixed() { return (sizeof(MfLds) + 15) & ~(size_t)15; }

// 16 bytes of LDS as an MFMA operand
__device__ __forceinline__ v4i lds16(const uint8_t *p) { return *(const v4i *)p; }

typedef __attribute__((address_space(3))) uint8_t lds_u8;

// One 16-byte-per-lane copy from global memory into LDS (lane L writes
// bytes 16L..16L+15 after the wave-uniform LDS address `lds`).  Written as
// asm so that the compiler's wait-count model, which counts these copies
// as LDS operations too and then waits for every outstanding LDS read,
// keeps exact counts for the operand reads; completion is awaited
// explicitly (s_waitcnt vmcnt) before the barrier that publishes a stage.
// The instruction has no register result.
__device__ __forceinline__ void lds_copy16(const uint8_t *src, uint32_t lds)
{
    asm volatile("s_mov_b32 m0, %1\n\tglobal_load_lds_dwordx4 %0, off"
                 :
                 : "v"(src), "s"(lds)
                 : "memory", "m0");
}

// LDS byte address of a pointer into LDS.
__device__ __forceinline__ uint32_t lds_addr(const void *p)
{
    return (uint32_t)(uintptr_t)(const lds_u8 *)p;
}

// Copies one stage from global memory straight into an LDS stage buffer
// (global_load_lds_dwordx4: lane L of a copy writes bytes 16L..16L+15 of a
// 1 KiB block given by M0).  The j and k rows (32 KiB) are 32 copies of 8
// rows each, 8 per wavefront; lane L takes row L/8 of its copy and reads
// the global chunk that the swizzle puts in its slot L%8.  SNP
// i's 256 bytes are one 4-lane copy per wavefront.  Every wavefront issues
// kStageCopies + 1 copies, so one vmcnt value covers a stage on all.
// `stage` is the LDS byte address of the MfStage to fill (wave-uniform).
__device__ __forceinline__ void mf_issue_stage(uint32_t stage, const uint8_t *E, uint32_t Lc,
                                               uint32_t s0, uint32_t J0, uint32_t K0, uint32_t i,
                                               uint32_t wave, uint32_t lane)
{
    const uint32_t rsub = lane >> 3;
#pragma unroll
    for (int q = 0; q < kStageCopies; ++q) {
        // rows first..first+7 of the stage's 256 (j rows, then k rows)
        const uint32_t first = (wave * kStageCopies + q) * 8;
        const uint32_t blk = first >> 7, rowi = (first & 127u) + rsub;
        const uint32_t chunk = (lane & 7u) ^ ((rowi >> 1) & 7u);
        const uint32_t snp = (blk ? K0 : J0) + (rowi >> 1);
        const uint8_t *src = E + ((size_t)snp * 2 + (rowi & 1u)) * Lc + s0 + 16 * chunk;
        lds_copy16(src, stage + first * kRowBytes);  // krows follows jrows
    }
    if (lane < 4) {
        const uint32_t e = 4 * wave + lane;  // 16-byte chunk of SNP i's two rows
        const uint8_t *src = E + ((size_t)i * 2 + (e >> 3)) * Lc + s0 + 16 * (e & 7u);
        lds_copy16(src, stage + (uint32_t)__builtin_offsetof(MfStage, irow) + 64 * wave);
    }
}

// Stages one j half of the item's (j, k) pair margins, and with half 0 also
// the (i, j) and (i, k) margins and the single-SNP counts of both blocks.
// Entries with j >= k or i >= j belong to lanes that are scored but never
// offered; all indices stay below Mp.
__device__ __forceinline__ void mf_stage_margins(MfMargins &G, const EPI_CONST CohortDev *C,
                                                 uint32_t i, uint32_t J0, uint32_t K0,
                                                 uint32_t Mp, int fa, uint32_t tid)
{
    constexpr int kLoads = 2 * 32 * 64 / (kMfWaves * 64);  // uint4 per thread
    uint4 v[kLoads];
#pragma unroll
    for (int q = 0; q < kLoads; ++q) {
        const uint32_t idx = tid + 256u * q;  // (c, row, k_local)
        const uint32_t c = idx >> 11, row = (idx >> 6) & 31u, kl = idx & 63u;
        const uint32_t jl = 16 * (row >> 3) + 8 * fa + (row & 7u);
        v[q] = C->pair[c][(size_t)(J0 + jl) * Mp + K0 + kl];
    }
    uint4 x = make_uint4(0, 0, 0, 0);
    uint2 y = make_uint2(0, 0);
    if (fa == 0) {
        const uint32_t c = (tid >> 7) & 1u, e = tid & 63u;
        if (tid & 64u)
            x = C->pair[c][(size_t)i * Mp + K0 + e];
        else
            x = C->pair[c][(size_t)i * Mp + J0 + e];
        y = C->single[c][((tid & 64u) ? K0 : J0) + e];
    }
#pragma unroll
    for (int q = 0; q < kLoads; ++q) {
        const uint32_t idx = tid + 256u * q;
        G.jk[idx >> 11][(idx >> 6) & 31u][idx & 63u] = v[q];
    }
    if (fa == 0) {
        const uint32_t c = (tid >> 7) & 1u, e = tid & 63u;
        if (tid & 64u) {
            G.ik[c][e] = x;
            G.sk[c][e] = y;
        } else {
            G.ij[c][e] = x;
            G.sj[c][e] = y;
        }
    }
}

template <bool TLDS, bool DUMP, bool FP4>
__global__ void __launch_bounds__(kMfWaves * 64, 1) mfma_scan_kernel(const ScanParams P)
{
    extern __shared__ __align__(16) unsigned char lds_raw[];
    MfLds &L = *reinterpret_cast<MfLds *>(lds_raw);
    const uint32_t stage_lds[2] = {lds_addr(&L.stage[0]), lds_addr(&L.stage[1])};
    float *Tl = reinterpret_cast<float *>(lds_raw + mf_lds_fixed());

    const uint32_t tid = threadIdx.x;
    const uint32_t lane = tid & 63u;
    const uint32_t wave = mf::rfl(tid >> 6);
    topk::WaveTop Wt;
    Wt.buf = L.cand[wave];
    Wt.keep = Wt.buf + kWaveCap;
    Wt.reset();

    if (TLDS)
        for (uint32_t t = tid; t < P.table_len; t += kMfWaves * 64)
            Tl[t] = P.table[t];

    uint32_t cur = 0xFFFFFFFFu;
    uint32_t sg = 0, r = 0;
    bool fresh = true;
    const EPI_CONST Segment *segs = mf::as_const(P.segs);
    const uint32_t n_segs = mf::as_const(P.hdr)->n_segments;
    const uint32_t n_items = mf::as_const(P.hdr)->n_items;

    // lane roles in the MFMA tiles
    const uint32_t half = lane >> 5;   // which 16 bytes of a 32-sample step
    const uint32_t l32 = lane & 31u;   // A row / B column within a tile

    // lane-constant LDS offsets of the MFMA operands (bytes, within a stage
    // buffer): row offsets, and per 32-sample step the swizzled slot
    const uint32_t a_lane = (l32 >> 1) & 1u;  // g_i plane of this lane's A rows
    uint32_t ja_off[2];
#pragma unroll
    for (int fa = 0; fa < 2; ++fa) {
        // A row l32 of tile fa = (j_local, a, b), j_local = 16w + 8fa + l32/4
        const uint32_t jl = 16 * wave + 8 * fa + (l32 >> 2);
        ja_off[fa] = (jl * 2 + (l32 & 1u)) * kRowBytes;
    }
    // B column l32 of tile fb = (k_local, c), k_local = 16fb + l32/2: row 32fb + l32
    const uint32_t kb_off = l32 * kRowBytes;
    // swizzle keys (row >> 1) & 7 of this lane's A and B rows
    // (A row 2 j_local + b: (row >> 1) & 7 = j_local & 7 = (l32 / 4) & 7;
    //  B row 32 fb + l32: (l32 / 2) & 7)
    const uint32_t jx = (l32 >> 2) & 7u;
    const uint32_t kx = (l32 >> 1) & 7u;
    uint32_t ja_slot[kStage / 32], kb_slot[kStage / 32];
#pragma unroll
    for (int kk = 0; kk < kStage / 32; ++kk) {
        ja_slot[kk] = 16 * ((2 * kk + half) ^ jx);
        kb_slot[kk] = 16 * ((2 * kk + half) ^ kx);
    }
    const uint32_t ei_off = a_lane * kStage + 16 * half;

    // Work-queue items: the next item is claimed while this one runs.
    if (tid == 0)
        L.misc[0] = atomicAdd(&P.counters[0], 1u);
    __syncthreads();
    uint32_t item = mf::rfl(L.misc[0]);
    for (;;) {
        __syncthreads();  // every wavefront has read misc[0]
        uint32_t claimed = 0;
        if (tid == 0 && item < n_items)
            claimed = atomicAdd(&P.counters[0], 1u);
        uint32_t cohort = 0xFFFFFFFFu, u = 0;
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
                topk::wave_flush(Wt, P, cur, lane);
            cur = cohort;
        }
        if (item >= n_items)
            break;
        if (!DUMP) {
            Wt.gthr = fminf(Wt.gthr, topk::load_thr(P.thr, cohort));
            Wt.thr_slot = P.thr + cohort;
        }

        const uint32_t K0 = rows[r].kb * 64;
        const uint32_t J0 = rows[r].jg * 64;
        const uint32_t i = u - rows[r].item_start;  // one i per item
        const EPI_CONST CohortDev *C = mf::as_const(P.cohorts) + cohort;
        const uint32_t M = C->M, Mp = C->Mp;

        typename MfAcc<FP4>::type acc[2][2][4];
#pragma unroll
        for (int c = 0; c < 2; ++c)
#pragma unroll
            for (int fa = 0; fa < 2; ++fa)
#pragma unroll
                for (int fb = 0; fb < 4; ++fb)
#pragma unroll
                    for (int e = 0; e < 16; ++e)
                        acc[c][fa][fb][e] = 0;

        // Stage loop (see the pipelining note at the top of this file).
#pragma unroll
        for (int c = 0; c < 2; ++c) {
            const uint8_t *E = FP4 ? C->F[c] : C->E[c];
            const uint32_t Lc = FP4 ? C->FL[c] : C->EL[c];  // bytes per (snp, plane) row, a multiple of kStage
            if (Lc == 0)
                continue;
            const uint32_t n_st = Lc / kStage;
            // Both buffers are free here: the previous stage loop ended with
            // a barrier, and the previous item's epilogue (which reuses them
            // for margins) ended with the barrier before the next claim.
            mf_issue_stage(stage_lds[0], E, Lc, 0u, J0, K0, i, wave, lane);
            __builtin_amdgcn_s_waitcnt(0);
            __syncthreads();
            // The last step's MFMAs of a stage are issued after the barrier
            // that ends it, together with the next stage's first reads, so
            // the matrix cores stay busy across the barrier.
            v4i dA[2], dB[4];  // deferred step: A (masked by SNP i) and B
            for (uint32_t st = 0; st < n_st; ++st) {
                const MfStage &S = L.stage[st & 1];
                v4i ei[kStage / 32];
#pragma unroll
                for (int kk = 0; kk < kStage / 32; ++kk)
                    ei[kk] = lds16(S.irow + ei_off + 32 * kk);
                if (st + 1 < n_st)  // the next stage, into the buffer the last one read
                    mf_issue_stage(stage_lds[(st + 1) & 1], E, Lc, (st + 1) * kStage, J0, K0, i,
                                   wave, lane);
                v4i aj[2][2], bk[2][4];
#pragma unroll
                for (int fa = 0; fa < 2; ++fa)
                    aj[0][fa] = lds16(S.jrows + ja_off[fa] + ja_slot[0]);
#pragma unroll
                for (int fb = 0; fb < 4; ++fb)
                    bk[0][fb] = lds16(S.krows + kb_off + 32 * kRowBytes * fb + kb_slot[0]);
                if (st > 0) {  // the previous stage's last step
                    __builtin_amdgcn_sched_barrier(0);
#pragma unroll
                    for (int fb = 0; fb < 4; ++fb)
#pragma unroll
                        for (int fa = 0; fa < 2; ++fa)
                            acc[c][fa][fb] = mf_mma(dA[fa], dB[fb], acc[c][fa][fb]);
                    __builtin_amdgcn_sched_barrier(0);
                }
#pragma unroll
                for (int kk = 0; kk < kStage / 32 - 1; ++kk) {
                    const int cb = kk & 1, nb = cb ^ 1;
                    // operands of the next step (the last one into dA, dB)
#pragma unroll
                    for (int fa = 0; fa < 2; ++fa)
                        aj[nb][fa] = lds16(S.jrows + ja_off[fa] + ja_slot[kk + 1]);
#pragma unroll
                    for (int fb = 0; fb < 4; ++fb)
                        bk[nb][fb] = lds16(S.krows + kb_off + 32 * kRowBytes * fb + kb_slot[kk + 1]);
                    __builtin_amdgcn_sched_barrier(0);
                    v4i A[2];
#pragma unroll
                    for (int fa = 0; fa < 2; ++fa)
                        A[fa] = aj[cb][fa] & ei[kk];
#pragma unroll
                    for (int fb = 0; fb < 4; ++fb)
#pragma unroll
                        for (int fa = 0; fa < 2; ++fa)
                            acc[c][fa][fb] = mf_mma(A[fa], bk[cb][fb], acc[c][fa][fb]);
                    __builtin_amdgcn_sched_barrier(0);
                }
                constexpr int kl = kStage / 32 - 1, lb = kl & 1;
#pragma unroll
                for (int fa = 0; fa < 2; ++fa)
                    dA[fa] = aj[lb][fa] & ei[kl];
#pragma unroll
                for (int fb = 0; fb < 4; ++fb)
                    dB[fb] = bk[lb][fb];
                // this buffer is read and the next stage has landed in the other
                __builtin_amdgcn_s_waitcnt(0);
                __syncthreads();
            }
#pragma unroll
            for (int fb = 0; fb < 4; ++fb)
#pragma unroll
                for (int fa = 0; fa < 2; ++fa)
                    acc[c][fa][fb] = mf_mma(dA[fa], dB[fb], acc[c][fa][fb]);
        }

        // ---- epilogue: lane holds column (k, cc) and rows (j, ab) with
        // row = 8 (e/4) + 4 half + e%4, i.e. j_local = 16w + 8fa + 2 (e/4) + half,
        // ab = e%4, for accumulator element e.  The partner lane (lane ^ 1)
        // holds the other cc; even lanes score j-slots 0,1, odd lanes 2,3.
        // One tile at a time (not unrolled) keeps the scoring code single.
        const uint32_t n0 = C->n[0], n1 = C->n[1];
        const bool odd = (lane & 1u) != 0;
#pragma unroll 1
        for (int t = 0; t < 8; ++t) {
            if ((t & 3) == 0) {
                // the stage buffers (or the previous half) are no longer read
                __syncthreads();
                mf_stage_margins(L.mg, C, i, J0, K0, Mp, t >> 2, tid);
                __syncthreads();
            }
            int v[2][16];
            switch (t) {
#define EPI_MF_COPY(T, FA, FB)                      \
    case T:                                          \
        for (int e = 0; e < 16; ++e) {               \
            v[0][e] = (int)acc[0][FA][FB][e];        \
            v[1][e] = (int)acc[1][FA][FB][e];        \
        }                                            \
        break;
                EPI_MF_COPY(0, 0, 0)
                EPI_MF_COPY(1, 0, 1)
                EPI_MF_COPY(2, 0, 2)
                EPI_MF_COPY(3, 0, 3)
                EPI_MF_COPY(4, 1, 0)
                EPI_MF_COPY(5, 1, 1)
                EPI_MF_COPY(6, 1, 2)
                EPI_MF_COPY(7, 1, 3)
#undef EPI_MF_COPY
            }
            const uint32_t fa = (uint32_t)t >> 2, fb = (uint32_t)t & 3u;
            const uint32_t k = K0 + 16 * fb + (l32 >> 1);
            const uint32_t kl = k - K0;
            int w[2][16];  // the partner lane's values: swap lanes 2m and 2m+1 (DPP)
#pragma unroll
            for (int c = 0; c < 2; ++c)
#pragma unroll
                for (int e = 0; e < 16; ++e)
                    w[c][e] = __builtin_amdgcn_mov_dpp(v[c][e], 0xB1, 0xF, 0xF, false);
#pragma unroll
            for (int sl = 0; sl < 2; ++sl) {
                const uint32_t slot = (odd ? 2u : 0u) + (uint32_t)sl;  // e / 4
                const uint32_t j = J0 + 16 * wave + 8 * fa + 2 * slot + half;
                const bool valid = (i < j) && (j < k) && (k < M);
                const uint32_t jl = j - J0;
                // every lane is scored (lanes outside i < j < k score +inf):
                // no divergent branch around the batched table reads.  The
                // pair table holds every (p, q), so even those lanes' counts
                // are exact and every index below is at most N + 1, inside
                // the table.
                uint32_t x[2][27];
#pragma unroll
                for (int c = 0; c < 2; ++c) {
                    uint32_t t8[8];
#pragma unroll
                    for (int ab = 0; ab < 4; ++ab) {
                        // element of this lane's j-slot: even lanes slots 0,1, odd 2,3
                        const uint32_t mine = (uint32_t)(odd ? v[c][8 + 4 * sl + ab] : v[c][4 * sl + ab]);
                        const uint32_t theirs = (uint32_t)(odd ? w[c][8 + 4 * sl + ab] : w[c][4 * sl + ab]);
                        t8[2 * ab + 0] = odd ? theirs : mine;  // cc = 0 lives in the even lane
                        t8[2 * ab + 1] = odd ? mine : theirs;
                    }
                    const uint4 ij = L.mg.ij[c][jl];
                    const uint4 ik = L.mg.ik[c][kl];
                    const uint4 jk = L.mg.jk[c][8 * wave + 2 * slot + half][kl];
                    const uint2 si = mf::ldc2(C->single[c] + i);
                    const uint2 sj = L.mg.sj[c][jl];
                    const uint2 sk = L.mg.sk[c][kl];
                    const uint32_t aij[4] = {ij.x, ij.y, ij.z, ij.w};
                    const uint32_t aik[4] = {ik.x, ik.y, ik.z, ik.w};
                    const uint32_t ajk[4] = {jk.x, jk.y, jk.z, jk.w};
                    const uint32_t asi[2] = {si.x, si.y};
                    const uint32_t asj[2] = {sj.x, sj.y};
                    const uint32_t ask[2] = {sk.x, sk.y};
                    derive27(t8, aij, aik, ajk, asi, asj, ask, c ? n1 : n0, x[c]);
                }
                // K2 in the reference order (GPU/v4.cpp:579-583,
                // CPU/v2.cpp:516-519); the table reads of 9 cells are issued
                // ahead of their ordered sums
                float sc = 0.0f;
#pragma unroll
                for (int m0 = 0; m0 < 27; m0 += 9) {
                    float ta[9], tb[9], td[9];
#pragma unroll
                    for (int m = 0; m < 9; ++m) {
                        const uint32_t ib = x[0][m0 + m], id = x[1][m0 + m];
                        const uint32_t ia = ib + id + 1;
                        ta[m] = TLDS ? Tl[ia] : P.table[ia];
                        tb[m] = TLDS ? Tl[ib] : P.table[ib];
                        td[m] = TLDS ? Tl[id] : P.table[id];
                    }
                    __builtin_amdgcn_sched_barrier(0);
#pragma unroll
                    for (int m = 0; m < 9; ++m)
                        sc += (ta[m] - tb[m]) - td[m];
                }
                const float s = valid ? fabsf(sc) : INFINITY;
                if (DUMP && valid) {
                    const uint64_t MM = M;
                    auto c3 = [](uint64_t n) { return n < 3 ? 0 : n * (n - 1) * (n - 2) / 6; };
                    auto c2 = [](uint64_t n) { return n < 2 ? 0 : n * (n - 1) / 2; };
                    const uint64_t rk = c3(MM) - c3(MM - i) + c2(MM - i - 1) - c2(MM - j) +
                                        (uint64_t)(k - j - 1);
                    P.dump_scores[rk] = s;
                    if (P.dump_counts) {
#pragma unroll
                        for (int m = 0; m < 27; ++m) {
                            P.dump_counts[rk * 54 + m] = x[0][m];
                            P.dump_counts[rk * 54 + 27 + m] = x[1][m];
                        }
                    }
                }
                if (!DUMP)
                    topk::wave_offer(Wt, valid, s, make_key(i, j, k), lane);
            }
        }
        if (tid == 0)
            L.misc[0] = claimed;
        __syncthreads();
        item = mf::rfl(L.misc[0]);
    }
}

size_t mfma_lds_bytes(bool table_lds, uint32_t table_len)
{
    return mf_lds_fixed() + (table_lds ? (size_t)table_len * 4 : 0);
}

typedef void (*MfFn)(const ScanParams);

template <bool FP4>
static MfFn mf_pick2(bool tl, bool dump)
{
    if (tl)
        return dump ? mfma_scan_kernel<true, true, FP4> : mfma_scan_kernel<true, false, FP4>;
    return dump ? mfma_scan_kernel<false, true, FP4> : mfma_scan_kernel<false, false, FP4>;
}

static MfFn mf_pick(bool fp4, bool tl, bool dump)
{
    return fp4 ? mf_pick2<true>(tl, dump) : mf_pick2<false>(tl, dump);
}

hipError_t mfma_occupancy(bool fp4, bool table_lds, bool dump, size_t lds_bytes,
                          int *blocks_per_cu)
{
    return hipOccupancyMaxActiveBlocksPerMultiprocessor(
        blocks_per_cu, (const void *)mf_pick(fp4, table_lds, dump), kMfWaves * 64, lds_bytes);
}

hipError_t mfma_allow_lds(bool fp4, bool table_lds, bool dump, size_t lds_bytes)
{
    return hipFuncSetAttribute((const void *)mf_pick(fp4, table_lds, dump),
                               hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds_bytes);
}

hipError_t launch_mfma_scan(bool fp4, bool table_lds, bool dump, const ScanParams &p,
                            uint32_t grid, hipStream_t s)
{
    const size_t lds = mfma_lds_bytes(table_lds, p.table_len);
    hipLaunchKernelGGL(mf_pick(fp4, table_lds, dump), dim3(grid), dim3(kMfWaves * 64), lds, s,
                       p);
    return hipGetLastError();
}

}  // namespace epi
