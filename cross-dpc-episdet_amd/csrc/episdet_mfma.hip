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
//        v_mfma_scale_f32_32x32x64_f8f6f4 (block scale 2^0 on A, 2^2 on B),
//        FP32 accumulation.  Every product is 0 or 4 and every partial sum an
//        integer below 2^24 (class sizes are below 2^22, which cohort
//        creation enforces), so the FP32 sums are exact (4 x the counts);
//        the instruction does twice the int8 instruction's MACs in the same
//        issue time, and the operands take half the bytes.
// The operand bytes come from a per-cohort expansion of the bit planes
// (E[c] or F[c], [snp][plane][sample]).  Both formats use one LDS layout:
// 128-byte rows per stage, 32 bytes per MFMA step (32 samples as bytes, 64
// as nibbles); A and B map a byte position to the same sample, so the sum
// over the K dimension covers each sample once whatever that mapping is.
//
// Work item = (i, j-block of 64 SNPs, k-block of 64 SNPs), one workgroup of
// 8 wavefronts, two per SIMD.  Per stage the workgroup copies the j-block
// and k-block rows (and SNP i's) into LDS; wavefront w (jq = w % 4,
// kq = w / 4) computes 64 rows (16 j x 4 (a,b); its j are the block's SNPs
// jq, jq + 4, ..., jq + 60, so that the pairs outside i < j < k spread
// evenly over the wavefronts) against 64 columns (32 k of half kq, x 2 c):
// 2 x 2 MFMA tiles of 32 x 32 per class, 128 accumulators for both classes.
// Pipelining: two LDS stage buffers, one barrier per stage; the next
// stage's copies (global_load_lds, no registers) are issued at the start of
// a stage, and each step's operand reads a step ahead.  The epilogue
// rebuilds the 27-cell tables exactly from the 8 counted cells and the pair
// and single margins, sums K2 in the reference order and keeps a
// per-wavefront top-64.
#include <math.h>

#include "episdet_cells.h"
#include "episdet_internal.h"
#include "episdet_order.h"
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
// bit planes -> 0/1 bytes: E[c] row (snp, plane) holds EL bytes (64 W of
// the larger class rounded up to a whole 128-byte stage: both classes have
// one row pitch), byte t = bit t of the plane (sample t of the class);
// padding bytes are 0.  One thread per (snp, 64-bit word slot, plane) writes
// 64 bytes; slots w >= W write zeros.
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

// bit planes -> e2m1 nibbles: F[c] row (snp, plane) holds FL bytes (32 W of
// the larger class rounded up to a whole 128-byte stage, as for E); sample
// t is nibble t % 8 of dword t / 8, 0x2 (1.0) when bit t of the plane is
// set, else 0; padding is 0.
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
// instruction reads 4 of the 8 operand registers for e2m1, format code 4).
// The e2m1 form scales A by 2^0 and B by 2^2 (E8M0 exponents 127 and 129),
// so its sums are 4 x the counts: the epilogue's table addresses (below
// 2^24 as well, so still exact).
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
    return __builtin_amdgcn_mfma_scale_f32_32x32x64_f8f6f4(a8, b8, c, 4, 4, 0, 127, 0, 129);
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
// Eight wavefronts per workgroup, two per SIMD.  The matrix cores take this
// instruction from a single wavefront at about half the rate they reach
// with two (independent accumulators in both cases), and one wavefront's
// VALU and LDS work runs beside the other wavefront's MFMAs on their SIMD
// at almost no cost to them (both measured on MI355X); two wavefronts per
// SIMD leave each 256 registers, so each holds 128 accumulators: 16 j x 32 k
// of the 64 x 64 (j, k) tile, both classes.
constexpr int kMfWaves = kMfmaWaves;
constexpr int kStage = 128;           // bytes per row per LDS stage (four 32-byte MFMA steps)
constexpr int kRowBytes = kStage;     // LDS bytes per staged (snp, plane) row
constexpr int kStageRows = 64 * 2;    // 64 SNPs x 2 planes
constexpr int kStageCopies = 2 * kStageRows * kRowBytes / 1024 / kMfWaves;  // 1 KiB copies per wave
constexpr int kIChunks = 2 * kStage / 16 / kMfWaves;  // 16-byte chunks of SNP i's rows per wave
constexpr int kGather = 3;  // K2 cells whose table reads are issued ahead of their sums (divides 27)
static_assert(kStageCopies == 4 && kIChunks == 2, "copy split");

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

// The margins of SNP i and the single-SNP counts an item's epilogue reads,
// staged in LDS x 4, in an area of their own: the stage buffers stay free
// during the epilogue and receive the next item's first stage.  The (j, k)
// pair margins are read from global memory by each lane, a slot ahead.
struct MfMargins {
    uint4 ij[2][64];            // pair margins of (i, j)
    uint4 ik[2][64];            // pair margins of (i, k)
    uint2 sj[2][64];            // single-SNP counts of the j-block
    uint2 sk[2][64];            // and of the k-block
    uint32_t oj[64], ok[64];    // original indices of the j- and k-block SNPs
};

struct MfLds {
    Cand cand[kMfWaves][kWaveCap];
    uint32_t misc[4];
    uint32_t flush[kMfWaves + 1];  // a flush's per-wavefront counts, then its base
    MfMargins mg;      // the epilogue's margins
    MfStage stage[2];  // double buffer: one is read while the other is filled
};

// The kernel's LDS is one static block.  With the ln(n!) table in LDS the
// table comes first, at LDS address 0, so that 4 x a count is the address
// of its entry and the table reads need no address arithmetic; it holds
// up to kMfTabMax entries (cohorts of up to kMfTabMax - 2 samples), the
// rest of the 160 KiB after MfLds.
constexpr size_t kMfLdsBudget = 160 * 1024;
constexpr uint32_t kMfTabMax = (uint32_t)((kMfLdsBudget - sizeof(MfLds)) / 16 * 4);
template <bool TLDS>
struct MfShared {
    float tab[kMfTabMax];
    MfLds L;
};
template <>
struct MfShared<false> {
    MfLds L;
};
static_assert(sizeof(MfShared<true>) <= kMfLdsBudget, "LDS budget");

// Rank of triplet i < j < k among all triplets of M SNPs in lexicographic
// order (the dump layout).
__device__ __forceinline__ uint64_t dump_rank(uint32_t M, uint32_t i, uint32_t j, uint32_t k)
{
    auto c3 = [](uint64_t n) { return n < 3 ? 0 : n * (n - 1) * (n - 2) / 6; };
    auto c2 = [](uint64_t n) { return n < 2 ? 0 : n * (n - 1) / 2; };
    const uint64_t MM = M;
    return c3(MM) - c3(MM - i) + c2(MM - i - 1) - c2(MM - j) + (uint64_t)(k - j - 1);
}

// ln(n!) of the count whose 4 x value is `off` (a byte offset into the table)
__device__ __forceinline__ float tab_at(const unsigned char *tab, uint32_t off)
{
    return *reinterpret_cast<const float *>(tab + off);
}

typedef __attribute__((address_space(3))) uint8_t lds_u8;
typedef __attribute__((address_space(3))) v4i lds_v4i;

// 16 bytes of LDS at byte address `addr` as an MFMA operand
__device__ __forceinline__ v4i lds16a(uint32_t addr)
{
    return *reinterpret_cast<const lds_v4i *>((uintptr_t)addr);
}

// One 16-byte-per-lane copy from global memory into LDS (lane L writes
// bytes 16L..16L+15 after the wave-uniform LDS address `lds`), from the
// wave-uniform base `base` (SGPR pair) plus this lane's 32-bit byte offset
// `off`.  Written as asm so that the compiler's wait-count model, which
// counts these copies as LDS operations too and then waits for every
// outstanding LDS read, keeps exact counts for the operand reads;
// completion is awaited explicitly (s_waitcnt vmcnt) before the barrier
// that publishes a stage.  The instruction has no register result.  M0 is
// a reserved register the compiler may not preserve across the asm; no
// compiler-generated instruction of these kernels reads or writes it
// (tests/test_isa.py checks the gfx950 assembly).
__device__ __forceinline__ void lds_copy16(const uint8_t *base, uint32_t off, uint32_t lds)
{
    asm volatile("s_mov_b32 m0, %2\n\tglobal_load_lds_dwordx4 %0, %1"
                 :
                 : "v"(off), "s"(base), "s"(lds)
                 : "memory", "m0");
}

// LDS byte address of a pointer into LDS.
__device__ __forceinline__ uint32_t lds_addr(const void *p)
{
    return (uint32_t)(uintptr_t)(const lds_u8 *)p;
}

// One stage is copied from global memory straight into an LDS stage buffer
// (global_load_lds_dwordx4: lane L of a copy writes bytes 16L..16L+15 of a
// 1 KiB block given by M0).  The j and k rows (32 KiB) are 32 copies of 8
// rows each, 4 per wavefront (wavefronts 0-3: j rows; 4-7: k rows); lane L
// takes row L/8 of its copy and reads the global chunk that the swizzle
// puts in its slot L%8.  Row placement and swizzle key of row slot s
// (0..63) and plane p, and the SNP (local to its block) the slot holds:
//   j rows: row 2s + p,            key s & 7,        SNP 4 (s % 16) + s / 16
//   k rows: row 2s + (p ^ (s & 1)), key (s >> 1) & 7, SNP s
// (the B operand reads of one plane then alternate between the two bank
// halves; see the operand offsets in the kernel).  The j slots 16q..16q+15
// hold SNPs q, q + 4, ..., q + 60: each wavefront's j spread over the whole
// block, so the (j, k) pairs past the triangle of a block on the diagonal,
// or not above SNP i, fall evenly on the wavefronts and their epilogues
// take about equal time.  SNP i's 256 bytes are one 2-lane copy per
// wavefront.  Every copy lands before the barrier that ends the stage
// (s_waitcnt vmcnt(0) on every wavefront).
//
// Per-lane source offsets of the copies within the wavefront's block (and
// of SNP i's rows), for rows of Lc bytes; fixed for a class.
struct MfCopyOffsets {
    uint32_t row[kStageCopies];
    uint32_t irow;
};

__device__ __forceinline__ MfCopyOffsets mf_copy_offsets(uint32_t Lc, uint32_t wave,
                                                         uint32_t lane)
{
    MfCopyOffsets o;
    const uint32_t rsub = lane >> 3;
#pragma unroll
    for (int q = 0; q < kStageCopies; ++q) {
        const uint32_t first = (wave * kStageCopies + q) * 8;
        const uint32_t blk = first >> 7, rowi = (first & 127u) + rsub;
        const uint32_t sl = rowi >> 1;  // row slot within its block
        const uint32_t key = blk ? (sl >> 1) & 7u : sl & 7u;
        const uint32_t plane = (rowi & 1u) ^ (blk & sl & 1u);
        const uint32_t chunk = (lane & 7u) ^ key;
        const uint32_t snp = blk ? sl : 4 * (sl & 15u) + (sl >> 4);
        o.row[q] = (2 * snp + plane) * Lc + 16 * chunk;
    }
    // 16-byte chunk e of SNP i's two rows (plane e / 8, chunk e % 8)
    const uint32_t e = kIChunks * wave + (lane % (uint32_t)kIChunks);
    o.irow = (e >> 3) * Lc + 16 * (e & 7u);
    return o;
}

// Copy q of the wavefront's block rows of a stage (src: the block's rows at
// the stage's byte offset), and the 2-lane copy of SNP i's rows.
__device__ __forceinline__ void mf_issue_copy(uint32_t stage, const uint8_t *src,
                                              const MfCopyOffsets &o, uint32_t wave, int q)
{
    const uint32_t first = (wave * kStageCopies + q) * 8;
    lds_copy16(src, o.row[q], stage + first * kRowBytes);  // krows follows jrows
}
__device__ __forceinline__ void mf_issue_icopy(uint32_t stage, const uint8_t *src,
                                               const MfCopyOffsets &o, uint32_t wave,
                                               uint32_t lane)
{
    if (lane < (uint32_t)kIChunks)
        lds_copy16(src, o.irow,
                   stage + (uint32_t)__builtin_offsetof(MfStage, irow) + 16 * kIChunks * wave);
}

// 4 x each of four counts: the epilogue works on 4 x counts, the byte
// offsets of their ln(n!) entries.
__device__ __forceinline__ uint4 x4(uint4 v)
{
    return make_uint4(v.x << 2, v.y << 2, v.z << 2, v.w << 2);
}

__device__ __forceinline__ void mf_stage_margins(MfMargins &G, const EPI_CONST CohortDev *C,
                                                 uint32_t i, uint32_t J0, uint32_t K0,
                                                 uint32_t Mp, uint32_t tid)
{
    // thread tid (< 256): class (tid / 128), block (tid / 64 % 2: j or k),
    // SNP tid % 64 of the block
    const uint32_t c = (tid >> 7) & 1u, e = tid & 63u;
    const bool kb = (tid & 64u) != 0;
    const uint32_t q = (kb ? K0 : J0) + e;
    const uint4 x = C->pair[c][pair_at(C->pslot[q >> 6], Mp, i, q)];
    const uint2 y = C->single[c][(kb ? K0 : J0) + e];
    if (kb) {
        G.ik[c][e] = x4(x);
        G.sk[c][e] = make_uint2(y.x << 2, y.y << 2);
    } else {
        G.ij[c][e] = x4(x);
        G.sj[c][e] = make_uint2(y.x << 2, y.y << 2);
    }
    if (c == 0)
        (kb ? G.ok : G.oj)[e] = C->orig ? C->orig[q] : q;
}

// 4 x a count from an accumulator element: the FP4 sums are already 4 x
// the count (exact integers in FP32), the int8 ones are shifted.
template <typename T>
__device__ __forceinline__ uint32_t acc_x4(T v);
template <>
__device__ __forceinline__ uint32_t acc_x4<float>(float v)
{
    return (uint32_t)v;
}
template <>
__device__ __forceinline__ uint32_t acc_x4<int>(int v)
{
    return (uint32_t)v << 2;
}

template <bool TLDS, bool DUMP, bool FP4>
__global__ void __launch_bounds__(kMfWaves * 64, 1) mfma_scan_kernel(const ScanParams P)
{
    __shared__ MfShared<TLDS> sh;
    MfLds &L = sh.L;
    const uint32_t stage_lds[2] = {lds_addr(&L.stage[0]), lds_addr(&L.stage[1])};
    float *Tl = TLDS ? reinterpret_cast<float *>(&sh) : nullptr;  // the table (LDS address 0)
    const unsigned char *tab = TLDS ? reinterpret_cast<const unsigned char *>(&sh)
                                    : reinterpret_cast<const unsigned char *>(P.table);

    const uint32_t tid = threadIdx.x;
    const uint32_t lane = tid & 63u;
    const uint32_t wave = mf::rfl(tid >> 6);
    topk::WaveTop Wt;
    Wt.buf = L.cand[wave];
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
    const uint32_t jq = wave & 3u;     // j slots 16 jq .. 16 jq + 15 (SNPs jq, jq + 4, ...)
    const uint32_t kq = wave >> 2;     // k SNPs 32 kq .. 32 kq + 31 of the block

    // Lane-constant LDS offsets of the MFMA operands, in bytes from the
    // start of a stage buffer, per 32-byte step kk (the swizzled slot):
    //   A (j rows): row l32 of tile fa = (j slot, a, b), j slot =
    //     16 jq + 8 fa + l32 / 4, at row 2 slot + b; tile fa = 1 is 2048
    //     bytes on.
    //   B (k rows): column l32 of tile cc is k_local = 32 kq + l32, plane
    //     cc, at row 2 k_local + (cc ^ (k_local & 1)), so a lane holds both
    //     planes of its k.
    //   swizzle keys: A row slot & 7 = (l32 / 4) & 7; B row (k_local >> 1)
    //     & 7 = (l32 / 2) & 7.  16 lanes of one operand read then hit 8 rows
    //     in each bank half (B: alternating parity), at 8 different slots:
    //     no bank conflicts.
    // A stage's reads add the stage buffer's LDS address (wave-uniform) to
    // these offsets; the fa and kk parts that are constant fold into the
    // instructions' offset fields.
    const uint32_t a_lane = (l32 >> 1) & 1u;  // g_i plane of this lane's A rows
    const uint32_t jx = (l32 >> 2) & 7u;
    const uint32_t kx = (l32 >> 1) & 7u;
    uint32_t aoff[kStage / 32], boff[2][kStage / 32];
    {
        const uint32_t ja = (uint32_t)__builtin_offsetof(MfStage, jrows) +
                            ((16 * jq + (l32 >> 2)) * 2 + (l32 & 1u)) * kRowBytes;
        const uint32_t kbase = (uint32_t)__builtin_offsetof(MfStage, krows) + 8192 * kq;
        const uint32_t kb0 = kbase + (2 * l32 + (l32 & 1u)) * kRowBytes;
        const uint32_t kb1 = kbase + (2 * l32 + (1u ^ (l32 & 1u))) * kRowBytes;
#pragma unroll
        for (int kk = 0; kk < kStage / 32; ++kk) {
            aoff[kk] = ja + 16 * ((2 * kk + half) ^ jx);
            boff[0][kk] = kb0 + 16 * ((2 * kk + half) ^ kx);
            boff[1][kk] = kb1 + 16 * ((2 * kk + half) ^ kx);
        }
    }
    const uint32_t eoff = (uint32_t)__builtin_offsetof(MfStage, irow) + a_lane * kStage + 16 * half;

    // Work-queue items: the next item is claimed while this one runs.
    // Thread 0 publishes the claim in misc[0] after the item's stage loop
    // (every wavefront read misc[0] before the stage loop's barriers) and
    // the margin barrier makes it visible, so the epilogue knows the next
    // item and copies its first stage into the idle stage buffer 0 (`pre`).
    if (tid == 0)
        L.misc[0] = atomicAdd(&P.counters[0], 1u);
    __syncthreads();
    uint32_t item = mf::rfl(L.misc[0]);
    bool pre = false;
    for (;;) {
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
            if (!DUMP && cur != 0xFFFFFFFFu) {
                // The workgroup's wavefronts leave a cohort together (they
                // take the same items), so they append with one atomic: at
                // the end of a scan every workgroup flushes at once, and one
                // counter per cohort taken by every wavefront serializes.
                bool pass;
                const uint32_t nn = topk::flush_select(Wt, P, cur, lane, pass);
                if (lane == 0)
                    L.flush[wave] = nn;
                __syncthreads();
                if (tid == 0) {
                    uint32_t tot = 0;
#pragma unroll
                    for (int w = 0; w < kMfWaves; ++w)
                        tot += L.flush[w];
                    L.flush[kMfWaves] = tot ? atomicAdd(&P.counters[1 + cur], tot) : 0u;
                }
                __syncthreads();
                uint32_t base = L.flush[kMfWaves];
                for (uint32_t w = 0; w < wave; ++w)
                    base += L.flush[w];
                topk::flush_append(Wt, P, cur, lane, pass, mf::rfl(base), nn);
                Wt.reset();
            }
            cur = cohort;
        }
        if (item >= n_items)
            break;
        if (!DUMP)
            Wt.thr_slot = P.thr + cohort;

        const uint32_t K0 = rows[r].kb * 64;
        const uint32_t J0 = rows[r].jg * 64;
        const uint32_t i = u - rows[r].item_start;  // one i per item
        const EPI_CONST CohortDev *C = mf::as_const(P.cohorts) + cohort;
        const uint32_t M = C->M, Mp = C->Mp;

        // acc[class][fa][cc]: rows (j slot 16 jq + 8 fa + ..., a, b) x
        // columns (k = 32 kq + l32, plane cc)
        typename MfAcc<FP4>::type acc[2][2][2];
#pragma unroll
        for (int c = 0; c < 2; ++c)
#pragma unroll
            for (int fa = 0; fa < 2; ++fa)
#pragma unroll
                for (int cc = 0; cc < 2; ++cc)
#pragma unroll
                    for (int e = 0; e < 16; ++e)
                        acc[c][fa][cc][e] = 0;

        // Stage loop.  Both classes' rows have one pitch (bytes per (snp,
        // plane) row, a multiple of kStage, zero padded) and class 1's
        // planes follow class 0's in one allocation; class c has its own
        // stage count (256 samples per stage as nibbles, 128 as bytes).  The
        // last stage of class 0 copies class 1's first stage (a byte offset
        // jb further), so class 1 starts without waiting for memory.
        uint32_t nst[2];
#pragma unroll
        for (int c = 0; c < 2; ++c)
            nst[c] = FP4 ? (C->W[c] + 3) / 4 : (C->W[c] + 1) / 2;
        uint32_t par = 0;     // stage buffer of the class's first stage
        bool staged = false;  // class 1's first stage is in stage_lds[par]
#pragma unroll
        for (int c = 0; c < 2; ++c) {
            const uint8_t *E = FP4 ? C->F[c] : C->E[c];
            const uint32_t Lc = FP4 ? C->FL[c] : C->EL[c];
            const uint32_t n_st = nst[c];
            if (n_st == 0)
                continue;
            const uint32_t jb = c == 0 && nst[1] > 0 ? (FP4 ? C->F1off : C->E1off) : 0u;
            const MfCopyOffsets co = mf_copy_offsets(Lc, wave, lane);
            const uint8_t *blk_base = E + (size_t)((wave >> 2) ? K0 : J0) * 2 * Lc;
            const uint8_t *i_base = E + (size_t)i * 2 * Lc;
            if (pre) {
                // the previous item's epilogue copied this stage into buffer 0
                pre = false;
                par = 0;
            } else if (!(c == 1 && staged)) {
                // Both buffers are free here: the previous stage loop ended
                // with a barrier, and the stage buffers are not read during
                // an epilogue.
                par = 0;
#pragma unroll
                for (int q = 0; q < kStageCopies; ++q)
                    mf_issue_copy(stage_lds[0], blk_base, co, wave, q);
                mf_issue_icopy(stage_lds[0], i_base, co, wave, lane);
            }
            __builtin_amdgcn_s_waitcnt(0);
            __syncthreads();
            // Step 3 of a stage runs after the barrier that ends it (its
            // operands are in registers by then): 4 MFMAs that keep the
            // matrix cores busy while the wavefronts meet at the barrier
            // and while the next stage's copies are issued, one after each
            // of them.  Steps 0-2 run before the barrier, each with the
            // reads of the following step's operands issued ahead.
            v4i dA[2], dB[2];  // the deferred step's A (masked by SNP i) and B
            for (uint32_t st = 0; st < n_st; ++st) {
                const uint32_t sb = stage_lds[(st + par) & 1];
                // The next stage's copies go to the buffer the last stage
                // read.  In the last stage they copy class 1's first stage
                // (class 0, jb != 0) or this stage again (valid rows, an idle
                // buffer) rather than branch around the copies; like every
                // copy they land before this stage's barrier.
                const uint32_t nsb = stage_lds[(st + 1 + par) & 1];
                const uint32_t ns0 = st + 1 < n_st ? (st + 1) * kStage : (jb ? jb : st * kStage);
                v4i ei[kStage / 32], aj[kStage / 32][2], bk[kStage / 32][2];
                auto read_step = [&](int kk) {
                    ei[kk] = lds16a(sb + eoff + 32 * kk);
#pragma unroll
                    for (int fa = 0; fa < 2; ++fa)
                        aj[kk][fa] = lds16a(sb + aoff[kk] + 2048 * fa);
#pragma unroll
                    for (int cc = 0; cc < 2; ++cc)
                        bk[kk][cc] = lds16a(sb + boff[cc][kk]);
                };
                read_step(0);
                if (st > 0) {
                    // the previous stage's step 3 with the copies of the next stage
#pragma unroll
                    for (int m = 0; m < 4; ++m) {
                        __builtin_amdgcn_sched_barrier(0);
                        acc[c][m & 1][m >> 1] = mf_mma(dA[m & 1], dB[m >> 1], acc[c][m & 1][m >> 1]);
                        __builtin_amdgcn_sched_barrier(0);
                        mf_issue_copy(nsb, blk_base + ns0, co, wave, m);
                        if (m == 0)
                            mf_issue_icopy(nsb, i_base + ns0, co, wave, lane);
                    }
                } else {
#pragma unroll
                    for (int q = 0; q < kStageCopies; ++q)
                        mf_issue_copy(nsb, blk_base + ns0, co, wave, q);
                    mf_issue_icopy(nsb, i_base + ns0, co, wave, lane);
                }
                // steps 0-2; step 3's operands are read during step 2
#pragma unroll
                for (int kk = 0; kk < 3; ++kk) {
                    read_step(kk + 1);
                    __builtin_amdgcn_sched_barrier(0);
                    v4i A[2];
#pragma unroll
                    for (int fa = 0; fa < 2; ++fa)
                        A[fa] = aj[kk][fa] & ei[kk];
#pragma unroll
                    for (int cc = 0; cc < 2; ++cc)
#pragma unroll
                        for (int fa = 0; fa < 2; ++fa)
                            acc[c][fa][cc] = mf_mma(A[fa], bk[kk][cc], acc[c][fa][cc]);
                    __builtin_amdgcn_sched_barrier(0);
                }
#pragma unroll
                for (int fa = 0; fa < 2; ++fa)
                    dA[fa] = aj[3][fa] & ei[3];
#pragma unroll
                for (int cc = 0; cc < 2; ++cc)
                    dB[cc] = bk[3][cc];
                // this buffer is read and the next stage has landed in the other
                __builtin_amdgcn_s_waitcnt(0);
                __syncthreads();
            }
#pragma unroll
            for (int cc = 0; cc < 2; ++cc)
#pragma unroll
                for (int fa = 0; fa < 2; ++fa)
                    acc[c][fa][cc] = mf_mma(dA[fa], dB[cc], acc[c][fa][cc]);
            par = jb ? (n_st & 1) : par;
            staged = jb != 0;
        }

        // ---- epilogue.  Accumulator element e of tile (fa, cc) in lane
        // (half, l32) is the count of rows (j, ab) x column (k, cc): j slot
        // 16 jq + 8 fa + 2 (e / 4) + half, which holds SNP j_local = 32 fa +
        // 8 (e / 4) + 4 half + jq, ab = e % 4, k_local = 32 kq + l32.  So
        // each lane holds all 8 cells of 8 (j, k) pairs: slot s8 = 4 fa + sl
        // (sl = e / 4) scores j_local = 32 fa + 8 sl + 4 half + jq against
        // the lane's own k.  Counts are handled as 4 x counts (byte offsets
        // into the ln(n!) table); margins are staged x 4.  The slot loop is
        // not unrolled (one copy of the scoring code); a switch reads the
        // slot's 16 accumulators.
        const uint32_t n04 = C->n[0] << 2, n14 = C->n[1] << 2;
        // The previous item's margins were read before this item's first
        // stage barrier, so they are overwritten freely.
        if (tid == 0)
            L.misc[0] = claimed;
        if (tid < 256)
            mf_stage_margins(L.mg, C, i, J0, K0, Mp, tid);
        __syncthreads();
        const uint32_t next = mf::rfl(L.misc[0]);
        if (!DUMP)
            Wt.gthr = fminf(Wt.gthr, topk::load_thr(P.thr, cohort));
        // The next item's first stage (of its first class with samples)
        // goes to stage buffer 0 now: no stage buffer is read until the
        // next item's first stage barrier, and these copies land while the
        // slot loop below waits for its first pair margins.  The walk to
        // the next item works on copies of the walk state.
        if (next < n_items) {
            uint32_t sg2 = sg, r2 = r;
            bool fresh2 = false;
            while (sg2 + 1 < n_segs && segs[sg2 + 1].item_base <= next) {
                ++sg2;
                fresh2 = true;
            }
            const EPI_CONST Row *rows2 = (const EPI_CONST Row *)segs[sg2].rows;
            if (fresh2)
                r2 = segs[sg2].row_begin;
            const uint32_t u2 = next - segs[sg2].item_base + segs[sg2].item_first;
            const uint32_t row_end2 = segs[sg2].row_end;
            while (r2 + 1 < row_end2 && rows2[r2 + 1].item_start <= u2)
                ++r2;
            const EPI_CONST CohortDev *C2 = mf::as_const(P.cohorts) + segs[sg2].cohort;
            const uint32_t c2 = C2->W[0] > 0 ? 0u : 1u;
            if (C2->W[c2] > 0) {
                const uint8_t *E2 = FP4 ? C2->F[c2] : C2->E[c2];
                const uint32_t Lc2 = FP4 ? C2->FL[c2] : C2->EL[c2];
                const MfCopyOffsets co2 = mf_copy_offsets(Lc2, wave, lane);
                const uint32_t b2 = ((wave >> 2) ? rows2[r2].kb : rows2[r2].jg) * 64;
                const uint32_t i2 = u2 - rows2[r2].item_start;
#pragma unroll
                for (int q = 0; q < kStageCopies; ++q)
                    mf_issue_copy(stage_lds[0], E2 + (size_t)b2 * 2 * Lc2, co2, wave, q);
                mf_issue_icopy(stage_lds[0], E2 + (size_t)i2 * 2 * Lc2, co2, wave, lane);
                pre = true;
            }
        }
        const uint32_t kl = 32 * kq + l32;
        const uint32_t k = K0 + kl;
        uint4 ik[2];
        uint2 sk[2], si[2];
#pragma unroll
        for (int c = 0; c < 2; ++c) {
            ik[c] = L.mg.ik[c][kl];
            sk[c] = L.mg.sk[c][kl];
            const uint2 s1 = mf::ldc2(C->single[c] + i);
            si[c] = make_uint2(s1.x << 2, s1.y << 2);
        }
        // (j, k) pair margins of the lane's slots, read a slot ahead: row
        // J0 + j_local, column k of the k block's strip; every row p < Mp is
        // in the strip
        const uint32_t jl0 = 4 * half + jq;  // slot 0's j_local
        const uint32_t slot_k = mf::rfl(C->pslot[K0 >> 6]);
        const uint4 *jkp[2];
#pragma unroll
        for (int c = 0; c < 2; ++c)
            jkp[c] = C->pair[c] + pair_at(slot_k, Mp, J0 + jl0, k);
        uint4 jkn[2];
#pragma unroll
        for (int c = 0; c < 2; ++c)
            jkn[c] = jkp[c][0];
        uint32_t redo_mask = 0;  // bit s8: the lane's slot s8 is summed again below
#pragma unroll 1
        for (uint32_t s8 = 0; s8 < 8; ++s8) {
            uint4 jkc[2];
#pragma unroll
            for (int c = 0; c < 2; ++c)
                jkc[c] = x4(jkn[c]);
            // slot s8 + 1, at its j_local offset from slot 0 (rows of a strip
            // are 64 entries apart).  Both loads are unconditional (the last
            // slot reloads its own row) and follow both copies above, so the
            // wait for them sits at the top of the next slot, a slot's work
            // after their issue, instead of on this slot's path.
            {
                const uint32_t nx = s8 + 1 < 8 ? s8 + 1 : s8;
                const size_t off = (size_t)(32 * (nx >> 2) + 8 * (nx & 3)) * 64;
#pragma unroll
                for (int c = 0; c < 2; ++c)
                    jkn[c] = jkp[c][off];
            }
            const uint32_t jl = jl0 + 32 * (s8 >> 2) + 8 * (s8 & 3);
            const uint32_t j = J0 + jl;
            const bool valid = (i < j) && (j < k) && (k < M);
            // Only lanes with i < j < k are offered; a slot with none is
            // skipped by a wave-uniform branch.  The pair table holds every
            // (p, q), so counts are exact for every lane and every offset
            // below is at most 4 (N + 1), inside the table.
            if (__ballot(valid) == 0)
                continue;
            // t8[class][4a + 2b + c]: the slot's 16 accumulators
#define EPI_MF_T8(t8)                                                                     \
    switch (s8) {                                                                         \
        EPI_MF_SLOT(t8, 0)                                                                \
        EPI_MF_SLOT(t8, 1)                                                                \
        EPI_MF_SLOT(t8, 2)                                                                \
        EPI_MF_SLOT(t8, 3)                                                                \
        EPI_MF_SLOT(t8, 4)                                                                \
        EPI_MF_SLOT(t8, 5)                                                                \
        EPI_MF_SLOT(t8, 6)                                                                \
        EPI_MF_SLOT(t8, 7)                                                                \
    }
#define EPI_MF_SLOT(t8, S8)                                                               \
    case S8:                                                                              \
        for (int c = 0; c < 2; ++c)                                                       \
            for (int ab = 0; ab < 4; ++ab)                                                \
                for (int cc = 0; cc < 2; ++cc)                                            \
                    t8[c][2 * ab + cc] = acc_x4(acc[c][(S8) >> 2][cc][4 * ((S8) & 3) + ab]); \
        break;
            // x[class][9a + 3b + c]: the slot's 27 cells (x 4) from t8 and
            // the margins, (j, k) pair margins jkv
#define EPI_MF_CELLS(x, t8, jkv)                                                          \
    for (int c = 0; c < 2; ++c) {                                                         \
        const uint4 ij = L.mg.ij[c][jl];                                                  \
        const uint4 jk = jkv[c];                                                          \
        const uint2 sj = L.mg.sj[c][jl];                                                  \
        const uint32_t aij[4] = {ij.x, ij.y, ij.z, ij.w};                                 \
        const uint32_t aik[4] = {ik[c].x, ik[c].y, ik[c].z, ik[c].w};                     \
        const uint32_t ajk[4] = {jk.x, jk.y, jk.z, jk.w};                                 \
        const uint32_t asi[2] = {si[c].x, si[c].y};                                       \
        const uint32_t asj[2] = {sj.x, sj.y};                                             \
        const uint32_t ask[2] = {sk[c].x, sk[c].y};                                       \
        derive27(t8[c], aij, aik, ajk, asi, asj, ask, c ? n14 : n04, x[c]);               \
    }
            uint32_t t8[2][8];
            EPI_MF_T8(t8)
            float sc = 0.0f;
            {
                uint32_t x[2][27];
                EPI_MF_CELLS(x, t8, jkc)
                // K2 in the reference order (GPU/v4.cpp:579-583,
                // CPU/v2.cpp:516-519): ((T[n0 + n1 + 1] - T[n0]) - T[n1])
                // summed over the cells in order from 0 (the first term
                // needs no add: no term is -0, so 0 + t == t bit for bit);
                // the table reads of kGather cells are issued ahead of their
                // ordered sums.
#pragma unroll
                for (int m0 = 0; m0 < 27; m0 += kGather) {
                    float ta[kGather], tb[kGather], td[kGather];
#pragma unroll
                    for (int m = 0; m < kGather; ++m) {
                        const uint32_t ob = x[0][m0 + m], od = x[1][m0 + m];
                        ta[m] = tab_at(tab + 4, ob + od);
                        tb[m] = tab_at(tab, ob);
                        td[m] = tab_at(tab, od);
                    }
                    __builtin_amdgcn_sched_barrier(0);
#pragma unroll
                    for (int m = 0; m < kGather; ++m) {
                        const float term = (ta[m] - tb[m]) - td[m];
                        sc = (m0 + m == 0) ? term : sc + term;
                    }
                }
            }
            const float s = valid ? fabsf(sc) : INFINITY;
            // Lanes that may enter the top list: the stored-order sum is
            // within kOrderSlack of the reference-order one, so a lane whose
            // scaled sum exceeds the wave's and the cohort's thresholds
            // cannot pass them in either order.  Most slots have none.
            const bool near = valid && (DUMP || s * kOrderSlack <= fminf(Wt.gthr, Wt.ts));
            if (__ballot(near) == 0)
                continue;
            // Rows in another order than the original one (C->orig): the key
            // holds the sorted original indices, and the reference sums the
            // cells in the order of those; a lane whose axis order is not the
            // identity is summed again in that order after the slot loop
            // (every dump lane goes there).
            uint64_t key = make_key(i, j, k);
            bool redo = DUMP && valid;
            if (C->orig != nullptr) {
                const uint32_t oi = C->orig[i], oj = L.mg.oj[jl], ok = L.mg.ok[kl];
                uint32_t t3[3];
                sort3(oi, oj, ok, t3);
                key = make_key(t3[0], t3[1], t3[2]);
                redo = redo || (near && axis_order(oi, oj, ok) != 0);
            }
            redo_mask |= (uint32_t)redo << s8;
            if (!DUMP)
                topk::wave_offer(Wt, near && !redo, s, key, lane);
        }
        // Lanes summed in the reference's cell order: the slot's cells are
        // rebuilt from the accumulators and margins (the stored-order ones
        // are not kept), with the axes in increasing original index.
        if (__ballot(redo_mask != 0) != 0) {
#pragma unroll 1
            for (uint32_t s8 = 0; s8 < 8; ++s8) {
                const bool rd = ((redo_mask >> s8) & 1u) != 0;
                if (__ballot(rd) == 0)
                    continue;
                const uint32_t jl = jl0 + 32 * (s8 >> 2) + 8 * (s8 & 3);
                const uint32_t j = J0 + jl;
                uint32_t t8[2][8];
                EPI_MF_T8(t8)
                const size_t off = (size_t)(32 * (s8 >> 2) + 8 * (s8 & 3)) * 64;
                uint32_t o = 0, t3[3] = {i, j, k};
                if (C->orig != nullptr) {
                    const uint32_t oi = C->orig[i], oj = L.mg.oj[jl], ok = L.mg.ok[kl];
                    sort3(oi, oj, ok, t3);
                    o = axis_order(oi, oj, ok);
                }
                // the cells in the reference's order: derive27 of the
                // margins rearranged into the axes of increasing original index
                uint32_t x[2][27];
#pragma unroll
                for (int c = 0; c < 2; ++c) {
                    const uint4 ij = L.mg.ij[c][jl];
                    const uint4 jk = x4(jkp[c][off]);
                    const uint2 sj = L.mg.sj[c][jl];
                    Axes3 st, rf;
#pragma unroll
                    for (int e = 0; e < 8; ++e)
                        st.t8[e] = t8[c][e];
                    st.ij[0] = ij.x, st.ij[1] = ij.y, st.ij[2] = ij.z, st.ij[3] = ij.w;
                    st.ik[0] = ik[c].x, st.ik[1] = ik[c].y, st.ik[2] = ik[c].z, st.ik[3] = ik[c].w;
                    st.jk[0] = jk.x, st.jk[1] = jk.y, st.jk[2] = jk.z, st.jk[3] = jk.w;
                    st.si[0] = si[c].x, st.si[1] = si[c].y;
                    st.sj[0] = sj.x, st.sj[1] = sj.y;
                    st.sk[0] = sk[c].x, st.sk[1] = sk[c].y;
                    to_ref_axes(o, st, rf);
                    derive27(rf.t8, rf.ij, rf.ik, rf.jk, rf.si, rf.sj, rf.sk, c ? n14 : n04, x[c]);
                }
                float sc = 0.0f;
#pragma unroll
                for (int m0 = 0; m0 < 27; m0 += kGather) {
                    float ta[kGather], tb[kGather], td[kGather];
#pragma unroll
                    for (int m = 0; m < kGather; ++m) {
                        const uint32_t ob = x[0][m0 + m], od = x[1][m0 + m];
                        ta[m] = tab_at(tab + 4, ob + od);
                        tb[m] = tab_at(tab, ob);
                        td[m] = tab_at(tab, od);
                    }
                    __builtin_amdgcn_sched_barrier(0);
#pragma unroll
                    for (int m = 0; m < kGather; ++m) {
                        const float term = (ta[m] - tb[m]) - td[m];
                        sc = (m0 + m == 0) ? term : sc + term;
                    }
                }
                const float s = fabsf(sc);
                if (DUMP && rd) {
                    const uint64_t rk = dump_rank(M, t3[0], t3[1], t3[2]);
                    P.dump_scores[rk] = s;
                    if (P.dump_counts) {
#pragma unroll
                        for (int m = 0; m < 27; ++m) {
                            P.dump_counts[rk * 54 + m] = x[0][m] >> 2;
                            P.dump_counts[rk * 54 + 27 + m] = x[1][m] >> 2;
                        }
                    }
                }
                if (!DUMP)
                    topk::wave_offer(Wt, rd, s, make_key(t3[0], t3[1], t3[2]), lane);
            }
        }
#undef EPI_MF_CELLS
#undef EPI_MF_SLOT
#undef EPI_MF_T8
        // The wavefronts leave the epilogue together (measured faster than
        // letting the early ones run ahead into the next stage loop).
        __syncthreads();
        item = next;
    }
}
// Static LDS of the instance; a table too large for LDS reports more than
// any device has, so that the caller places it in global memory.
size_t mfma_lds_bytes(bool table_lds, uint32_t table_len)
{
    if (!table_lds)
        return sizeof(MfShared<false>);
    return table_len <= kMfTabMax ? sizeof(MfShared<true>) : ((size_t)1 << 40);
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

// The MFMA kernels' LDS is static (no dynamic LDS at launch).
hipError_t mfma_occupancy(bool fp4, bool table_lds, bool dump, size_t lds_bytes,
                          int *blocks_per_cu)
{
    (void)lds_bytes;
    return hipOccupancyMaxActiveBlocksPerMultiprocessor(
        blocks_per_cu, (const void *)mf_pick(fp4, table_lds, dump), kMfWaves * 64, 0);
}

hipError_t mfma_allow_lds(bool fp4, bool table_lds, bool dump, size_t lds_bytes)
{
    (void)fp4, (void)table_lds, (void)dump, (void)lds_bytes;
    return hipSuccess;
}

hipError_t launch_mfma_scan(bool fp4, bool table_lds, bool dump, const ScanParams &p,
                            uint32_t grid, hipStream_t s)
{
    if (table_lds && p.table_len > kMfTabMax)
        return hipErrorInvalidValue;
    hipLaunchKernelGGL(mf_pick(fp4, table_lds, dump), dim3(grid), dim3(kMfWaves * 64), 0, s, p);
    return hipGetLastError();
}

}  // namespace epi
