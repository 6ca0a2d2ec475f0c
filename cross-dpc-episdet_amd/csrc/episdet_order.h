// episdet_order.h -- the stored SNP order of a cohort and the reference's
// cell order of a triplet.
//
// A cohort may store its SNPs in order of allele frequency (stored position
// p holds original SNP orig[p]): neighbouring SNPs then have similar
// genotype counts, so the 64 lanes of a slot, which hold 64 SNPs k, look up
// nearby ln(n!) entries and the lookups hit fewer LDS bank conflicts.
//
// The reference scores triplet i < j < k (original indices) with its 27
// cells in the order 9 g_i + 3 g_j + g_k (GPU/v4.cpp:317-319, :579-583),
// summed in FP32.  The kernels count a stored triplet p < q < r with its
// cells in stored axis order (9 g_p + 3 g_q + g_r); its original indices
// orig[p], orig[q], orig[r] are in one of six orders (axis_order), and the
// reference's cell m is the stored cell stored_cell(order, m).  The terms
// are the same numbers in both orders, only the order of the FP32 sum
// differs: order 0 (the original indices increase with the stored
// positions) gives the reference's score bit for bit; for the other five
// the kernels rebuild the cells in the reference's order (to_ref_axes, or
// a recount) for the lanes that may reach the top 64 and sum them again.
//
// Two FP32 sums of the same 27 terms, each term >= 0, differ by at most
// 2 * 26 * u / (1 - 26 u) of the sum (recursive summation, u = 2^-24):
// 3.1e-6 of the stored-order sum.  A stored-order sum s therefore bounds
// the reference's sum from below by s * (1 - 3.2e-6); kOrderSlack keeps a
// margin over that.
//
// Plain C++ so the same code is compiled for the GPU and, in the CPU test
// suite, for the host.
#pragma once
#include <stdint.h>

#include <utility>

#ifndef EPI_HD
#if defined(__HIPCC__)
#define EPI_HD __host__ __device__ inline __attribute__((always_inline))
#else
#define EPI_HD inline
#endif
#endif

namespace epi {

// s * kOrderSlack <= threshold is necessary for the reference's sum of a
// triplet whose stored-order sum is s to be <= threshold.
constexpr float kOrderSlack = 1.0f - 1.0e-5f;

// Axis order of a stored triplet p < q < r with original indices op, oq,
// orr (distinct): with rp, rq, rr the ranks (0 = smallest original index)
// of axes p, q, r, the order is 2 rp + (rq > rr), so 0 is rp, rq, rr =
// 0, 1, 2 (no reordering) and 5 is 2, 1, 0.
EPI_HD uint32_t axis_order(uint32_t op, uint32_t oq, uint32_t orr)
{
    const uint32_t rp = (uint32_t)(op > oq) + (uint32_t)(op > orr);
    const uint32_t rq = (uint32_t)(oq > op) + (uint32_t)(oq > orr);
    const uint32_t rr = 3u - rp - rq;
    return 2u * rp + (uint32_t)(rq > rr);
}

// Stored cell (9 g_p + 3 g_q + g_r) of the reference's cell m (9 g of the
// smallest original index + 3 g of the middle one + g of the largest)
// under axis order o.
constexpr int stored_cell(int o, int m)
{
    // ranks of axes p, q, r for order o = 2 rp + (rq > rr)
    const int rp = o / 2;
    const int lo = rp == 0 ? 1 : 0, hi = rp == 2 ? 1 : 2;  // the two ranks left for q, r
    const int rq = (o & 1) ? hi : lo, rr = (o & 1) ? lo : hi;
    const int d[3] = {m / 9, (m / 3) % 3, m % 3};  // genotype of rank 0, 1, 2
    return 9 * d[rp] + 3 * d[rq] + d[rr];
}

// Rank (0 = smallest original index) of stored axis a (0 = p, 1 = q,
// 2 = r) under axis order o, and the stored axis of rank t.
constexpr int axis_rank(int o, int a)
{
    const int rp = o / 2;
    const int lo = rp == 0 ? 1 : 0, hi = rp == 2 ? 1 : 2;
    const int rq = (o & 1) ? hi : lo, rr = (o & 1) ? lo : hi;
    return a == 0 ? rp : a == 1 ? rq : rr;
}
constexpr int rank_axis(int o, int t)
{
    return axis_rank(o, 0) == t ? 0 : axis_rank(o, 1) == t ? 1 : 2;
}

// The inputs of derive27 for one class (episdet_cells.h: 8 cells with
// genotypes in {0,1}, pair margins of (first, second), (first, third),
// (second, third) axes, single margins of each axis).
struct Axes3 {
    uint32_t t8[8], ij[4], ik[4], jk[4], si[2], sj[2], sk[2];
};

// Stored 8-cell index (4 g_p + 2 g_q + g_r) of cell e = 4 d0 + 2 d1 + d2 in
// axes of rank 0, 1, 2.
constexpr int stored_cell8(int o, int e)
{
    const int d[3] = {(e >> 2) & 1, (e >> 1) & 1, e & 1};
    return 4 * d[axis_rank(o, 0)] + 2 * d[axis_rank(o, 1)] + d[axis_rank(o, 2)];
}
// Stored pair table (0: p,q  1: p,r  2: q,r) holding the pair of ranks
// t0 < t1, and the entry of that table for e = 2 d_t0 + d_t1.
constexpr int pair_table(int o, int t0, int t1)
{
    const int a = rank_axis(o, t0), b = rank_axis(o, t1);
    const int lo = a < b ? a : b, hi = a < b ? b : a;
    return lo == 0 ? (hi == 1 ? 0 : 1) : 2;
}
constexpr int pair_entry(int o, int t0, int t1, int e)
{
    return rank_axis(o, t0) < rank_axis(o, t1) ? e : 2 * (e & 1) + (e >> 1);
}

template <int O, int E>
EPI_HD void ref_t8(const Axes3 &s, Axes3 &r)
{
    constexpr int c = stored_cell8(O, E);
    r.t8[E] = s.t8[c];
}
template <int O, int T0, int T1, int E>
EPI_HD uint32_t ref_pair(const Axes3 &s)
{
    constexpr int tb = pair_table(O, T0, T1), e = pair_entry(O, T0, T1, E);
    return tb == 0 ? s.ij[e] : tb == 1 ? s.ik[e] : s.jk[e];
}
template <int O, int T, int G>
EPI_HD uint32_t ref_single(const Axes3 &s)
{
    constexpr int a = rank_axis(O, T);
    return a == 0 ? s.si[G] : a == 1 ? s.sj[G] : s.sk[G];
}
template <int O, int... E>
EPI_HD void ref_axes_seq(const Axes3 &s, Axes3 &r, std::integer_sequence<int, E...>)
{
    (ref_t8<O, E>(s, r), ...);
    ((r.ij[E & 3] = ref_pair<O, 0, 1, (E & 3)>(s)), ...);
    ((r.ik[E & 3] = ref_pair<O, 0, 2, (E & 3)>(s)), ...);
    ((r.jk[E & 3] = ref_pair<O, 1, 2, (E & 3)>(s)), ...);
    ((r.si[E & 1] = ref_single<O, 0, (E & 1)>(s)), ...);
    ((r.sj[E & 1] = ref_single<O, 1, (E & 1)>(s)), ...);
    ((r.sk[E & 1] = ref_single<O, 2, (E & 1)>(s)), ...);
}

// The derive27 inputs of a stored triplet rearranged into the reference's
// axes (increasing original index) for axis order o: derive27 of the result
// gives the cells in the reference's order 9 d0 + 3 d1 + d2, so that the
// reference-order sum is the ordinary sum over m.  o is per lane; each case
// moves registers by constant indices.
EPI_HD void to_ref_axes(uint32_t o, const Axes3 &s, Axes3 &r)
{
    using Seq = std::make_integer_sequence<int, 8>;
    switch (o) {
    case 1:
        ref_axes_seq<1>(s, r, Seq{});
        break;
    case 2:
        ref_axes_seq<2>(s, r, Seq{});
        break;
    case 3:
        ref_axes_seq<3>(s, r, Seq{});
        break;
    case 4:
        ref_axes_seq<4>(s, r, Seq{});
        break;
    case 5:
        ref_axes_seq<5>(s, r, Seq{});
        break;
    default:
        ref_axes_seq<0>(s, r, Seq{});
        break;
    }
}

// The three original indices of a stored triplet in increasing order.
EPI_HD void sort3(uint32_t a, uint32_t b, uint32_t c, uint32_t out[3])
{
    const uint32_t lo = a < b ? (a < c ? a : c) : (b < c ? b : c);
    const uint32_t hi = a > b ? (a > c ? a : c) : (b > c ? b : c);
    out[0] = lo;
    out[1] = a + b + c - lo - hi;
    out[2] = hi;
}

}  // namespace epi
