// Host build of the product's cell-derivation and SNP-order code and ABI
// layout probe, for the CPU test suite (tests/test_cells.py,
// tests/test_abi.py).
#include <stddef.h>
#include <stdint.h>

#include "../../cross-dpc-episdet_amd/csrc/episdet_cells.h"
#include "../../cross-dpc-episdet_amd/csrc/episdet_order.h"
#include "../../cross-dpc-episdet_amd/csrc/episdet_sample.h"
#include "../../include/episdet.h"

extern "C" void cells_derive(const uint32_t *t8, const uint32_t *ij, const uint32_t *ik,
                             const uint32_t *jk, const uint32_t *si, const uint32_t *sj,
                             const uint32_t *sk, uint32_t n, uint32_t *x)
{
    epi::derive27(t8, ij, ik, jk, si, sj, sk, n, x);
}

extern "C" void abi_sizes(size_t *out)
{
    out[0] = sizeof(epi_candidate);
    out[1] = sizeof(epi_result);
    out[2] = sizeof(epi_tile);
    out[3] = offsetof(epi_result, top);
    out[4] = offsetof(epi_result, n_top);
    out[5] = offsetof(epi_candidate, counts);
}

extern "C" uint32_t order_axis(uint32_t op, uint32_t oq, uint32_t orr)
{
    return epi::axis_order(op, oq, orr);
}

extern "C" int order_stored_cell(int o, int m) { return epi::stored_cell(o, m); }

// derive27 of stored-axis inputs rearranged into the reference's axes
extern "C" void order_ref_derive(uint32_t o, const uint32_t *t8, const uint32_t *ij,
                                 const uint32_t *ik, const uint32_t *jk, const uint32_t *si,
                                 const uint32_t *sj, const uint32_t *sk, uint32_t n, uint32_t *x)
{
    epi::Axes3 st, rf;
    for (int e = 0; e < 8; ++e)
        st.t8[e] = t8[e];
    for (int e = 0; e < 4; ++e) {
        st.ij[e] = ij[e];
        st.ik[e] = ik[e];
        st.jk[e] = jk[e];
    }
    for (int e = 0; e < 2; ++e) {
        st.si[e] = si[e];
        st.sj[e] = sj[e];
        st.sk[e] = sk[e];
    }
    epi::to_ref_axes(o, st, rf);
    epi::derive27(rf.t8, rf.ij, rf.ik, rf.jk, rf.si, rf.sj, rf.sk, n, x);
}

// the seed sample (episdet_sample.h): triplet of colex rank r, t-th rank
extern "C" void sample_triplet(uint64_t r, uint32_t M, uint32_t *out) { epi::colex_triplet(r, M, out); }

extern "C" uint64_t sample_rank(uint64_t total, uint32_t N, uint32_t t)
{
    return epi::SeedSampler(total, N).rank(t);
}
