// Host build of the product's cell-derivation code and ABI layout probe,
// for the CPU test suite (tests/test_cells.py, tests/test_abi.py).
#include <stddef.h>
#include <stdint.h>

#include "../../cross-dpc-episdet_amd/csrc/episdet_cells.h"
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
