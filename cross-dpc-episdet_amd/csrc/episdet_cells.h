// episdet_cells.h -- exact 27-cell contingency table of one SNP triplet in
// one class, rebuilt from 8 popcounted cells plus pair and single margins.
//
// The reference counts all 27 cells per sample word, with genotype 2
// derived as ~(g0|g1) under the padding mask (GPU/v4.cpp:321-448): 27
// three-way ANDs and 27 popcounts per word.  Every cell that involves
// genotype 2 is a difference of counts that do not depend on the third
// SNP's word loop:
//   x[a][b][2] = IJ[a][b] - x[a][b][0] - x[a][b][1]
// and likewise for the other positions, where IJ, IK, JK are pair tables
// and si, sj, sk single-SNP counts.  So the per-word work is the 8 cells
// with a, b, c in {0,1}; the rest is integer arithmetic once per triplet.
// All quantities are exact unsigned counts (intermediate wrap-around
// cancels), so the result equals the reference's table bit for bit.
//
// This header is plain C++ so the same code is compiled for the GPU and,
// in the CPU test suite, for the host.
#pragma once
#include <stdint.h>

#ifndef EPI_HD
#if defined(__HIPCC__)
#define EPI_HD __host__ __device__ inline __attribute__((always_inline))
#else
#define EPI_HD inline
#endif
#endif

namespace epi {

// t8[4a + 2b + c]  : popcounted cells, a (SNP i), b (SNP j), c (SNP k) in {0,1}
// ij[2a + b]       : count(g_i = a, g_j = b), a, b in {0,1}
// ik[2a + c], jk[2b + c] likewise
// si[a], sj[b], sk[c] : count(g = a) for a in {0,1};  n : class size
// x[9a + 3b + c]   : output, a, b, c in {0,1,2}
EPI_HD void derive27(const uint32_t t8[8], const uint32_t ij[4],
                     const uint32_t ik[4], const uint32_t jk[4],
                     const uint32_t si[2], const uint32_t sj[2],
                     const uint32_t sk[2], uint32_t n, uint32_t x[27])
{
    // extended pair margins including genotype 2
    uint32_t IK2[2], JK2[2], JK20, JK21, JK22;
    IK2[0] = si[0] - ik[0] - ik[1];        // count(g_i = 0, g_k = 2)
    IK2[1] = si[1] - ik[2] - ik[3];        // count(g_i = 1, g_k = 2)
    JK2[0] = sj[0] - jk[0] - jk[1];        // count(g_j = 0, g_k = 2)
    JK2[1] = sj[1] - jk[2] - jk[3];        // count(g_j = 1, g_k = 2)
    JK20 = sk[0] - jk[0] - jk[2];          // count(g_j = 2, g_k = 0)
    JK21 = sk[1] - jk[1] - jk[3];          // count(g_j = 2, g_k = 1)
    JK22 = n - sj[0] - sj[1] - JK20 - JK21;// count(g_j = 2, g_k = 2)

#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b) {
            const uint32_t c0 = t8[4 * a + 2 * b + 0];
            const uint32_t c1 = t8[4 * a + 2 * b + 1];
            x[9 * a + 3 * b + 0] = c0;
            x[9 * a + 3 * b + 1] = c1;
            x[9 * a + 3 * b + 2] = ij[2 * a + b] - c0 - c1;
        }
    // x[a][2][c], a, c in {0,1}
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int c = 0; c < 2; ++c)
            x[9 * a + 6 + c] = ik[2 * a + c] - x[9 * a + c] - x[9 * a + 3 + c];
    // x[2][b][c], b, c in {0,1}
#pragma unroll
    for (int b = 0; b < 2; ++b)
#pragma unroll
        for (int c = 0; c < 2; ++c)
            x[18 + 3 * b + c] = jk[2 * b + c] - x[3 * b + c] - x[9 + 3 * b + c];
    // x[a][2][2]
#pragma unroll
    for (int a = 0; a < 2; ++a)
        x[9 * a + 8] = IK2[a] - x[9 * a + 2] - x[9 * a + 5];
    // x[2][b][2]
#pragma unroll
    for (int b = 0; b < 2; ++b)
        x[18 + 3 * b + 2] = JK2[b] - x[3 * b + 2] - x[9 + 3 * b + 2];
    // x[2][2][c]
    x[24] = JK20 - x[6] - x[15];
    x[25] = JK21 - x[7] - x[16];
    // x[2][2][2]
    x[26] = JK22 - x[8] - x[17];
}

}  // namespace epi
