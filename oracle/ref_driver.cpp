/*
 * ref_driver.cpp -- file-input driver for the compiled reference CPU scans.
 *
 * TEST INFRASTRUCTURE ONLY.  The reference programs generate their own
 * data (generate_data, CPU/v2.cpp:56-84) with exactly N/2 cases, so they
 * cannot scan a cohort read from disk or one whose case fraction is not
 * one half.  This driver is linked with a reference program compiled from
 * its own source under the reference tree (its `main` renamed by the
 * preprocessor, the source unchanged) and stands in for that `main` only:
 * it reads a sample-major genotype matrix and a phenotype vector, then
 * calls the reference's own routines in the order its `main` does
 *   transpose_data       CPU/v2.cpp:86-101, CPU/v4_popc_avx512.cpp:104-119
 *   ln(n!) table         CPU/v2.cpp:580-583, CPU/v4_popc_avx512.cpp:4607-4611
 *   transposed_to_binary CPU/v2.cpp:103-164, CPU/v4_popc_avx512.cpp:121-182
 *   transform_data_2_block (blocked build) CPU/v4_popc_avx512.cpp:184-210
 *   the scan             CPU/v2.cpp:197-557 or CPU/v4_popc_avx512.cpp:243-4581
 * which prints the reference's "Time:" and "Best SNPs: i, j, k - Score: %f"
 * lines.
 *
 *   driver GENO PHENO N M [block_pac block_snp]
 *
 * GENO holds N*M bytes (sample-major, values 0/1/2), PHENO N bytes
 * (1 = case).  block_pac and block_snp are read by the blocked build only.
 */
#include <immintrin.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>

// Routines and the table pointer defined by the reference translation unit.
extern float *addlogtable;
float my_factorial(int n);
uint8_t *transpose_data(long long N, long long M, uint8_t *A);
void transposed_to_binary(uint8_t *original, uint8_t *original_ph, uint32_t **data_zeros,
                          uint32_t **data_ones, long long *phen_ones, long long num_snp,
                          long long num_pac);
#ifdef EPISDET_REF_BLOCKED
uint32_t *transform_data_2_block(long long N, long long M, uint32_t *A, int block_pac,
                                 int block_snp);
void process_epi_bin_no_phen_nor_block(uint32_t *data_zeros, uint32_t *data_ones,
                                       long long phen_ones, int dim_epi, long long num_snp,
                                       long long num_pac, int num_combs, int block_snp,
                                       int block_pac);
#else
void process_epi_bin_no_phen_nor(uint32_t *data_zeros, uint32_t *data_ones, long long phen_ones,
                                 int dim_epi, long long num_snp, long long num_pac,
                                 int num_combs);
#endif

static void read_exact(const char *path, void *dst, size_t bytes)
{
    FILE *f = fopen(path, "rb");
    if (!f) {
        fprintf(stderr, "ref_driver: cannot open %s\n", path);
        exit(3);
    }
    if (fread(dst, 1, bytes, f) != bytes) {
        fprintf(stderr, "ref_driver: %s is shorter than %zu bytes\n", path, bytes);
        exit(3);
    }
    fclose(f);
}

int main(int argc, char **argv)
{
#ifdef EPISDET_REF_BLOCKED
    const int need = 7;
#else
    const int need = 5;
#endif
    if (argc < need) {
        fprintf(stderr, "usage: %s GENO PHENO N M%s\n", argv[0],
                need == 7 ? " block_pac block_snp" : "");
        return 2;
    }
    const long long N = atoll(argv[3]);
    const long long M = atoll(argv[4]);
    if (N <= 0 || M < 3) {
        fprintf(stderr, "ref_driver: bad shape\n");
        return 2;
    }
    uint8_t *A = (uint8_t *)_mm_malloc((size_t)(N * M), 64);
    uint8_t *B = (uint8_t *)_mm_malloc((size_t)N, 64);
    read_exact(argv[1], A, (size_t)(N * M));
    read_exact(argv[2], B, (size_t)N);

    // the reference's main builds a 748-entry table (TABLE_MAX_SIZE)
    const int addlogsize = 748;
    addlogtable = new float[addlogsize];
    for (int i = 0; i < addlogsize; i++)
        addlogtable[i] = my_factorial(i);

    uint32_t *zeros = nullptr, *ones = nullptr;
    long long phen_ones = 0;
    uint8_t *At = transpose_data(N, M, A);  // frees A
    transposed_to_binary(At, B, &zeros, &ones, &phen_ones, M, N);
    _mm_free(At);
    _mm_free(B);
#ifdef EPISDET_REF_BLOCKED
    const int block_pac = atoi(argv[5]);
    const int block_snp = atoi(argv[6]);
    zeros = transform_data_2_block(N - phen_ones, M, zeros, block_pac, block_snp);
    ones = transform_data_2_block(phen_ones, M, ones, block_pac, block_snp);
    process_epi_bin_no_phen_nor_block(zeros, ones, phen_ones, 3, M, N, 27, block_snp, block_pac);
#else
    process_epi_bin_no_phen_nor(zeros, ones, phen_ones, 3, M, N, 27);
#endif
    _mm_free(zeros);
    _mm_free(ones);
    delete[] addlogtable;
    return 0;
}
