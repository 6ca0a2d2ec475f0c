// v4_cli.cpp -- drop-in for the reference's GPU program
//
//   v4 <num_samples> <num_snps> <block_size> [options]
//
// Same positional arguments (GPU/README.md:29-30, GPU/v4.cpp:210-212), the
// same generated data (srand(100) generator, GPU/v4.cpp:81-110) and the same
// four output lines (GPU/v4.cpp:251, :610-612):
//   Selected device: <name>
//   Time: <seconds>
//   Score: <best K2 score>
//   Solution: <i> <j> <k>
// Differences, all deliberate:
//   * the triplet indices are exact for every num_snps (the reference packs
//     them into a uint32 i*M*M + j*M + t, GPU/v4.cpp:589, which wraps above
//     1625 SNPs);
//   * block_size must be a positive integer; it is accepted for command-line
//     compatibility, the tile shape itself is chosen by the library;
//   * malformed arguments print a usage line and exit with status 2.
// Options (new): --json (one machine-readable result line), --hwe SEED /
// --planted SEED with --cases N (seeded Hardy-Weinberg data instead of the
// reference generator), --device D, --snps-j/--waves (tile override).
#include <chrono>
#include <cinttypes>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <iostream>
#include <string>
#include <vector>

#include "../../include/episdet.h"

static int usage()
{
    std::fprintf(stderr,
                 "usage: v4 <num_samples> <num_snps> <block_size> [--json] [--device D]\n"
                 "          [--hwe SEED | --planted SEED] [--cases N] [--snps-j J] [--waves W]\n");
    return 2;
}

static bool parse_u64(const char *s, uint64_t &out)
{
    if (!s || !*s)
        return false;
    char *end = nullptr;
    unsigned long long v = std::strtoull(s, &end, 10);
    if (*end != '\0' || s[0] == '-')
        return false;
    out = v;
    return true;
}

int main(int argc, char **argv)
{
    if (argc < 4)
        return usage();
    uint64_t n_samples = 0, n_snps = 0, block = 0;
    if (!parse_u64(argv[1], n_samples) || !parse_u64(argv[2], n_snps) ||
        !parse_u64(argv[3], block) || n_samples == 0 || n_snps < 3 || block == 0 ||
        n_samples > 0xFFFFFFFFull || n_snps > 0xFFFFFull)
        return usage();
    bool json = false;
    int device = 0;
    int mode = 0;  // 0 reference generator, 1 hwe, 2 planted
    uint64_t seed = 0, cases = n_samples / 2;
    epi_tile tile;
    std::memset(&tile, 0, sizeof(tile));
    for (int a = 4; a < argc; ++a) {
        std::string opt = argv[a];
        uint64_t v = 0;
        auto need = [&](uint64_t &dst) {
            if (a + 1 >= argc || !parse_u64(argv[a + 1], dst))
                return false;
            ++a;
            return true;
        };
        if (opt == "--json")
            json = true;
        else if (opt == "--device" && need(v))
            device = (int)v;
        else if (opt == "--hwe" && need(seed))
            mode = 1;
        else if (opt == "--planted" && need(seed))
            mode = 2;
        else if (opt == "--cases" && need(cases))
            ;
        else if (opt == "--snps-j" && need(v))
            tile.snps_j = (uint32_t)v;
        else if (opt == "--waves" && need(v))
            tile.waves = (uint32_t)v;
        else
            return usage();
    }
    if (cases > n_samples)
        return usage();

    const size_t cells = (size_t)n_samples * (size_t)n_snps;
    std::vector<uint8_t> geno(cells), pheno(n_samples);
    uint32_t planted[3] = {0, 0, 0};
    int rc = EPI_OK;
    if (mode == 0)
        rc = epi_generate_reference((uint32_t)n_samples, (uint32_t)n_snps, geno.data(), pheno.data());
    else if (mode == 1)
        rc = epi_generate_hwe(seed, (uint32_t)n_samples, (uint32_t)n_snps, 0.05, 0.5,
                              (uint32_t)cases, geno.data(), pheno.data());
    else
        rc = epi_generate_planted(seed, (uint32_t)n_samples, (uint32_t)n_snps, 0.05, 0.5,
                                  (uint32_t)cases, geno.data(), pheno.data(), planted);
    if (rc != EPI_OK) {
        std::fprintf(stderr, "v4: data generation failed: %s\n", epi_strerror(rc));
        return 1;
    }

    epi_ctx *ctx = nullptr;
    rc = epi_open(device, &ctx);
    if (rc != EPI_OK) {
        std::fprintf(stderr, "v4: cannot open device %d: %s\n", device, epi_strerror(rc));
        return 1;
    }
    char name[256];
    epi_device_name(ctx, name, sizeof(name));
    std::cout << "Selected device: " << name << std::endl;

    epi_cohort *cohort = nullptr;
    rc = epi_cohort_create(ctx, geno.data(), pheno.data(), (uint32_t)n_samples,
                           (uint32_t)n_snps, &cohort);
    if (rc != EPI_OK) {
        std::fprintf(stderr, "v4: cohort upload failed: %s\n", epi_strerror(rc));
        epi_close(ctx);
        return 1;
    }
    epi_result res;
    rc = epi_scan(ctx, cohort, &tile, &res);
    if (rc != EPI_OK) {
        std::fprintf(stderr, "v4: scan failed: %s\n", epi_strerror(rc));
        epi_cohort_destroy(cohort);
        epi_close(ctx);
        return 1;
    }
    std::cout << "Time: " << res.scan_ms / 1000.0 << std::endl;
    std::cout << "Score: " << res.score << std::endl;
    std::cout << "Solution: " << res.i << " " << res.j << " " << res.k << std::endl;
    if (json) {
        const double tps = (double)res.triplets * (double)res.n_samples / (res.scan_ms / 1000.0);
        std::printf("{\"n_samples\": %u, \"n_snps\": %u, \"cases\": %u, \"controls\": %u, "
                    "\"i\": %u, \"j\": %u, \"k\": %u, \"score\": %.9g, \"score_fp64\": %.17g, "
                    "\"scan_ms\": %.6f, \"kernel_ms\": %.6f, \"triplets\": %" PRIu64 ", "
                    "\"triplets_x_samples_per_s\": %.6e, \"max_abs_err_fp32\": %.3e, "
                    "\"fp64_order_differs\": %d",
                    res.n_samples, res.n_snps, res.n_cases, res.n_controls, res.i, res.j, res.k,
                    (double)res.score, res.score_fp64, res.scan_ms, res.kernel_ms, res.triplets,
                    tps, res.max_abs_err_fp32, res.fp64_order_differs);
        if (mode == 2)
            std::printf(", \"planted\": [%u, %u, %u]", planted[0], planted[1], planted[2]);
        std::printf("}\n");
    }
    epi_cohort_destroy(cohort);
    epi_close(ctx);
    return 0;
}
