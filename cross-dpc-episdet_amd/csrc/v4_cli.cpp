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
// reference generator), --device D, --variant V / --snps-j J / --waves W
// (kernel and tile override), --top K (the K best triplets after the four
// lines, K <= 64).
// File input (new; the reference only generates data):
//   v4 --bfile PREFIX [options]              PLINK 1 PREFIX.bed/.bim/.fam
//   v4 --raw GENO PHENO <num_samples> [options]
//                                            sample-major uint8 genotypes and
//                                            uint8 phenotypes
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
                 "usage: v4 <num_samples> <num_snps> <block_size> [options]\n"
                 "       v4 --bfile PREFIX [options]\n"
                 "       v4 --raw GENO PHENO <num_samples> [options]\n"
                 "options: [--json] [--device D] [--hwe SEED | --planted SEED] [--cases N]\n"
                 "         [--variant V] [--snps-j J] [--waves W] [--top K]\n");
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
    if (argc < 2)
        return usage();
    uint64_t n_samples = 0, n_snps = 0, block = 1;
    int first_opt = 4;
    std::string bfile, raw_geno, raw_pheno;
    if (std::strcmp(argv[1], "--bfile") == 0) {
        if (argc < 3)
            return usage();
        bfile = argv[2];
        first_opt = 3;
    } else if (std::strcmp(argv[1], "--raw") == 0) {
        if (argc < 5 || !parse_u64(argv[4], n_samples) || n_samples == 0 ||
            n_samples > 0xFFFFFFFFull)
            return usage();
        raw_geno = argv[2];
        raw_pheno = argv[3];
        first_opt = 5;
    } else {
        if (argc < 4 || !parse_u64(argv[1], n_samples) || !parse_u64(argv[2], n_snps) ||
            !parse_u64(argv[3], block) || n_samples == 0 || n_snps < 3 || block == 0 ||
            n_samples > 0xFFFFFFFFull || n_snps > 0xFFFFFull)
            return usage();
    }
    const bool from_file = !bfile.empty() || !raw_geno.empty();
    bool json = false;
    int device = 0;
    int mode = 0;  // 0 reference generator, 1 hwe, 2 planted
    uint64_t seed = 0, cases = 0, top = 0;
    bool cases_given = false;
    epi_tile tile;
    std::memset(&tile, 0, sizeof(tile));
    for (int a = first_opt; a < argc; ++a) {
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
        else if (opt == "--hwe" && !from_file && need(seed))
            mode = 1;
        else if (opt == "--planted" && !from_file && need(seed))
            mode = 2;
        else if (opt == "--cases" && !from_file && need(cases))
            cases_given = true;
        else if (opt == "--variant" && need(v) && v <= 3)
            tile.variant = (uint32_t)v;
        else if (opt == "--snps-j" && need(v))
            tile.snps_j = (uint32_t)v;
        else if (opt == "--waves" && need(v))
            tile.waves = (uint32_t)v;
        else if (opt == "--top" && need(top) && top <= EPI_TOPK)
            ;
        else
            return usage();
    }
    if (!cases_given)
        cases = n_samples / 2;
    if (cases > n_samples)
        return usage();

    // ---- input: generated, PLINK fileset or raw bytes
    std::vector<uint8_t> geno, pheno, bed;
    uint32_t planted[3] = {0, 0, 0};
    int rc = EPI_OK;
    if (!bfile.empty()) {
        uint32_t n = 0, m = 0;
        rc = epi_read_plink(bfile.c_str(), &n, &m, nullptr, nullptr);
        if (rc == EPI_OK) {
            bed.resize((size_t)m * ((n + 3) / 4));
            pheno.resize(n);
            rc = epi_read_plink(bfile.c_str(), &n, &m, bed.data(), pheno.data());
        }
        n_samples = n;
        n_snps = m;
    } else if (!raw_geno.empty()) {
        uint32_t m = 0;
        rc = epi_read_raw(raw_geno.c_str(), raw_pheno.c_str(), (uint32_t)n_samples, &m, nullptr,
                          nullptr);
        if (rc == EPI_OK) {
            geno.resize((size_t)n_samples * m);
            pheno.resize(n_samples);
            rc = epi_read_raw(raw_geno.c_str(), raw_pheno.c_str(), (uint32_t)n_samples, &m,
                              geno.data(), pheno.data());
        }
        n_snps = m;
    } else {
        geno.resize((size_t)n_samples * (size_t)n_snps);
        pheno.resize(n_samples);
        if (mode == 0)
            rc = epi_generate_reference((uint32_t)n_samples, (uint32_t)n_snps, geno.data(),
                                        pheno.data());
        else if (mode == 1)
            rc = epi_generate_hwe(seed, (uint32_t)n_samples, (uint32_t)n_snps, 0.05, 0.5,
                                  (uint32_t)cases, geno.data(), pheno.data());
        else
            rc = epi_generate_planted(seed, (uint32_t)n_samples, (uint32_t)n_snps, 0.05, 0.5,
                                      (uint32_t)cases, geno.data(), pheno.data(), planted);
    }
    if (rc != EPI_OK) {
        std::fprintf(stderr, "v4: %s failed: %s\n", from_file ? "reading input" : "data generation",
                     epi_strerror(rc));
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
    if (!bfile.empty())
        rc = epi_cohort_create_bed(ctx, bed.data(), pheno.data(), (uint32_t)n_samples,
                                   (uint32_t)n_snps, &cohort);
    else
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
    for (uint32_t r = 0; r < top && r < res.n_top; ++r) {
        const epi_candidate &c = res.top[r];
        std::printf("Top %u: %u %u %u %.9g %.17g\n", r + 1, c.i, c.j, c.k, (double)c.score,
                    c.score_fp64);
    }
    std::fflush(stdout);
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
