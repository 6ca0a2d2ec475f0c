// io.cpp -- on-disk genotype input and output.
//
// The reference only generates its data in memory (GPU/v4.cpp:81-110,
// called at :219); these readers bring real cohorts to the same cohort
// constructors.  Two formats:
//   raw    <geno> holds n_samples x n_snps uint8 genotypes, sample-major, and
//          <pheno> n_samples uint8 phenotypes: byte for byte the reference's
//          in-memory A[N][M] and B[N] (GPU/v4.cpp:86-108);
//   PLINK 1 binary fileset <prefix>.bed/.bim/.fam: .fam has one line per
//          sample (6th column phenotype: 1 control, 2 case, anything else
//          missing), .bim one line per SNP, .bed the magic bytes 6c 1b 01
//          (SNP-major) and then one row of (n_samples + 3) / 4 bytes per SNP,
//          2-bit codes, sample t in bits 2(t%4)..2(t%4)+1 of byte t/4:
//          00 homozygous first allele, 01 missing, 10 heterozygous,
//          11 homozygous second allele.  The .bed body goes to the device
//          as it is (epi_cohort_create_bed unpacks it there).
#include <algorithm>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/episdet.h"

namespace {

struct File {
    FILE *f = nullptr;
    File(const std::string &path, const char *mode) { f = std::fopen(path.c_str(), mode); }
    ~File()
    {
        if (f)
            std::fclose(f);
    }
    File(const File &) = delete;
    File &operator=(const File &) = delete;
};

int file_size(FILE *f, uint64_t &size)
{
    if (std::fseek(f, 0, SEEK_END) != 0)
        return EPI_EIO;
    const long n = std::ftell(f);
    if (n < 0 || std::fseek(f, 0, SEEK_SET) != 0)
        return EPI_EIO;
    size = (uint64_t)n;
    return EPI_OK;
}

// Splits a line into whitespace-separated fields.
std::vector<std::string> fields(const char *line)
{
    std::vector<std::string> out;
    const char *p = line;
    while (*p) {
        while (*p == ' ' || *p == '\t' || *p == '\r' || *p == '\n')
            ++p;
        if (!*p)
            break;
        const char *q = p;
        while (*q && *q != ' ' && *q != '\t' && *q != '\r' && *q != '\n')
            ++q;
        out.emplace_back(p, q);
        p = q;
    }
    return out;
}

// Reads a text file line by line; calls fn(fields) for each non-blank line.
template <typename Fn>
int for_each_line(const std::string &path, Fn fn)
{
    File f(path, "r");
    if (!f.f)
        return EPI_EIO;
    std::string line;
    char buf[4096];
    while (std::fgets(buf, sizeof(buf), f.f)) {
        line += buf;
        if (line.empty() || (line.back() != '\n' && !std::feof(f.f)))
            continue;  // a line longer than the buffer: keep reading
        std::vector<std::string> fs = fields(line.c_str());
        line.clear();
        if (fs.empty())
            continue;
        const int st = fn(fs);
        if (st != EPI_OK)
            return st;
    }
    if (std::ferror(f.f))
        return EPI_EIO;
    return EPI_OK;
}

const uint8_t kBedMagic[3] = {0x6c, 0x1b, 0x01};

}  // namespace

extern "C" int epi_read_plink(const char *prefix, uint32_t *n_samples, uint32_t *n_snps,
                              uint8_t *bed, uint8_t *pheno)
{
    if (!prefix || !n_samples || !n_snps || (!bed != !pheno))
        return EPI_EINVAL;
    const std::string pre = prefix;
    // .fam: samples and phenotypes
    std::vector<uint8_t> ph;
    int st = for_each_line(pre + ".fam", [&](const std::vector<std::string> &fs) {
        if (fs.size() < 6)
            return EPI_EFORMAT;
        ph.push_back(fs[5] == "2" ? 1 : fs[5] == "1" ? 0 : 0xFF);
        return ph.size() > 0xFFFFFFFFull ? EPI_EFORMAT : EPI_OK;
    });
    if (st != EPI_OK)
        return st;
    uint64_t m = 0;
    st = for_each_line(pre + ".bim", [&](const std::vector<std::string> &fs) {
        if (fs.size() < 4)
            return EPI_EFORMAT;
        ++m;
        return EPI_OK;
    });
    if (st != EPI_OK)
        return st;
    if (ph.empty() || m < 3 || m > (1u << 21) - 1)
        return EPI_EFORMAT;
    const uint64_t n = ph.size();
    const uint64_t bpr = (n + 3) / 4;
    File f(pre + ".bed", "rb");
    if (!f.f)
        return EPI_EIO;
    uint64_t size = 0;
    st = file_size(f.f, size);
    if (st != EPI_OK)
        return st;
    uint8_t magic[3];
    if (size != 3 + bpr * m || std::fread(magic, 1, 3, f.f) != 3 ||
        std::memcmp(magic, kBedMagic, 3) != 0)
        return EPI_EFORMAT;
    if (!bed) {  // size query
        *n_samples = (uint32_t)n;
        *n_snps = (uint32_t)m;
        return EPI_OK;
    }
    if (*n_samples != n || *n_snps != m)
        return EPI_EINVAL;  // buffers were sized for another fileset
    if (std::fread(bed, 1, bpr * m, f.f) != bpr * m)
        return EPI_EIO;
    std::memcpy(pheno, ph.data(), n);
    return EPI_OK;
}

extern "C" int epi_write_plink(const char *prefix, const uint8_t *geno, const uint8_t *pheno,
                               uint32_t n_samples, uint32_t n_snps)
{
    if (!prefix || !geno || !pheno || n_samples == 0 || n_snps == 0)
        return EPI_EINVAL;
    const std::string pre = prefix;
    {
        File f(pre + ".fam", "w");
        if (!f.f)
            return EPI_EIO;
        for (uint32_t s = 0; s < n_samples; ++s)
            std::fprintf(f.f, "f%u s%u 0 0 0 %d\n", s, s, pheno[s] == 1 ? 2 : 1);
        if (std::ferror(f.f))
            return EPI_EIO;
    }
    {
        File f(pre + ".bim", "w");
        if (!f.f)
            return EPI_EIO;
        for (uint32_t j = 0; j < n_snps; ++j)
            std::fprintf(f.f, "1\tsnp%u\t0\t%u\tA\tB\n", j, j + 1);
        if (std::ferror(f.f))
            return EPI_EIO;
    }
    File f(pre + ".bed", "wb");
    if (!f.f)
        return EPI_EIO;
    if (std::fwrite(kBedMagic, 1, 3, f.f) != 3)
        return EPI_EIO;
    const uint32_t bpr = (n_samples + 3) / 4;
    std::vector<uint8_t> row(bpr);
    for (uint32_t j = 0; j < n_snps; ++j) {
        std::fill(row.begin(), row.end(), 0);
        for (uint32_t s = 0; s < n_samples; ++s) {
            const uint8_t g = geno[(size_t)s * n_snps + j];
            // 0 -> 11, 1 -> 10, 2 -> 00, anything else -> 01 (missing)
            const uint32_t code = g == 0 ? 3u : g == 1 ? 2u : g == 2 ? 0u : 1u;
            row[s / 4] |= (uint8_t)(code << (2 * (s % 4)));
        }
        if (std::fwrite(row.data(), 1, bpr, f.f) != bpr)
            return EPI_EIO;
    }
    return std::fflush(f.f) == 0 ? EPI_OK : EPI_EIO;
}

extern "C" int epi_read_raw(const char *geno_path, const char *pheno_path, uint32_t n_samples,
                            uint32_t *n_snps, uint8_t *geno, uint8_t *pheno)
{
    if (!geno_path || !pheno_path || !n_snps || n_samples == 0 || (!geno != !pheno))
        return EPI_EINVAL;
    File g(geno_path, "rb"), p(pheno_path, "rb");
    if (!g.f || !p.f)
        return EPI_EIO;
    uint64_t gs = 0, ps = 0;
    int st = file_size(g.f, gs);
    if (st == EPI_OK)
        st = file_size(p.f, ps);
    if (st != EPI_OK)
        return st;
    if (ps != n_samples || gs % n_samples != 0)
        return EPI_EFORMAT;
    const uint64_t m = gs / n_samples;
    if (m < 3 || m > (1u << 21) - 1)
        return EPI_EFORMAT;
    if (!geno) {  // size query
        *n_snps = (uint32_t)m;
        return EPI_OK;
    }
    if (*n_snps != m)
        return EPI_EINVAL;
    if (std::fread(geno, 1, gs, g.f) != gs || std::fread(pheno, 1, ps, p.f) != ps)
        return EPI_EIO;
    return EPI_OK;
}
