// tune.cpp -- tile-shape autotuner and the tile table consulted by scans.
//
// The VALU scan kernel is instantiated for snps_j (j SNPs per wavefront) in
// {1, 2, 4, 8} and waves per workgroup in {4, 8, 16}; the int8 MFMA kernel
// has one shape.  Which is fastest depends on the sample count (loop length
// per item vs. scoring and queue overhead, LDS footprint of the ln(n!)
// table, register pressure, 64 x 64 block waste on the triangle).
// epi_autotune times them all on a synthetic cohort; the results are kept
// per context as a table of (max_samples -> tile).
#include <hip/hip_runtime.h>
#include <dlfcn.h>

#include <algorithm>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/episdet.h"
#include "episdet_internal.h"
#include "episdet_lib.h"

namespace epi {

// Loads "max_samples snps_j waves" lines; ignores blank and '#' lines.
int tune_load_file(epi_ctx *ctx, const std::string &path)
{
    FILE *f = std::fopen(path.c_str(), "r");
    if (!f)
        return EPI_ENOENT;
    std::vector<TuneEntry> t;
    char line[256];
    while (std::fgets(line, sizeof(line), f)) {
        if (line[0] == '#' || line[0] == '\n')
            continue;
        unsigned ms = 0, sj = 0, w = 0, var = 1;
        const int got = std::sscanf(line, "%u %u %u %u", &ms, &sj, &w, &var);
        if (got >= 3) {
            TuneEntry e;
            e.max_samples = ms;
            e.cfg.variant = got == 4 ? (int)var : 1;
            e.cfg.snps_j = (int)sj;
            // the MFMA kernels have one shape (64 x 64 tiles, kMfmaWaves wavefronts)
            e.cfg.waves = e.cfg.variant >= 2 ? kMfmaWaves : (int)w;
            e.cfg.words = 1;
            if (scan_config_supported(e.cfg))
                t.push_back(e);
        }
    }
    std::fclose(f);
    std::sort(t.begin(), t.end(),
              [](const TuneEntry &a, const TuneEntry &b) { return a.max_samples < b.max_samples; });
    ctx->tune = t;
    return EPI_OK;
}

// tune/<arch>.txt beside the directory holding this library (lib/..).
void tune_load_default(epi_ctx *ctx)
{
    Dl_info info;
    if (!dladdr((void *)&tune_load_default, &info) || !info.dli_fname)
        return;
    std::string so = info.dli_fname;
    const size_t slash = so.rfind('/');
    std::string dir = slash == std::string::npos ? "." : so.substr(0, slash);
    std::string arch = ctx->arch.substr(0, ctx->arch.find(':'));
    (void)tune_load_file(ctx, dir + "/../tune/" + arch + ".txt");
}

bool tune_lookup(const epi_ctx *ctx, uint32_t n_samples, ScanConfig &cfg)
{
    for (const TuneEntry &e : ctx->tune)
        if (n_samples <= e.max_samples) {
            cfg = e.cfg;
            return true;
        }
    return false;
}

}  // namespace epi

using namespace epi;

extern "C" int epi_tune_set(epi_ctx *ctx, const uint32_t *max_samples, const epi_tile *tiles,
                            uint32_t n)
{
    if (!ctx || (n && (!max_samples || !tiles)))
        return EPI_EINVAL;
    std::vector<TuneEntry> t;
    for (uint32_t e = 0; e < n; ++e) {
        TuneEntry x;
        x.max_samples = max_samples[e];
        const bool mf = tiles[e].variant == 2 || tiles[e].variant == 3;
        x.cfg.variant = mf ? (int)tiles[e].variant : 1;
        x.cfg.snps_j = mf ? 64 : (int)tiles[e].snps_j;
        x.cfg.waves = mf ? kMfmaWaves : (int)tiles[e].waves;
        x.cfg.words = 1;
        if (!scan_config_supported(x.cfg) || (e && max_samples[e] <= max_samples[e - 1]))
            return EPI_EINVAL;
        t.push_back(x);
    }
    std::lock_guard<epi::FairMutex> lk(ctx->mu);
    ctx->tune = t;
    return EPI_OK;
}

extern "C" int epi_tune_get(epi_ctx *ctx, uint32_t n_samples, epi_tile *tile)
{
    if (!ctx || !tile)
        return EPI_EINVAL;
    std::lock_guard<epi::FairMutex> lk(ctx->mu);
    const ScanConfig c = resolve_config(ctx, nullptr, n_samples);
    std::memset(tile, 0, sizeof(*tile));
    tile->variant = (uint32_t)c.variant;
    tile->snps_j = (uint32_t)c.snps_j;
    tile->waves = (uint32_t)c.waves;
    tile->words = (uint32_t)c.words;
    return EPI_OK;
}

extern "C" int epi_autotune(epi_ctx *ctx, uint32_t n_samples, uint32_t n_snps, epi_tile *best,
                            double *best_ms)
{
    if (!ctx || !best || !cohort_shape_ok(n_samples, n_snps))
        return EPI_EINVAL;
    const uint32_t M = std::min<uint32_t>(n_snps, 1024);
    std::vector<uint8_t> g((size_t)n_samples * M), p(n_samples);
    EPI_TRY(epi_generate_hwe(0x7475E5u, n_samples, M, 0.05, 0.5, n_samples / 2, g.data(),
                             p.data()));
    epi_cohort *co = nullptr;
    EPI_TRY(epi_cohort_create(ctx, g.data(), p.data(), n_samples, M, &co));
    double best_t = 1e300;
    epi_tile best_tile;
    std::memset(&best_tile, 0, sizeof(best_tile));
    int st = EPI_OK;
    epi_result ref, r;
    bool have_ref = false;
    std::vector<epi_tile> shapes;
    for (int sj : {1, 2, 4, 8})
        for (int w : {4, 8, 16}) {
            epi_tile t;
            std::memset(&t, 0, sizeof(t));
            t.variant = 1;
            t.snps_j = (uint32_t)sj;
            t.waves = (uint32_t)w;
            t.words = 1;
            shapes.push_back(t);
        }
    for (uint32_t v : {2u, 3u}) {
        epi_tile t;
        std::memset(&t, 0, sizeof(t));
        t.variant = v;
        t.snps_j = 64;
        t.waves = kMfmaWaves;
        t.words = 1;
        shapes.push_back(t);
    }
    for (const epi_tile &t : shapes) {
        double tmin = 1e300;
        for (int rep = 0; rep < 3 && st == EPI_OK; ++rep) {
            st = epi_scan(ctx, co, &t, &r);
            if (st == EPI_OK)
                tmin = std::min(tmin, r.kernel_ms);
        }
        if (st != EPI_OK)
            break;
        // every shape must give the same answer
        if (!have_ref) {
            ref = r;
            have_ref = true;
        } else if (r.i != ref.i || r.j != ref.j || r.k != ref.k ||
                   std::memcmp(&r.score, &ref.score, sizeof(float)) != 0) {
            st = EPI_EINTERNAL;
            break;
        }
        if (tmin < best_t) {
            best_t = tmin;
            best_tile = t;
        }
    }
    epi_cohort_destroy(co);
    if (st != EPI_OK)
        return st;
    *best = best_tile;
    if (best_ms)
        *best_ms = best_t;
    return EPI_OK;
}
