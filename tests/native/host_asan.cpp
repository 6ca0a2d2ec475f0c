// Host-code checks built with AddressSanitizer and UBSan (g++, no GPU):
// the product's data generators and file readers/writers
// (cross-dpc-episdet_amd/csrc/datagen.cpp, io.cpp, compiled into this
// program), the 27-cell derivation (episdet_cells.h), and the server's poll
// protocol (server_poll.h) driven by real threads.  Prints one line per
// check and "host_asan ok" at the end; any failure or sanitizer report
// exits nonzero.  Built and run by tests/test_host_asan.py.
#include <sys/stat.h>
#include <unistd.h>

#include <atomic>
#include <condition_variable>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <mutex>
#include <random>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

#include "../../cross-dpc-episdet_amd/csrc/episdet_cells.h"
#include "../../cross-dpc-episdet_amd/csrc/server_poll.h"
#include "../../include/episdet.h"

#define CHECK(c)                                                          \
    do {                                                                  \
        if (!(c)) {                                                       \
            std::fprintf(stderr, "FAILED %s:%d: %s\n", __FILE__, __LINE__, #c); \
            std::exit(1);                                                 \
        }                                                                 \
    } while (0)

static std::string g_dir;

static void check_generators()
{
    const uint32_t N = 333, M = 70;
    std::vector<uint8_t> g((size_t)N * M), p(N);
    CHECK(epi_generate_reference(N, M, g.data(), p.data()) == EPI_OK);
    uint32_t cases = 0;
    for (uint32_t s = 0; s < N; ++s)
        cases += p[s] == 1;
    for (uint8_t v : g)
        CHECK(v <= 2);
    std::printf("reference generator: %u cases of %u\n", cases, N);
    for (uint32_t nc : {0u, 1u, 100u, N}) {
        CHECK(epi_generate_hwe(7, N, M, 0.05, 0.5, nc, g.data(), p.data()) == EPI_OK);
        uint32_t c = 0;
        for (uint32_t s = 0; s < N; ++s)
            c += p[s];
        CHECK(c == nc);
        for (uint8_t v : g)
            CHECK(v <= 2);
    }
    uint32_t pl[3] = {0, 0, 0};
    CHECK(epi_generate_planted(11, N, M, 0.05, 0.5, N / 2, g.data(), p.data(), pl) == EPI_OK);
    CHECK(pl[0] < pl[1] && pl[1] < pl[2] && pl[2] < M);
    // rejected arguments
    CHECK(epi_generate_hwe(7, N, M, 0.05, 0.5, N + 1, g.data(), p.data()) == EPI_EINVAL);
    CHECK(epi_generate_hwe(7, N, M, 0.05, 0.5, 1, nullptr, p.data()) == EPI_EINVAL);
    CHECK(epi_generate_reference(N, M, g.data(), nullptr) == EPI_EINVAL);
    std::printf("generators ok\n");
}

static void write_file(const std::string &path, const void *data, size_t n)
{
    FILE *f = std::fopen(path.c_str(), "wb");
    CHECK(f);
    CHECK(std::fwrite(data, 1, n, f) == n);
    std::fclose(f);
}

static void check_plink()
{
    const uint32_t N = 37, M = 9;  // N % 4 != 0: a partly filled last byte
    std::vector<uint8_t> g((size_t)N * M), p(N);
    std::mt19937 rng(5);
    for (auto &v : g)
        v = (uint8_t)(rng() % 5);  // 3 and 4 are written as missing
    for (auto &v : p)
        v = (uint8_t)(rng() % 2);
    const std::string pre = g_dir + "/c";
    CHECK(epi_write_plink(pre.c_str(), g.data(), p.data(), N, M) == EPI_OK);
    uint32_t n = 0, m = 0;
    CHECK(epi_read_plink(pre.c_str(), &n, &m, nullptr, nullptr) == EPI_OK);
    CHECK(n == N && m == M);
    const uint32_t bpr = (N + 3) / 4;
    std::vector<uint8_t> bed((size_t)bpr * M), ph(N);
    CHECK(epi_read_plink(pre.c_str(), &n, &m, bed.data(), ph.data()) == EPI_OK);
    for (uint32_t j = 0; j < M; ++j)
        for (uint32_t s = 0; s < N; ++s) {
            const uint32_t code = (bed[(size_t)j * bpr + s / 4] >> (2 * (s % 4))) & 3u;
            const uint8_t v = g[(size_t)s * M + j];
            const uint32_t want = v == 0 ? 3u : v == 1 ? 2u : v == 2 ? 0u : 1u;
            CHECK(code == want);
        }
    for (uint32_t s = 0; s < N; ++s)
        CHECK(ph[s] == p[s]);
    // buffers sized for another fileset
    uint32_t n2 = N + 1;
    CHECK(epi_read_plink(pre.c_str(), &n2, &m, bed.data(), ph.data()) == EPI_EINVAL);
    // a truncated .bed, a bad magic, a short .fam line, a missing file
    std::vector<uint8_t> body(3 + (size_t)bpr * M - 1, 0);
    body[0] = 0x6c, body[1] = 0x1b, body[2] = 0x01;
    write_file(pre + ".bed", body.data(), body.size());
    CHECK(epi_read_plink(pre.c_str(), &n, &m, nullptr, nullptr) == EPI_EFORMAT);
    body.push_back(0);
    body[2] = 0x00;
    write_file(pre + ".bed", body.data(), body.size());
    CHECK(epi_read_plink(pre.c_str(), &n, &m, nullptr, nullptr) == EPI_EFORMAT);
    const char fam[] = "f0 s0 0 0 0\n";
    write_file(pre + ".fam", fam, sizeof(fam) - 1);
    CHECK(epi_read_plink(pre.c_str(), &n, &m, nullptr, nullptr) == EPI_EFORMAT);
    CHECK(epi_read_plink((g_dir + "/none").c_str(), &n, &m, nullptr, nullptr) == EPI_EIO);
    CHECK(epi_read_plink(pre.c_str(), &n, &m, bed.data(), nullptr) == EPI_EINVAL);
    std::printf("plink ok\n");
}

static void check_raw()
{
    const uint32_t N = 21, M = 5;
    std::vector<uint8_t> g((size_t)N * M), p(N);
    for (size_t t = 0; t < g.size(); ++t)
        g[t] = (uint8_t)(t % 3);
    for (uint32_t s = 0; s < N; ++s)
        p[s] = (uint8_t)(s & 1);
    const std::string gp = g_dir + "/g.raw", pp = g_dir + "/p.raw";
    write_file(gp, g.data(), g.size());
    write_file(pp, p.data(), p.size());
    uint32_t m = 0;
    CHECK(epi_read_raw(gp.c_str(), pp.c_str(), N, &m, nullptr, nullptr) == EPI_OK);
    CHECK(m == M);
    std::vector<uint8_t> g2(g.size()), p2(N);
    CHECK(epi_read_raw(gp.c_str(), pp.c_str(), N, &m, g2.data(), p2.data()) == EPI_OK);
    CHECK(g2 == g && p2 == p);
    CHECK(epi_read_raw(gp.c_str(), pp.c_str(), N + 1, &m, nullptr, nullptr) == EPI_EFORMAT);
    write_file(gp, g.data(), g.size() - 1);
    CHECK(epi_read_raw(gp.c_str(), pp.c_str(), N, &m, nullptr, nullptr) == EPI_EFORMAT);
    std::printf("raw ok\n");
}

// 27 cells of one class by textbook counting (genotypes 0/1/2).
static void brute27(const std::vector<uint8_t> &gi, const std::vector<uint8_t> &gj,
                    const std::vector<uint8_t> &gk, uint32_t x[27])
{
    std::memset(x, 0, 27 * sizeof(uint32_t));
    for (size_t s = 0; s < gi.size(); ++s)
        x[9 * gi[s] + 3 * gj[s] + gk[s]]++;
}

static void check_cells()
{
    std::mt19937 rng(9);
    for (int rep = 0; rep < 200; ++rep) {
        const uint32_t n = rng() % 300;
        std::vector<uint8_t> gi(n), gj(n), gk(n);
        for (uint32_t s = 0; s < n; ++s)
            gi[s] = (uint8_t)(rng() % 3), gj[s] = (uint8_t)(rng() % 3), gk[s] = (uint8_t)(rng() % 3);
        // the inputs the kernels hold: 8 cells with no genotype 2, the
        // 2x2 pair margins and the single margins of genotypes 0 and 1
        uint32_t t8[8] = {}, ij[4] = {}, ik[4] = {}, jk[4] = {}, si[2] = {}, sj[2] = {}, sk[2] = {};
        for (uint32_t s = 0; s < n; ++s) {
            const uint32_t a = gi[s], b = gj[s], c = gk[s];
            if (a < 2 && b < 2 && c < 2)
                t8[4 * a + 2 * b + c]++;
            if (a < 2 && b < 2)
                ij[2 * a + b]++;
            if (a < 2 && c < 2)
                ik[2 * a + c]++;
            if (b < 2 && c < 2)
                jk[2 * b + c]++;
            if (a < 2)
                si[a]++;
            if (b < 2)
                sj[b]++;
            if (c < 2)
                sk[c]++;
        }
        uint32_t got[27], want[27];
        epi::derive27(t8, ij, ik, jk, si, sj, sk, n, got);
        brute27(gi, gj, gk, want);
        CHECK(std::memcmp(got, want, sizeof(got)) == 0);
    }
    std::printf("cells ok\n");
}

// The server's request table, as far as the poll protocol sees it.
struct Req {
    bool done = false, polling = false;
    int status = EPI_OK;
    std::vector<int> payload = std::vector<int>(64, 7);  // freed with the entry
};

static void check_poll()
{
    std::mutex m;
    std::condition_variable cv;
    std::unordered_map<uint64_t, std::unique_ptr<Req>> all;
    uint32_t inside = 0;
    std::mt19937 rng(3);
    int collected_total = 0;
    for (int rep = 0; rep < 300; ++rep) {
        const uint64_t id = (uint64_t)rep + 1;
        const int status = rep % 3 == 0 ? EPI_ECANCELED : EPI_OK;
        {
            std::lock_guard<std::mutex> lk(m);
            all.emplace(id, std::unique_ptr<Req>(new Req()));
            // other entries come and go meanwhile (rehashing the table)
            for (uint64_t e = 0; e < 8; ++e)
                all.emplace(1000000 + (uint64_t)rep * 8 + e, std::unique_ptr<Req>(new Req()));
        }
        std::atomic<int> collected{0}, noent{0};
        const int delay_us = (int)(rng() % 2000);
        std::vector<std::thread> th;
        for (int t = 0; t < 4; ++t)
            th.emplace_back([&, t] {
                for (;;) {
                    std::unique_lock<std::mutex> lk(m);
                    const int32_t timeout = t == 0 ? -1 : t == 1 ? 0 : 1;
                    int sum = 0;
                    const int rc = epi::poll_request(all, id, lk, cv, timeout, inside,
                                                     [&](const Req &r) {
                                                         for (int v : r.payload)
                                                             sum += v;
                                                     });
                    if (rc == EPI_EAGAIN || rc == EPI_EBUSY)
                        continue;
                    if (rc == EPI_ENOENT) {
                        noent++;
                    } else {
                        CHECK(rc == status && sum == 64 * 7);
                        collected++;
                    }
                    return;
                }
            });
        std::thread completer([&] {
            std::this_thread::sleep_for(std::chrono::microseconds(delay_us));
            std::lock_guard<std::mutex> lk(m);
            Req *r = all.at(id).get();
            r->status = status;
            r->done = true;
            for (uint64_t e = 0; e < 8; ++e)
                all.erase(1000000 + (uint64_t)rep * 8 + e);
            cv.notify_all();
        });
        completer.join();
        for (auto &t : th)
            t.join();
        CHECK(collected == 1 && noent == 3);
        collected_total += collected;
        std::unique_lock<std::mutex> lk(m);
        CHECK(inside == 0 && all.count(id) == 0);
    }
    // an owner waiting for the pollers before freeing the table
    {
        std::unique_lock<std::mutex> lk(m);
        all.emplace(1, std::unique_ptr<Req>(new Req()));
    }
    std::thread poller([&] {
        std::unique_lock<std::mutex> lk(m);
        CHECK(epi::poll_request(all, 1, lk, cv, -1, inside, [](const Req &) {}) == EPI_ECANCELED);
    });
    std::this_thread::sleep_for(std::chrono::milliseconds(2));
    {
        std::unique_lock<std::mutex> lk(m);
        all.at(1)->status = EPI_ECANCELED;
        all.at(1)->done = true;
        cv.notify_all();
        cv.wait(lk, [&] { return inside == 0; });
        all.clear();
    }
    poller.join();
    std::printf("poll protocol ok (%d collected)\n", collected_total);
}

int main()
{
    char tmpl[] = "/tmp/host_asan.XXXXXX";
    CHECK(mkdtemp(tmpl) != nullptr);
    g_dir = tmpl;
    check_generators();
    check_plink();
    check_raw();
    check_cells();
    check_poll();
    for (const char *f : {"c.bed", "c.bim", "c.fam", "g.raw", "p.raw"})
        (void)unlink((g_dir + "/" + f).c_str());
    (void)rmdir(g_dir.c_str());
    std::printf("host_asan ok\n");
    return 0;
}
