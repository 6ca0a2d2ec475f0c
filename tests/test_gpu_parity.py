"""Parity of the HIP scan (through the C ABI) with the CPU oracle and the
reference's recorded outputs.  Integer counts and FP32 scores are compared
bit for bit; the reported triplet must be identical."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


def oracle_top(orc, A, B, topk=64):
    pc = orc.PackedCohort(A, B)
    _, top = pc.scan(topk=topk)
    return pc, top


def assert_top_matches(res, top):
    n = min(len(res.top), len([t for t in top if t[1] != 0xFFFFFFFF]))
    assert n == len(res.top)
    for e in range(n):
        c = res.top[e]
        s, i, j, k = top[e]
        assert (c.i, c.j, c.k) == (i, j, k), e
        assert np.float32(c.score).tobytes() == np.float32(s).tobytes(), e


# ------------------------------------------------------------ known answers
def test_reference_known_answers(epi, ctx, golden):
    """The reference generator + scan: best triplet and printed score as the
    reference's own programs print them (all sizes, incl. M > 1625 where
    GPU/v4.cpp:589 would wrap its index)."""
    for c in golden("kat_reference.json")["cases"]:
        say("KAT", c["N"], c["M"])
        g, p = epi.generate_reference(c["N"], c["M"])
        with ctx.cohort(g, p) as co:
            r = ctx.scan(co)
        assert [r.i, r.j, r.k] == c["runs"][0]["best"], c
        assert "%f" % r.score == c["runs"][0]["score"], c
        assert r.triplets == c["M"] * (c["M"] - 1) * (c["M"] - 2) // 6


def test_reference_fed_vectors(epi, ctx, golden):
    """Seeded HWE data fed to the reference programs (cells reach the
    Stirling branch): same best triplet and printed score."""
    for v in golden("ref_feed_vectors.json")["vectors"]:
        say("feed", v["seed"], v["N"], v["M"])
        g, p = epi.generate_hwe(v["seed"], v["N"], v["M"], v["cases"], *v["maf"])
        with ctx.cohort(g, p) as co:
            r = ctx.scan(co)
        for run in v["runs"]:
            assert [r.i, r.j, r.k] == run["best"], (v, run)
            assert "%f" % r.score == run["score"], (v, run)


# ------------------------------------------------------------ every triplet
@pytest.mark.parametrize("seed,N,M,cases", [(101, 1024, 256, 512), (102, 777, 70, 300),
                                            (103, 4096, 96, 2048), (104, 65, 67, 30),
                                            (105, 40000, 24, 20000)])
def test_all_scores_bit_exact(epi, orc, ctx, seed, N, M, cases):
    """Scan kernel in dump mode: the FP32 score of every triplet equals the
    oracle's, bit for bit (first case = the 256-SNP x 1024-sample config;
    the last one has a ln(n!) table too large for LDS)."""
    say("dump", seed, N, M)
    g, p = epi.generate_hwe(seed, N, M, cases)
    with ctx.cohort(g, p) as co:
        scores, _ = ctx.scan_dump(co)
    ref = orc.PackedCohort(g, p).all_scores()
    bad = np.nonzero(scores.view(np.uint32) != ref.view(np.uint32))[0]
    assert bad.size == 0, (bad[:10], scores[bad[:10]], ref[bad[:10]])


@pytest.mark.parametrize("N,M,p_case,extra", [(1, 5, 0.5, False), (63, 40, 0.5, True),
                                              (64, 33, 0.0, False), (65, 34, 1.0, False),
                                              (128, 64, 0.3, True), (500, 65, 0.6, True)])
def test_all_counts_bit_exact(epi, orc, ctx, N, M, p_case, extra):
    """54 counts of every triplet from the scan kernel (8 popcounted cells
    + margins) equal textbook counting: empty classes, word boundaries,
    genotype values > 2, non-binary phenotype values."""
    say("counts", N, M, p_case)
    rng = np.random.default_rng(N + M)
    A = rng.integers(0, 3, size=(N, M), dtype=np.uint8)
    B = (rng.random(N) < p_case).astype(np.uint8)
    if extra:
        A[rng.random((N, M)) < 0.05] = 3
        A[rng.random((N, M)) < 0.02] = 255
        B[rng.random(N) < 0.05] = 9
    with ctx.cohort(A, B) as co:
        scores, counts = ctx.scan_dump(co, with_counts=True)
    pc = orc.PackedCohort(A, B)
    r = 0
    for i in range(M):
        for j in range(i + 1, M):
            for k in range(j + 1, M):
                ref = pc.counts(i, j, k)
                assert np.array_equal(counts[r], ref), (i, j, k, counts[r], ref)
                assert scores[r] == np.float32(pc.score(ref))
                r += 1


def test_top64_matches_oracle(epi, orc, ctx):
    g, p = epi.generate_hwe(7, 1500, 150, 700)
    with ctx.cohort(g, p) as co:
        res = ctx.scan(co)
    pc, top = oracle_top(orc, g, p)
    assert len(res.top) == 64
    assert_top_matches(res, top)
    for c in res.top:
        ref = pc.counts(c.i, c.j, c.k)
        assert np.array_equal(c.counts, ref)
        assert c.score_fp64 == pc.score_f64(ref)
    assert res.max_abs_err_fp32 == max(abs(c.score - c.score_fp64) for c in res.top)


def test_count_entry_point(epi, orc, ctx):
    g, p = epi.generate_hwe(3, 3001, 200, 1400, 0.01, 0.2)
    rng = np.random.default_rng(0)
    tr = np.sort(np.stack([rng.choice(200, 3, replace=False) for _ in range(500)]), 1)
    with ctx.cohort(g, p) as co:
        counts, s32, s64 = ctx.count(co, tr)
    pc = orc.PackedCohort(g, p)
    for t in range(len(tr)):
        ref = pc.counts(*tr[t])
        assert np.array_equal(counts[t], ref)
        assert s32[t] == np.float32(pc.score(ref))
        assert s64[t] == pc.score_f64(ref)


@pytest.mark.parametrize("N,M,cases", [(1, 3, 0), (1, 3, 1), (2, 3, 1), (10, 4, 0), (10, 4, 10),
                                       (200, 63, 100), (200, 64, 100), (200, 65, 100),
                                       (129, 130, 64)])
def test_edge_shapes(epi, orc, ctx, N, M, cases):
    say("edge", N, M, cases)
    g, p = epi.generate_hwe(N * 1000 + M, N, M, cases)
    with ctx.cohort(g, p) as co:
        res = ctx.scan(co)
    _, top = oracle_top(orc, g, p)
    assert (res.i, res.j, res.k) == top[0][1:]
    assert_top_matches(res, top)


@pytest.mark.parametrize("tile", [dict(snps_j=1, waves=4), dict(snps_j=2, waves=8),
                                  dict(snps_j=4, waves=8), dict(snps_j=4, waves=16),
                                  dict(snps_j=8, waves=4), dict(snps_j=2, waves=16)])
def test_tile_shapes_identical(epi, orc, ctx, tile):
    g, p = epi.generate_hwe(44, 2000, 140, 1000)
    with ctx.cohort(g, p) as co:
        res = ctx.scan(co, tile=tile)
    _, top = oracle_top(orc, g, p)
    assert_top_matches(res, top)


def test_batch_equals_separate_scans(epi, ctx):
    rng = np.random.default_rng(77)
    cohorts = []
    try:
        for b in range(6):
            M = int(rng.integers(20, 200))
            N = int(rng.integers(100, 3000))
            g, p = epi.generate_hwe(500 + b, N, M, int(N * rng.uniform(0.3, 0.7)))
            cohorts.append(ctx.cohort(g, p))
        batch = ctx.scan_batch(cohorts)
        for co, rb in zip(cohorts, batch):
            rs = ctx.scan(co)
            assert (rb.i, rb.j, rb.k) == (rs.i, rs.j, rs.k)
            assert [(c.i, c.j, c.k, c.score) for c in rb.top] == \
                   [(c.i, c.j, c.k, c.score) for c in rs.top]
    finally:
        for co in cohorts:
            co.close()


def test_invalid_arguments(epi, ctx):
    g, p = epi.generate_hwe(1, 10, 5, 5)
    with pytest.raises(epi.EpiError):
        ctx.cohort(g[:, :2], p)  # fewer than three SNPs
    with ctx.cohort(g, p) as co:
        with pytest.raises(epi.EpiError):
            ctx.count(co, [[0, 0, 1]])
        with pytest.raises(epi.EpiError):
            ctx.count(co, [[0, 1, 5]])
        with pytest.raises(epi.EpiError):
            ctx.scan(co, tile=dict(snps_j=3))
