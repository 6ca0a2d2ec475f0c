"""The matrix-core formulations of the 8 popcounted cells, int8 (tile
variant 2) and FP4 (variant 3): every count and FP32 score bit-exact against
the oracle, the reference's known answers, and the same top-64 as the VALU
kernel."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu
VARIANTS = pytest.mark.parametrize("var", [2, 3], ids=["int8", "fp4"])


def say(*a):
    print(*a, flush=True)


@pytest.mark.parametrize("seed,N,M,cases", [(201, 1024, 256, 512), (202, 777, 70, 300),
                                            (203, 4096, 130, 2048), (204, 65, 67, 30),
                                            (205, 40000, 24, 20000), (206, 200, 3, 100),
                                            # unequal stage counts per class (the
                                            # classes share one row pitch): cases the
                                            # larger class, then the smaller
                                            (207, 3000, 90, 2200), (208, 2500, 100, 400)])
@VARIANTS
def test_mfma_all_scores_bit_exact(epi, orc, ctx, seed, N, M, cases, var):
    say("mfma dump", var, seed, N, M)
    g, p = epi.generate_hwe(seed, N, M, cases)
    with ctx.cohort(g, p) as co:
        scores, _ = ctx.scan_dump(co, tile={"variant": var})
    ref = orc.PackedCohort(g, p).all_scores()
    bad = np.nonzero(scores.view(np.uint32) != ref.view(np.uint32))[0]
    assert bad.size == 0, (bad[:10], scores[bad[:10]], ref[bad[:10]])


@pytest.mark.parametrize("N,M,p_case", [(1, 5, 0.5), (63, 40, 0.5), (64, 33, 0.0), (65, 34, 1.0),
                                        (129, 66, 0.3), (500, 65, 0.6)])
@VARIANTS
def test_mfma_all_counts_bit_exact(epi, orc, ctx, N, M, p_case, var):
    say("mfma counts", var, N, M, p_case)
    rng = np.random.default_rng(N * 3 + M)
    A = rng.integers(0, 3, size=(N, M), dtype=np.uint8)
    A[rng.random((N, M)) < 0.03] = 7
    B = (rng.random(N) < p_case).astype(np.uint8)
    with ctx.cohort(A, B) as co:
        scores, counts = ctx.scan_dump(co, tile={"variant": var}, with_counts=True)
    pc = orc.PackedCohort(A, B)
    r = 0
    for i in range(M):
        for j in range(i + 1, M):
            for k in range(j + 1, M):
                ref = pc.counts(i, j, k)
                assert np.array_equal(counts[r], ref), (i, j, k, counts[r], ref)
                r += 1


@VARIANTS
def test_mfma_reference_known_answers(epi, ctx, golden, var):
    for c in golden("kat_reference.json")["cases"]:
        say("mfma KAT", var, c["N"], c["M"])
        g, p = epi.generate_reference(c["N"], c["M"])
        with ctx.cohort(g, p) as co:
            r = ctx.scan(co, tile={"variant": var})
        assert [r.i, r.j, r.k] == c["runs"][0]["best"], c
        assert "%f" % r.score == c["runs"][0]["score"], c


@VARIANTS
def test_mfma_top64_equals_valu(epi, ctx, var):
    for seed, N, M in [(7, 4096, 300), (8, 1500, 500), (9, 16384, 160)]:
        g, p = epi.generate_hwe(seed, N, M, N // 2)
        with ctx.cohort(g, p) as co:
            a = ctx.scan(co, tile={"variant": 1})
            b = ctx.scan(co, tile={"variant": var})
        assert [(c.i, c.j, c.k, c.score) for c in a.top] == [(c.i, c.j, c.k, c.score) for c in b.top]


@VARIANTS
def test_mfma_batch_equals_separate(epi, ctx, var):
    rng = np.random.default_rng(31)
    cos = []
    try:
        for b in range(5):
            M = int(rng.integers(60, 300))
            N = int(rng.integers(100, 5000))
            g, p = epi.generate_hwe(900 + b, N, M, N // 3)
            cos.append(ctx.cohort(g, p))
        batch = ctx.scan_batch(cos, tile={"variant": var})
        for co, rb in zip(cos, batch):
            rs = ctx.scan(co, tile={"variant": 1})
            assert [(c.i, c.j, c.k, c.score) for c in rb.top] == [(c.i, c.j, c.k, c.score) for c in rs.top]
    finally:
        for co in cos:
            co.close()


def test_kernels_alternate_on_one_cohort(epi, ctx):
    """One cohort scanned by the three kernels in turn: its cached tile rows
    are rebuilt whenever the tile width changes, and every scan is exact."""
    g, p = epi.generate_hwe(55, 3000, 200, 1400)
    with ctx.cohort(g, p) as co:
        runs = [ctx.scan(co, tile={"variant": v}) for v in (3, 1, 2, 3, 1)]
    ref = [(c.i, c.j, c.k, c.score) for c in runs[0].top]
    for r in runs[1:]:
        assert [(c.i, c.j, c.k, c.score) for c in r.top] == ref


@VARIANTS
def test_mfma_cohorts_in_reused_blocks(epi, orc, ctx, var):
    """Device blocks of destroyed cohorts are kept and handed to later
    cohorts without being cleared: cohorts created into blocks that held
    other cohorts (larger, smaller, other class balance) scan exactly."""
    shapes = [(4000, 200, 2000), (1200, 120, 300), (4000, 200, 3200), (700, 190, 350),
              (2600, 160, 1300)]
    for b, (N, M, cases) in enumerate(shapes):
        g, p = epi.generate_hwe(700 + b, N, M, cases)
        with ctx.cohort(g, p) as co:
            r = ctx.scan(co, tile={"variant": var})
        n, top = orc.PackedCohort(g, p).scan()
        say("reused blocks", var, N, M, cases, (r.i, r.j, r.k), top[0][1:])
        assert [(c.i, c.j, c.k) for c in r.top] == [t[1:] for t in top]
        assert [c.score for c in r.top] == [t[0] for t in top]
