"""Cohorts stored in allele-frequency order (epi_set_snp_order) against
cohorts stored in the original order and the oracle.  The stored
order changes which rows the kernels read and the order of their FP32 sums;
every result must still be that of the original order bit for bit: indices,
scores, counts, tie order, dumps, epi_count, batches, strip rounds."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

VARIANTS = [1, 2, 3]


def say(*a):
    print(*a, flush=True)


def same(a, b):
    return ((a.i, a.j, a.k, a.score) == (b.i, b.j, b.k, b.score) and
            [(c.i, c.j, c.k, c.score, c.score_fp64, tuple(c.counts)) for c in a.top] ==
            [(c.i, c.j, c.k, c.score, c.score_fp64, tuple(c.counts)) for c in b.top])


@pytest.fixture
def plain_ctx(epi):
    """A context of its own whose cohorts keep the original SNP order."""
    c = epi.Context(0)
    c.set_snp_order(0)
    yield c
    c.close()


@pytest.fixture
def ord_ctx(epi):
    """A context of its own whose cohorts are stored in allele-frequency
    order whatever their size."""
    c = epi.Context(0)
    c.set_snp_order(1)
    yield c
    c.close()


def tied_cohort(epi, seed, N, M):
    """HWE genotypes (MAF 0.05..0.5) in three copies of M // 3 columns, the
    middle one reversed: copies in the same column order give triplets with
    identical cells (exact ties, decided by the original indices), the
    reversed copy the same cells in another order (scores that may differ
    only in the FP32 summation order)."""
    g, p = epi.generate_hwe(seed, N, M // 3, N // 2)
    return np.ascontiguousarray(np.concatenate([g, g[:, ::-1], g], 1)), p


@pytest.mark.parametrize("var", VARIANTS)
def test_dump_equals_original_order(epi, orc, ord_ctx, plain_ctx, var):
    """Every triplet's FP32 score (against the oracle) and 54 counts
    (against the original-order cohort) at its original lexicographic rank."""
    for seed, N, M in [(31, 700, 90), (32, 1500, 130)]:
        say("order dump", var, seed, N, M)
        g, p = tied_cohort(epi, seed, N, M) if seed == 31 else epi.generate_hwe(seed, N, M, N // 3)
        with ord_ctx.cohort(g, p) as co:
            s_ord, c_ord = ord_ctx.scan_dump(co, tile={"variant": var}, with_counts=True)
        with plain_ctx.cohort(g, p) as co:
            s_pl, c_pl = plain_ctx.scan_dump(co, tile={"variant": var}, with_counts=True)
        ref = orc.PackedCohort(g, p).all_scores()
        assert np.array_equal(s_ord.view(np.uint32), ref.view(np.uint32))
        assert np.array_equal(s_pl.view(np.uint32), ref.view(np.uint32))
        assert np.array_equal(c_ord, c_pl)


@pytest.mark.parametrize("var", VARIANTS)
def test_top64_ties_in_original_order(epi, orc, ord_ctx, plain_ctx, var):
    """Exact ties (repeated columns): the top 64 by (score, i, j, k) in
    original indices, equal to the oracle's and the original-order scan."""
    g, p = tied_cohort(epi, 33, 1200, 159)
    say("order ties", var)
    with ord_ctx.cohort(g, p) as co:
        a = ord_ctx.scan(co, tile={"variant": var})
    with plain_ctx.cohort(g, p) as co:
        b = plain_ctx.scan(co, tile={"variant": var})
    assert same(a, b)
    pc = orc.PackedCohort(g, p)
    _, top = pc.scan(topk=64)
    assert len(a.top) == 64
    scores = [t[0] for t in top]
    assert len(set(np.float32(s).tobytes() for s in scores)) < 64  # ties present
    for e, c in enumerate(a.top):
        s, i, j, k = top[e]
        assert (c.i, c.j, c.k) == (i, j, k), e
        assert np.float32(c.score).tobytes() == np.float32(s).tobytes(), e
        assert np.array_equal(c.counts, pc.counts(i, j, k))


def test_large_ordered_scan_equals_original_order(epi, ctx, plain_ctx):
    """2048 SNPs x 4096 samples, FP4 kernel (the scan's default; the default
    order mode stores this cohort in allele-frequency order): identical
    results from both stored orders."""
    g, p = epi.generate_hwe(34, 4096, 2048, 2048)
    say("order large")
    with ctx.cohort(g, p) as co:
        a = ctx.scan(co)
    with plain_ctx.cohort(g, p) as co:
        b = plain_ctx.scan(co)
    assert same(a, b)


def test_count_and_batch_on_ordered_cohorts(epi, orc, ord_ctx):
    """epi_count takes and returns original indices; a batch mixing both
    stored orders equals the separate scans."""
    ctx = ord_ctx
    rng = np.random.default_rng(5)
    data = [epi.generate_hwe(40 + b, 600 + 150 * b, 100 + 30 * b, 300) for b in range(3)]
    ctx.set_snp_order(0)
    try:
        plain = ctx.cohort(*data[1])
    finally:
        ctx.set_snp_order(1)
    cos = [ctx.cohort(*data[0]), plain, ctx.cohort(*data[2])]
    try:
        g, p = data[0]
        tr = np.sort(np.stack([rng.choice(100, 3, replace=False) for _ in range(300)]), 1)
        counts, s32, s64 = ctx.count(cos[0], tr)
        pc = orc.PackedCohort(g, p)
        for t in range(len(tr)):
            ref = pc.counts(*tr[t])
            assert np.array_equal(counts[t], ref)
            assert s32[t] == np.float32(pc.score(ref))
            assert s64[t] == pc.score_f64(ref)
        for var in VARIANTS:
            say("order batch", var)
            batch = ctx.scan_batch(cos, tile={"variant": var})
            for co, r in zip(cos, batch):
                assert same(r, ctx.scan(co, tile={"variant": var}))
    finally:
        for co in cos:
            co.close()


@pytest.mark.parametrize("var", VARIANTS)
def test_strip_rounds_on_ordered_cohort(epi, plain_ctx, var):
    """A cohort past its pair budget (strips refilled per round) in stored
    order: identical to the full-table original-order scan."""
    g, p = epi.generate_hwe(35, 2000, 260, 900)
    say("order strips", var)
    Mp = (260 + 63) // 64 * 64
    own = epi.Context(0)
    try:
        own.set_snp_order(1)
        own.set_pair_budget(3 * 2 * Mp * 64 * 16)
        with own.cohort(g, p) as co:
            a = own.scan(co, tile={"variant": var})
    finally:
        own.close()
    with plain_ctx.cohort(g, p) as co:
        b = plain_ctx.scan(co, tile={"variant": var})
    assert same(a, b)


@pytest.mark.parametrize("var", VARIANTS)
def test_noalloc_on_ordered_cohort(epi, ord_ctx, plain_ctx, var):
    """epi_scan_noalloc of a cohort in stored order (its rows shuffled as
    epi_scan's are) equals the original-order scan."""
    g, p = epi.generate_hwe(36, 1800, 300, 800)
    say("order noalloc", var)
    with ord_ctx.cohort(g, p) as co:
        ord_ctx.scan(co, tile={"variant": var})  # builds the expanded planes
        a = ord_ctx.scan_noalloc(co, tile={"variant": var})
    with plain_ctx.cohort(g, p) as co:
        b = plain_ctx.scan(co, tile={"variant": var})
    assert same(a, b)


def test_bed_cohort_in_stored_order(epi, orc, ord_ctx, plain_ctx):
    """A .bed cohort (unpacked on the device) in stored order: scan and
    epi_count equal the sample-major cohort in original order."""
    N, M = 1500, 200
    geno, pheno = epi.generate_hwe(37, N, M, N // 2)
    rng = np.random.default_rng(37)
    geno[rng.random(geno.shape) < 0.02] = 3  # missing calls count as genotype 2
    bed = orc.bed_encode(geno)
    say("order bed")
    tr = np.array([[0, 1, 2], [3, 17, M - 1], [M - 3, M - 2, M - 1], [5, 100, 150]], np.uint32)
    with ord_ctx.cohort_bed(bed, pheno, N) as b:
        rb = ord_ctx.scan(b)
        cb, sb, _ = ord_ctx.count(b, tr)
    with plain_ctx.cohort(geno, pheno) as a:
        ra = plain_ctx.scan(a)
        ca, sa, _ = plain_ctx.count(a, tr)
    assert same(ra, rb)
    assert np.array_equal(ca, cb) and np.array_equal(sa, sb)


@pytest.mark.parametrize("var", VARIANTS)
def test_seed_of_every_triplet_with_ties(epi, orc, plain_ctx, var):
    """A cohort of 24 SNPs has 2024 triplets, so its threshold seed is the
    exact 64th best score of all of them; three copies of 8 SNPs make many
    triplets tie with it.  The top 64 (scores, ties in lexicographic order)
    equal the oracle's."""
    g8, p = epi.generate_hwe(38, 900, 8, 450)
    g = np.ascontiguousarray(np.concatenate([g8, g8, g8], 1))
    say("seed ties", var)
    with plain_ctx.cohort(g, p) as co:
        a = plain_ctx.scan(co, tile={"variant": var})
    _, top = orc.PackedCohort(g, p).scan(topk=64)
    assert len(a.top) == 64
    assert len(set(np.float32(t[0]).tobytes() for t in top)) < 64  # ties present
    for e, c in enumerate(a.top):
        s_, i, j, k = top[e]
        assert (c.i, c.j, c.k) == (i, j, k), e
        assert np.float32(c.score).tobytes() == np.float32(s_).tobytes(), e
