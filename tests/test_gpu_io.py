"""Cohorts built from PLINK 1 .bed bodies (unpacked on the device) scan
exactly like the same cohort given as sample-major bytes, and like the
oracle's scan of the decoded genotypes."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


def same(a, b):
    return ((a.i, a.j, a.k) == (b.i, b.j, b.k) and
            [(c.i, c.j, c.k, c.score, c.score_fp64, tuple(c.counts)) for c in a.top] ==
            [(c.i, c.j, c.k, c.score, c.score_fp64, tuple(c.counts)) for c in b.top])


def oracle_top(orc, geno, pheno, topk=64):
    _, top = orc.PackedCohort(geno, pheno).scan(topk=topk)
    return [(i, j, k, np.float32(s)) for s, i, j, k in top]


@pytest.mark.parametrize("N,M,seed", [(777, 150, 1), (1024, 96, 2), (4099, 70, 3)])
def test_bed_cohort_equals_sample_major(epi, orc, ctx, N, M, seed):
    say(f"bed cohort N={N} M={M} seed={seed}")
    geno, pheno = epi.generate_hwe(seed, N, M, N // 2)
    rng = np.random.default_rng(seed)
    geno[rng.random(geno.shape) < 0.02] = 3  # missing calls count as genotype 2
    bed = orc.bed_encode(geno)
    with ctx.cohort(geno, pheno) as a, ctx.cohort_bed(bed, pheno, N) as b:
        ra, rb = ctx.scan(a), ctx.scan(b)
        assert b.n_samples == N and b.n_snps == M
        tr = np.array([[0, 1, 2], [3, 17, M - 1], [M - 3, M - 2, M - 1]], np.uint32)
        ca, sa, _ = ctx.count(a, tr)
        cb, sb, _ = ctx.count(b, tr)
    assert same(ra, rb)
    assert np.array_equal(ca, cb) and np.array_equal(sa, sb)
    top = [(c.i, c.j, c.k, np.float32(c.score)) for c in rb.top]
    assert top == oracle_top(orc, geno, pheno)


def test_bed_missing_phenotype_left_out(epi, orc, ctx):
    N, M = 900, 80
    geno, pheno = epi.generate_hwe(9, N, M, 450)
    pheno = pheno.copy()
    rng = np.random.default_rng(9)
    pheno[rng.random(N) < 0.1] = 255
    keep = pheno <= 1
    bed = orc.bed_encode(geno)
    with ctx.cohort_bed(bed, pheno, N) as b, ctx.cohort(geno[keep], pheno[keep]) as a:
        rb, ra = ctx.scan(b), ctx.scan(a)
    assert rb.n_samples == int(keep.sum())
    assert (rb.n_controls, rb.n_cases) == (int((pheno == 0).sum()), int((pheno == 1).sum()))
    assert same(ra, rb)


def test_bed_all_phenotypes_missing_is_rejected(epi, orc, ctx):
    geno, _ = epi.generate_hwe(4, 40, 10, 20)
    with pytest.raises(epi.EpiError) as e:
        ctx.cohort_bed(orc.bed_encode(geno), np.full(40, 255, np.uint8), 40)
    assert e.value.code == epi.EPI_EINVAL


def test_plink_fileset_end_to_end(epi, orc, ctx, tmp_path):
    N, M = 2050, 120
    geno, pheno, planted = epi.generate_planted(21, N, M, N // 2)
    pre = str(tmp_path / "cohort")
    epi.write_plink(pre, geno, pheno)
    bed, ph = epi.read_plink(pre)
    with ctx.cohort_bed(bed, ph, N) as co:
        r = ctx.scan(co)
    say("planted", planted, "found", r.best)
    top = [(c.i, c.j, c.k, np.float32(c.score)) for c in r.top]
    assert top == oracle_top(orc, geno, pheno)
