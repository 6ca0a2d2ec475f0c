"""Cohorts past their pair budget (epi_set_pair_budget): the pair margins
live in a few 64-column strip slots that each scan round refills, and the
scan runs in rounds over groups of column blocks.  Every count and score
must equal the oracle's and the full-table cohort's bit for bit, for the
VALU, int8 and FP4 kernels."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


def same(a, b):
    return ((a.i, a.j, a.k) == (b.i, b.j, b.k) and
            [(c.i, c.j, c.k, c.score, c.score_fp64, tuple(c.counts)) for c in a.top] ==
            [(c.i, c.j, c.k, c.score, c.score_fp64, tuple(c.counts)) for c in b.top])


def strip_bytes(M):
    Mp = (M + 63) // 64 * 64
    return 2 * Mp * 64 * 16  # one 64-column strip, both classes


@pytest.fixture
def budget_ctx(epi):
    """A context of its own, so that the session context keeps its budget."""
    c = epi.Context(0)
    yield c
    c.close()


@pytest.mark.parametrize("M,N,slots", [(300, 3000, 2), (300, 3000, 5), (200, 777, 3),
                                       (65, 500, 2)])
def test_strip_rounds_equal_full_table(epi, orc, ctx, budget_ctx, M, N, slots):
    """Budgets of 2..5 strips (groups of 1 or 2 column blocks, a ragged last
    group): top-64 counts and FP32/FP64 scores equal the full-table scan and
    the oracle for all three kernels."""
    say("strips", M, N, slots)
    g, p = epi.generate_hwe(M * 7 + slots, N, M, N // 2)
    with ctx.cohort(g, p) as full_co:
        full = {v: ctx.scan(full_co, {"variant": v}) for v in (1, 2, 3)}
        full_bytes = full_co.device_bytes
    budget_ctx.set_pair_budget(slots * strip_bytes(M))
    with budget_ctx.cohort(g, p) as co:
        assert co.device_bytes < full_bytes
        for v in (1, 2, 3, 1):
            r = budget_ctx.scan(co, {"variant": v})
            assert same(full[v], r), (v, r.best, full[v].best)
    pc = orc.PackedCohort(g, p)
    _, top = pc.scan(topk=64)
    for c, (s, i, j, k) in zip(full[3].top, top):
        assert (c.i, c.j, c.k) == (i, j, k)
        assert np.float32(c.score).tobytes() == np.float32(s).tobytes()
        assert np.array_equal(c.counts, pc.counts(i, j, k))


@pytest.mark.parametrize("variant", [1, 3])
def test_strip_rounds_every_triplet(epi, orc, budget_ctx, variant):
    """Dump mode through strip rounds: every triplet's 54 counts and FP32
    score equal the oracle's (two strip slots, three column blocks)."""
    M, N = 150, 900
    g, p = epi.generate_hwe(41, N, M, 400)
    budget_ctx.set_pair_budget(2 * strip_bytes(M))
    with budget_ctx.cohort(g, p) as co:
        scores, counts = budget_ctx.scan_dump(co, tile={"variant": variant}, with_counts=True)
    pc = orc.PackedCohort(g, p)
    ref = pc.all_scores()
    assert np.array_equal(scores.view(np.uint32), ref.view(np.uint32))
    rng = np.random.default_rng(0)
    ranks = rng.choice(len(ref), 400, replace=False)
    trip = [(i, j, k) for i in range(M) for j in range(i + 1, M) for k in range(j + 1, M)]
    for r in ranks:
        assert np.array_equal(counts[r], pc.counts(*trip[r])), trip[r]


def test_strip_cohort_refused_where_every_strip_is_read(epi, budget_ctx):
    """One batch launch, the no-allocation scan and the server read any
    strip: a cohort past its budget is refused with EPI_EINVAL, and the
    budget's default comes back with 0."""
    g, p = epi.generate_hwe(5, 600, 200, 300)
    budget_ctx.set_pair_budget(2 * strip_bytes(200))
    with budget_ctx.cohort(g, p) as co:
        for call in (lambda: budget_ctx.scan_batch([co]), lambda: budget_ctx.scan_noalloc(co)):
            with pytest.raises(epi.EpiError) as e:
                call()
            assert e.value.code == epi.EPI_EINVAL
        with budget_ctx.server() as srv:
            with pytest.raises(epi.EpiError) as e:
                srv.submit_cohort(co)
            assert e.value.code == epi.EPI_EINVAL
            # submitted data makes a cohort of the server's own, every strip resident
            rid = srv.submit(g, p)
            r = srv.poll(rid, 60000)
        assert same(budget_ctx.scan(co), r)
    budget_ctx.set_pair_budget(0)
    with budget_ctx.cohort(g, p) as co:
        assert len(budget_ctx.scan_batch([co])) == 1
