"""epi_scan_noalloc (a static workspace per GPU, no device allocation)
against epi_scan, through the C ABI: identical results on two calls in a
row, the second with a smaller cohort; no change in free device memory
over a call on a fresh context; EPI_EINVAL beyond the workspace."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


def same(a, b):
    return ((a.i, a.j, a.k, a.score, a.score_fp64, a.n_controls, a.n_cases, a.triplets) ==
            (b.i, b.j, b.k, b.score, b.score_fp64, b.n_controls, b.n_cases, b.triplets) and
            [(c.i, c.j, c.k, c.score, c.score_fp64, tuple(c.counts)) for c in a.top] ==
            [(c.i, c.j, c.k, c.score, c.score_fp64, tuple(c.counts)) for c in b.top])


def free_bytes():
    hip = ctypes.CDLL("libamdhip64.so")
    free, total = ctypes.c_size_t(), ctypes.c_size_t()
    assert hip.hipMemGetInfo(ctypes.byref(free), ctypes.byref(total)) == 0
    return free.value


@pytest.mark.parametrize("tile", [None, {"variant": 1}])
def test_noalloc_equals_scan(epi, ctx, tile):
    """Large cohort, then a smaller one; the default tile (FP4 MFMA once the
    cohort holds its planes, else VALU) and the VALU tile."""
    for seed, N, M in ((61, 6000, 700), (62, 900, 150)):
        say("noalloc", tile, N, M)
        g, p = epi.generate_hwe(seed, N, M, N // 2)
        with ctx.cohort(g, p) as co:
            first = ctx.scan_noalloc(co, tile)  # before any plain scan of it
            plain = ctx.scan(co, tile)
            again = ctx.scan_noalloc(co, tile)
        assert same(plain, first) and same(plain, again), (seed, tile)


def test_noalloc_allocates_nothing(epi, ctx):
    """On a fresh context, a noalloc scan of a small cohort, then one of a
    larger cohort (more rows, a longer ln(n!) table: a path with buffers
    that grow would allocate here): free device memory is unchanged over
    the second call."""
    small = epi.generate_hwe(63, 1000, 100, 500)
    large = epi.generate_hwe(64, 4000, 600, 2000)
    refs = []
    for g, p in (small, large):
        with ctx.cohort(g, p) as co:
            refs.append(ctx.scan(co, {"variant": 1}))
    with epi.Context(0) as c2:
        with c2.cohort(*small) as cs, c2.cohort(*large) as cl:
            r0 = c2.scan_noalloc(cs)
            before = free_bytes()
            r1 = c2.scan_noalloc(cl)
            after = free_bytes()
    say("free before", before, "after", after)
    assert same(refs[0], r0) and same(refs[1], r1)
    assert after == before


def test_noalloc_rejects_beyond_workspace(epi, ctx):
    """ln(n!) tables of 2^20 samples alone take 12 MiB: with the slots the
    scan needs more than the 16 MiB workspace."""
    N, M = 1 << 20, 3
    g = np.zeros((N, M), np.uint8)
    p = np.zeros(N, np.uint8)
    p[::2] = 1
    with ctx.cohort(g, p) as co:
        with pytest.raises(epi.EpiError) as e:
            ctx.scan_noalloc(co)
        assert e.value.code == epi.EPI_EINVAL
        r = ctx.scan(co)  # the allocating scan still works
    assert r.triplets == 1
