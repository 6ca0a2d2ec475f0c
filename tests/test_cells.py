"""The product's exact 27-cell reconstruction (episdet_cells.h: 8 popcounted
cells with genotypes in {0,1} plus pair/single margins), built for the host,
against textbook per-sample counting."""
import ctypes
import os

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SO = os.path.join(ROOT, "tests", "native", "libcells_host.so")


@pytest.fixture(scope="module")
def derive():
    if not os.path.exists(SO):
        pytest.skip("tests/native/libcells_host.so not built (run __graft_entry__.build())")
    L = ctypes.CDLL(SO)
    P = ctypes.POINTER(ctypes.c_uint32)
    L.cells_derive.argtypes = [P, P, P, P, P, P, P, ctypes.c_uint32, P]

    def call(t8, ij, ik, jk, si, sj, sk, n):
        arrs = [np.ascontiguousarray(a, np.uint32) for a in (t8, ij, ik, jk, si, sj, sk)]
        x = np.zeros(27, np.uint32)
        L.cells_derive(*[a.ctypes.data_as(P) for a in arrs], n, x.ctypes.data_as(P))
        return x
    return call


def margins(g, i, j, k):
    def pair(a, b):
        return [int(((g[:, a] == x) & (g[:, b] == y)).sum()) for x in (0, 1) for y in (0, 1)]
    t8 = [int(((g[:, i] == a) & (g[:, j] == b) & (g[:, k] == c)).sum())
          for a in (0, 1) for b in (0, 1) for c in (0, 1)]
    single = lambda s: [int((g[:, s] == 0).sum()), int((g[:, s] == 1).sum())]
    return t8, pair(i, j), pair(i, k), pair(j, k), single(i), single(j), single(k)


@pytest.mark.parametrize("n,maf", [(0, 0.3), (1, 0.3), (7, 0.5), (100, 0.05), (1000, 0.5), (5000, 0.01)])
def test_derive27_matches_per_sample_counting(derive, n, maf):
    rng = np.random.default_rng(n + 1)
    p = np.array([maf, 0.5, 0.2, 0.9])
    g = np.stack([rng.choice(3, size=n, p=[(1 - q) ** 2, 2 * q * (1 - q), q * q]) for q in p], 1)
    for i, j, k in [(0, 1, 2), (0, 1, 3), (1, 2, 3), (0, 2, 3)]:
        x = derive(*margins(g, i, j, k), n)
        cell = 9 * g[:, i] + 3 * g[:, j] + g[:, k]
        assert np.array_equal(x, np.bincount(cell, minlength=27)), (i, j, k)


# ---------------------------------------------------------------------------
# Stored SNP order (episdet_order.h): a cohort may store its SNPs in another
# order than the original one; a stored triplet p < q < r then has original
# indices in one of six orders, and the reference's cells (axes by increasing
# original index) must come out of the stored-axis inputs.
# ---------------------------------------------------------------------------
@pytest.fixture(scope="module")
def order_lib():
    if not os.path.exists(SO):
        pytest.skip("tests/native/libcells_host.so not built (run __graft_entry__.build())")
    L = ctypes.CDLL(SO)
    P = ctypes.POINTER(ctypes.c_uint32)
    L.order_axis.restype = ctypes.c_uint32
    L.order_axis.argtypes = [ctypes.c_uint32] * 3
    L.order_stored_cell.restype = ctypes.c_int
    L.order_stored_cell.argtypes = [ctypes.c_int, ctypes.c_int]
    L.order_ref_derive.argtypes = [ctypes.c_uint32, P, P, P, P, P, P, P, ctypes.c_uint32, P]
    return L


def test_axis_order_is_the_rank_code(order_lib):
    import itertools
    seen = set()
    for perm in itertools.permutations([5, 17, 40]):
        o = order_lib.order_axis(*perm)
        ranks = [sorted(perm).index(v) for v in perm]
        assert o == 2 * ranks[0] + (ranks[1] > ranks[2]), perm
        seen.add(o)
    assert seen == set(range(6))
    assert order_lib.order_axis(1, 2, 3) == 0


@pytest.mark.parametrize("n,seed", [(0, 1), (1, 2), (63, 3), (64, 4), (700, 5), (4097, 6)])
def test_reference_axes_from_stored_axes(order_lib, n, seed):
    """Cells of a stored triplet rebuilt in the reference's axis order equal
    per-sample counting with the SNPs taken by increasing original index, for
    all six orders; stored_cell maps the stored cells to the same table."""
    import itertools
    rng = np.random.default_rng(seed)
    q = rng.uniform(0.05, 0.5, 3)
    g = np.stack([rng.choice(3, size=n, p=[(1 - a) ** 2, 2 * a * (1 - a), a * a]) for a in q], 1)
    P = ctypes.POINTER(ctypes.c_uint32)
    stored = np.bincount(9 * g[:, 0] + 3 * g[:, 1] + g[:, 2], minlength=27)
    inputs = [np.ascontiguousarray(a, np.uint32) for a in margins(g, 0, 1, 2)]
    for orig in itertools.permutations([3, 11, 29]):
        o = order_lib.order_axis(*orig)
        ax = np.argsort(orig)  # stored axes by increasing original index
        ref = np.bincount(9 * g[:, ax[0]] + 3 * g[:, ax[1]] + g[:, ax[2]], minlength=27)
        x = np.zeros(27, np.uint32)
        order_lib.order_ref_derive(o, *[a.ctypes.data_as(P) for a in inputs], n,
                                   x.ctypes.data_as(P))
        assert np.array_equal(x, ref), (orig, o)
        sc = [order_lib.order_stored_cell(o, m) for m in range(27)]
        assert np.array_equal(stored[sc], ref), (orig, o)


# ---------------------------------------------------------------------------
# The triplet sample of a cohort's threshold seed (episdet_sample.h): its 64th
# best score bounds the cohort's 64th best only if the triplets are distinct.
# ---------------------------------------------------------------------------
@pytest.fixture(scope="module")
def sample_lib():
    if not os.path.exists(SO):
        pytest.skip("tests/native/libcells_host.so not built (run __graft_entry__.build())")
    L = ctypes.CDLL(SO)
    L.sample_triplet.argtypes = [ctypes.c_uint64, ctypes.c_uint32, ctypes.POINTER(ctypes.c_uint32)]
    L.sample_rank.restype = ctypes.c_uint64
    L.sample_rank.argtypes = [ctypes.c_uint64, ctypes.c_uint32, ctypes.c_uint32]
    return L


def _triplet(L, r, M):
    out = (ctypes.c_uint32 * 3)()
    L.sample_triplet(r, M, out)
    return tuple(out)


def test_colex_unranking_enumerates_every_triplet_once(sample_lib):
    import itertools
    from math import comb
    for M in (3, 4, 5, 7, 16, 33, 64, 65):
        got = [_triplet(sample_lib, r, M) for r in range(comb(M, 3))]
        assert sorted(got) == list(itertools.combinations(range(M), 3)), M
        # colex order: the last index leads
        assert got == sorted(got, key=lambda t: (t[2], t[1], t[0])), M


def test_colex_unranking_inverts_the_rank_at_scale(sample_lib):
    from math import comb
    rng = np.random.default_rng(9)
    for M in (4096, 100000, (1 << 21) - 1):
        total = comb(M, 3)
        for r in [0, 1, total - 1] + [int(x) for x in rng.integers(0, total, 200)]:
            i, j, k = _triplet(sample_lib, r, M)
            assert 0 <= i < j < k < M
            assert comb(k, 3) + comb(j, 2) + i == r


def test_sample_ranks_are_distinct(sample_lib):
    from math import comb
    for M, N in ((10, 100), (40, 1000), (41, 1000), (300, 2048), (4096, 16384),
                 ((1 << 21) - 1, 77)):
        total = comb(M, 3)
        n = min(total, 8192)
        ranks = [sample_lib.sample_rank(total, N, t) for t in range(n)]
        assert all(0 <= r < total for r in ranks)
        assert len(set(ranks)) == n, (M, N)
        if total <= 8192:
            assert ranks == list(range(total))
