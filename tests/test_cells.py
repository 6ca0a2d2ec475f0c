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
