"""Pins the CPU oracle (oracle/oracle.c) before it is trusted as the checker.

Each check is chosen so that a plausible mistake in the restatement (a
dropped cell, a swapped class, a wrong padding mask, an off-by-one in the
ln(n!) table, a reordered K2 sum) fails it:
  * known answers of the reference's own programs (compiled from the
    reference sources into oracle/_ref and recorded in tests/golden/);
  * the same programs fed seeded HWE data whose cells reach n >= 748;
  * textbook per-sample counting in numpy and count invariants;
  * the ln(n!) table against lgamma, and its bits against a fixture;
  * brute-force rescans of tiny inputs.
"""
import hashlib
import math

import numpy as np
import pytest

RAND_MAX = 2147483647


def oracle_best(orc, A, B, topk=1):
    pc = orc.PackedCohort(A, B)
    _, top = pc.scan(topk=topk)
    return pc, top


# ---------------------------------------------------------------- known answers
def kat_cases(golden, max_work=6e10):
    out = []
    for c in golden("kat_reference.json")["cases"]:
        work = c["N"] * c["M"] * (c["M"] - 1) * (c["M"] - 2) / 6
        if work <= max_work:
            out.append(c)
    return out


def test_kat_reference_programs_agree(golden):
    for c in golden("kat_reference.json")["cases"]:
        bests = {tuple(r["best"]) for r in c["runs"]}
        scores = {r["score"] for r in c["runs"]}
        assert len(bests) == 1 and len(scores) == 1, c


def test_oracle_reproduces_reference_known_answers(orc, golden):
    cases = kat_cases(golden)
    assert len(cases) >= 9
    for c in cases:
        A, B = orc.ref_generate(c["N"], c["M"])
        assert int(B.sum()) == c["N"] // 2
        _, top = oracle_best(orc, A, B)
        score, i, j, k = top[0]
        assert [i, j, k] == c["runs"][0]["best"], c
        assert "%f" % score == c["runs"][0]["score"], c


def test_oracle_reproduces_reference_on_fed_hwe_data(orc, golden):
    vecs = golden("ref_feed_vectors.json")["vectors"]
    assert any(v["max_cell_plus1"] >= 748 for v in vecs), "Stirling branch not exercised"
    for v in vecs:
        A, B, _ = orc.hwe_generate(v["seed"], v["N"], v["M"], v["cases"], *v["maf"])
        _, top = oracle_best(orc, A, B)
        score, i, j, k = top[0]
        for r in v["runs"]:
            assert [i, j, k] == r["best"], (v, r)
            assert "%f" % score == r["score"], (v, r)


# ---------------------------------------------------------------- counting
def numpy_counts(A, B, i, j, k):
    g = np.where(A > 2, 2, A).astype(np.int64)
    g = np.where((A != 0) & (A != 1), 2, g)
    cell = 9 * g[:, i] + 3 * g[:, j] + g[:, k] + 27 * (B == 1)
    return np.bincount(cell, minlength=54).astype(np.uint32)


def random_cohort(rng, N, M, p_case=0.5, extra_values=False):
    A = rng.integers(0, 3, size=(N, M), dtype=np.uint8)
    if extra_values:
        mask = rng.random((N, M)) < 0.05
        A[mask] = rng.integers(3, 256, size=int(mask.sum()), dtype=np.uint8)
    B = (rng.random(N) < p_case).astype(np.uint8)
    if extra_values:
        B[rng.random(N) < 0.05] = 7  # anything but 1 is a control
    return A, B


@pytest.mark.parametrize("N,p_case,extra", [(1, 0.5, False), (31, 0.5, False), (64, 0.5, True),
                                            (65, 0.3, True), (100, 0.0, False), (100, 1.0, False),
                                            (333, 0.7, True)])
def test_counts_brute_packed_numpy_agree(orc, N, p_case, extra):
    rng = np.random.default_rng(N * 7 + int(p_case * 10))
    M = 9
    A, B = random_cohort(rng, N, M, p_case, extra)
    pc = orc.PackedCohort(A, B)
    n_case = int((B == 1).sum())
    for i in range(M):
        for j in range(i + 1, M):
            for k in range(j + 1, M):
                ref = numpy_counts(A, B, i, j, k)
                assert np.array_equal(orc.count_brute(A, B, i, j, k), ref)
                c = pc.counts(i, j, k)
                assert np.array_equal(c, ref), (i, j, k)
                assert c[:27].sum() == N - n_case and c[27:].sum() == n_case


def test_count_marginals(orc):
    rng = np.random.default_rng(5)
    A, B = random_cohort(rng, 500, 6)
    pc = orc.PackedCohort(A, B)
    c = pc.counts(0, 2, 5).astype(np.int64).reshape(2, 3, 3, 3)
    for cls, sel in ((0, B != 1), (1, B == 1)):
        for g in range(3):
            assert c[cls, g].sum() == (A[sel, 0] == g).sum()
            assert c[cls, :, g].sum() == (A[sel, 2] == g).sum()
            assert c[cls, :, :, g].sum() == (A[sel, 5] == g).sum()


# ---------------------------------------------------------------- ln(n!)
def test_log_table_against_lgamma(orc):
    T = orc.log_table(70000)
    n = np.arange(70001)
    lg = np.array([math.lgamma(x + 1) for x in n])
    err = np.abs(T.astype(np.float64) - lg)
    # float recurrence below 748: accumulated rounding only
    assert err[:748].max() < 2e-3
    # Stirling with the constant -0.0810: (n + 0.5) ln n - (n - 1) - 0.0810
    # differs from ln n! by 0.9190 - 0.5 ln(2 pi) - 1/(12 n) + ... < 2e-4 for
    # n >= 748; rounding to float adds at most half a unit in the last place.
    half_ulp = np.spacing(T[748:]).astype(np.float64) / 2
    assert (err[748:] <= half_ulp + 2e-4).all()
    # the first Stirling entries: T[748] is not the recurrence value
    assert T[748] != np.float32(np.float64(T[747]) + math.log(748))
    assert T[0] == 0 and T[1] == 0 and T[2] == np.float32(math.log(2))


def test_log_table_f64_against_lgamma(orc):
    """FP64 twin (used by the top-64 rescoring): exact log sums below 748,
    the Stirling form's truncation error only (no float rounding) above."""
    T = orc.log_table_f64(70000)
    lg = np.array([math.lgamma(x + 1) for x in range(70001)])
    err = np.abs(T - lg)
    assert err[:748].max() < 1e-9
    # (n + 1/2) ln n - (n - 1) - 0.0810 - ln n! = 0.9190 - ln(2 pi)/2 - 1/(12 n)
    # + 1/(360 n^3) - ... (Stirling series)
    n = np.arange(748, 70001, dtype=np.float64)
    series = 0.919 - 0.5 * math.log(2 * math.pi) - 1 / (12 * n) + 1 / (360 * n ** 3)
    assert np.abs((T[748:] - lg[748:]) - series).max() < 1e-8
    T32 = orc.log_table(70000)
    assert np.array_equal(T32[748:], T[748:].astype(np.float32))


def test_log_table_fixture_bits(orc, golden):
    fx = golden("log_table.json")
    for n, digest in fx["sha256_float32_le"].items():
        T = orc.log_table(int(n) - 1)
        assert hashlib.sha256(T.tobytes()).hexdigest() == digest, n
    T = orc.log_table(20000)
    for n, hexv in fx["samples"].items():
        assert float(T[int(n)]).hex() == hexv, n


def test_k2_against_lgamma_closed_form(orc):
    rng = np.random.default_rng(9)
    A, B = random_cohort(rng, 3000, 5, 0.4)
    pc = orc.PackedCohort(A, B)
    c = pc.counts(0, 1, 4).astype(np.int64)
    n0, n1 = c[:27], c[27:]
    exact = sum(math.lgamma(a + b + 2) - math.lgamma(a + 1) - math.lgamma(b + 1)
                for a, b in zip(n0, n1))
    s = pc.score(c)
    # 81 table lookups, each within 8e-3 of lgamma at these sizes
    assert abs(s - exact) < 81 * 8e-3
    assert s > 0
    # FP64 twin of the same approximation
    assert abs(pc.score_f64(c) - s) < 1e-2


def test_scan_equals_brute_force_argmin(orc):
    rng = np.random.default_rng(3)
    A, B = random_cohort(rng, 150, 11, 0.5, True)
    pc = orc.PackedCohort(A, B)
    allsc = []
    M = A.shape[1]
    for i in range(M):
        for j in range(i + 1, M):
            for k in range(j + 1, M):
                allsc.append((pc.score(orc.count_brute(A, B, i, j, k)), i, j, k))
    allsc.sort()
    n, top = pc.scan(topk=20)
    assert n == len(allsc)
    assert top == [(np.float32(s).item(), i, j, k) for s, i, j, k in allsc[:20]]
    scores = pc.all_scores()
    for s, i, j, k in allsc:
        assert scores[orc.rank3(M, i, j, k)] == np.float32(s)


def test_rank3_enumeration(orc):
    M = 13
    r = 0
    for i in range(M):
        for j in range(i + 1, M):
            for k in range(j + 1, M):
                assert orc.rank3(M, i, j, k) == r
                r += 1


# ---------------------------------------------------------------- generators
def test_hwe_generator_statistics(orc):
    N, M = 20000, 40
    A, B, maf = orc.hwe_generate(7, N, M, 6000)
    assert int(B.sum()) == 6000 and set(np.unique(B)) <= {0, 1}
    assert ((maf >= 0.05) & (maf <= 0.5)).all()
    for j in range(M):
        p = maf[j]
        exp = np.array([(1 - p) ** 2, 2 * p * (1 - p), p * p]) * N
        obs = np.bincount(A[:, j], minlength=3)
        chi2 = (((obs - exp) ** 2) / exp).sum()
        assert chi2 < 25, (j, obs, exp)


def test_planted_generator_signal(orc):
    N, M = 2048, 24
    A, B, pl = orc.planted_generate(21, N, M, N // 2)
    assert int(B.sum()) == N // 2
    odd = ((A[:, pl[0]].astype(int) + A[:, pl[1]] + A[:, pl[2]]) & 1) == 1
    # cases are enriched for odd genotype sums (penetrance 0.5 vs 0.05)
    assert odd[B == 1].mean() > 0.6 and odd[B == 0].mean() < 0.5
    _, top = oracle_best(orc, A, B)
    assert (top[0][1], top[0][2], top[0][3]) == pl
