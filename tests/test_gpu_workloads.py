"""The benchmark workloads themselves, scanned through the C ABI and checked
against the reference program's results on the same cohorts
(tests/golden/workload_reference.json, written by make_workload_golden.py
from CPU/v4_popc_avx512.cpp built here) and against the oracle:

  planted 2048 x 4096     best triplet = the reference's = the planted one
  4096 x 16384            FP4 MFMA and VALU scans: best triplet and score
  64-cohort batch         one epi_scan_batch of all 64 cohorts
  64-request stream       every request through the server

Each result's printed score must equal the reference's ("%f", as
GPU/v4.cpp:619 prints it) and its triplet must be the reference's, or a
triplet whose FP32 score ties it bit for bit (the reference keeps the first
minimum of its own loop order, CPU/v4_popc_avx512.cpp:4583-4630; this scan
keeps the lexicographically smallest).  Every returned top-64 candidate is
recounted by the oracle: 54 counts, FP32 and FP64 scores, and each class's
27 cells summing to the class size."""
import os
import sys

import numpy as np
import pytest

from conftest import PKG

sys.path.insert(0, PKG)
import workloads  # noqa: E402

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


@pytest.fixture(scope="module")
def ref(golden):
    return golden("workload_reference.json")["groups"]


def check_result(ctx, co, res, entry, spec):
    """Printed score and triplet against the reference program's."""
    assert "%f" % res.score == entry["score"], (spec["name"], res.best, res.score, entry)
    if list(res.best) != entry["best"]:
        # a tie: the reference's triplet scores the same FP32 value here
        _, s32, _ = ctx.count(co, [entry["best"]])
        assert np.float32(s32[0]).tobytes() == np.float32(res.score).tobytes(), \
            (spec["name"], res.best, entry["best"], float(s32[0]), res.score)
    n = spec["n_snps"]
    assert res.triplets == n * (n - 1) * (n - 2) // 6
    assert (res.n_controls, res.n_cases) == (spec["n_samples"] - spec["n_cases"], spec["n_cases"])


def recount(orc, A, B, res):
    """Every top candidate recounted on the CPU."""
    pc = orc.PackedCohort(A, B)
    assert len(res.top) == min(64, res.triplets)
    prev = -np.inf
    for c in res.top:
        want = pc.counts(c.i, c.j, c.k)
        assert np.array_equal(c.counts, want), (c.i, c.j, c.k)
        assert int(want[:27].sum()) == res.n_controls and int(want[27:].sum()) == res.n_cases
        assert np.float32(c.score).tobytes() == np.float32(pc.score(want)).tobytes()
        assert c.score_fp64 == pc.score_f64(want)
        assert c.score >= prev  # ascending
        prev = c.score
    assert res.max_abs_err_fp32 == max(abs(c.score - c.score_fp64) for c in res.top)


def test_planted_2048x4096(epi, orc, ctx, ref):
    spec = workloads.planted_2048x4096()
    entry = ref["planted_2048x4096"][0]
    assert entry["spec"] == spec
    say("planted", spec)
    A, B, planted = workloads.materialize(spec, epi)
    with ctx.cohort(A, B) as co:
        res = ctx.scan(co)
        check_result(ctx, co, res, entry, spec)
    assert tuple(res.best) == planted == tuple(entry["planted"])
    recount(orc, A, B, res)


@pytest.mark.parametrize("variant", [3, 1])
def test_scan_4096x16384(epi, orc, ctx, ref, variant):
    """The headline scan with the FP4 MFMA kernel (the default tile) and the
    VALU popcount kernel."""
    spec = workloads.scan_4096x16384()
    entry = ref["scan_4096x16384"][0]
    assert entry["spec"] == spec
    say("scan 4096x16384 variant", variant)
    A, B, _ = workloads.materialize(spec, epi)
    with ctx.cohort(A, B) as co:
        res = ctx.scan(co, tile=None if variant == 3 else {"variant": variant})
        say("  kernel ms", res.kernel_ms, res.best, res.score)
        check_result(ctx, co, res, entry, spec)
    recount(orc, A, B, res)


def test_batch_64(epi, orc, ctx, ref):
    """One epi_scan_batch over the 64 cohorts of uneven shape."""
    specs = workloads.batch_64()
    entries = ref["batch_64"]
    assert [e["spec"] for e in entries] == specs
    data = [workloads.materialize(s, epi) for s in specs]
    cos = [ctx.cohort(A, B) for A, B, _ in data]
    try:
        say("batch of", len(cos))
        res = ctx.scan_batch(cos)
        for b, (r, e, s) in enumerate(zip(res, entries, specs)):
            check_result(ctx, cos[b], r, e, s)
    finally:
        for co in cos:
            co.close()
    for (A, B, _), r in zip(data, res):
        recount(orc, A, B, r)


def test_stream_64(epi, orc, ctx, ref):
    """The 64 requests of the serving stream, all submitted to one server
    and polled in submission order."""
    specs, _ = workloads.stream(64)
    entries = ref["stream_64"]
    assert [e["spec"] for e in entries] == specs
    data = [workloads.materialize(s, epi) for s in specs]
    res = []
    with ctx.server() as srv:
        ids = [srv.submit(A, B) for A, B, _ in data]
        say("submitted", len(ids))
        for rid in ids:
            r = srv.poll(rid, 60000)
            assert r is not None, rid
            res.append(r)
        st = srv.stats()
    assert st["completed"] >= len(ids)
    for (A, B, _), r, e, s in zip(data, res, entries, specs):
        with ctx.cohort(A, B) as co:
            check_result(ctx, co, r, e, s)
        recount(orc, A, B, r)
