"""Batched scans (graph-captured, replayed), the continuous-batching server
and the tile autotuner, all through the C ABI.  Every result must equal a
plain scan of the same cohort bit for bit."""
import time

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


def same(a, b):
    return ((a.i, a.j, a.k) == (b.i, b.j, b.k) and
            [(c.i, c.j, c.k, c.score, c.score_fp64) for c in a.top] ==
            [(c.i, c.j, c.k, c.score, c.score_fp64) for c in b.top])


def make_cohorts(epi, ctx, n, seed, m_range=(20, 300), n_range=(100, 4000)):
    rng = np.random.default_rng(seed)
    out = []
    for b in range(n):
        M = int(rng.integers(*m_range))
        N = int(rng.integers(*n_range))
        g, p = epi.generate_hwe(seed * 100 + b, N, M, int(round(N * rng.uniform(0.3, 0.7))))
        out.append((g, p))
    return out


def test_batch_graph_replay(epi, ctx):
    """Same batch twice (the second call replays the captured graph), then a
    batch of other cohorts with the same shape, then a different shape."""
    data = make_cohorts(epi, ctx, 12, 5)
    cos = [ctx.cohort(g, p) for g, p in data]
    try:
        single = [ctx.scan(c) for c in cos]
        for rep in range(3):
            batch = ctx.scan_batch(cos)
            assert all(same(a, b) for a, b in zip(single, batch)), rep
        rev = ctx.scan_batch(cos[::-1])
        assert all(same(a, b) for a, b in zip(single[::-1], rev))
        part = ctx.scan_batch(cos[3:7])
        assert all(same(a, b) for a, b in zip(single[3:7], part))
    finally:
        for c in cos:
            c.close()


def test_batch_repeated_cohort(epi, ctx):
    """A cohort may appear more than once in a batch: its cached row table is
    shared, each entry gets its own running list, every result is exact."""
    data = make_cohorts(epi, ctx, 3, 21)
    cos = [ctx.cohort(g, p) for g, p in data]
    try:
        single = [ctx.scan(c) for c in cos]
        order = [0, 1, 0, 2, 1, 0]
        batch = ctx.scan_batch([cos[b] for b in order])
        assert all(same(single[b], r) for b, r in zip(order, batch))
    finally:
        for c in cos:
            c.close()


def test_server_results_equal_plain_scans(epi, ctx):
    data = make_cohorts(epi, ctx, 16, 9, (10, 400), (64, 6000))
    plain = []
    for g, p in data:
        with ctx.cohort(g, p) as co:
            plain.append(ctx.scan(co))
    with ctx.server(max_active=6, round_ms=2.0) as srv:
        ids = [srv.submit(g, p) for g, p in data]
        got = [srv.poll(i, 120000) for i in ids]
        st = srv.stats()
    say(st)
    assert st["completed"] == len(data) and st["failed"] == 0
    assert st["rounds"] > 1 and st["bytes_in_use"] == 0
    for a, b in zip(plain, got):
        assert b is not None and same(a, b)
        assert b.latency_ms >= b.queue_ms >= 0


def test_server_late_arrivals_join_running_work(epi, ctx):
    """A large request is running when small ones arrive; the small ones
    complete before it (continuous batching) and all results are exact."""
    g, p = epi.generate_hwe(71, 16384, 2048, 8192)
    small = make_cohorts(epi, ctx, 4, 13, (20, 60), (200, 800))
    with ctx.cohort(g, p) as co:
        big_plain = ctx.scan(co)
    say("big plain scan", big_plain.scan_ms, "ms")
    with ctx.server(max_active=8, round_ms=5.0) as srv:
        t_big = time.perf_counter()
        big = srv.submit(g, p)
        time.sleep(0.05)
        t_small = time.perf_counter()
        sids = [srv.submit(a, b) for a, b in small]
        sres = [srv.poll(i, 120000) for i in sids]
        bres = srv.poll(big, 300000)
    assert same(big_plain, bres)
    done_big = t_big + bres.latency_ms / 1e3
    for r in sres:
        assert t_small + r.latency_ms / 1e3 < done_big, (r.latency_ms, bres.latency_ms)
    for (a, b), r in zip(small, sres):
        with ctx.cohort(a, b) as co:
            assert same(ctx.scan(co), r)


def test_server_memory_admission_and_cancel(epi, ctx):
    data = make_cohorts(epi, ctx, 6, 17, (200, 300), (1000, 2000))
    # budget for about one cohort at a time: requests are admitted one by one
    with ctx.cohort(*data[0]) as co:
        one = co.device_bytes
    with ctx.server(memory_budget=int(one * 1.6), max_active=8, round_ms=2.0) as srv:
        ids = [srv.submit(g, p) for g, p in data]
        srv.cancel(ids[-1])
        res = []
        for i in ids[:-1]:
            res.append(srv.poll(i, 120000))
        with pytest.raises(epi.EpiError) as e:
            srv.poll(ids[-1], 120000)
        assert e.value.code == epi.EPI_ECANCELED
        st = srv.stats()
        # a cohort larger than the whole budget is refused at submission
        g, p = epi.generate_hwe(3, 4000, 2000, 2000)
        with pytest.raises(epi.EpiError) as e:
            srv.submit(g, p)
        assert e.value.code == epi.EPI_ENOMEM
    assert st["completed"] == 5 and st["cancelled"] == 1
    for (g, p), r in zip(data, res):
        with ctx.cohort(g, p) as co:
            assert same(ctx.scan(co), r)


def test_server_cancel_while_running_and_tight_memory(epi, ctx):
    """Requests cancelled while their rounds run and a budget that admits
    only a few cohorts at a time: device blocks are released and reused by
    later cohorts while other rounds are on the device, and every request
    that is not cancelled completes with its plain-scan result."""
    data = make_cohorts(epi, ctx, 14, 23, (60, 400), (200, 6000))
    plain, one = [], 0
    for g, p in data:
        with ctx.cohort(g, p) as co:
            plain.append(ctx.scan(co))
            one = max(one, co.device_bytes)
    # the largest cohort fits (its expansion included); a few run at a time
    with ctx.server(memory_budget=int(one * 3), max_active=8, round_ms=1.0) as srv:
        ids = [srv.submit(g, p) for g, p in data]
        time.sleep(0.02)
        dropped = set(range(1, len(ids), 4))
        for b in dropped:
            srv.cancel(ids[b])
        got = {}
        for b, rid in enumerate(ids):
            try:
                got[b] = srv.poll(rid, 120000)
            except epi.EpiError as e:
                assert b in dropped and e.code == epi.EPI_ECANCELED, (b, e.code)
                got[b] = None
        st = srv.stats()
    say(st)
    assert st["failed"] == 0 and st["completed"] + st["cancelled"] == len(ids)
    for b, r in got.items():
        if r is None:
            assert b in dropped
        else:
            assert same(plain[b], r), b


def test_server_poll_unknown_and_timeout(epi, ctx):
    with ctx.server() as srv:
        with pytest.raises(epi.EpiError) as e:
            srv.poll(12345, 0)
        assert e.value.code == epi.EPI_ENOENT
        g, p = epi.generate_hwe(1, 4096, 800, 2048)
        rid = srv.submit(g, p)
        r = srv.poll(rid, 0)  # almost certainly not finished yet
        if r is None:
            r = srv.poll(rid, 120000)
        assert r is not None and r.n_snps == 800


def test_server_two_pollers_one_request(epi, ctx):
    """Two threads poll the same request: one gets the result, the other
    EPI_EBUSY while the first waits and EPI_ENOENT once it is collected."""
    import threading
    g, p = epi.generate_hwe(31, 4096, 700, 2048)
    with ctx.cohort(g, p) as co:
        plain = ctx.scan(co)
    with ctx.server() as srv:
        for rep in range(3):
            rid = srv.submit(g, p)
            got, codes = [], []

            def poller():
                while True:
                    try:
                        r = srv.poll(rid, 5)
                    except epi.EpiError as e:
                        if e.code == epi.EPI_EBUSY:
                            continue
                        codes.append(e.code)
                        return
                    if r is not None:
                        got.append(r)
                        return

            th = [threading.Thread(target=poller) for _ in range(3)]
            for t in th:
                t.start()
            for t in th:
                t.join(120)
            assert len(got) == 1 and codes == [epi.EPI_ENOENT] * 2, (rep, len(got), codes)
            assert same(plain, got[0])


def test_submit_resident_cohort(epi, ctx):
    """epi_submit_cohort scans cohorts already on the device: results equal
    epi_scan, the cohort cannot be freed while its request runs, and the same
    cohort may be queued several times and scanned between requests with
    another tile shape."""
    # cohort 0 takes about a hundred milliseconds, so its requests are
    # still queued or running when the test tries to free it
    data = [epi.generate_hwe(40, 8192, 1200, 4096)] + make_cohorts(epi, ctx, 5, 41, (30, 500),
                                                                     (100, 5000))
    cos = [ctx.cohort(g, p) for g, p in data]
    try:
        plain = [ctx.scan(c) for c in cos]
        with ctx.server() as srv:
            ids = [srv.submit_cohort(c) for c in cos] + [srv.submit_cohort(cos[0])]
            # still queued or running: the cohort is held
            with pytest.raises(epi.EpiError) as e:
                cos[0].close()
            assert e.value.code == epi.EPI_EBUSY and cos[0].handle
            # a plain scan of a held cohort with the VALU tile meanwhile
            assert same(plain[1], ctx.scan(cos[1], tile={"variant": 1}))
            res = [srv.poll(rid, 120000) for rid in ids]
            assert all(same(a, b) for a, b in zip(plain, res[:-1]))
            assert same(plain[0], res[-1])
            st = srv.stats()
            assert st["completed"] == len(ids) and st["failed"] == 0
        # every request collected: the cohorts can go
    finally:
        for c in cos:
            c.close()


def test_plain_scan_beside_busy_server(epi, ctx):
    """A plain scan on the context of a busy server waits for the server's
    round in flight, not for its queue to drain: the context lock is taken
    in arrival order (the worker takes it again after every round)."""
    g, p = epi.generate_hwe(60, 8192, 1200, 4096)  # about 9 ms of scanning
    gs, ps = epi.generate_hwe(61, 1000, 60, 500)
    big, small = ctx.cohort(g, p), ctx.cohort(gs, ps)
    try:
        ref = ctx.scan(small)
        with ctx.server() as srv:
            ids = [srv.submit_cohort(big) for _ in range(12)]
            time.sleep(0.005)
            got = ctx.scan(small)
            st = srv.stats()
            say("plain scan beside the server:", st)
            assert st["completed"] < len(ids)
            assert same(ref, got)
            res = [srv.poll(rid, 120000) for rid in ids]
            assert all(same(res[0], r) for r in res[1:])
    finally:
        big.close()
        small.close()


def test_autotune_and_tile_table(epi):
    # a context of its own: the tile table is per context and is replaced here
    with epi.Context(0) as ctx:
        tile, ms = ctx.autotune(4096, 400)
        say("autotune 4096 x 400:", tile, ms)
        assert ms > 0
        if tile["variant"] in (2, 3):
            assert tile["snps_j"] == 64 and tile["waves"] == 8
        else:
            assert tile["variant"] == 1
            assert tile["snps_j"] in (1, 2, 4, 8) and tile["waves"] in (4, 8, 16)
        ctx.tune_set([(2048, {"variant": 1, "snps_j": 2, "waves": 4}), (1 << 20, tile)])
        assert ctx.tune_get(1000) == {"variant": 1, "snps_j": 2, "waves": 4}
        assert ctx.tune_get(5000) == tile
        ctx.tune_set([(1 << 20, {"variant": 2})])
        assert ctx.tune_get(5000) == {"variant": 2, "snps_j": 64, "waves": 8}
        g, p = epi.generate_hwe(2, 1500, 100, 700)
        with ctx.cohort(g, p) as co:
            a = ctx.scan(co)  # the table's MFMA entry
            b = ctx.scan(co, tile={"snps_j": 8, "waves": 16})  # an explicit VALU shape
            c = ctx.scan(co, tile={"variant": 1})
        assert same(a, b) and same(a, c)
        ctx.tune_set([])
        assert ctx.tune_get(1000)["variant"] == 1  # untuned default
