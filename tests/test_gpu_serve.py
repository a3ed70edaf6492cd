"""Serving form (gk_serve_*): requests submitted while others are on the GPU
are admitted by footprint, stepped together with them, and each result must
equal, bit for bit, the same graph partitioned alone (gk_partition on another
context) and the oracle's partition of it."""
import time

import numpy as np
import pytest

from test_gpu_batch import mixed_set

pytestmark = pytest.mark.gpu


def check_alone(ctx, g, k, part, st):
    # alone on another context, and against the oracle
    import _oracle
    op, ost = _oracle.partition(g, k)
    assert np.array_equal(part, op), f"oracle n={g.n} k={k}"
    assert st.cut == ost.cut
    p1, s1 = ctx.partition(g, k)
    assert np.array_equal(part, p1), f"n={g.n} k={k}"
    assert (st.cut, st.nlevels, st.coarsest_n, st.init_cut) == (s1.cut, s1.nlevels, s1.coarsest_n, s1.init_cut)
    for l in range(s1.nlevels):
        a, b = st.levels[l], s1.levels[l]
        assert (a.n, a.nnz, a.rounds, a.moves, a.cut_after) == (b.n, b.nnz, b.rounds, b.moves, b.cut_after)


def test_serve_mix_equals_alone(gk, ctx):
    graphs, ks = mixed_set(gk, 12, 11, lo=500, hi=40000, ks=(8, 16, 64, 256))
    with gk.Server() as srv:
        ids = [srv.submit(g, k) for g, k in zip(graphs, ks)]
        res = [srv.wait(r) for r in ids]
        s = srv.stats()
        assert s.completed == len(graphs) and s.failed == 0
        assert 0 < s.p50_ms <= s.p99_ms <= s.max_ms
    for (part, st), g, k in zip(res, graphs, ks):
        check_alone(ctx, g, k, part, st)


def test_serve_staggered_arrivals(gk, ctx):
    # Requests arrive while earlier ones are in flight and join at step
    # boundaries (continuous batching); results are unchanged.
    graphs, ks = mixed_set(gk, 10, 12, lo=2000, hi=60000, ks=(8, 64))
    with gk.Server() as srv:
        ids = []
        for i, (g, k) in enumerate(zip(graphs, ks)):
            ids.append(srv.submit(g, k))
            if i % 3 == 2:
                time.sleep(0.002)
        res = [srv.wait(r) for r in ids]
        assert srv.stats().steps > 0
    for (part, st), g, k in zip(res, graphs, ks):
        check_alone(ctx, g, k, part, st)


def test_serve_budget_and_slots(gk, ctx):
    graphs, ks = mixed_set(gk, 8, 13, lo=5000, hi=30000, ks=(16,))
    feet = []
    for g, k in zip(graphs, ks):
        p = gk.params(k)
        feet.append(gk.lib().gk_serve_footprint(g.n, g.nnz, p))
    budget = int(max(feet) * 1.5)
    with gk.Server(hbm_budget=budget, max_inflight=3) as srv:
        ids = [srv.submit(g, k) for g, k in zip(graphs, ks)]
        res = [srv.wait(r) for r in ids]
        s = srv.stats()
        assert s.footprint_peak <= budget
        assert s.footprint_now == 0 and s.inflight == 0 and s.queued == 0
    for (part, st), g, k in zip(res, graphs, ks):
        check_alone(ctx, g, k, part, st)


def test_serve_poll_and_empty(gk, ctx):
    g = gk.delaunay(3000, 21)
    empty = gk.Graph([0], [])
    with gk.Server() as srv:
        r0 = srv.submit(empty, 4)
        r1 = srv.submit(g, 8)
        got = None
        for _ in range(20000):
            got = srv.poll(r1)
            if got is not None:
                break
            time.sleep(0.001)
        assert got is not None
        part, st = got
        p0, s0 = srv.wait(r0)
        assert len(p0) == 0
    check_alone(ctx, g, 8, part, st)


def test_serve_rejects_other_shape(gk):
    g = gk.delaunay(1000, 3)
    with gk.Server() as srv:
        srv.wait(srv.submit(g, 4))
        with pytest.raises(RuntimeError):
            srv.submit(g, 4, max_moves=256)


def test_serve_submit_many_all_or_none(gk, ctx):
    """A burst is validated before anything is queued: one bad request (k = 0)
    rejects the burst and queues none of it; a good burst is served in
    order, each result equal to its graph partitioned alone."""
    gs = [gk.delaunay(n, 40 + i) for i, n in enumerate((1500, 3000, 2200))]
    with gk.Server() as srv:
        with pytest.raises(RuntimeError, match="bad params"):
            srv.submit_many(gs, [8, 0, 16])
        assert srv.stats().submitted == 0
        ids = srv.submit_many(gs, [8, 4, 16])
        assert ids == sorted(ids) and len(set(ids)) == 3
        res = [srv.wait(r) for r in ids]
        s = srv.stats()
    assert s.submitted == 3 and s.completed == 3 and s.failed == 0
    for (part, st), g, k in zip(res, gs, (8, 4, 16)):
        check_alone(ctx, g, k, part, st)


# Failures stay with the request that causes them.  A device pool limit below
# what the large graph needs makes its allocations fail (admission at 0.9 x
# its CSR bytes, its first coarsening step at 1.3 x) while several requests
# share the device (all are queued in one burst, so the loop admits them
# together); the loop then runs each of them again alone: the small ones
# finish (equal to alone and the oracle), only the large one reports the
# error.
@pytest.mark.parametrize("factor", [0.9, 1.3])
def test_serve_failure_isolated(gk, ctx, factor):
    smalls, ks = mixed_set(gk, 6, 14, lo=2000, hi=8000, ks=(8, 16))
    big = gk.delaunay(400000, 15)
    limit = int(factor * 4 * (2 * big.n + 1 + 2 * big.nnz))
    print(f"pool limit {limit} bytes", flush=True)
    with gk.Server(pool_limit=limit) as srv:
        *ids, bid = srv.submit_many(smalls + [big], list(ks) + [64])
        res = [srv.wait(r) for r in ids]
        with pytest.raises(RuntimeError, match="pool limit"):
            srv.wait(bid)
        s = srv.stats()
    assert s.completed == len(smalls) and s.failed == 1
    assert s.isolated >= 2
    for (part, st), g, k in zip(res, smalls, ks):
        check_alone(ctx, g, k, part, st)


def test_serve_over_budget_runs_alone(gk, ctx):
    # Every request is bigger than the budget: each is admitted only onto an
    # empty device, so the admitted footprint never exceeds the largest one.
    graphs, ks = mixed_set(gk, 4, 16, lo=20000, hi=40000, ks=(16,))
    feet = [gk.lib().gk_serve_footprint(g.n, g.nnz, gk.params(k)) for g, k in zip(graphs, ks)]
    with gk.Server(hbm_budget=min(feet) // 2) as srv:
        res = [srv.wait(r) for r in [srv.submit(g, k) for g, k in zip(graphs, ks)]]
        s = srv.stats()
    assert s.completed == len(graphs) and s.failed == 0
    assert s.footprint_peak == max(feet)
    for (part, st), g, k in zip(res, graphs, ks):
        check_alone(ctx, g, k, part, st)


def test_serve_one_taker_per_request(gk):
    # Two callers wait on one request: one takes the result, the other is
    # refused (GK_ERR_ARG), and the result cannot be taken again.
    import ctypes
    import threading
    g = gk.delaunay(30000, 5)
    L = gk.lib()
    rcs = {}
    with gk.Server() as srv:
        rid = srv.submit(g, 8)

        def waiter(i):
            part = np.zeros(g.n, np.int32)
            st = gk.Stats()
            rcs[i] = L.gk_serve_wait(srv._h, rid, part.ctypes.data_as(gk.I32P), ctypes.byref(st))

        ts = [threading.Thread(target=waiter, args=(i,)) for i in range(2)]
        for t in ts:
            t.start()
        for t in ts:
            t.join()
        assert sorted(rcs.values()) == [-1, 0]
        assert L.gk_serve_poll(srv._h, rid, None, None) == -1
