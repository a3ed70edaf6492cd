"""Serving form (gk_serve_*): requests submitted while others are on the GPU
are admitted by footprint, stepped together with them, and each result must
equal, bit for bit, the same graph partitioned alone (gk_partition on another
context)."""
import time

import numpy as np
import pytest

from test_gpu_batch import mixed_set

pytestmark = pytest.mark.gpu


def check_alone(ctx, g, k, part, st):
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
