"""Batched form (gk_partition_batch / gk_partition_packed): each graph's result
from one packed run must equal, bit for bit, its own single-graph run
(gk_partition), which test_gpu_parity.py ties to the oracle -- and a few
graphs are compared with the oracle here directly.  Per-level sizes, rounds,
moves and cuts must agree too, not only the final partition."""
import ctypes
import math

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def mixed_set(gk, count, seed, lo=300, hi=30000, ks=(2, 4, 8, 16, 32)):
    rng = np.random.default_rng(seed)
    graphs, kk = [], []
    for i in range(count):
        n = int(round(math.exp(rng.uniform(math.log(lo), math.log(hi)))))
        if i % 3 == 0:
            g = gk.delaunay(n, 100 + i)
        elif i % 3 == 1:
            g = gk.dag(n, max(2, min(200, int(round(math.sqrt(n))))), 200 + i)
        else:
            g = gk.road(n, 300 + i)
        graphs.append(g)
        kk.append(int(rng.choice(ks)))
    return graphs, kk


def check_same(ctx, graphs, ks, parts, stats, **kw):
    for i, (g, k) in enumerate(zip(graphs, ks)):
        p1, s1 = ctx.partition(g, k, **kw)
        b = stats[i]
        tag = f"graph {i}: n={g.n} nnz={g.nnz} k={k}"
        assert np.array_equal(parts[i], p1), tag
        assert (b.nlevels, b.coarsest_n) == (s1.nlevels, s1.coarsest_n), tag
        for l in range(s1.nlevels):
            x, y = b.levels[l], s1.levels[l]
            got = (x.n, x.nnz, x.rounds, x.moves, x.cut_after)
            exp = (y.n, y.nnz, y.rounds, y.moves, y.cut_after)
            assert got == exp, f"{tag} level {l}: {got} != {exp}"
        got = (b.cut, b.cut_tracked, b.init_cut, b.max_part_wgt, b.total_vwgt, b.cap)
        exp = (s1.cut, s1.cut_tracked, s1.init_cut, s1.max_part_wgt, s1.total_vwgt, s1.cap)
        assert got == exp, f"{tag}: {got} != {exp}"


def test_batch_mixed_equals_alone(gk, ctx):
    graphs, ks = mixed_set(gk, 24, 4)
    print(f"batch of {len(graphs)}: n={[g.n for g in graphs]}", flush=True)
    parts, stats = ctx.partition_batch(graphs, ks)
    check_same(ctx, graphs, ks, parts, stats)
    for s in stats:
        assert s.max_part_wgt <= s.cap


def test_batch_lockstep_copies(gk, ctx):
    # Identical graphs advance in lockstep: every coarsening pack after the
    # first is the previous contraction's output taken whole (no copy).
    g = gk.delaunay(6000, 7)
    graphs, ks = [g] * 12, [8] * 12
    parts, stats = ctx.partition_batch(graphs, ks)
    check_same(ctx, graphs, ks, parts, stats)
    for p in parts[1:]:
        assert np.array_equal(p, parts[0])


def test_batch_edge_cases(gk, ctx):
    path = gk.Graph([0, 1, 3, 5, 7, 8], [1, 0, 2, 1, 3, 2, 4, 3])
    empty = gk.Graph([0], [])
    isolated = gk.Graph(np.zeros(51, np.int32), [])
    d = gk.delaunay(1500, 9)
    graphs = [path, d, empty, isolated, d, gk.delaunay(4000, 10)]
    ks = [2, 1, 4, 4, 16, 64]
    parts, stats = ctx.partition_batch(graphs, ks)
    check_same(ctx, graphs, ks, parts, stats)
    assert stats[2].nlevels == 0 and len(parts[2]) == 0
    assert np.all(parts[1] == 0)


def test_batch_weighted_ref_policy(gk, ctx):
    graphs, ks = mixed_set(gk, 6, 5, lo=20000, hi=60000, ks=(8, 64, 256))
    parts, stats = ctx.partition_batch(graphs, ks, coarsen_policy=gk.GK_COARSEN_REF)
    check_same(ctx, graphs, ks, parts, stats, coarsen_policy=gk.GK_COARSEN_REF)


def test_batch_against_oracle(gk, oracle, ctx):
    graphs, ks = mixed_set(gk, 5, 6, lo=1000, hi=12000)
    parts, stats = ctx.partition_batch(graphs, ks)
    for i, (g, k) in enumerate(zip(graphs, ks)):
        opart, ost = oracle.partition(g, k)
        assert np.array_equal(parts[i], opart), f"graph {i}"
        assert stats[i].cut == ost.cut


def test_packed_device_entry(gk, ctx):
    import torch
    from _dev import dev, host, ptr

    graphs, ks = mixed_set(gk, 10, 7, lo=500, hi=8000)
    ref_parts, ref_stats = ctx.partition_batch(graphs, ks)
    pg, v_off, e_off = gk.pack_graphs(graphs)
    x, a, w, vw = dev(pg.xadj), dev(pg.adjncy), dev(pg.adjwgt), dev(pg.vwgt)
    out = torch.full((pg.n,), -1, dtype=torch.int32, device="cuda")
    torch.cuda.synchronize()
    pk = gk.PackedCsr()
    pk.ngraphs, pk.n, pk.nnz = len(graphs), pg.n, pg.nnz
    pk.xadj, pk.adjncy, pk.adjwgt, pk.vwgt = ptr(x), ptr(a), ptr(w), ptr(vw)
    pk.v_off = v_off.ctypes.data_as(gk.I32P)
    pk.e_off = e_off.ctypes.data_as(gk.I32P)
    prm = (gk.Params * len(graphs))(*[gk.params(k) for k in ks])
    st = (gk.Stats * len(graphs))()
    rc = gk.lib().gk_partition_packed(ctx.handle, ctypes.byref(pk), prm, ptr(out), st, None)
    assert rc == gk.GK_OK, gk.lib().gk_last_error().decode()
    got = host(out)
    for i in range(len(graphs)):
        assert np.array_equal(got[v_off[i]:v_off[i + 1]], ref_parts[i]), f"graph {i}"
        assert st[i].cut == ref_stats[i].cut


def test_batch_repeat_is_deterministic(gk, ctx):
    graphs, ks = mixed_set(gk, 8, 8)
    p1, s1 = ctx.partition_batch(graphs, ks)
    p2, s2 = ctx.partition_batch(graphs, ks)
    for a, b, x, y in zip(p1, p2, s1, s2):
        assert np.array_equal(a, b)
        assert x.cut == y.cut


def test_batch_rejects_mixed_moves(gk, ctx):
    g = gk.delaunay(1000, 1)
    n = 2
    csrs = (gk.Csr * n)(g.csr(), g.csr())
    prm = (gk.Params * n)(gk.params(4), gk.params(4, max_moves=512))
    parts = [np.zeros(g.n, np.int32) for _ in range(n)]
    outs = (gk.I32P * n)(*[p.ctypes.data_as(gk.I32P) for p in parts])
    rc = gk.lib().gk_partition_batch(ctx.handle, n, csrs, prm, outs, None)
    assert rc == -1  # GK_ERR_ARG


def test_batch_small_move_cap(gk, ctx):
    # every round selects far more than M, so thresholds and their fallback
    # run for many jobs at once
    graphs, ks = mixed_set(gk, 24, 11)
    parts, stats = ctx.partition_batch(graphs, ks, max_moves=24)
    check_same(ctx, graphs, ks, parts, stats, max_moves=24)


def test_batch_graph_rounds(gk, ctx):
    # Rounds replayed from a captured hipGraph: each graph's result equals its
    # eager single-graph run; two packed calls in a row (two captures), with
    # eager single-graph runs in between.
    graphs, ks = mixed_set(gk, 16, 9)
    for rep in range(2):
        print(f"graph-round batch rep {rep}: n={[g.n for g in graphs]}", flush=True)
        parts, stats = ctx.partition_batch(graphs, ks, flags=gk.GK_FLAG_GRAPH_ROUNDS, max_moves=64)
        check_same(ctx, graphs, ks, parts, stats, max_moves=64)


def test_batch_persistent(gk, ctx):
    # Persistent work-stealing launches over the pack: each graph's result
    # equals its plain run (direct launches).
    graphs, ks = mixed_set(gk, 24, 17, lo=2000, hi=60000)
    print(f"persistent batch: n={[g.n for g in graphs]}", flush=True)
    parts, stats = ctx.partition_batch(graphs, ks, flags=gk.GK_FLAG_PERSISTENT)
    check_same(ctx, graphs, ks, parts, stats)


@pytest.mark.parametrize("ks", [(2, 4, 8, 16, 32), (300, 400)])
def test_batch_all_launch_options(gk, ctx, ks):
    # Rounds replayed from captured hipGraphs, autotuned launch shapes and
    # persistent work-stealing launches together, twice in a row: each
    # graph's result equals its plain run; with k > 256 the long rows go
    # through the heavy list and its own kernel.
    big_k = ks[0] > 256
    graphs, kk = mixed_set(gk, 9 if big_k else 24, 18, lo=5000, hi=30000 if big_k else 60000, ks=ks)
    flags = gk.GK_FLAG_GRAPH_ROUNDS | gk.GK_FLAG_AUTOTUNE | gk.GK_FLAG_PERSISTENT
    for rep in range(2):
        print(f"all launch options rep {rep}: n={[g.n for g in graphs]} k={kk}", flush=True)
        parts, stats = ctx.partition_batch(graphs, kk, flags=flags)
        check_same(ctx, graphs, kk, parts, stats)


def test_batch_autotune(gk, ctx):
    # A pack of over 1 Mi entries is tuned on its first coarsening level; each
    # graph's result still equals its plain run.
    graphs, ks = mixed_set(gk, 24, 12, lo=20000, hi=60000)
    print(f"autotune batch: n={[g.n for g in graphs]}", flush=True)
    parts, stats = ctx.partition_batch(graphs, ks, flags=gk.GK_FLAG_AUTOTUNE)
    check_same(ctx, graphs, ks, parts, stats)
