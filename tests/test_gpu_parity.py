"""End-to-end parity: the HIP V-cycle against the CPU oracle, bit for bit.

Every level's candidate array, clustering map and coarse graph (FNV-1a
hashes), every level's refinement trace (rounds, moves, cut after the level,
partition hash) and the final partition must agree exactly.
"""
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def compare(gk, oracle, ctx, g, k, **kw):
    print(f"case n={g.n} nnz={g.nnz} k={k} {kw}", flush=True)
    part, st = ctx.partition(g, k, flags=gk.GK_FLAG_TRACE, **kw)
    opart, ost = oracle.partition(g, k, flags=gk.GK_FLAG_TRACE, **kw)
    assert st.nlevels == ost.nlevels
    for l in range(st.nlevels):
        a, b = st.levels[l], ost.levels[l]
        got = (a.n, a.nnz, a.sweeps, a.h_cand, a.h_cmap, a.h_coarse)
        exp = (b.n, b.nnz, b.sweeps, b.h_cand, b.h_cmap, b.h_coarse)
        assert got == exp, f"coarsening level {l}: {got} != {exp}"
    assert st.init_cut == ost.init_cut
    for l in reversed(range(st.nlevels)):
        a, b = st.levels[l], ost.levels[l]
        got = (a.rounds, a.moves, a.cut_after, a.h_part)
        exp = (b.rounds, b.moves, b.cut_after, b.h_part)
        assert got == exp, f"refinement level {l}: {got} != {exp}"
    assert np.array_equal(part, opart)
    assert st.cut == ost.cut == st.cut_tracked
    c, pw = oracle.cut(g, k, part)
    assert c == st.cut
    assert pw.max() == st.max_part_wgt
    return st


@pytest.mark.parametrize("k", [2, 4, 8, 64])
def test_delaunay_n11(gk, oracle, ctx, k):
    g = gk.read_metis(os.path.join(GOLDEN, "delaunay_n11.graph"))
    st = compare(gk, oracle, ctx, g, k)
    assert st.imbalance <= 1.03 + 1e-12


@pytest.mark.parametrize("k", [4, 16])
def test_delaunay_n11_ref_policy(gk, oracle, ctx, k):
    g = gk.read_metis(os.path.join(GOLDEN, "delaunay_n11.graph"))
    compare(gk, oracle, ctx, g, k, coarsen_policy=gk.GK_COARSEN_REF)


@pytest.mark.parametrize("n,k,seed", [(2048, 4, 1), (20000, 16, 2), (60000, 64, 3)])
def test_delaunay_generated(gk, oracle, ctx, n, k, seed):
    compare(gk, oracle, ctx, gk.delaunay(n, seed), k)


@pytest.mark.parametrize("n,layers,k", [(20000, 40, 32), (50000, 100, 256)])
def test_dag_weighted(gk, oracle, ctx, n, layers, k):
    compare(gk, oracle, ctx, gk.dag(n, layers, 3), k)


@pytest.mark.parametrize("n,k", [(30000, 8), (60000, 256)])
def test_road(gk, oracle, ctx, n, k):
    compare(gk, oracle, ctx, gk.road(n, 2), k)


# Small move caps: the best-M radix select, the per-job gain threshold of the
# next round (set when more than 2M are selected) and its fallback (fewer than
# M pass the threshold, the deferred vertices are tested by the round kernel)
# all run in most rounds.
@pytest.mark.parametrize("family,n,k,moves", [("delaunay", 20000, 16, 8), ("delaunay", 20000, 16, 37),
                                              ("dag", 20000, 32, 16), ("dag", 50000, 256, 100),
                                              ("road", 30000, 8, 64)])
def test_small_move_cap(gk, oracle, ctx, family, n, k, moves):
    g = {"delaunay": lambda: gk.delaunay(n, 5), "dag": lambda: gk.dag(n, 40, 3), "road": lambda: gk.road(n, 2)}[family]()
    compare(gk, oracle, ctx, g, k, max_moves=moves)


# Part counts that are not powers of two (uneven recursive splits) and large
# k: the gain pass's LDS histogram path and balanced prefixes over thousands
# of parts (k = 8192 is the largest the kernels take).
@pytest.mark.parametrize("family,n,k", [("delaunay", 20000, 3), ("dag", 30000, 7), ("delaunay", 60000, 5),
                                        ("road", 30000, 100), ("delaunay", 400000, 1000), ("dag", 300000, 3000),
                                        ("road", 200000, 8192)])
def test_any_k(gk, oracle, ctx, family, n, k):
    g = {"delaunay": lambda: gk.delaunay(n, 4), "dag": lambda: gk.dag(n, min(100, n // 300), 4),
         "road": lambda: gk.road(n, 4)}[family]()
    st = compare(gk, oracle, ctx, g, k)
    assert st.max_part_wgt <= st.cap


# Refinement rounds replayed from a captured hipGraph: the same trace and
# partition as the oracle, on levels with many replays (small move cap) and
# across calls that re-capture the graph, with eager calls in between; also
# with persistent work-stealing launches, whose queue-clearing fill kernel
# and gain launch are then captured nodes.
@pytest.mark.parametrize("persistent", [False, True])
@pytest.mark.parametrize("family,n,k,moves", [("delaunay", 60000, 64, 1024), ("dag", 20000, 32, 16),
                                              ("road", 30000, 8, 64)])
def test_graph_rounds(gk, oracle, ctx, family, n, k, moves, persistent):
    g = {"delaunay": lambda: gk.delaunay(n, 3), "dag": lambda: gk.dag(n, 40, 3), "road": lambda: gk.road(n, 2)}[family]()
    flags = gk.GK_FLAG_TRACE | gk.GK_FLAG_GRAPH_ROUNDS | (gk.GK_FLAG_PERSISTENT if persistent else 0)
    for rep in range(2):
        print(f"graph rounds rep {rep} n={g.n} k={k} moves={moves} persistent={persistent}", flush=True)
        part, st = ctx.partition(g, k, flags=flags, max_moves=moves)
        eager, se = ctx.partition(g, k, flags=gk.GK_FLAG_TRACE, max_moves=moves)
        assert np.array_equal(part, eager)
        for l in range(st.nlevels):
            a, b = st.levels[l], se.levels[l]
            assert (a.rounds, a.moves, a.cut_after, a.h_part) == (b.rounds, b.moves, b.cut_after, b.h_part), l
        assert (st.cut, st.cut_tracked, st.max_part_wgt) == (se.cut, se.cut_tracked, se.max_part_wgt)
    opart, _ = oracle.partition(g, k, max_moves=moves)
    assert np.array_equal(part, opart)


# Launch shapes chosen by the autotuner: identical results, on a first call
# (candidates timed, each writing the output) and a second call (cached
# shapes), for a graph large enough to be tuned (over 1 Mi entries, row
# classes over 16 Ki rows).
def test_autotune_same_result(gk, oracle, ctx):
    g = gk.delaunay(300000, 11)
    print(f"autotune n={g.n} nnz={g.nnz}", flush=True)
    base, sb = ctx.partition(g, 64, flags=gk.GK_FLAG_TRACE)
    for rep in range(2):
        part, st = ctx.partition(g, 64, flags=gk.GK_FLAG_TRACE | gk.GK_FLAG_AUTOTUNE)
        assert np.array_equal(part, base), rep
        for l in range(st.nlevels):
            a, b = st.levels[l], sb.levels[l]
            assert (a.h_cand, a.h_cmap, a.h_coarse, a.h_part) == (b.h_cand, b.h_cmap, b.h_coarse, b.h_part), (rep, l)
    opart, _ = oracle.partition(g, 64)
    assert np.array_equal(base, opart)


# Persistent work-stealing launches of matching, the contraction rows and the
# gain pass: the schedule changes, the results do not.  Every level's hashes
# equal the direct launches' and the partition equals the oracle's, on a mesh
# large enough that its first levels hand out a tail through the queues, a
# DAG with heavy rows (k > 256: the heavy list) and a road network (static
# schedule only), twice in a row (the queue counters are cleared per launch).
@pytest.mark.parametrize("kind,k", [("delaunay", 64), ("dag", 300), ("road", 16)])
def test_persistent_same_result(gk, oracle, ctx, kind, k):
    g = {"delaunay": lambda: gk.delaunay(600000, 31), "dag": lambda: gk.dag(60000, 100, 32),
         "road": lambda: gk.road(150000, 33)}[kind]()
    print(f"persistent {kind} n={g.n} nnz={g.nnz} k={k}", flush=True)
    base, sb = ctx.partition(g, k, flags=gk.GK_FLAG_TRACE)
    for rep in range(2):
        part, st = ctx.partition(g, k, flags=gk.GK_FLAG_TRACE | gk.GK_FLAG_PERSISTENT)
        assert np.array_equal(part, base), rep
        assert (st.cut, st.nlevels) == (sb.cut, sb.nlevels)
        for l in range(st.nlevels):
            a, b = st.levels[l], sb.levels[l]
            assert (a.h_cand, a.h_cmap, a.h_coarse, a.h_part, a.rounds, a.moves) == \
                (b.h_cand, b.h_cmap, b.h_coarse, b.h_part, b.rounds, b.moves), (rep, l)
    opart, _ = oracle.partition(g, k)
    assert np.array_equal(base, opart)


# A circuit-timing-shaped DAG at 2^21 vertices (200 layers, edge weights
# 1..100, vertex weights 1..4, k = 256, 1024 moves per round): thousands of
# rounds, most selecting more than 1024 moves, so the top-1024 radix select of
# the round kernel runs throughout; every level's trace equals the oracle's.
@pytest.mark.slow
def test_dag_2m_k256(gk, oracle, ctx):
    g = gk.dag(1 << 21, 200, 3)
    st = compare(gk, oracle, ctx, g, 256)
    paths = list(st.round_paths)
    print(f"round paths (top-M select, 64-bit keys, deferred tests): {paths}", flush=True)
    assert paths[0] > 1000
    assert st.max_part_wgt <= st.cap


# Group sizes other than the reference's 6: pairs only, and groups of up to 64
# members (coarse rows gathering up to 64 member rows, many of them over 64
# entries), each level checked against the oracle.
@pytest.mark.parametrize("family,n,k,group", [("delaunay", 30000, 16, 2), ("delaunay", 30000, 16, 64),
                                              ("dag", 30000, 32, 64), ("road", 40000, 8, 17)])
def test_group_size(gk, oracle, ctx, family, n, k, group):
    g = {"delaunay": lambda: gk.delaunay(n, 8), "dag": lambda: gk.dag(n, 40, 4), "road": lambda: gk.road(n, 3)}[family]()
    compare(gk, oracle, ctx, g, k, max_group=group)
