"""Stage-level parity: every kernel stage of the HIP path against the
oracle's stage function on the same inputs (bit-exact).

Graphs cover the degree regimes of the kernels: road-like (deg ~2), Delaunay
(deg ~6), weighted layered DAG (power-law fanout), and hub graphs whose
coarse rows exceed 64 entries (big-row contraction) and whose vertices exceed
the small-row gain path (LDS histogram path).
"""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def hub_graph(gk, n=6000, hubs=6, fan=700, seed=7):
    """Sparse random graph plus a few hubs adjacent to `fan` vertices each."""
    rng = np.random.default_rng(seed)
    edges = set()
    for v in range(n):
        for _ in range(2):
            u = int(rng.integers(0, n))
            if u != v:
                edges.add((min(u, v), max(u, v)))
    for h in range(hubs):
        for u in rng.choice(n, size=fan, replace=False):
            u = int(u)
            if u != h:
                edges.add((min(u, h), max(u, h)))
    e = np.array(sorted(edges), dtype=np.int64)
    w = rng.integers(1, 20, size=len(e))
    src = np.concatenate([e[:, 0], e[:, 1]])
    dst = np.concatenate([e[:, 1], e[:, 0]])
    ww = np.concatenate([w, w])
    order = np.lexsort((dst, src))
    src, dst, ww = src[order], dst[order], ww[order]
    xadj = np.zeros(n + 1, np.int64)
    np.add.at(xadj, src + 1, 1)
    xadj = np.cumsum(xadj)
    vw = rng.integers(1, 5, size=n)
    return gk.Graph(xadj, dst, ww, vw)


GRAPHS = {
    "road": lambda gk: gk.road(20000, 5),
    "delaunay": lambda gk: gk.delaunay(20000, 6),
    "dag": lambda gk: gk.dag(30000, 60, 7),
    "hubs": lambda gk: hub_graph(gk),
}


@pytest.fixture(scope="module", params=sorted(GRAPHS))
def graph(request, gk):
    return GRAPHS[request.param](gk)


def _I(t):
    from _dev import ptr
    return ptr(t)


def test_match_group_cluster_contract(gk, oracle, ctx, graph):
    from _dev import DevGraph, dev, host
    L = gk.lib()
    g = graph
    dg = DevGraph(g)
    cand = torch.zeros(g.n, dtype=torch.int32, device="cuda")
    assert L.gk_k_match(ctx.handle, dg.ref(), _I(cand), None) == 0
    ocand = oracle.match(g)
    assert np.array_equal(host(cand), ocand)

    label = torch.zeros_like(cand)
    stamp = torch.zeros_like(cand)
    rounds = ctypes.c_int32(0)
    dcand = dev(ocand)
    assert L.gk_k_group(ctx.handle, g.n, _I(dcand), _I(label), _I(stamp), ctypes.cast(ctypes.pointer(rounds), gk.I32P), None) == 0
    olabel, ostamp, orounds = oracle.group(ocand)
    assert np.array_equal(host(label), olabel)
    assert np.array_equal(host(stamp), ostamp)
    assert rounds.value == orounds

    order = torch.zeros_like(cand)
    cmap = torch.zeros_like(cand)
    cvw = torch.zeros_like(cand)
    nc = ctypes.c_int32(0)
    for G in (6, 2, 64):
        assert L.gk_k_cluster(ctx.handle, g.n, _I(dev(olabel)), _I(dev(ostamp)), _I(dg.vw), G, _I(order),
                              _I(cmap), _I(cvw), ctypes.cast(ctypes.pointer(nc), gk.I32P), None) == 0
        oorder, ocmap, ocvw, onc = oracle.cluster(olabel, ostamp, g.vwgt, G)
        assert nc.value == onc
        assert np.array_equal(host(order), oorder)
        assert np.array_equal(host(cmap), ocmap)
        assert np.array_equal(host(cvw)[:onc], ocvw)

    # contraction of the reference's groups of at most 6, and of groups of up
    # to 64 members (max_group = 64: up to 64 members per coarse row)
    for G in (6, 64):
        _, ocmap, _, onc = oracle.cluster(olabel, ostamp, g.vwgt, G)
        cx = torch.zeros(onc + 1, dtype=torch.int32, device="cuda")
        ca = torch.zeros(max(g.nnz, 1), dtype=torch.int32, device="cuda")
        cw = torch.zeros_like(ca)
        cnnz = ctypes.c_int32(0)
        assert L.gk_k_contract(ctx.handle, dg.ref(), _I(dev(ocmap)), onc, _I(cx), _I(ca), _I(cw),
                               ctypes.cast(ctypes.pointer(cnnz), gk.I32P), None) == 0
        ox, oa, ow = oracle.contract(g, ocmap, onc)
        assert cnnz.value == len(oa), G
        assert np.array_equal(host(cx), ox), G
        assert np.array_equal(host(ca)[: len(oa)], oa), G
        assert np.array_equal(host(cw)[: len(ow)], ow), G


def test_contract_big_rows(gk, oracle, ctx):
    """Coarse vertices gathering > 64 entries (hub members, large G)."""
    from _dev import DevGraph, dev, host
    L = gk.lib()
    g = hub_graph(gk, n=3000, hubs=10, fan=400, seed=11)
    rng = np.random.default_rng(3)
    nc = 200
    cmap = np.concatenate([np.arange(nc), rng.integers(0, nc, size=g.n - nc)]).astype(np.int32)
    dg = DevGraph(g)
    cx = torch.zeros(nc + 1, dtype=torch.int32, device="cuda")
    ca = torch.zeros(g.nnz, dtype=torch.int32, device="cuda")
    cw = torch.zeros_like(ca)
    cnnz = ctypes.c_int32(0)
    assert L.gk_k_contract(ctx.handle, dg.ref(), _I(dev(cmap)), nc, _I(cx), _I(ca), _I(cw),
                           ctypes.cast(ctypes.pointer(cnnz), gk.I32P), None) == 0
    ox, oa, ow = oracle.contract(g, cmap, nc)
    assert np.array_equal(host(cx), ox)
    assert np.array_equal(host(ca)[: len(oa)], oa)
    assert np.array_equal(host(cw)[: len(ow)], ow)


@pytest.mark.parametrize("wide", [False, True])
def test_contract_row_classes(gk, oracle, ctx, wide, monkeypatch):
    """Coarse rows of every lane class (16, 32, 64, big) from random groups
    of a weighted mesh, rows whose members outnumber the lanes (isolated
    members without entries), rows with no entries at all; both forms of the
    row kernels (lane number packed beside the coarse id, and (id, weight)
    pairs, forced by GK_CONTRACT_WIDE=1)."""
    from _dev import DevGraph, dev, host
    if wide:
        monkeypatch.setenv("GK_CONTRACT_WIDE", "1")
    L = gk.lib()
    m = gk.delaunay(6000, 9)
    iso = 400
    rng = np.random.default_rng(21)
    xadj = np.concatenate([m.xadj, np.full(iso, m.xadj[-1])]).astype(np.int64)
    # symmetric weights 1..49: one draw per undirected edge
    rows = np.repeat(np.arange(m.n), np.diff(m.xadj))
    key = np.minimum(rows, m.adjncy) * np.int64(m.n) + np.maximum(rows, m.adjncy)
    _, inv = np.unique(key, return_inverse=True)
    w = rng.integers(1, 50, size=inv.max() + 1)[inv]
    g = gk.Graph(xadj, m.adjncy, w, np.ones(m.n + iso, np.int64))
    # mesh vertices: ids 0..1499 once each, then half spread evenly and half
    # on ids 0..49 (big rows); isolated vertices: 100 on id 1400 and 100 on
    # ids 1401..1404 (more members than lanes), 200 on ids 1500..1502 (no
    # entries)
    nm = 1500
    rest = m.n - nm
    cmap = np.concatenate([np.arange(nm), rng.integers(0, nm, size=rest // 2), rng.integers(0, 50, size=rest - rest // 2),
                           np.full(100, 1400), rng.integers(1401, 1405, size=100),
                           np.full(100, nm), np.full(50, nm + 1), np.full(50, nm + 2)]).astype(np.int32)
    nc = nm + 3
    dg = DevGraph(g)
    cx = torch.zeros(nc + 1, dtype=torch.int32, device="cuda")
    ca = torch.zeros(g.nnz, dtype=torch.int32, device="cuda")
    cw = torch.zeros_like(ca)
    cnnz = ctypes.c_int32(0)
    assert L.gk_k_contract(ctx.handle, dg.ref(), _I(dev(cmap)), nc, _I(cx), _I(ca), _I(cw),
                           ctypes.cast(ctypes.pointer(cnnz), gk.I32P), None) == 0
    ox, oa, ow = oracle.contract(g, cmap, nc)
    ub = np.zeros(nc, np.int64)
    np.add.at(ub, cmap, np.diff(xadj))
    assert (ub <= 16).any() and ((ub > 16) & (ub <= 32)).any() and ((ub > 32) & (ub <= 64)).any() and (ub > 64).any()
    assert cnnz.value == len(oa)
    assert np.array_equal(host(cx), ox)
    assert np.array_equal(host(ca)[: len(oa)], oa)
    assert np.array_equal(host(cw)[: len(ow)], ow)


# (k, moves, seed, lowest slack below the cap): the oracle's prefix ends at
# 0, 79, 9, 3, 0, 6, 1, 9 for the tight cases, 501, 1012, 729, 3, 1023 for
# the roomier ones, and nowhere (-1) for (8192, 1024, 14, -12).
@pytest.mark.parametrize("k,nmoves,seed,lo", [(2, 1, 0, -60), (2, 300, 1, -60), (7, 1024, 2, -60), (64, 1024, 3, -60),
                                              (64, 37, 4, -60), (1000, 1024, 5, -60), (8192, 1024, 6, -60),
                                              (8192, 500, 7, -60), (64, 1024, 8, -600), (7, 1024, 10, -3000),
                                              (64, 1024, 9, -700), (1000, 1024, 12, -30), (8192, 1024, 14, -12),
                                              (16, 1024, 11, -1500)])
def test_balance_adversarial(gk, oracle, ctx, k, nmoves, seed, lo):
    """Balanced prefix on move lists the selection never produces: part
    weights straddling the cap (some parts already above it), many moves into
    the same parts, moves whose destination is their own part, k up to the
    kernel's 8192.  Position and part weights must equal the oracle's."""
    from _dev import DevGraph, dev, host
    L = gk.lib()
    rng = np.random.default_rng(seed)
    n = 4096
    xadj = np.arange(n + 1, dtype=np.int32)  # a ring: the balance stage reads only vertex weights
    adj = ((np.arange(n) + 1) % n).astype(np.int32)
    g = gk.Graph(xadj, adj, None, rng.integers(1, 20, size=n).astype(np.int32))
    dg = DevGraph(g)
    kk = min(k, 16) if seed % 2 else k  # odd seeds: few distinct parts, heavy collisions
    part = rng.integers(0, kk, size=n).astype(np.int32)
    dest = rng.integers(0, kk, size=n).astype(np.int32)
    same = rng.random(n) < 0.05
    dest[same] = part[same]
    cap = 1000
    pw = (cap + rng.integers(lo, 8, size=k)).astype(np.int32)
    moves = rng.choice(n, size=nmoves, replace=False).astype(np.int32)
    dpw = dev(pw)
    pos = ctypes.c_int32(-7)
    assert L.gk_k_balance(ctx.handle, dg.ref(), k, _I(dev(moves)), nmoves, _I(dev(part)), _I(dev(dest)),
                          _I(dpw), cap, ctypes.cast(ctypes.pointer(pos), gk.I32P), None) == 0
    opos, opw = oracle.balance(g, k, moves, part, dest, pw, cap)
    assert pos.value == opos
    assert np.array_equal(host(dpw), opw)


@pytest.mark.parametrize("k", [2, 16, 256])
def test_refinement_stages(gk, oracle, ctx, graph, k):
    from _dev import DevGraph, dev, host
    L = gk.lib()
    g = graph
    rng = np.random.default_rng(k)
    # A partition with spatial coherence: initial partitioner on the graph,
    # then 5% random vertices reassigned (many boundary vertices, positive gains).
    part, _ = gk.initpart(g, k)
    flip = rng.random(g.n) < 0.05
    part = part.copy()
    part[flip] = rng.integers(0, k, size=int(flip.sum()))
    upd = (rng.random(g.n) < 0.7).astype(np.int32)
    junk = rng.integers(-5, 5, size=(3, g.n)).astype(np.int32)
    dg = DevGraph(g)

    bnd, gain, dest = dev(junk[0]), dev(junk[1]), dev(junk[2])
    assert L.gk_k_gain(ctx.handle, dg.ref(), k, _I(dev(part)), _I(dev(upd)), _I(bnd), _I(gain), _I(dest), None) == 0
    obnd, ogain, odest = oracle.gain(g, k, part, upd, junk[0], junk[1], junk[2])
    assert np.array_equal(host(bnd), obnd)
    assert np.array_equal(host(gain), ogain)
    assert np.array_equal(host(dest), odest)

    # Full gain state for selection.
    ones = np.ones(g.n, np.int32)
    obnd, ogain, odest = oracle.gain(g, k, part, ones)
    pw = np.bincount(part, weights=g.vwgt, minlength=k).astype(np.int32)
    T = int(g.vwgt.sum())
    for cap in (gk.lib().gk_part_cap(T, k, 30), int(pw.max()) + 2):
        dupd = dev(upd)
        moves = torch.zeros(1024, dtype=torch.int32, device="cuda")
        nsel = ctypes.c_int32(0)
        assert L.gk_k_select(ctx.handle, dg.ref(), _I(dev(part)), _I(dev(obnd)), _I(dev(ogain)), _I(dev(odest)),
                             _I(dev(pw)), cap, 1024, _I(dupd), _I(moves),
                             ctypes.cast(ctypes.pointer(nsel), gk.I32P), None) == 0
        omoves, onsel, oupd = oracle.select(g, part, obnd, ogain, odest, pw, cap, 1024, upd)
        assert nsel.value == onsel
        assert np.array_equal(host(moves)[: len(omoves)], omoves)
        assert np.array_equal(host(dupd), oupd)
        if len(omoves) == 0:
            continue
        # Balanced prefix on the selected moves.
        dpw = dev(pw)
        pos = ctypes.c_int32(0)
        assert L.gk_k_balance(ctx.handle, dg.ref(), k, _I(dev(omoves)), len(omoves), _I(dev(part)), _I(dev(odest)),
                              _I(dpw), cap, ctypes.cast(ctypes.pointer(pos), gk.I32P), None) == 0
        opos, opw = oracle.balance(g, k, omoves, part, odest, pw, cap)
        assert pos.value == opos
        assert np.array_equal(host(dpw), opw)
        # Apply.
        dpart = dev(part)
        dupd2 = dev(np.zeros(g.n, np.int32))
        gs = ctypes.c_int64(0)
        assert L.gk_k_apply(ctx.handle, dg.ref(), _I(dev(omoves)), opos, _I(dev(ogain)), _I(dev(odest)), _I(dpart),
                            _I(dupd2), ctypes.byref(gs), None) == 0
        ogs, opart2, oupd2 = oracle.apply(g, omoves, opos, ogain, odest, part, np.zeros(g.n, np.int32))
        assert gs.value == ogs
        assert np.array_equal(host(dpart), opart2)
        assert np.array_equal(host(dupd2), oupd2)
        # The cut falls by exactly the applied gains (independent set).
        c0, _ = oracle.cut(g, k, part)
        c1, _ = oracle.cut(g, k, opart2)
        assert c0 - c1 == ogs

    # Cut and part weights.
    cut = ctypes.c_int64(0)
    pw64 = np.zeros(k, np.int64)
    assert L.gk_k_cut(ctx.handle, dg.ref(), k, _I(dev(part)), ctypes.byref(cut),
                      pw64.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)), None) == 0
    oc, opw64 = oracle.cut(g, k, part)
    assert cut.value == oc
    assert np.array_equal(pw64, opw64)


def test_project(gk, oracle, ctx):
    from _dev import dev, host
    L = gk.lib()
    rng = np.random.default_rng(5)
    n, nc = 5000, 900
    cmap = np.concatenate([np.arange(nc), rng.integers(0, nc, size=n - nc)]).astype(np.int32)
    cpart = rng.integers(0, 8, size=nc).astype(np.int32)
    cbnd = rng.integers(0, 2, size=nc).astype(np.int32)
    outs = [torch.full((n,), 7, dtype=torch.int32, device="cuda") for _ in range(5)]
    assert L.gk_k_project(ctx.handle, n, _I(dev(cmap)), _I(dev(cpart)), _I(dev(cbnd)), *[_I(o) for o in outs], None) == 0
    exp = [cpart[cmap], cbnd[cmap], np.zeros(n, np.int32), cpart[cmap], cbnd[cmap]]
    for o, e in zip(outs, exp):
        assert np.array_equal(host(o), e)


@pytest.mark.parametrize("k", [8, 64])
def test_end_to_end_hubs(gk, oracle, ctx, k):
    from test_gpu_parity import compare
    compare(gk, oracle, ctx, hub_graph(gk, n=20000, hubs=12, fan=1500, seed=2), k)
