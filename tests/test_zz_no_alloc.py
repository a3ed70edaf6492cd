"""gk_k_cut_noalloc: the allocation-free cut (static device scratch) against
gk_k_cut (pool scratch) and the oracle, on consecutive calls that reuse the
same scratch array, the second with a smaller input."""
import ctypes

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu

GK_ERR_ARG, GK_ERR_LIMIT = -1, -7


def _cut_pool(gk, ctx, dg, k, dpart):
    from _dev import ptr
    cut = ctypes.c_int64(0)
    pw = np.zeros(k, np.int64)
    rc = gk.lib().gk_k_cut(ctx.handle, dg.ref(), k, ptr(dpart), ctypes.byref(cut),
                           pw.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)), None)
    assert rc == 0, gk.lib().gk_last_error()
    return cut.value, pw


def _cut_noalloc(gk, ctx, dg, k, dpart):
    from _dev import ptr
    # Outputs start as garbage: the entry point must overwrite all of them.
    dcut = torch.full((1,), -7, dtype=torch.int64, device="cuda")
    dpw = torch.full((k,), -7, dtype=torch.int64, device="cuda")
    torch.cuda.synchronize()
    rc = gk.lib().gk_k_cut_noalloc(ctx.handle, dg.ref(), k, ptr(dpart), ctypes.c_void_p(dcut.data_ptr()),
                                   ctypes.c_void_p(dpw.data_ptr()), None)
    assert rc == 0, gk.lib().gk_last_error()
    torch.cuda.synchronize()
    return int(dcut.cpu()[0]), dpw.cpu().numpy()


def _case(gk, oracle, ctx, g, k, seed):
    from _dev import DevGraph, dev
    rng = np.random.default_rng(seed)
    part = rng.integers(0, k, size=g.n).astype(np.int32)
    dg = DevGraph(g)
    dpart = dev(part)
    c0, pw0 = _cut_pool(gk, ctx, dg, k, dpart)
    c1, pw1 = _cut_noalloc(gk, ctx, dg, k, dpart)
    oc, opw = oracle.cut(g, k, part)
    print(f"noalloc cut n={g.n} nnz={g.nnz} k={k}: {c1} (pool {c0}, oracle {oc})", flush=True)
    assert c1 == c0 == oc
    assert np.array_equal(pw1, pw0)
    assert np.array_equal(pw1, opw)


def test_noalloc_cut_two_calls(gk, oracle, ctx):
    # First a large weighted graph at k = 256, then a smaller one at k = 5:
    # the second call reuses the scratch array the first one filled.
    _case(gk, oracle, ctx, gk.dag(200000, 100, 3), 256, 1)
    _case(gk, oracle, ctx, gk.delaunay(3000, 2), 5, 2)


def test_noalloc_cut_largest_k(gk, oracle, ctx):
    # k = 8192 leaves room for 31 blocks of slots; the sums do not depend on
    # the block count.
    _case(gk, oracle, ctx, gk.delaunay(60000, 4), 8192, 3)
    _case(gk, oracle, ctx, gk.road(20000, 5), 3, 4)


def test_noalloc_cut_limits(gk, ctx):
    from _dev import DevGraph, dev, ptr
    g = gk.delaunay(2048, 1)
    dg = DevGraph(g)
    dpart = dev(np.zeros(g.n, np.int32))
    dcut = torch.zeros(1, dtype=torch.int64, device="cuda")
    dpw = torch.zeros(8193, dtype=torch.int64, device="cuda")
    L = gk.lib()
    vp = ctypes.c_void_p
    # k beyond the scratch's limit: rejected before any launch.
    assert L.gk_k_cut_noalloc(ctx.handle, dg.ref(), 8193, ptr(dpart), vp(dcut.data_ptr()), vp(dpw.data_ptr()),
                              None) == GK_ERR_LIMIT
    assert b"8192" in L.gk_last_error()
    assert L.gk_k_cut_noalloc(ctx.handle, dg.ref(), 4, ptr(dpart), None, vp(dpw.data_ptr()), None) == GK_ERR_ARG
    assert L.gk_k_cut_noalloc(ctx.handle, dg.ref(), 0, ptr(dpart), vp(dcut.data_ptr()), vp(dpw.data_ptr()),
                              None) == GK_ERR_ARG


def test_noalloc_cut_empty_graph(gk, ctx):
    from _dev import ptr
    c = gk.Csr()
    xadj = torch.zeros(1, dtype=torch.int32, device="cuda")
    one = torch.zeros(1, dtype=torch.int32, device="cuda")
    c.n, c.nnz = 0, 0
    c.xadj, c.adjncy, c.adjwgt, c.vwgt = ptr(xadj), ptr(one), ptr(one), ptr(one)
    dcut = torch.full((1,), -7, dtype=torch.int64, device="cuda")
    dpw = torch.full((3,), -7, dtype=torch.int64, device="cuda")
    torch.cuda.synchronize()
    assert gk.lib().gk_k_cut_noalloc(ctx.handle, ctypes.byref(c), 3, None, ctypes.c_void_p(dcut.data_ptr()),
                                     ctypes.c_void_p(dpw.data_ptr()), None) == 0
    torch.cuda.synchronize()
    assert int(dcut.cpu()[0]) == 0
    assert (dpw.cpu().numpy() == 0).all()
