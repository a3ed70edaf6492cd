"""Python-side handle on libgkway.so (the MI355X partitioner) and libgkgen.so
(the seeded workload generators).

The product is the C ABI declared in include/gkway.h; this module only mirrors
its structs with ctypes and wraps a few calls for the tests and the benchmark.
It never computes a partition itself: if the native library is missing,
`lib()` raises.

The directory name contains a hyphen, so load it with `load_package()` from a
path (tests/conftest.py, bench.py and __graft_entry__.py do this).
"""
from __future__ import annotations

import ctypes
import os
import subprocess
from ctypes import POINTER, Structure, c_char_p, c_double, c_int, c_int32, c_int64, c_uint64, c_void_p

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIBGKWAY = os.path.join(HERE, "libgkway.so")
LIBGKGEN = os.path.join(HERE, "libgkgen.so")
CLI = os.path.join(HERE, "g-kway")

GK_OK = 0
GK_COARSEN_TARGET = 0
GK_COARSEN_REF = 1
GK_FLAG_CHECK = 1
GK_FLAG_TRACE = 2
GK_MAX_LEVELS = 64

I32P = POINTER(c_int32)


class Csr(Structure):
    _fields_ = [("n", c_int32), ("nnz", c_int32), ("xadj", I32P), ("adjncy", I32P),
                ("adjwgt", I32P), ("vwgt", I32P)]


class Params(Structure):
    _fields_ = [("k", c_int32), ("ubfactor", c_int32), ("max_group", c_int32),
                ("max_moves", c_int32), ("coarsen_policy", c_int32), ("coarsen_to", c_int32),
                ("max_levels", c_int32), ("seed", c_int32), ("initpart_threads", c_int32),
                ("flags", c_int32)]


class LevelStats(Structure):
    _fields_ = [("n", c_int32), ("nnz", c_int32), ("sweeps", c_int32), ("rounds", c_int32),
                ("moves", c_int64), ("cut_after", c_int64), ("h_cand", c_uint64),
                ("h_cmap", c_uint64), ("h_coarse", c_uint64), ("h_part", c_uint64)]


GK_FLAG_STAGE_TIME = 4
GK_FLAG_GRAPH_ROUNDS = 8
GK_FLAG_AUTOTUNE = 16
GK_FLAG_PERSISTENT = 32
GK_ST_COUNT = 16
STAGE_NAMES = ["match", "group", "sort", "cluster", "contract", "contract_big", "project", "gain",
               "select", "msort", "balance", "apply", "cut", "gain_heavy", "contract_aux", "pack"]


class Stats(Structure):
    _fields_ = [("nlevels", c_int32), ("coarsest_n", c_int32), ("cut", c_int64),
                ("cut_tracked", c_int64), ("init_cut", c_int64), ("max_part_wgt", c_int64),
                ("total_vwgt", c_int64), ("cap", c_int64), ("imbalance", c_double),
                ("t_coarsen_ms", c_double), ("t_init_ms", c_double), ("t_refine_ms", c_double),
                ("t_total_ms", c_double), ("alg_bytes", c_int64),
                ("levels", LevelStats * GK_MAX_LEVELS),
                ("stage_ms", c_double * GK_ST_COUNT), ("stage_launches", c_int32 * GK_ST_COUNT),
                ("stage_bytes", c_int64 * GK_ST_COUNT), ("round_paths", c_int64 * 3)]

    def stages(self):
        return {name: (self.stage_ms[i], self.stage_launches[i], self.stage_bytes[i])
                for i, name in enumerate(STAGE_NAMES)}

    def level_list(self):
        return [self.levels[i] for i in range(self.nlevels)]


class PackedCsr(Structure):
    _fields_ = [("ngraphs", c_int32), ("n", c_int32), ("nnz", c_int32), ("xadj", I32P),
                ("adjncy", I32P), ("adjwgt", I32P), ("vwgt", I32P), ("v_off", I32P),
                ("e_off", I32P)]


class ServeConfig(Structure):
    _fields_ = [("hbm_budget", c_int64), ("max_inflight", c_int32), ("workers", c_int32)]


class ServeStats(Structure):
    _fields_ = [("submitted", c_int64), ("completed", c_int64), ("failed", c_int64),
                ("queued", c_int64), ("inflight", c_int64), ("edges_done", c_int64),
                ("steps", c_int64), ("footprint_now", c_int64), ("footprint_peak", c_int64),
                ("device_peak", c_int64), ("p50_ms", c_double), ("p99_ms", c_double),
                ("max_ms", c_double), ("edges_per_s", c_double), ("isolated", c_int64)]


class GenGraph(Structure):
    _fields_ = [("n", c_int32), ("nnz", c_int32), ("xadj", I32P), ("adjncy", I32P),
                ("adjwgt", I32P), ("vwgt", I32P)]


def params(k, **kw) -> Params:
    p = Params()
    p.k = k
    p.ubfactor = kw.pop("ubfactor", 30)
    p.max_group = kw.pop("max_group", 6)
    p.max_moves = kw.pop("max_moves", 1024)
    p.coarsen_policy = kw.pop("coarsen_policy", GK_COARSEN_TARGET)
    p.coarsen_to = kw.pop("coarsen_to", 0)
    p.max_levels = kw.pop("max_levels", 0)
    p.seed = kw.pop("seed", 1)
    p.initpart_threads = kw.pop("initpart_threads", 0)
    p.flags = kw.pop("flags", 0)
    if kw:
        raise TypeError(f"unknown params {sorted(kw)}")
    return p


class Graph:
    """Host CSR held as numpy int32 arrays (0-based adjncy)."""

    def __init__(self, xadj, adjncy, adjwgt=None, vwgt=None):
        self.xadj = np.ascontiguousarray(xadj, dtype=np.int32)
        self.adjncy = np.ascontiguousarray(adjncy, dtype=np.int32)
        self.n = len(self.xadj) - 1
        self.nnz = int(self.xadj[-1]) if self.n >= 0 else 0
        self.adjwgt = (np.ascontiguousarray(adjwgt, dtype=np.int32) if adjwgt is not None
                       else np.ones(self.nnz, np.int32))
        self.vwgt = (np.ascontiguousarray(vwgt, dtype=np.int32) if vwgt is not None
                     else np.ones(self.n, np.int32))

    @property
    def m(self):
        return self.nnz // 2

    def csr(self) -> Csr:
        c = Csr()
        c.n = self.n
        c.nnz = self.nnz
        c.xadj = self.xadj.ctypes.data_as(I32P)
        c.adjncy = self.adjncy.ctypes.data_as(I32P)
        c.adjwgt = self.adjwgt.ctypes.data_as(I32P)
        c.vwgt = self.vwgt.ctypes.data_as(I32P)
        c._keep = self  # keep arrays alive
        return c

    def degrees(self):
        return np.diff(self.xadj)


_LIB = None
_GEN = None


def lib():
    """libgkway.so, loaded once; raises if it has not been built."""
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIBGKWAY):
            raise RuntimeError(f"{LIBGKWAY} is missing: run __graft_entry__.build() first")
        L = ctypes.CDLL(LIBGKWAY)
        L.gk_strerror.restype = c_char_p
        L.gk_last_error.restype = c_char_p
        L.gk_part_cap.restype = c_int64
        L.gk_part_cap.argtypes = [c_int64, c_int32, c_int32]
        L.gk_coarsen_floor.restype = c_int32
        L.gk_ctx_create.argtypes = [c_int, POINTER(c_void_p)]
        L.gk_ctx_destroy.argtypes = [c_void_p]
        L.gk_ctx_set_stage_mask.argtypes = [c_void_p, ctypes.c_uint32]
        L.gk_ctx_set_pool_limit.argtypes = [c_void_p, c_int64]
        L.gk_partition.argtypes = [c_void_p, POINTER(Csr), POINTER(Params), I32P, POINTER(Stats)]
        L.gk_partition_device.argtypes = [c_void_p, POINTER(Csr), POINTER(Params), I32P,
                                          POINTER(Stats), c_void_p]
        L.gk_initpart.argtypes = [POINTER(Csr), c_int32, c_int64, c_int32, c_int32, I32P,
                                  POINTER(c_int64)]
        VP = c_void_p
        L.gk_k_match.argtypes = [VP, POINTER(Csr), I32P, VP]
        L.gk_k_group.argtypes = [VP, c_int32, I32P, I32P, I32P, I32P, VP]
        L.gk_k_cluster.argtypes = [VP, c_int32, I32P, I32P, I32P, c_int32, I32P, I32P, I32P, I32P, VP]
        L.gk_k_contract.argtypes = [VP, POINTER(Csr), I32P, c_int32, I32P, I32P, I32P, I32P, VP]
        L.gk_k_project.argtypes = [VP, c_int32, I32P, I32P, I32P, I32P, I32P, I32P, I32P, I32P, VP]
        L.gk_k_gain.argtypes = [VP, POINTER(Csr), c_int32, I32P, I32P, I32P, I32P, I32P, VP]
        L.gk_k_select.argtypes = [VP, POINTER(Csr), I32P, I32P, I32P, I32P, I32P, c_int64, c_int32,
                                  I32P, I32P, I32P, VP]
        L.gk_k_balance.argtypes = [VP, POINTER(Csr), c_int32, I32P, c_int32, I32P, I32P, I32P, c_int64,
                                   I32P, VP]
        L.gk_k_apply.argtypes = [VP, POINTER(Csr), I32P, c_int32, I32P, I32P, I32P, I32P,
                                 POINTER(c_int64), VP]
        L.gk_k_cut.argtypes = [VP, POINTER(Csr), c_int32, I32P, POINTER(c_int64), POINTER(c_int64), VP]
        L.gk_k_cut_noalloc.argtypes = [VP, POINTER(Csr), c_int32, I32P, VP, VP, VP]
        L.gk_k_round.argtypes = [VP, POINTER(Csr), c_int32, c_int64, c_int32, c_int32, I32P, I32P, I32P, I32P,
                                 I32P, I32P, I32P, POINTER(c_int64), I32P, POINTER(c_int64), VP]
        L.gk_graph_read.argtypes = [c_char_p, POINTER(Csr)]
        L.gk_graph_write.argtypes = [c_char_p, POINTER(Csr)]
        L.gk_graph_check.argtypes = [POINTER(Csr), c_int]
        L.gk_graph_free.argtypes = [POINTER(Csr)]
        L.gk_graph_write_bin.argtypes = [c_char_p, POINTER(Csr)]
        L.gk_graph_map_bin.argtypes = [c_char_p, c_int, POINTER(Csr), POINTER(c_void_p)]
        L.gk_graph_unmap.argtypes = [c_void_p]
        L.gk_write_partition.argtypes = [c_char_p, I32P, c_int32]
        L.gk_partition_file.argtypes = [VP, c_char_p, c_char_p, c_int32, POINTER(Stats)]
        L.gk_partition_packed.argtypes = [VP, POINTER(PackedCsr), POINTER(Params), I32P,
                                          POINTER(Stats), VP]
        L.gk_partition_batch.argtypes = [VP, c_int32, POINTER(Csr), POINTER(Params),
                                         POINTER(I32P), POINTER(Stats)]
        L.gk_serve_open.argtypes = [VP, POINTER(ServeConfig), POINTER(c_void_p)]
        L.gk_serve_submit.argtypes = [VP, POINTER(Csr), POINTER(Params), POINTER(c_int64)]
        L.gk_serve_submit_many.argtypes = [VP, c_int32, POINTER(Csr), POINTER(Params), POINTER(c_int64)]
        L.gk_serve_poll.argtypes = [VP, c_int64, I32P, POINTER(Stats)]
        L.gk_serve_wait.argtypes = [VP, c_int64, I32P, POINTER(Stats)]
        L.gk_serve_query.argtypes = [VP, POINTER(ServeStats)]
        L.gk_serve_close.argtypes = [VP]
        L.gk_serve_footprint.restype = c_int64
        L.gk_serve_footprint.argtypes = [c_int32, c_int32, POINTER(Params)]
        _LIB = L
    return _LIB


def gen():
    global _GEN
    if _GEN is None:
        if not os.path.exists(LIBGKGEN):
            raise RuntimeError(f"{LIBGKGEN} is missing: run __graft_entry__.build() first")
        G = ctypes.CDLL(LIBGKGEN)
        G.gkgen_delaunay.argtypes = [c_int32, c_uint64, POINTER(GenGraph)]
        G.gkgen_road.argtypes = [c_int32, c_uint64, c_int32, POINTER(GenGraph)]
        G.gkgen_dag.argtypes = [c_int32, c_int32, c_uint64, c_int32, c_double, c_int32, c_int32,
                                c_int32, POINTER(GenGraph)]
        G.gkgen_points.argtypes = [c_int32, c_uint64, I32P, I32P]
        G.gkgen_delaunay_points.argtypes = [c_int32, I32P, I32P, POINTER(GenGraph)]
        G.gkgen_fnv1a_i32.restype = c_uint64
        G.gkgen_fnv1a_i32.argtypes = [I32P, c_int64, c_uint64]
        _GEN = G
    return _GEN


def _take(gg: GenGraph) -> Graph:
    n, nnz = gg.n, gg.nnz
    xadj = np.ctypeslib.as_array(gg.xadj, (n + 1,)).copy()
    adj = np.ctypeslib.as_array(gg.adjncy, (max(nnz, 1),))[:nnz].copy()
    w = np.ctypeslib.as_array(gg.adjwgt, (max(nnz, 1),))[:nnz].copy() if gg.adjwgt else None
    vw = np.ctypeslib.as_array(gg.vwgt, (max(n, 1),))[:n].copy() if gg.vwgt else None
    gen().gkgen_free(ctypes.byref(gg))
    return Graph(xadj, adj, w, vw)


def delaunay(n, seed) -> Graph:
    gg = GenGraph()
    rc = gen().gkgen_delaunay(n, seed, ctypes.byref(gg))
    if rc != 0:
        raise RuntimeError(f"gkgen_delaunay failed: {rc}")
    return _take(gg)


def hilbert_index(x, y, bits=24):
    """Position of each point (x, y in [0, 2^bits)) along the Hilbert curve
    of the grid (the classic iterative xy -> d mapping)."""
    x = np.asarray(x, np.int64).copy()
    y = np.asarray(y, np.int64).copy()
    d = np.zeros(len(x), np.int64)
    side = 1 << bits
    s = side >> 1
    while s > 0:
        rx = (x & s) > 0
        ry = (y & s) > 0
        d += s * s * ((3 * rx.astype(np.int64)) ^ ry.astype(np.int64))
        flip = ~ry & rx
        x[flip] = side - 1 - x[flip]
        y[flip] = side - 1 - y[flip]
        swap = ~ry
        x[swap], y[swap] = y[swap], x[swap].copy()
        s >>= 1
    return d


def delaunay_hilbert(n, seed) -> Graph:
    """The triangulation of the same points as delaunay(n, seed), with the
    vertices numbered in Hilbert-curve order of their points instead of in
    drawing order (an isomorphic graph whose neighbours have nearby ids)."""
    x = np.zeros(n, np.int32)
    y = np.zeros(n, np.int32)
    rc = gen().gkgen_points(n, seed, x.ctypes.data_as(I32P), y.ctypes.data_as(I32P))
    if rc != 0:
        raise RuntimeError(f"gkgen_points failed: {rc}")
    order = np.argsort(hilbert_index(x, y), kind="stable")
    xs, ys = np.ascontiguousarray(x[order]), np.ascontiguousarray(y[order])
    gg = GenGraph()
    rc = gen().gkgen_delaunay_points(n, xs.ctypes.data_as(I32P), ys.ctypes.data_as(I32P), ctypes.byref(gg))
    if rc != 0:
        raise RuntimeError(f"gkgen_delaunay_points failed: {rc}")
    return _take(gg)


def road(n, seed, avg_deg=2.1) -> Graph:
    gg = GenGraph()
    rc = gen().gkgen_road(n, seed, int(round(avg_deg * 100)), ctypes.byref(gg))
    if rc != 0:
        raise RuntimeError(f"gkgen_road failed: {rc}")
    return _take(gg)


def dag(n, layers, seed, max_fanout=64, exponent=2.5, window=3, wmax=100, vwmax=4) -> Graph:
    gg = GenGraph()
    rc = gen().gkgen_dag(n, layers, seed, max_fanout, exponent, window, wmax, vwmax, ctypes.byref(gg))
    if rc != 0:
        raise RuntimeError(f"gkgen_dag failed: {rc}")
    return _take(gg)


def fnv1a(a, h=0xcbf29ce484222325) -> int:
    a = np.ascontiguousarray(a, dtype=np.int32)
    x = np.uint64(h)
    # vectorised-in-python is slow for big arrays; use the native helper.
    return int(gen().gkgen_fnv1a_i32(a.ctypes.data_as(I32P), len(a), h))


def write_bin(g: "Graph", path):
    """Binary CSR file (gk_graph_write_bin)."""
    L = lib()
    rc = L.gk_graph_write_bin(path.encode(), ctypes.byref(g.csr()))
    if rc != GK_OK:
        raise RuntimeError(f"gk_graph_write_bin: {L.gk_strerror(rc).decode()}: {L.gk_last_error().decode()}")


def map_bin(path, verify=True) -> Graph:
    """Maps a binary CSR file (gk_graph_map_bin) and returns a copy of it."""
    L = lib()
    c = Csr()
    h = c_void_p()
    rc = L.gk_graph_map_bin(path.encode(), 1 if verify else 0, ctypes.byref(c), ctypes.byref(h))
    if rc != GK_OK:
        raise RuntimeError(f"gk_graph_map_bin: {L.gk_strerror(rc).decode()}: {L.gk_last_error().decode()}")
    n, nnz = c.n, c.nnz
    try:
        return Graph(np.ctypeslib.as_array(c.xadj, (n + 1,)).copy(),
                     np.ctypeslib.as_array(c.adjncy, (max(nnz, 1),))[:nnz].copy() if nnz else np.zeros(0, np.int32),
                     np.ctypeslib.as_array(c.adjwgt, (max(nnz, 1),))[:nnz].copy() if c.adjwgt else None,
                     np.ctypeslib.as_array(c.vwgt, (max(n, 1),))[:n].copy() if c.vwgt else None)
    finally:
        L.gk_graph_unmap(h)


def read_metis(path) -> Graph:
    L = lib()
    c = Csr()
    rc = L.gk_graph_read(path.encode(), ctypes.byref(c))
    if rc != GK_OK:
        raise RuntimeError(f"gk_graph_read: {L.gk_strerror(rc).decode()}: {L.gk_last_error().decode()}")
    n, nnz = c.n, c.nnz
    g = Graph(np.ctypeslib.as_array(c.xadj, (n + 1,)).copy(),
              np.ctypeslib.as_array(c.adjncy, (max(nnz, 1),))[:nnz].copy(),
              np.ctypeslib.as_array(c.adjwgt, (max(nnz, 1),))[:nnz].copy() if c.adjwgt else None,
              np.ctypeslib.as_array(c.vwgt, (max(n, 1),))[:n].copy() if c.vwgt else None)
    L.gk_graph_free(ctypes.byref(c))
    return g


def initpart(g: Graph, k, cap=None, seed=1, threads=0):
    L = lib()
    if cap is None:
        cap = L.gk_part_cap(int(g.vwgt.sum()), k, 30)
    part = np.zeros(g.n, np.int32)
    cut = c_int64(0)
    rc = L.gk_initpart(ctypes.byref(g.csr()), k, cap, seed, threads, part.ctypes.data_as(I32P),
                       ctypes.byref(cut))
    if rc != GK_OK:
        raise RuntimeError(f"gk_initpart: {rc}")
    return part, cut.value


class Context:
    """One gk_ctx (device arena + stream) on `device`."""

    def __init__(self, device=0):
        self._h = c_void_p()
        rc = lib().gk_ctx_create(device, ctypes.byref(self._h))
        if rc != GK_OK:
            raise RuntimeError(f"gk_ctx_create: {lib().gk_strerror(rc).decode()}: "
                               f"{lib().gk_last_error().decode()}")

    @property
    def handle(self):
        return self._h

    def close(self):
        if self._h:
            lib().gk_ctx_destroy(self._h)
            self._h = c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def partition(self, g: Graph, k, **kw):
        p = params(k, **kw)
        part = np.zeros(g.n, np.int32)
        st = Stats()
        rc = lib().gk_partition(self._h, ctypes.byref(g.csr()), ctypes.byref(p),
                                part.ctypes.data_as(I32P), ctypes.byref(st))
        if rc != GK_OK:
            raise RuntimeError(f"gk_partition: {lib().gk_strerror(rc).decode()}: "
                               f"{lib().gk_last_error().decode()}")
        return part, st

    def partition_batch(self, graphs, ks, **kw):
        """gk_partition_batch over host graphs; ks[i] is graph i's part count.
        Returns (list of partitions, list of Stats)."""
        n = len(graphs)
        csrs = (Csr * n)(*[g.csr() for g in graphs])
        prm = (Params * n)(*[params(int(k), **kw) for k in ks])
        parts = [np.zeros(max(g.n, 1), np.int32) for g in graphs]
        outs = (I32P * n)(*[p.ctypes.data_as(I32P) for p in parts])
        st = (Stats * n)()
        rc = lib().gk_partition_batch(self._h, n, csrs, prm, outs, st)
        if rc != GK_OK:
            raise RuntimeError(f"gk_partition_batch: {lib().gk_strerror(rc).decode()}: "
                               f"{lib().gk_last_error().decode()}")
        return [p[:g.n] for p, g in zip(parts, graphs)], list(st)


class Server:
    """gk_serve_* request loop on its own gk_ctx (the server owns it)."""

    def __init__(self, device=0, hbm_budget=0, max_inflight=0, workers=0, pool_limit=0):
        self._ctx = Context(device)
        if pool_limit:
            lib().gk_ctx_set_pool_limit(self._ctx.handle, int(pool_limit))
        cfg = ServeConfig(hbm_budget, max_inflight, workers)
        self._h = c_void_p()
        rc = lib().gk_serve_open(self._ctx.handle, ctypes.byref(cfg), ctypes.byref(self._h))
        if rc != GK_OK:
            self._ctx.close()
            raise RuntimeError(f"gk_serve_open: {lib().gk_strerror(rc).decode()}")
        self._n = {}

    def submit(self, g: Graph, k, **kw) -> int:
        p = params(k, **kw)
        rid = c_int64(0)
        rc = lib().gk_serve_submit(self._h, ctypes.byref(g.csr()), ctypes.byref(p), ctypes.byref(rid))
        if rc != GK_OK:
            raise RuntimeError(f"gk_serve_submit: {lib().gk_strerror(rc).decode()}: "
                               f"{lib().gk_last_error().decode()}")
        self._n[rid.value] = g.n
        return rid.value

    def submit_many(self, graphs, ks, **kw) -> list:
        """Queues the graphs (graph i at k = ks[i]) together; their ids."""
        cnt = len(graphs)
        csrs = (Csr * cnt)(*[g.csr() for g in graphs])
        ps = (Params * cnt)(*[params(k, **kw) for k in ks])
        ids = (c_int64 * cnt)()
        rc = lib().gk_serve_submit_many(self._h, cnt, csrs, ps, ids)
        if rc != GK_OK:
            raise RuntimeError(f"gk_serve_submit_many: {lib().gk_strerror(rc).decode()}: "
                               f"{lib().gk_last_error().decode()}")
        for g, rid in zip(graphs, ids):
            self._n[rid] = g.n
        return list(ids)

    def poll(self, rid):
        """None while pending, else (part, Stats)."""
        n = self._n[rid]
        part = np.zeros(max(n, 1), np.int32)
        st = Stats()
        rc = lib().gk_serve_poll(self._h, rid, part.ctypes.data_as(I32P), ctypes.byref(st))
        if rc == 0:
            return None
        del self._n[rid]
        if rc < 0:
            raise RuntimeError(f"request {rid}: {lib().gk_strerror(rc).decode()}: "
                               f"{lib().gk_last_error().decode()}")
        return part[:n], st

    def wait(self, rid):
        n = self._n.pop(rid)
        part = np.zeros(max(n, 1), np.int32)
        st = Stats()
        rc = lib().gk_serve_wait(self._h, rid, part.ctypes.data_as(I32P), ctypes.byref(st))
        if rc != GK_OK:
            raise RuntimeError(f"request {rid}: {lib().gk_strerror(rc).decode()}: "
                               f"{lib().gk_last_error().decode()}")
        return part[:n], st

    def stats(self) -> ServeStats:
        s = ServeStats()
        lib().gk_serve_query(self._h, ctypes.byref(s))
        return s

    def close(self):
        rc = GK_OK
        if self._h:
            rc = lib().gk_serve_close(self._h)
            self._h = c_void_p()
            self._ctx.close()
        return rc

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()


def pack_graphs(graphs):
    """Concatenate host graphs into one CSR with global ids (the layout of
    gk_packed_csr); returns (Graph, v_off, e_off)."""
    v_off = np.zeros(len(graphs) + 1, np.int64)
    e_off = np.zeros(len(graphs) + 1, np.int64)
    for i, g in enumerate(graphs):
        v_off[i + 1] = v_off[i] + g.n
        e_off[i + 1] = e_off[i] + g.nnz
    xs = [g.xadj[:-1].astype(np.int64) + e_off[i] for i, g in enumerate(graphs)]
    xadj = np.concatenate(xs + [np.array([e_off[-1]])]) if graphs else np.zeros(1, np.int64)
    adj = np.concatenate([g.adjncy.astype(np.int64) + v_off[i] for i, g in enumerate(graphs)])
    w = np.concatenate([g.adjwgt for g in graphs])
    vw = np.concatenate([g.vwgt for g in graphs])
    return Graph(xadj, adj, w, vw), v_off.astype(np.int32), e_off.astype(np.int32)


def build(force=True, jobs=8):
    """Compile libgkway.so (hipcc, gfx950), libgkgen.so and the g-kway CLI."""
    cmd = ["make", "-C", HERE, f"-j{jobs}"] + (["-B"] if force else [])
    subprocess.run(cmd, check=True)
