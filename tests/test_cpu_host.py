"""Host-side components on CPU: the C ABI library (symbols, status strings,
parameters), METIS file I/O and validation, the initial partitioner, and the
workload generators.  No GPU calls."""
import ctypes
import os
import re
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(ROOT, "tests", "golden")


# ------------------------------------------------------------------ C ABI
def declared(header):
    txt = open(os.path.join(ROOT, "include", header)).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b((?:gk|gkgen)_[a-z0-9_]+)\s*\(", txt)))


@pytest.mark.parametrize("header,lib", [("gkway.h", "libgkway.so"), ("gkgen.h", "libgkgen.so")])
def test_library_exports_every_declared_symbol(header, lib):
    path = os.path.join(ROOT, "dac_g-kway_amd", lib)
    ctypes.CDLL(path)  # loads without a GPU
    out = subprocess.run(["nm", "-D", "--defined-only", path], capture_output=True, text=True).stdout
    exported = set(re.findall(r"\bT (\w+)", out))
    names = declared(header)
    assert names, header
    missing = [s for s in names if s not in exported]
    assert not missing, missing


def test_header_constants_match_binding(gk):
    # Flag bits, stage ids and coarsening policies: the Python binding's
    # values are the header's.
    txt = open(os.path.join(ROOT, "include", "gkway.h")).read()
    defs = dict(re.findall(r"#define\s+(GK_(?:FLAG|ST|COARSEN)_[A-Z_]+)\s+(-?\d+)", txt))
    assert {"GK_FLAG_CHECK", "GK_FLAG_TRACE", "GK_FLAG_STAGE_TIME", "GK_FLAG_GRAPH_ROUNDS",
            "GK_FLAG_AUTOTUNE", "GK_FLAG_PERSISTENT"} <= set(defs)
    checked = 0
    for name, val in defs.items():
        if hasattr(gk, name):
            assert getattr(gk, name) == int(val), name
            checked += 1
    assert checked >= 6
    flags = [int(v) for k, v in defs.items() if k.startswith("GK_FLAG_")]
    assert all(f & (f - 1) == 0 for f in flags) and len(set(flags)) == len(flags)


def test_status_strings_and_defaults(gk):
    L = gk.lib()
    for code in range(0, -10, -1):
        assert L.gk_strerror(code)
    p = gk.Params()
    L.gk_params_default(ctypes.byref(p), 64)
    assert (p.k, p.ubfactor, p.max_group, p.max_moves, p.coarsen_policy, p.seed) == (64, 30, 6, 1024, 0, 1)


def test_cap_and_floor_match_oracle(gk, oracle):
    L, O = gk.lib(), oracle.lib()
    for T, k, ub in [(2048, 4, 30), (4194304, 64, 30), (10, 3, 30), (7, 8, 30), (10 ** 9, 256, 50)]:
        assert L.gk_part_cap(T, k, ub) == O.oracle_part_cap(T, k, ub)
    # cap is at least ceil(T/k) and exactly floor(T*1.03/k) when that is larger
    assert L.gk_part_cap(2048, 4, 30) == 527
    assert L.gk_part_cap(7, 8, 30) == 1
    for n0, k, pol, ov in [(4194304, 64, 0, 0), (2048, 4, 0, 0), (10 ** 6, 256, 0, 0), (5000, 4, 1, 0), (100, 2, 0, 77)]:
        assert L.gk_coarsen_floor(n0, k, pol, ov) == O.oracle_coarsen_floor(n0, k, pol, ov)
    assert L.gk_coarsen_floor(4194304, 64, 0, 0) == 240 * 64
    assert L.gk_coarsen_floor(1 << 20, 64, 0, 0) == (1 << 20) // 120


def test_serve_footprint(gk):
    L = gk.lib()
    p = gk.params(8)
    assert L.gk_serve_footprint(0, 0, ctypes.byref(p)) == 0
    # At or below the coarsening floor there is one level: CSR + scratch.
    n, nnz = 200, 1000
    assert L.gk_coarsen_floor(n, 8, 0, 0) >= n
    csr = 4 * (n + 1) + 8 * nnz + 4 * n
    assert L.gk_serve_footprint(n, nnz, ctypes.byref(p)) == csr + 96 * n + 4096
    # Above it, the CSR term is multiplied by the estimated level count
    # (levels shrink at least 3x down to the floor).
    n, nnz = 1 << 20, 6 << 20
    fl = L.gk_coarsen_floor(n, 8, 0, 0)
    lv, m = 1, n
    while m > fl:
        m //= 3
        lv += 1
    csr = 4 * (n + 1) + 8 * nnz + 4 * n
    assert L.gk_serve_footprint(n, nnz, ctypes.byref(p)) == csr * lv + 96 * n + 4096
    assert L.gk_serve_footprint(n, nnz + 10, ctypes.byref(p)) > L.gk_serve_footprint(n, nnz, ctypes.byref(p))


def test_pack_graphs_layout(gk):
    gs = [gk.Graph([0, 1, 2], [1, 0]), gk.Graph([0], []), gk.Graph([0, 2, 3, 4], [1, 2, 0, 0], [5, 6, 5, 6], [1, 2, 3])]
    pg, v_off, e_off = gk.pack_graphs(gs)
    assert list(v_off) == [0, 2, 2, 5] and list(e_off) == [0, 2, 2, 6]
    assert list(pg.xadj) == [0, 1, 2, 4, 5, 6]
    assert list(pg.adjncy) == [1, 0, 3, 4, 2, 2]
    assert list(pg.adjwgt) == [1, 1, 5, 6, 5, 6] and list(pg.vwgt) == [1, 1, 1, 2, 3]
    for i, g in enumerate(gs):
        xs = pg.xadj[v_off[i]:v_off[i + 1] + 1] - e_off[i]
        assert np.array_equal(xs, g.xadj)
        assert np.array_equal(pg.adjncy[e_off[i]:e_off[i + 1]] - v_off[i], g.adjncy)


def test_batch_and_serve_argument_checks(gk):
    L = gk.lib()
    assert L.gk_partition_batch(None, 1, None, None, None, None) == -1
    assert L.gk_partition_packed(None, None, None, None, None, None) == -1
    h = ctypes.c_void_p()
    assert L.gk_serve_open(None, None, ctypes.byref(h)) == -1
    assert L.gk_serve_poll(None, 0, None, None) == -1


def test_no_device_is_a_status_not_a_crash(gk):
    L = gk.lib()
    h = ctypes.c_void_p()
    rc = L.gk_ctx_create(0, ctypes.byref(h))
    if rc == 0:
        L.gk_ctx_destroy(h)
        pytest.skip("a HIP device is present")
    assert rc == -9  # GK_ERR_NODEVICE
    assert b"device" in L.gk_last_error()


# ------------------------------------------------------------------ METIS I/O
def write(tmp_path, name, text):
    p = tmp_path / name
    p.write_text(text)
    return str(p)


def test_read_sample_graph(gk):
    g = gk.read_metis(os.path.join(GOLDEN, "delaunay_n11.graph"))
    assert (g.n, g.nnz) == (2048, 12254)
    d = g.degrees()
    assert d.min() == 3 and d.max() == 13
    assert gk.lib().gk_graph_check(ctypes.byref(g.csr()), 1) == 0


def test_read_formats(gk, tmp_path):
    # triangle 1-2-3 plus isolated vertex 4, with comments and both weight kinds
    plain = "% comment\n4 3\n2 3\n1 3\n1 2\n\n"
    g = gk.read_metis(write(tmp_path, "a.graph", plain))
    assert list(g.xadj) == [0, 2, 4, 6, 6] and list(g.adjncy) == [1, 2, 0, 2, 0, 1]
    ew = "4 3 1\n2 5 3 7\n1 5 3 2\n1 7 2 2\n\n"
    g = gk.read_metis(write(tmp_path, "b.graph", ew))
    assert list(g.adjwgt) == [5, 7, 5, 2, 7, 2]
    vwew = "4 3 011\n3 2 5 3 7\n1 1 5 3 2\n2 1 7 2 2\n4\n"
    g = gk.read_metis(write(tmp_path, "c.graph", vwew))
    assert list(g.vwgt) == [3, 1, 2, 4] and list(g.adjwgt) == [5, 7, 5, 2, 7, 2]
    vsize = "3 2 100\n9 2\n9 1 3\n9 2\n"
    g = gk.read_metis(write(tmp_path, "d.graph", vsize))
    assert list(g.adjncy) == [1, 0, 2, 1]


@pytest.mark.parametrize("text", [
    "",                          # empty
    "3\n",                       # header without edge count
    "3 2\n2\n1 3\n",             # fewer lines than vertices
    "2 1\n2\n1\n1\n",            # extra non-blank line
    "2 1\n3\n1\n",               # neighbour out of range
    "2 2\n2\n1\n",               # entry count != 2m
    "2 1 1\n2\n1 1\n",           # missing edge weight
    "2 1 2\n2\n1\n",             # bad fmt digit
    "2 1\n2 x\n1\n",             # junk token
])
def test_read_rejects_malformed(gk, tmp_path, text):
    with pytest.raises(RuntimeError):
        gk.read_metis(write(tmp_path, "bad.graph", text))


def test_check_rejects_invalid_graphs(gk):
    L = gk.lib()

    def check(xadj, adj, w=None, vw=None, sym=1):
        return L.gk_graph_check(ctypes.byref(gk.Graph(xadj, adj, w, vw).csr()), sym)

    assert check([0, 1, 2], [1, 0]) == 0
    assert check([0, 1, 1], [1]) == -4          # asymmetric
    assert check([0, 1, 2], [0, 1]) == -4       # self loops
    assert check([0, 1, 2], [1, 0], [2, 3]) == -4  # unequal weights
    assert check([0, 1, 2], [1, 0], [0, 0]) == -4  # zero weight
    assert check([0, 2, 2], [1, 1], sym=0) == 0    # duplicates only caught with symmetry check
    assert check([0, 2, 2], [1, 1]) == -4


def test_write_read_round_trip(gk, tmp_path):
    g = gk.dag(500, 10, 4)
    path = str(tmp_path / "rt.graph")
    assert gk.lib().gk_graph_write(path.encode(), ctypes.byref(g.csr())) == 0
    h = gk.read_metis(path)
    for a in ("xadj", "adjncy", "adjwgt", "vwgt"):
        assert np.array_equal(getattr(g, a), getattr(h, a))


def _same_graph(a, b, weights=True):
    assert (a.n, a.nnz) == (b.n, b.nnz)
    assert np.array_equal(a.xadj, b.xadj) and np.array_equal(a.adjncy, b.adjncy)
    if weights:
        assert np.array_equal(a.adjwgt, b.adjwgt) and np.array_equal(a.vwgt, b.vwgt)


def test_binary_csr_round_trip(gk, tmp_path):
    # weighted graph: mapped without copying, and through gk_graph_read
    g = gk.dag(3000, 12, 5)
    path = str(tmp_path / "g.gkcsr")
    gk.write_bin(g, path)
    _same_graph(gk.map_bin(path), g)
    _same_graph(gk.map_bin(path, verify=False), g)
    _same_graph(gk.read_metis(path), g)
    # the METIS sample graph, text -> binary -> back
    s = gk.read_metis(os.path.join(GOLDEN, "delaunay_n11.graph"))
    p2 = str(tmp_path / "s.gkcsr")
    gk.write_bin(s, p2)
    _same_graph(gk.read_metis(p2), s)
    # sections start on 64-byte boundaries
    raw = open(path, "rb").read()
    assert raw[:5] == b"GKCSR" and len(raw) >= 64 + 4 * (g.n + 1) + 8 * g.nnz + 4 * g.n


def test_binary_csr_without_weights(gk, tmp_path):
    L = gk.lib()
    x = np.array([0, 1, 3, 4], np.int32)
    a = np.array([1, 0, 2, 1], np.int32)
    c = gk.Csr()
    c.n, c.nnz = 3, 4
    c.xadj, c.adjncy = x.ctypes.data_as(gk.I32P), a.ctypes.data_as(gk.I32P)
    path = str(tmp_path / "u.gkcsr")
    assert L.gk_graph_write_bin(path.encode(), ctypes.byref(c)) == 0
    m = gk.Csr()
    h = ctypes.c_void_p()
    assert L.gk_graph_map_bin(path.encode(), 1, ctypes.byref(m), ctypes.byref(h)) == 0
    assert (m.n, m.nnz) == (3, 4) and not m.adjwgt and not m.vwgt
    assert list(np.ctypeslib.as_array(m.adjncy, (4,))) == [1, 0, 2, 1]
    L.gk_graph_unmap(h)
    r = gk.read_metis(path)  # gk_graph_read fills unit weights
    assert list(r.adjwgt) == [1, 1, 1, 1] and list(r.vwgt) == [1, 1, 1]


def test_binary_csr_rejects_damage(gk, tmp_path):
    g = gk.delaunay(500, 3)
    path = str(tmp_path / "d.gkcsr")
    gk.write_bin(g, path)
    raw = bytearray(open(path, "rb").read())
    flipped = bytearray(raw)
    adj_off = (64 + 4 * (g.n + 1) + 63) // 64 * 64  # sections start on 64-byte boundaries
    flipped[adj_off + 13] ^= 0x40  # inside adjncy
    open(path, "wb").write(bytes(flipped))
    with pytest.raises(RuntimeError, match="checksum"):
        gk.map_bin(path)
    gk.map_bin(path, verify=False)  # unchecked mapping still works
    open(path, "wb").write(bytes(raw[: len(raw) - 40]))
    with pytest.raises(RuntimeError, match="truncated"):
        gk.map_bin(path)
    bad = bytearray(raw)
    bad[0] = ord("X")
    open(path, "wb").write(bytes(bad))
    with pytest.raises(RuntimeError, match="magic"):
        gk.map_bin(path)
    with pytest.raises(RuntimeError):
        gk.map_bin(str(tmp_path / "missing.gkcsr"))


def test_write_partition_format(gk, tmp_path):
    part = np.array([3, 0, 2, 1], np.int32)
    path = str(tmp_path / "x.out")
    assert gk.lib().gk_write_partition(path.encode(), part.ctypes.data_as(gk.I32P), 4) == 0
    assert open(path).read() == "3\n0\n2\n1\n"


# ------------------------------------------------------------------ initial partitioner
@pytest.mark.parametrize("k", [2, 3, 7, 16, 64])
def test_initpart_valid_balanced_deterministic(gk, k):
    g = gk.delaunay(4000, 3)
    cap = gk.lib().gk_part_cap(int(g.vwgt.sum()), k, 30)
    p1, c1 = gk.initpart(g, k, cap, threads=1)
    p8, c8 = gk.initpart(g, k, cap, threads=8)
    assert np.array_equal(p1, p8) and c1 == c8
    assert p1.min() == 0 and p1.max() == k - 1
    pw = np.bincount(p1, weights=g.vwgt, minlength=k)
    assert pw.max() <= cap
    src = np.repeat(np.arange(g.n), g.degrees())
    assert c1 == int(g.adjwgt[(p1[src] != p1[g.adjncy]) & (src < g.adjncy)].sum())


def test_initpart_weighted(gk):
    g = gk.dag(6000, 30, 8)
    k = 32
    cap = gk.lib().gk_part_cap(int(g.vwgt.sum()), k, 30)
    p, c = gk.initpart(g, k, cap)
    assert np.bincount(p, weights=g.vwgt, minlength=k).max() <= cap


def test_initpart_k1_and_tiny(gk):
    g = gk.Graph([0, 1, 2], [1, 0])
    p, c = gk.initpart(g, 1)
    assert list(p) == [0, 0] and c == 0
    p, c = gk.initpart(g, 2)
    assert sorted(p) == [0, 1] and c == 1


def _initpart_corpus(gk):
    """Graphs of every family and size class the V-cycle hands to the initial
    partitioner: tiny, mesh-like, weighted DAGs, road chains, the sample graph."""
    out = [gk.read_metis(os.path.join(GOLDEN, "delaunay_n11.graph"))]
    for n in (3, 40, 700, 5000):
        out += [gk.delaunay(n, n + 1), gk.dag(n, max(2, int(n ** 0.5)), n + 2), gk.road(n, n + 3)]
    return out


@pytest.mark.parametrize("threads", [1, 8])
def test_initpart_equals_oracle_restatement(gk, oracle, threads):
    # The product's initial partitioner (indexed heaps, fork-join pool,
    # parallel contraction) against the oracle's single-threaded restatement
    # of its rules: identical partitions and cuts for every graph, k and seed.
    for g in _initpart_corpus(gk):
        for k in (2, 3, 8, 13, 64):
            for seed in (1, 7):
                cap = gk.lib().gk_part_cap(int(g.vwgt.sum()), k, 30)
                p, c = gk.initpart(g, k, cap, seed=seed, threads=threads)
                op, oc = oracle.initpart(g, k, cap, seed)
                assert np.array_equal(p, op), (g.n, k, seed)
                assert c == oc


# Recorded METIS 5.2.1 cuts on the sample graph (see test_cpu_oracle.py).
@pytest.mark.parametrize("k,metis", [(2, 91), (4, 205), (8, 349), (64, 2083)])
def test_initpart_quality_on_sample_graph(gk, k, metis):
    g = gk.read_metis(os.path.join(GOLDEN, "delaunay_n11.graph"))
    _, c = gk.initpart(g, k)
    assert c <= 1.05 * metis


# ------------------------------------------------------------------ generators
def test_delaunay_matches_scipy(gk):
    spatial = pytest.importorskip("scipy.spatial")
    G = gk.gen()
    for n, seed in [(50, 1), (2048, 1), (30000, 2)]:
        x = np.zeros(n, np.int32)
        y = np.zeros(n, np.int32)
        G.gkgen_points(n, seed, x.ctypes.data_as(gk.I32P), y.ctypes.data_as(gk.I32P))
        tri = spatial.Delaunay(np.stack([x, y], 1).astype(np.float64))
        ref = set()
        for s in tri.simplices:
            for a, b in ((s[0], s[1]), (s[1], s[2]), (s[0], s[2])):
                ref.add((min(a, b), max(a, b)))
        g = gk.delaunay(n, seed)
        src = np.repeat(np.arange(g.n), g.degrees())
        m = src < g.adjncy
        assert set(zip(src[m].tolist(), g.adjncy[m].tolist())) == ref


def test_generators_structure(gk):
    L = gk.lib()
    d = gk.delaunay(5000, 4)
    assert L.gk_graph_check(ctypes.byref(d.csr()), 1) == 0
    assert abs(d.m - 3 * 5000) < 200  # planar triangulation: m ~ 3n - hull
    r = gk.road(20000, 2)
    assert L.gk_graph_check(ctypes.byref(r.csr()), 1) == 0
    assert r.m == round(1.05 * 20000)
    # connected (the EMST spans every point)
    seen = np.zeros(r.n, bool)
    stack = [0]
    seen[0] = True
    while stack:
        v = stack.pop()
        for u in r.adjncy[r.xadj[v]:r.xadj[v + 1]]:
            if not seen[u]:
                seen[u] = True
                stack.append(u)
    assert seen.all()
    g = gk.dag(20000, 100, 3)
    assert L.gk_graph_check(ctypes.byref(g.csr()), 1) == 0
    assert g.adjwgt.min() >= 1 and g.adjwgt.max() <= 100
    assert g.vwgt.min() >= 1 and g.vwgt.max() <= 4
    layer = np.arange(g.n) * 100 // g.n
    src = np.repeat(np.arange(g.n), g.degrees())
    gap = np.abs(layer[src] - layer[g.adjncy])
    assert gap.min() >= 1 and gap.max() <= 3
    # same seed, same graph
    g2 = gk.dag(20000, 100, 3)
    assert np.array_equal(g.adjncy, g2.adjncy) and np.array_equal(g.adjwgt, g2.adjwgt)


# ------------------------------------------------------------------ C++ mirror of the reference API
GRAPH_PROBE = r'''
#include <cstdio>
#include "gkway/graph.hpp"
int main(int argc, char** argv) {
  gk::Graph g(argv[1]);
  const size_t n = g.get_num_vertex(), e = g.get_num_edge();
  std::printf("%zu %zu %d\n", n, e, g.get_total_vertex_weight());
  for (size_t v = 0; v <= n; ++v) std::printf("%d ", g.get_adjp()[v]);
  std::printf("\n");
  for (size_t j = 0; j < e; ++j) std::printf("%d ", g.get_adjncy()[j]);
  std::printf("\n");
  for (size_t j = 0; j < e; ++j) std::printf("%d ", g.get_adjncy_source()[j]);
  std::printf("\n");
  for (size_t j = 0; j < e; ++j) std::printf("%d ", g.get_adjwgt()[j]);
  std::printf("\n");
  for (size_t v = 0; v < n; ++v) std::printf("%d ", g.get_vwgt()[v]);
  std::printf("\n");
  std::vector<int> part(n);
  for (size_t v = 0; v < n; ++v) part[v] = int(v % 3);
  g.dump_result(part, argv[2]);
  return 0;
}
'''


@pytest.fixture(scope="module")
def graph_probe(tmp_path_factory):
    d = tmp_path_factory.mktemp("probe")
    src = d / "probe.cpp"
    src.write_text(GRAPH_PROBE)
    exe = d / "probe"
    lib = os.path.join(ROOT, "dac_g-kway_amd")
    subprocess.run(["g++", "-std=c++17", "-O1", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe),
                    "-L", lib, "-lgkway", f"-Wl,-rpath,{lib}"], check=True)
    return exe


def test_cpp_graph_mirrors_reference_conventions(gk, graph_probe, tmp_path):
    """gk::Graph on the sample graph: the reference's getters (graph.hpp:146-183)
    -- xadj, 1-based neighbour ids as in the file, 1-based source ids per
    entry, unit edge and vertex weights, 2 x m entries -- and dump_result
    writing <cwd>/<name>.out."""
    path = os.path.join(GOLDEN, "delaunay_n11.graph")
    out = subprocess.run([str(graph_probe), path, "probe_out"], cwd=tmp_path, capture_output=True, text=True,
                         check=True).stdout.splitlines()
    n, e, tw = map(int, out[0].split())
    g = gk.read_metis(path)
    assert (n, e, tw) == (2048, 2 * 6127, 2048) == (g.n, g.nnz, int(g.vwgt.sum()))
    arr = [np.array(line.split(), dtype=np.int64) for line in out[1:6]]
    assert np.array_equal(arr[0], g.xadj)
    assert np.array_equal(arr[1], g.adjncy + 1)
    assert np.array_equal(arr[2], np.repeat(np.arange(g.n), g.degrees()) + 1)
    assert (arr[3] == 1).all() and (arr[4] == 1).all()
    # the file's own first row, 1-based as written
    with open(path) as f:
        rows = [ln for ln in f.read().splitlines() if not ln.startswith("%")]
    assert list(arr[1][: g.xadj[1]]) == [int(t) for t in rows[1].split()]
    assert (tmp_path / "probe_out.out").read_text() == "".join(f"{v % 3}\n" for v in range(n))


def test_cpp_graph_weighted_file(gk, graph_probe, tmp_path):
    # fmt 011: vertex and edge weights are read and returned as given
    p = tmp_path / "w.graph"
    p.write_text("3 2 011\n5 2 7\n1 1 7 3 9\n2 2 9\n")
    out = subprocess.run([str(graph_probe), str(p), "w"], cwd=tmp_path, capture_output=True, text=True,
                         check=True).stdout.splitlines()
    assert out[0].split() == ["3", "4", "8"]
    assert out[2].split() == ["2", "1", "3", "2"]
    assert out[3].split() == ["1", "2", "2", "3"]
    assert out[4].split() == ["7", "7", "9", "9"]
    assert out[5].split() == ["5", "1", "2"]


@pytest.mark.skipif(not os.path.exists("/dev/full"), reason="needs /dev/full")
def test_write_failures_are_reported(gk):
    # Writes that fail when the buffer is flushed (no space left) return GK_ERR_IO.
    part = np.zeros(1 << 20, np.int32)
    L = gk.lib()
    assert L.gk_write_partition(b"/dev/full", part.ctypes.data_as(gk.I32P), len(part)) == -2
    g = gk.delaunay(3000, 1)
    assert L.gk_graph_write(b"/dev/full", ctypes.byref(g.csr())) == -2
