"""The reference's own entry points on the GPU path, compared with the oracle:
the CLI with the reference's argument order (`g-kway graph k out`,
main/main.cu:11-17), which calls graph_partitioner() (graph_partitioner.hpp:222)
through the C++ mirror, and gk_partition_file; both write <cwd>/<out>.out as
gk::Graph::dump_result does (graph.hpp:227-240)."""
import ctypes
import json
import os
import subprocess

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def read_out(path, n):
    with open(path) as f:
        lines = f.read().splitlines()
    assert len(lines) == n
    return np.array([int(x) for x in lines], np.int32)


@pytest.mark.parametrize("k", [2, 4, 64])
def test_cli_sample_graph(gk, oracle, tmp_path, k):
    path = os.path.join(GOLDEN, "delaunay_n11.graph")
    print(f"cli delaunay_n11 k={k}", flush=True)
    r = subprocess.run([gk.CLI, path, str(k), "out", "--json"], cwd=tmp_path, capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 0, r.stderr
    info = json.loads(r.stdout.strip().splitlines()[-1])
    g = gk.read_metis(path)
    part = read_out(tmp_path / "out.out", g.n)
    opart, ost = oracle.partition(g, k)
    assert np.array_equal(part, opart)
    assert info["cut"] == ost.cut and info["k"] == k and info["n"] == g.n
    assert info["imbalance"] <= 1.03 + 1e-9


def test_cli_weighted_file(gk, oracle, tmp_path):
    # A weighted DAG written as METIS fmt 011, partitioned by the CLI.
    g = gk.dag(20000, 40, 5)
    path = str(tmp_path / "dag.graph")
    assert gk.lib().gk_graph_write(path.encode(), ctypes.byref(g.csr())) == 0
    r = subprocess.run([gk.CLI, path, "8", "dag"], cwd=tmp_path, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    part = read_out(tmp_path / "dag.out", g.n)
    opart, _ = oracle.partition(g, 8)
    assert np.array_equal(part, opart)


def test_cli_errors(gk, tmp_path):
    r = subprocess.run([gk.CLI, str(tmp_path / "missing.graph"), "4", "x"], cwd=tmp_path, capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 1 and "missing.graph" in r.stderr
    r = subprocess.run([gk.CLI, "only-one-arg"], capture_output=True, text=True, timeout=120)
    assert r.returncode == 2


def test_partition_file(gk, oracle, ctx, tmp_path):
    path = os.path.join(GOLDEN, "delaunay_n11.graph")
    out = str(tmp_path / "pf")
    st = gk.Stats()
    L = gk.lib()
    assert L.gk_partition_file(ctx.handle, path.encode(), out.encode(), 8, ctypes.byref(st)) == 0
    g = gk.read_metis(path)
    opart, ost = oracle.partition(g, 8)
    assert np.array_equal(read_out(out + ".out", g.n), opart)
    assert st.cut == ost.cut
    assert L.gk_partition_file(ctx.handle, b"/nonexistent/x.graph", out.encode(), 8, None) == -2


def test_mapped_binary_csr(gk, oracle, ctx, tmp_path):
    # A graph stored as a binary CSR file, mapped without a copy
    # (gk_graph_map_bin) and handed to gk_partition straight from the
    # mapping; the same file through gk_partition_file and the CLI, which
    # read binary files as they read METIS text.  All equal the oracle.
    g = gk.dag(30000, 60, 9)
    path = str(tmp_path / "dag.gkb")
    gk.write_bin(g, path)
    c = gk.Csr()
    h = ctypes.c_void_p()
    L = gk.lib()
    assert L.gk_graph_map_bin(path.encode(), 1, ctypes.byref(c), ctypes.byref(h)) == 0
    try:
        assert c.n == g.n and c.nnz == g.nnz
        part = np.zeros(g.n, np.int32)
        st = gk.Stats()
        p = gk.params(16)
        print("mapped binary csr: dag 30000 k=16", flush=True)
        assert L.gk_partition(ctx.handle, ctypes.byref(c), ctypes.byref(p), part.ctypes.data_as(gk.I32P),
                              ctypes.byref(st)) == 0
    finally:
        L.gk_graph_unmap(h)
    opart, ost = oracle.partition(g, 16)
    assert np.array_equal(part, opart) and st.cut == ost.cut
    out = str(tmp_path / "pf")
    assert L.gk_partition_file(ctx.handle, path.encode(), out.encode(), 16, None) == 0
    assert np.array_equal(read_out(out + ".out", g.n), opart)
    r = subprocess.run([gk.CLI, path, "16", "cli"], cwd=tmp_path, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    assert np.array_equal(read_out(tmp_path / "cli.out", g.n), opart)
