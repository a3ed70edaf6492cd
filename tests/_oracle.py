"""ctypes bindings of the CPU oracle (oracle/liboracle.so) -- test code only.

The oracle is the checker: tests compare the HIP path with it.  The initial
partition of the coarsest graph is passed to the oracle as a callback (the
reference calls METIS there); the oracle's own restatement of the build's
initial partitioner, oracle_initpart, is the default, so an oracle run uses
no product code.  The product's gk_initpart is compared with it directly.
"""
from __future__ import annotations

import ctypes
import os
import subprocess
from ctypes import POINTER, c_int32, c_int64

import numpy as np

from _pkg import gk

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIBORACLE = os.path.join(ROOT, "oracle", "liboracle.so")
I32P = gk.I32P
INITPART_FN = ctypes.CFUNCTYPE(ctypes.c_int, POINTER(gk.Csr), c_int32, c_int64, c_int32, c_int32,
                               I32P, POINTER(c_int64))

_O = None


def lib():
    global _O
    if _O is None:
        if not os.path.exists(LIBORACLE):
            subprocess.run(["make", "-C", os.path.join(ROOT, "oracle")], check=True)
        O = ctypes.CDLL(LIBORACLE)
        O.oracle_partition.argtypes = [POINTER(gk.Csr), POINTER(gk.Params), INITPART_FN, I32P,
                                       POINTER(gk.Stats)]
        O.oracle_part_cap.restype = c_int64
        O.oracle_part_cap.argtypes = [c_int64, c_int32, c_int32]
        O.oracle_coarsen_floor.restype = c_int32
        O.oracle_select.argtypes = [POINTER(gk.Csr), I32P, I32P, I32P, I32P, I32P, c_int64, c_int32,
                                    I32P, I32P, I32P]
        O.oracle_balance.argtypes = [POINTER(gk.Csr), c_int32, I32P, c_int32, I32P, I32P, I32P, c_int64]
        O.oracle_apply.restype = c_int64
        O.oracle_apply.argtypes = [POINTER(gk.Csr), I32P, c_int32, I32P, I32P, I32P, I32P]
        O.oracle_cut.argtypes = [POINTER(gk.Csr), c_int32, I32P, POINTER(c_int64), POINTER(c_int64)]
        O.oracle_initpart.argtypes = [POINTER(gk.Csr), c_int32, c_int64, c_int32, c_int32, I32P, POINTER(c_int64)]
        _O = O
    return _O


def _p(a):
    return np.ascontiguousarray(a, dtype=np.int32).ctypes.data_as(I32P)


def initpart_callback(product=False):
    """The oracle's restatement of the initial partitioner as an oracle
    callback (product=True: the product's gk_initpart instead)."""
    if product:
        return ctypes.cast(gk.lib().gk_initpart, INITPART_FN)
    return ctypes.cast(lib().oracle_initpart, INITPART_FN)


def initpart(g, k, cap, seed=1):
    """oracle_initpart on a host graph: (part, cut)."""
    part = np.zeros(max(g.n, 1), np.int32)
    cut = c_int64(0)
    rc = lib().oracle_initpart(ctypes.byref(g.csr()), k, cap, seed, 1, part.ctypes.data_as(I32P), ctypes.byref(cut))
    if rc != 0:
        raise RuntimeError(f"oracle_initpart failed: {rc}")
    return part[:g.n], cut.value


def partition(g: "gk.Graph", k, product_initpart=False, **kw):
    p = gk.params(k, **kw)
    part = np.zeros(g.n, np.int32)
    st = gk.Stats()
    cb = initpart_callback(product_initpart)
    rc = lib().oracle_partition(ctypes.byref(g.csr()), ctypes.byref(p), cb, part.ctypes.data_as(I32P),
                                ctypes.byref(st))
    if rc != 0:
        raise RuntimeError(f"oracle_partition failed: {rc}")
    return part, st


def match(g):
    cand = np.zeros(g.n, np.int32)
    lib().oracle_match(ctypes.byref(g.csr()), cand.ctypes.data_as(I32P))
    return cand


def group(cand):
    n = len(cand)
    label = np.zeros(n, np.int32)
    stamp = np.zeros(n, np.int32)
    cand = np.ascontiguousarray(cand, np.int32)
    r = lib().oracle_group(n, cand.ctypes.data_as(I32P), label.ctypes.data_as(I32P), stamp.ctypes.data_as(I32P))
    return label, stamp, r


def cluster(label, stamp, vwgt, G):
    n = len(label)
    order = np.zeros(n, np.int32)
    cmap = np.zeros(n, np.int32)
    cvw = np.zeros(max(n, 1), np.int32)
    label = np.ascontiguousarray(label, np.int32)
    stamp = np.ascontiguousarray(stamp, np.int32)
    vwgt = np.ascontiguousarray(vwgt, np.int32)
    nc = lib().oracle_cluster(n, label.ctypes.data_as(I32P), stamp.ctypes.data_as(I32P),
                              vwgt.ctypes.data_as(I32P), G, order.ctypes.data_as(I32P),
                              cmap.ctypes.data_as(I32P), cvw.ctypes.data_as(I32P))
    return order, cmap, cvw[:nc], nc


def contract(g, cmap, nc):
    cx = np.zeros(nc + 1, np.int32)
    ca = np.zeros(max(g.nnz, 1), np.int32)
    cw = np.zeros(max(g.nnz, 1), np.int32)
    cmap = np.ascontiguousarray(cmap, np.int32)
    e = lib().oracle_contract(ctypes.byref(g.csr()), cmap.ctypes.data_as(I32P), nc, cx.ctypes.data_as(I32P),
                              ca.ctypes.data_as(I32P), cw.ctypes.data_as(I32P))
    return cx, ca[:e], cw[:e]


def gain(g, k, part, upd, bnd=None, gain_=None, dest=None):
    n = g.n
    bnd = np.zeros(n, np.int32) if bnd is None else bnd.copy()
    gain_ = np.zeros(n, np.int32) if gain_ is None else gain_.copy()
    dest = np.zeros(n, np.int32) if dest is None else dest.copy()
    part = np.ascontiguousarray(part, np.int32)
    upd = np.ascontiguousarray(upd, np.int32)
    lib().oracle_gain(ctypes.byref(g.csr()), k, part.ctypes.data_as(I32P), upd.ctypes.data_as(I32P),
                      bnd.ctypes.data_as(I32P), gain_.ctypes.data_as(I32P), dest.ctypes.data_as(I32P))
    return bnd, gain_, dest


def select(g, part, bnd, gain_, dest, pw, cap, M, upd):
    upd = upd.copy()
    moves = np.zeros(max(M, 1), np.int32)
    nsel = c_int32(0)
    arrs = [np.ascontiguousarray(a, np.int32) for a in (part, bnd, gain_, dest, pw)]
    B = lib().oracle_select(ctypes.byref(g.csr()), *[a.ctypes.data_as(I32P) for a in arrs], cap, M,
                            upd.ctypes.data_as(I32P), moves.ctypes.data_as(I32P), ctypes.byref(nsel))
    return moves[:B], nsel.value, upd


def balance(g, k, moves, part, dest, pw, cap):
    pw = np.ascontiguousarray(pw, np.int32).copy()
    moves = np.ascontiguousarray(moves, np.int32)
    part = np.ascontiguousarray(part, np.int32)
    dest = np.ascontiguousarray(dest, np.int32)
    pos = lib().oracle_balance(ctypes.byref(g.csr()), k, moves.ctypes.data_as(I32P), len(moves),
                               part.ctypes.data_as(I32P), dest.ctypes.data_as(I32P), pw.ctypes.data_as(I32P),
                               cap)
    return pos, pw


def apply(g, moves, pos, gain_, dest, part, upd):
    part = np.ascontiguousarray(part, np.int32).copy()
    upd = np.ascontiguousarray(upd, np.int32).copy()
    moves = np.ascontiguousarray(moves, np.int32)
    gain_ = np.ascontiguousarray(gain_, np.int32)
    dest = np.ascontiguousarray(dest, np.int32)
    s = lib().oracle_apply(ctypes.byref(g.csr()), moves.ctypes.data_as(I32P), pos, gain_.ctypes.data_as(I32P),
                           dest.ctypes.data_as(I32P), part.ctypes.data_as(I32P), upd.ctypes.data_as(I32P))
    return s, part, upd


def cut(g, k, part):
    c = c_int64(0)
    pw = np.zeros(k, np.int64)
    part = np.ascontiguousarray(part, np.int32)
    lib().oracle_cut(ctypes.byref(g.csr()), k, part.ctypes.data_as(I32P), ctypes.byref(c),
                     pw.ctypes.data_as(POINTER(c_int64)))
    return c.value, pw
