"""Device-memory plumbing for GPU tests: torch tensors hold the buffers, the
C ABI receives their addresses."""
import ctypes

import numpy as np
import torch

from _pkg import gk

I32P = gk.I32P


def dev(a):
    return torch.from_numpy(np.ascontiguousarray(a, dtype=np.int32)).cuda()


def ptr(t):
    return ctypes.cast(t.data_ptr(), I32P)


def host(t):
    torch.cuda.synchronize()
    return t.cpu().numpy()


class DevGraph:
    def __init__(self, g):
        self.g = g
        self.xadj = dev(g.xadj)
        self.adj = dev(g.adjncy if g.nnz else np.zeros(1, np.int32))
        self.w = dev(g.adjwgt if g.nnz else np.zeros(1, np.int32))
        self.vw = dev(g.vwgt if g.n else np.zeros(1, np.int32))
        torch.cuda.synchronize()
        c = gk.Csr()
        c.n = g.n
        c.nnz = g.nnz
        c.xadj = ptr(self.xadj)
        c.adjncy = ptr(self.adj)
        c.adjwgt = ptr(self.w)
        c.vwgt = ptr(self.vw)
        self.csr = c

    def ref(self):
        return ctypes.byref(self.csr)
