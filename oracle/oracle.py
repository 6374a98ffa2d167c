"""ctypes wrapper of the FP64 CPU oracle (oracle/ace_oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, by ``__graft_entry__.smoke()`` and by the CPU-baseline
leg of bench.py, as the checker.  The product path (ace_amd) never imports it.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "liboracle.so")
dp = C.POINTER(C.c_double)
ip = C.POINTER(C.c_int)
lp = C.POINTER(C.c_longlong)
_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("oracle/liboracle.so is not built (run __graft_entry__.build())")
        L = C.CDLL(LIB_PATH)
        sig = {
            "or_coscut": (C.c_double, [C.c_double]),
            "or_sincut": (C.c_double, [C.c_double]),
            "or_fmexp": (C.c_double, [C.c_double]),
            "or_powint": (C.c_double, [C.c_double, C.c_int]),
            "or_cutf": (C.c_double, [C.c_int, C.c_double, C.c_double]),
            "or_cutg": (C.c_double, [C.c_int, C.c_double, C.c_double]),
            "or_radial_f": (C.c_double, [C.c_int, C.c_int, C.c_double, C.c_double, C.c_double, C.c_double]),
            "or_radial_d": (C.c_double, [C.c_int, C.c_int, C.c_double, C.c_double, C.c_double, C.c_double]),
            "or_angular_f": (C.c_double, [C.c_int, C.c_int, C.c_double, C.c_double, C.c_double, C.c_int,
                                          C.c_double, C.c_double, C.c_double]),
            "or_angular_d": (None, [C.c_int, C.c_int, C.c_double, C.c_double, C.c_double, C.c_int, C.c_double,
                                    C.c_double, C.c_double, dp, dp]),
            "or_input_width": (C.c_int, [C.c_void_p, C.c_int]),
            "or_param_count": (C.c_int, [C.c_void_p, C.c_int]),
            "or_neighbors": (C.c_longlong, [dp, C.c_int, dp, C.c_double, C.c_longlong, lp, ip, ip]),
            "or_symm": (None, [C.c_void_p, dp, C.c_int, dp, ip, lp, ip, ip, C.c_int, dp]),
            "or_network": (C.c_double, [C.c_void_p, C.c_int, dp, dp]),
            "or_energy": (C.c_double, [C.c_void_p, C.c_int, ip, C.c_int, dp]),
            "or_force": (C.c_double, [C.c_void_p, dp, C.c_int, dp, ip, lp, ip, ip, C.c_int, dp, dp, dp]),
            "or_train_struct": (None, [C.c_void_p, C.c_void_p, dp, C.c_int, dp, ip, lp, ip, ip, C.c_int, dp,
                                       C.c_double, dp, dp, dp]),
            "or_adam_step": (None, [C.c_int, dp, dp, C.c_double, C.c_double, dp, dp, dp]),
        }
        for n, (r, a) in sig.items():
            f = getattr(L, n)
            f.restype = r
            f.argtypes = a
        _lib = L
    return _lib


def _d(a):
    return a.ctypes.data_as(dp)


def _i(a):
    return a.ctypes.data_as(ip)


def cellR(cell):
    """Column-major flattening of the cell (columns = lattice vectors)."""
    return np.ascontiguousarray(np.asarray(cell, dtype=np.float64).T.reshape(-1))


class Neighbors:
    def __init__(self, off, j, img):
        self.off, self.j, self.img = off, j, img


def neighbors(cell, pos, rc):
    """NeighborList::build restated (brute force, FP64)."""
    L = lib()
    R = cellR(cell)
    p = np.ascontiguousarray(pos, dtype=np.float64)
    n = len(p)
    off = np.zeros(n + 1, dtype=np.int64)
    cnt = L.or_neighbors(_d(R), n, _d(p), rc, 0, off.ctypes.data_as(lp), None, None)
    j = np.zeros(max(cnt, 1), dtype=np.int32)
    img = np.zeros((max(cnt, 1), 3), dtype=np.int32)
    L.or_neighbors(_d(R), n, _d(p), rc, cnt, off.ctypes.data_as(lp), _i(j), _i(img))
    return Neighbors(off, j[:cnt], img[:cnt])


class Oracle:
    """Oracle bound to one model description."""

    def __init__(self, model):
        self.model = model
        self.desc = model.desc()
        self.pdesc = C.cast(C.pointer(self.desc), C.c_void_p)
        self.L = lib()
        self.kmax = max(model.input_width(I) for I in range(model.ntypes))

    def refresh(self):
        self.desc = self.model.desc()
        self.pdesc = C.cast(C.pointer(self.desc), C.c_void_p)

    def neighbors(self, s, rc=None):
        return neighbors(s.cell, s.pos, self.model.rc if rc is None else rc)

    def symm(self, s, nl=None):
        nl = nl or self.neighbors(s)
        G = np.zeros((s.natoms, self.kmax))
        p = np.ascontiguousarray(s.pos, dtype=np.float64)
        t = np.ascontiguousarray(s.type, dtype=np.int32)
        self.L.or_symm(self.pdesc, _d(cellR(s.cell)), s.natoms, _d(p), _i(t), nl.off.ctypes.data_as(lp),
                       _i(nl.j), _i(np.ascontiguousarray(nl.img)), self.kmax, _d(G))
        return G

    def energy(self, s, G):
        t = np.ascontiguousarray(s.type, dtype=np.int32)
        return self.L.or_energy(self.pdesc, s.natoms, _i(t), self.kmax, _d(np.ascontiguousarray(G)))

    def force(self, s, G, nl=None):
        """(sum of NN outputs without E0, forces N x 3, virial 3 x 3)."""
        nl = nl or self.neighbors(s)
        F = np.zeros((s.natoms, 3))
        W = np.zeros(9)
        p = np.ascontiguousarray(s.pos, dtype=np.float64)
        t = np.ascontiguousarray(s.type, dtype=np.int32)
        pe = self.L.or_force(self.pdesc, _d(cellR(s.cell)), s.natoms, _d(p), _i(t), nl.off.ctypes.data_as(lp),
                             _i(nl.j), _i(np.ascontiguousarray(nl.img)), self.kmax, _d(np.ascontiguousarray(G)),
                             _d(F), _d(W))
        return pe, F, W.reshape(3, 3)

    def compute(self, s):
        """NNP::compute: energy with E0, forces, virial."""
        nl = self.neighbors(s)
        G = self.symm(s, nl)
        pe, F, W = self.force(s, G, nl)
        e0 = np.asarray(self.model.E0 if self.model.E0 is not None else np.zeros(self.model.ntypes))
        return pe + e0[s.type].sum(), F, W

    def network(self, I, g):
        d = np.zeros(len(g))
        out = self.L.or_network(self.pdesc, I, _d(np.ascontiguousarray(g, dtype=np.float64)), _d(d))
        return out, d

    def train_struct(self, loss_desc, s, G, E_ref, F_ref, grad, nl=None):
        nl = nl or self.neighbors(s)
        out = np.zeros(3)
        p = np.ascontiguousarray(s.pos, dtype=np.float64)
        t = np.ascontiguousarray(s.type, dtype=np.int32)
        Fr = None if F_ref is None else np.ascontiguousarray(F_ref, dtype=np.float64)
        self.L.or_train_struct(self.pdesc, C.cast(C.pointer(loss_desc), C.c_void_p), _d(cellR(s.cell)), s.natoms,
                               _d(p), _i(t), nl.off.ctypes.data_as(lp), _i(nl.j), _i(np.ascontiguousarray(nl.img)),
                               self.kmax, _d(np.ascontiguousarray(G)), E_ref, None if Fr is None else _d(Fr),
                               _d(grad), _d(out))
        return out
