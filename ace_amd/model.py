"""Behler-Parrinello potential description (the content of an ace ``NNP``) and its C layout.

An ``NNP`` (src/nnp/nnp.hpp:64-121) holds one ``NNH`` per species (src/nnp/nnh.hpp:46-155): a
radial basis per neighbour species J (``BasisR``, src/nnp/basis_radial.hpp:23-89), an angular
basis per unordered neighbour-species pair in TensorTri order (``BasisA``,
src/nnp/basis_angular.hpp:23-103, src/nnp/tensortri.hpp:99-111), and an atomic network
(``NN::ANN``, src/ml/nn.hpp:243-369) whose parameter vector is all biases then all weights
column-major (src/ml/nn.cpp:881-896).  :meth:`Model.desc` packs it into ``ace_model_desc``
(include/ace_hip.h), the struct both the HIP library and the CPU oracle read.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np

MAX_TYPES = 8
MAX_LAYERS = 6

CUT_STEP, CUT_COS, CUT_TANH, CUT_POLY3 = 1, 2, 3, 4
RAD_GAUSSIAN, RAD_SECH, RAD_LOGISTIC, RAD_TANH, RAD_LOGCOSH, RAD_LOGCOSH2 = 1, 2, 3, 4, 5, 6
ANG_GAUSS, ANG_GAUSS2, ANG_SECH, ANG_STUDENT3, ANG_STUDENT4, ANG_STUDENT5 = 1, 2, 3, 4, 5, 6
ACT_LINEAR, ACT_SIGMOID, ACT_TANH, ACT_SWISH, ACT_SOFTPLUS = 1, 2, 3, 10, 15
LOSS_MSE, LOSS_MAE, LOSS_HUBER, LOSS_ASINH = 0, 1, 2, 3
NORM_IDENTITY, NORM_LINEAR, NORM_SQRT, NORM_CBRT, NORM_LOG = 0, 1, 2, 3, 4


def tri_index(j: int, k: int) -> int:
    """TensorTri<2> linear index of the unordered pair (j, k) (tensortri.hpp:99-111)."""
    a, b = (j, k) if j > k else (k, j)
    return a * (a + 1) // 2 + b


class ModelDesc(C.Structure):
    _fields_ = [
        ("ntypes", C.c_int), ("rc", C.c_double),
        ("r_form", C.POINTER(C.c_int)), ("r_cut", C.POINTER(C.c_int)), ("r_rc", C.POINTER(C.c_double)),
        ("r_n", C.POINTER(C.c_int)), ("r_off", C.POINTER(C.c_int)), ("r_rs", C.POINTER(C.c_double)),
        ("r_eta", C.POINTER(C.c_double)),
        ("a_form", C.POINTER(C.c_int)), ("a_cut", C.POINTER(C.c_int)), ("a_rc", C.POINTER(C.c_double)),
        ("a_n", C.POINTER(C.c_int)), ("a_off", C.POINTER(C.c_int)), ("a_eta", C.POINTER(C.c_double)),
        ("a_zeta", C.POINTER(C.c_double)), ("a_lambda", C.POINTER(C.c_int)),
        ("nn_nlayer", C.POINTER(C.c_int)), ("nn_width", C.POINTER(C.c_int)), ("nn_act", C.POINTER(C.c_int)),
        ("nn_c", C.POINTER(C.c_double)), ("in_off", C.POINTER(C.c_int)), ("inpw", C.POINTER(C.c_double)),
        ("inpb", C.POINTER(C.c_double)), ("outw", C.POINTER(C.c_double)), ("outb", C.POINTER(C.c_double)),
        ("E0", C.POINTER(C.c_double)), ("par_off", C.POINTER(C.c_int)), ("params", C.POINTER(C.c_double)),
    ]


class LossDesc(C.Structure):
    _fields_ = [("loss", C.c_int), ("norm", C.c_int), ("delta", C.c_double), ("w_energy", C.c_double),
                ("w_force", C.c_double)]


@dataclass
class Radial:
    """One BasisR block: form, cutoff, radius, centres rs and widths eta (reference units)."""
    form: int
    cut: int
    rc: float
    rs: np.ndarray
    eta: np.ndarray


@dataclass
class Angular:
    """One BasisA block: form, cutoff, radius, eta (a length), zeta, lambda (+1/-1)."""
    form: int
    cut: int
    rc: float
    eta: np.ndarray
    zeta: np.ndarray
    lam: np.ndarray


@dataclass
class Network:
    widths: list            # [nInput, h1, ..., 1]
    act: int = ACT_TANH
    c: float = 1.0
    params: np.ndarray = None
    inpw: np.ndarray = None
    inpb: np.ndarray = None
    outw: float = 1.0
    outb: float = 0.0

    @property
    def nlayer(self) -> int:
        return len(self.widths) - 1

    def nparams(self) -> int:
        return sum(self.widths[l + 1] + self.widths[l + 1] * self.widths[l] for l in range(self.nlayer))

    def unpack(self):
        """(biases, weights) per layer from the flat vector (nn.cpp:881-917)."""
        p = np.asarray(self.params, dtype=np.float64)
        bs, ws, off = [], [], 0
        for l in range(self.nlayer):
            bs.append(p[off:off + self.widths[l + 1]])
            off += self.widths[l + 1]
        for l in range(self.nlayer):
            no, ni = self.widths[l + 1], self.widths[l]
            ws.append(p[off:off + no * ni].reshape(ni, no).T)   # column-major W (no x ni)
            off += no * ni
        return bs, ws


@dataclass
class Model:
    ntypes: int
    rc: float
    radial: list            # radial[I][J] -> Radial
    angular: list           # angular[I][b] -> Angular, b = tri_index(J, K)
    nets: list              # nets[I] -> Network
    E0: np.ndarray = None
    names: list = None      # species names (NNH::Type::name); files written without them use X0, X1, ...
    types: list = None      # per species: the other NNH::Type properties (mass, charge, ...) as a dict

    def input_width(self, I: int) -> int:
        return sum(len(r.rs) for r in self.radial[I]) + sum(len(a.eta) for a in self.angular[I])

    def radial_offset(self, I: int, J: int) -> int:
        return sum(len(self.radial[I][j].rs) for j in range(J))

    def angular_offset(self, I: int, b: int) -> int:
        return sum(len(r.rs) for r in self.radial[I]) + sum(len(self.angular[I][q].eta) for q in range(b))

    def param_counts(self):
        return [n.nparams() for n in self.nets]

    def all_params(self) -> np.ndarray:
        return np.concatenate([np.asarray(n.params, dtype=np.float64) for n in self.nets])

    def set_all_params(self, x: np.ndarray):
        off = 0
        for n in self.nets:
            k = n.nparams()
            n.params = np.array(x[off:off + k], dtype=np.float64)
            off += k

    def desc(self) -> ModelDesc:
        """Pack into ``ace_model_desc``; the returned struct owns the arrays its pointers refer to."""
        nt = self.ntypes
        nab = nt * (nt + 1) // 2
        keep = []

        def arr(a, ct):
            npt = np.int32 if ct is C.c_int else np.float64
            x = np.ascontiguousarray(np.asarray(a, dtype=npt))
            if x.size == 0:
                x = np.zeros(1, dtype=npt)
            keep.append(x)
            return x.ctypes.data_as(C.POINTER(ct))

        r_form, r_cut, r_rc, r_n, r_off, r_rs, r_eta = [], [], [], [], [], [], []
        for I in range(nt):
            for J in range(nt):
                b = self.radial[I][J]
                r_form.append(b.form); r_cut.append(b.cut); r_rc.append(b.rc); r_n.append(len(b.rs))
                r_off.append(len(r_rs)); r_rs.extend(b.rs); r_eta.extend(b.eta)
        a_form, a_cut, a_rc, a_n, a_off, a_eta, a_zeta, a_lam = [], [], [], [], [], [], [], []
        for I in range(nt):
            for q in range(nab):
                b = self.angular[I][q]
                a_form.append(b.form); a_cut.append(b.cut); a_rc.append(b.rc); a_n.append(len(b.eta))
                a_off.append(len(a_eta)); a_eta.extend(b.eta); a_zeta.extend(b.zeta); a_lam.extend(b.lam)
        nl, wid, act, cc, ino, iw, ib, ow, ob, po, prm = [], [], [], [], [], [], [], [], [], [], []
        for I in range(nt):
            n = self.nets[I]
            K = self.input_width(I)
            assert n.widths[0] == K, "network input width must equal the basis size"
            nl.append(n.nlayer)
            w = list(n.widths) + [0] * (MAX_LAYERS + 1 - len(n.widths))
            wid.extend(w)
            act.append(n.act); cc.append(n.c)
            ino.append(len(iw))
            iw.extend(np.ones(K) if n.inpw is None else n.inpw)
            ib.extend(np.zeros(K) if n.inpb is None else n.inpb)
            ow.append(n.outw); ob.append(n.outb)
            po.append(len(prm)); prm.extend(n.params)
        e0 = np.zeros(nt) if self.E0 is None else self.E0
        d = ModelDesc(
            nt, self.rc,
            arr(r_form, C.c_int), arr(r_cut, C.c_int), arr(r_rc, C.c_double), arr(r_n, C.c_int),
            arr(r_off, C.c_int), arr(r_rs, C.c_double), arr(r_eta, C.c_double),
            arr(a_form, C.c_int), arr(a_cut, C.c_int), arr(a_rc, C.c_double), arr(a_n, C.c_int),
            arr(a_off, C.c_int), arr(a_eta, C.c_double), arr(a_zeta, C.c_double), arr(a_lam, C.c_int),
            arr(nl, C.c_int), arr(wid, C.c_int), arr(act, C.c_int), arr(cc, C.c_double), arr(ino, C.c_int),
            arr(iw, C.c_double), arr(ib, C.c_double), arr(ow, C.c_double), arr(ob, C.c_double),
            arr(e0, C.c_double), arr(po, C.c_int), arr(prm, C.c_double),
        )
        d._keep = keep          # the arrays live exactly as long as this descriptor
        return d


def xavier_params(widths, rng: np.random.Generator, sigma_w=1.0, sigma_b=1e-3) -> np.ndarray:
    """Xavier-normal weights, sigma = sigma_w*sqrt(2/(n_in+n_out)) (src/ml/nn.cpp:1193-1210), small
    normal biases; drawn from ``rng`` (the reference's std::mt19937 stream is not reproduced)."""
    bs, ws = [], []
    for l in range(len(widths) - 1):
        ni, no = widths[l], widths[l + 1]
        s = np.sqrt(2.0 / (ni + no))
        bs.append(rng.normal(0.0, sigma_b * s, size=no))
        ws.append(rng.normal(0.0, sigma_w * s, size=no * ni))
    return np.concatenate(bs + ws)
