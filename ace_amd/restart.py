"""The reference's binary restart file: NNPTEFR::write_restart / read_restart (src/nnp/nnptefr.cpp:509-574).

The reference writes a memory image of the trainer through its serialize:: pack functions and reads it back
with the matching unpack functions.  This module writes and reads the same layout, field by field in the
reference's order, little-endian, int = 4 bytes, bool = 1 byte, every enum = 4 bytes (plain `enum Type`
members of a wrapper class), double = 8 bytes:

  NNPTEFR (nnptefr.cpp:121-157): NNP; four strings (parameter, error, ann and restart file names); four
      bools (restart, reset, wparams, sparams); Batch (training), Batch (validation); Objective; Iterator;
      the optimiser (a shared_ptr); Decay; the Norm, PreScale, PreBias and Regularization enums; eight
      doubles inscale, inbias, delta, beta, betai, rmse_t_a, rmse_v_a, eta.
  NNP (nnp.cpp:525-537): int ntypes; Map<int,int> (name hash -> index, map.cpp:50-64); one NNH per species;
      double rc.
  NNH (nnh.cpp:275-298): NNH::Type (string name; doubles energy, mass, charge, z, radius, rcov, rvdw,
      amp); NN::ANN; int ntypes; BasisR per neighbour species; BasisA for j <= k (j outer, k inner).
  NN::ANN (nn.cpp:2264-2301): Neuron enum; double c; int nlayer; if nlayer > 0: int nInp; int nNodes[l];
      the biases of every layer, then every weight matrix column-major (the parameter vector,
      nn.cpp:881-896); inpw[nInp], inpb[nInp]; outw[nOut], outb[nOut].
  BasisR (basis_radial.cpp:402-413): int size; Name enum; Cutoff (Name enum, double rc); rs[size],
      eta[size].  BasisA (basis_angular.cpp:590-602): int size; Name; Cutoff; eta[size], zeta[size] doubles,
      lambda[size] ints.
  string (serialize.cpp:63-69): int length, then the characters.  Batch (batch.cpp:115-123): int count,
      size, capacity; elements[size], data[capacity] ints.  Objective (objective.cpp:85-105): doubles gamma,
      val, valOld, dv, dp; int dim; p, pOld, g, gOld.  Iterator (iter.cpp:77-92): ints nPrint, nWrite,
      step, count, max; Stop and Loss enums; double tol.  Decay (decay.cpp:118-125): Name enum; int period;
      double alpha.
  optimiser (algo.cpp:896-1110): int null flag; Name enum; then the algorithm's own record, which starts
      with int dim and the Name again, followed by its state (see ALGO_FIELDS).

Reading restates the reference's unpack, including one quirk: unpack re-creates the optimiser and resizes it
(algo.cpp:1116-1146, 1230-1244), which resets the bias-correction powers beta1^t, beta2^t of ADAM and the
other moment methods, so a resumed run starts its bias corrections afresh; those powers are not in the file.
"""
from __future__ import annotations

import struct
from dataclasses import dataclass, field

import numpy as np

from .model import Angular, Model, Network, Radial, tri_index

ALGO_NAMES = ("NONE", "SGD", "SDM", "NAG", "ADAGRAD", "ADADELTA", "RMSPROP", "ADAM", "ADAMW", "ADAB", "YOGI",
              "NOGI", "NADAM", "AMSGRAD", "BFGS", "RPROP", "CG")                    # algo.hpp:31-49
DECAY_NAMES = ("NONE", "CONST", "EXP", "GAUSS", "SQRT", "INV", "STEP")              # decay.hpp:28-36
LOSS_NAMES = ("NONE", "MSE", "MAE", "HUBER", "ASINH")                               # loss.hpp:18-24
STOP_NAMES = ("NONE", "FABS", "FREL", "XABS", "XREL")                               # stop.hpp:16-22
NORM_NAMES = ("IDENTITY", "LINEAR", "SQRT", "CBRT", "LOG", "NONE")                  # nnptefr.hpp:129-136
PRESCALE_NAMES = ("IDENTITY", "DEV", "MINMAX", "MAX", "NONE")                       # nnptefr.hpp:71-77
PREBIAS_NAMES = ("IDENTITY", "MEAN", "HAR", "RMS", "MID", "MIN", "NONE")            # nnptefr.hpp:99-107
REG_NAMES = ("LASSO", "RIDGE", "HUBER", "ASINH", "NONE")                            # nnptefr.hpp:185-191
TYPE_FIELDS = ("energy", "mass", "charge", "z", "radius", "rcov", "rvdw", "amp")    # nnh.cpp:260-272

# per optimiser, the record after (dim, name): scalars, then vectors of dim doubles ("B" entries: dim x dim,
# column-major) -- algo.cpp:926-1110; each entry is (reference field, attribute of ace_amd.optimizers)
ALGO_FIELDS = {
    "SGD": ((), ()),
    "SDM": ((("eta", "eta"),), (("dx", "dx"),)),
    "NAG": ((("eta", "eta"),), (("dx", "dx"),)),
    "ADAGRAD": ((("eps", "eps"),), (("mgrad2", "g2"),)),
    "ADADELTA": ((("eta", "eta"), ("eps", "eps")), (("mgrad2", "g2"), ("mdx2", "dx2"), ("dx", "dx"))),
    "RMSPROP": ((), (("mgrad2", "g2"),)),
    "ADAM": ((("eps", "eps"),), (("mgrad", "m"), ("mgrad2", "v"))),
    "ADAMW": ((("w", "w"), ("eps", "eps")), (("mgrad", "m"), ("mgrad2", "v"))),
    "ADAB": ((("eps", "eps"),), (("mgrad", "m"), ("mgrad2", "v"))),
    "YOGI": ((("eps", "eps"),), (("mgrad", "m"), ("mgrad2", "v"))),
    "NOGI": ((("eps", "eps"),), (("mgrad", "m"), ("mgrad2", "v"))),
    "NADAM": ((("eps", "eps"),), (("mgrad", "m"), ("mgrad2", "v"))),
    "AMSGRAD": ((("eps", "eps"),), (("mgrad", "m"), ("mgrad2", "v"), ("mgrad2m", "vmax"))),
    "BFGS": ((), (("s", "s"), ("y", "y"), ("B", "B"), ("BOld", "B_old"))),
    "RPROP": ((), (("dx", "dx"), ("delta", "delta"))),
    "CG": ((("eps", "eps"),), (("cgd", "d"),)),
}


def name_hash(name: str) -> int:
    """string::hash (src/str/string.cpp:18-33): sum of (c - 'a' + 1) 31^i mod 2^31 - 1 in C int arithmetic
    (products and sums wrap to 32 bits, % truncates toward zero), the key of the NNP's species map."""
    def i32(x):
        x &= 0xFFFFFFFF
        return x - (1 << 32) if x & 0x80000000 else x

    def cmod(a, m):
        r = abs(a) % m
        return -r if a < 0 else r

    p, m, val, pw = 31, 2147483647, 0, 1
    for ch in name.encode():
        val = cmod(i32(val + i32((ch - ord("a") + 1) * pw)), m)
        pw = cmod(i32(pw * p), m)
    return val


@dataclass
class Batch:
    """ml/batch.hpp:18-55: count (position of the cyclic walk), size, capacity, elements, data."""
    count: int = 0
    size: int = 0
    capacity: int = 0
    elements: np.ndarray = field(default_factory=lambda: np.zeros(0, np.int32))
    data: np.ndarray = field(default_factory=lambda: np.zeros(0, np.int32))


@dataclass
class RestartState:
    """The content of a reference restart file (the NNPTEFR members it serialises)."""
    model: Model
    files: tuple = ("", "", "", "")          # file_params_, file_error_, file_ann_, file_restart_
    flags: tuple = (False, False, False, False)   # restart_, reset_, wparams_, sparams_
    batch_train: Batch = field(default_factory=Batch)
    batch_val: Batch = field(default_factory=Batch)
    gamma: float = 0.0
    val: float = 0.0
    val_old: float = 0.0
    dv: float = 0.0
    dp: float = 0.0
    p: np.ndarray = None
    p_old: np.ndarray = None
    g: np.ndarray = None
    g_old: np.ndarray = None
    n_print: int = 1
    n_write: int = 1
    step: int = 0
    count: int = 0
    max_iter: int = 0
    stop: str = "NONE"
    loss: str = "MSE"
    tol: float = 0.0
    algo: str = None                          # ALGO_NAMES entry, or None for an empty pointer
    algo_state: dict = field(default_factory=dict)   # reference field -> value (scalars, vectors, B matrices)
    decay: str = "CONST"
    decay_period: int = 1
    decay_alpha: float = 0.0
    norm: str = "IDENTITY"
    prescale: str = "DEV"
    prebias: str = "MEAN"
    reg: str = "NONE"
    inscale: float = 1.0
    inbias: float = 0.0
    delta: float = 1.0
    beta: float = 0.0
    betai: float = 0.0
    rmse_t_a: float = 0.0
    rmse_v_a: float = 0.0
    eta: float = 0.0


class _Out:
    def __init__(self):
        self.b = bytearray()

    def i32(self, v):
        self.b += struct.pack("<i", int(v))

    def f64(self, v):
        self.b += struct.pack("<d", float(v))

    def flag(self, v):
        self.b += struct.pack("<?", bool(v))

    def s(self, text):
        raw = text.encode()
        self.i32(len(raw))
        self.b += raw

    def f64s(self, a):
        self.b += np.ascontiguousarray(a, dtype="<f8").tobytes()

    def i32s(self, a):
        self.b += np.ascontiguousarray(a, dtype="<i4").tobytes()


class _In:
    def __init__(self, data: bytes):
        self.d, self.p = data, 0

    def take(self, n):
        if n < 0 or self.p + n > len(self.d):
            raise ValueError("restart file: truncated at byte %i" % self.p)
        v = self.d[self.p:self.p + n]
        self.p += n
        return v

    def i32(self):
        return struct.unpack("<i", self.take(4))[0]

    def f64(self):
        return struct.unpack("<d", self.take(8))[0]

    def flag(self):
        return struct.unpack("<?", self.take(1))[0]

    def s(self):
        return self.take(self.count()).decode()

    def count(self, what="size"):
        n = self.i32()
        if n < 0:
            raise ValueError("restart file: negative %s at byte %i" % (what, self.p - 4))
        return n

    def f64s(self, n):
        return np.frombuffer(self.take(8 * n), dtype="<f8").astype(np.float64)

    def i32s(self, n):
        return np.frombuffer(self.take(4 * n), dtype="<i4").astype(np.int32)

    def enum(self, names, what):
        v = self.i32()
        if not 0 <= v < len(names):
            raise ValueError("restart file: invalid %s %i" % (what, v))
        return names[v]


def _pack_model(w: _Out, m: Model):
    nt = m.ntypes
    names = m.names or ["X%i" % I for I in range(nt)]
    types = m.types or [{} for _ in range(nt)]
    e0 = np.zeros(nt) if m.E0 is None else np.asarray(m.E0, dtype=float)
    w.i32(nt)
    w.i32(nt)                                            # Map<int,int>: size, keys, values
    for I in range(nt):
        w.i32(name_hash(names[I]))
    for I in range(nt):
        w.i32(I)
    for I in range(nt):
        w.s(names[I])                                    # NNH::Type
        for f in TYPE_FIELDS:
            w.f64(e0[I] if f == "energy" else types[I].get(f, 0.0))
        n = m.nets[I]                                    # NN::ANN
        w.i32(n.act)
        w.f64(n.c)
        w.i32(n.nlayer)
        if n.nlayer > 0:
            w.i32(n.widths[0])
            for l in range(n.nlayer):
                w.i32(n.widths[l + 1])
            w.f64s(n.params)
            K = n.widths[0]
            w.f64s(np.ones(K) if n.inpw is None else n.inpw)
            w.f64s(np.zeros(K) if n.inpb is None else n.inpb)
            w.f64s([n.outw])
            w.f64s([n.outb])
        w.i32(nt)
        for J in range(nt):                              # BasisR
            b = m.radial[I][J]
            w.i32(len(b.rs))
            w.i32(b.form)
            w.i32(b.cut)
            w.f64(b.rc)
            if len(b.rs):
                w.f64s(b.rs)
                w.f64s(b.eta)
        for j in range(nt):                              # BasisA, j outer, k inner
            for k in range(j, nt):
                b = m.angular[I][tri_index(j, k)]
                w.i32(len(b.eta))
                w.i32(b.form)
                w.i32(b.cut)
                w.f64(b.rc)
                if len(b.eta):
                    w.f64s(b.eta)
                    w.f64s(b.zeta)
                    w.i32s(b.lam)
    w.f64(m.rc)


def _unpack_model(r: _In) -> Model:
    nt = r.count("species count")
    if nt == 0 or nt > 8:
        raise ValueError("restart file: %i species" % nt)
    nmap = r.count("map size")
    keys = r.i32s(nmap)
    r.i32s(nmap)
    nab = nt * (nt + 1) // 2
    names, types, E0, nets = [], [], np.zeros(nt), []
    radial = [[None] * nt for _ in range(nt)]
    angular = [[None] * nab for _ in range(nt)]
    for I in range(nt):
        names.append(r.s())
        t = {f: r.f64() for f in TYPE_FIELDS}
        E0[I] = t.pop("energy")
        types.append(t)
        act = r.i32()
        c = r.f64()
        nlayer = r.count("layer count")
        if nlayer == 0:
            raise ValueError("restart file: species %s has no network" % names[I])
        widths = [r.count("input count")] + [r.count("layer width") for _ in range(nlayer)]
        npar = sum(widths[l + 1] * (widths[l] + 1) for l in range(nlayer))
        params = r.f64s(npar)
        inpw, inpb = r.f64s(widths[0]), r.f64s(widths[0])
        outw, outb = r.f64s(widths[-1]), r.f64s(widths[-1])
        if widths[-1] != 1:
            raise ValueError("restart file: the network of %s has %i outputs" % (names[I], widths[-1]))
        nets.append(Network(widths, act, c, params, inpw, inpb, float(outw[0]), float(outb[0])))
        if r.i32() != nt:
            raise ValueError("restart file: species count of %s differs" % names[I])
        for J in range(nt):
            n = r.count("basis size")
            form, cut, rc = r.i32(), r.i32(), r.f64()
            rs, eta = (r.f64s(n), r.f64s(n)) if n else (np.zeros(0), np.zeros(0))
            radial[I][J] = Radial(form, cut, rc, rs, eta)
        for j in range(nt):
            for k in range(j, nt):
                n = r.count("basis size")
                form, cut, rc = r.i32(), r.i32(), r.f64()
                if n:
                    eta, zeta, lam = r.f64s(n), r.f64s(n), r.i32s(n)
                else:
                    eta, zeta, lam = np.zeros(0), np.zeros(0), np.zeros(0, np.int32)
                angular[I][tri_index(j, k)] = Angular(form, cut, rc, eta, zeta, lam)
    rc = r.f64()
    if nmap != nt or any(int(keys[I]) != name_hash(names[I]) for I in range(nt)):
        raise ValueError("restart file: the species map does not match the species names")
    m = Model(nt, rc, radial, angular, nets, E0=E0, names=names, types=types)
    for I in range(nt):
        if nets[I].widths[0] != m.input_width(I):
            raise ValueError("restart file: network input width of %s does not match its basis" % names[I])
    return m


def _pack_batch(w: _Out, b: Batch):
    w.i32(b.count)
    w.i32(b.size)
    w.i32(b.capacity)
    w.i32s(np.asarray(b.elements, dtype=np.int32)[:b.size])
    w.i32s(np.asarray(b.data, dtype=np.int32)[:b.capacity])


def _unpack_batch(r: _In) -> Batch:
    count, size, cap = r.i32(), r.count("batch size"), r.count("batch capacity")
    return Batch(count, size, cap, r.i32s(size), r.i32s(cap))


def _vec(x, dim):
    a = np.zeros(dim) if x is None else np.asarray(x, dtype=np.float64).reshape(-1)
    if a.size != dim:
        raise ValueError("restart: vector of %i entries, expected %i" % (a.size, dim))
    return a


def pack_restart(st: RestartState) -> bytes:
    """NNPTEFR serialise::pack (nnptefr.cpp:121-157) of the given state."""
    w = _Out()
    _pack_model(w, st.model)
    for f in st.files:
        w.s(f)
    for f in st.flags:
        w.flag(f)
    _pack_batch(w, st.batch_train)
    _pack_batch(w, st.batch_val)
    dim = 0 if st.p is None else len(st.p)
    for v in (st.gamma, st.val, st.val_old, st.dv, st.dp):   # Objective
        w.f64(v)
    w.i32(dim)
    if dim:
        for a in (st.p, st.p_old, st.g, st.g_old):
            w.f64s(_vec(a, dim))
    for v in (st.n_print, st.n_write, st.step, st.count, st.max_iter):   # Iterator
        w.i32(v)
    w.i32(STOP_NAMES.index(st.stop))
    w.i32(LOSS_NAMES.index(st.loss))
    w.f64(st.tol)
    w.i32(st.algo is None)                                  # optimiser
    if st.algo is not None:
        code = ALGO_NAMES.index(st.algo)
        w.i32(code)
        w.i32(dim)
        w.i32(code)
        scal, vecs = ALGO_FIELDS[st.algo]
        for ref, _ in scal:
            w.f64(st.algo_state[ref])
        if dim:
            for ref, _ in vecs:
                if ref in ("B", "BOld"):
                    w.f64s(np.asarray(st.algo_state[ref], dtype=np.float64).reshape(dim, dim).T.reshape(-1))
                else:
                    w.f64s(_vec(st.algo_state.get(ref), dim))
    w.i32(DECAY_NAMES.index(st.decay))
    w.i32(st.decay_period)
    w.f64(st.decay_alpha)
    w.i32(NORM_NAMES.index(st.norm))
    w.i32(PRESCALE_NAMES.index(st.prescale))
    w.i32(PREBIAS_NAMES.index(st.prebias))
    w.i32(REG_NAMES.index(st.reg))
    for v in (st.inscale, st.inbias, st.delta, st.beta, st.betai, st.rmse_t_a, st.rmse_v_a, st.eta):
        w.f64(v)
    return bytes(w.b)


def unpack_restart(data: bytes) -> RestartState:
    """NNPTEFR serialise::unpack (nnptefr.cpp:163-201); raises ValueError on a malformed or truncated image."""
    r = _In(data)
    st = RestartState(model=_unpack_model(r))
    st.files = tuple(r.s() for _ in range(4))
    st.flags = tuple(r.flag() for _ in range(4))
    st.batch_train = _unpack_batch(r)
    st.batch_val = _unpack_batch(r)
    st.gamma, st.val, st.val_old, st.dv, st.dp = (r.f64() for _ in range(5))
    dim = r.count("parameter count")
    if dim:
        st.p, st.p_old, st.g, st.g_old = (r.f64s(dim) for _ in range(4))
    st.n_print, st.n_write, st.step, st.count, st.max_iter = (r.i32() for _ in range(5))
    st.stop = r.enum(STOP_NAMES, "stop condition")
    st.loss = r.enum(LOSS_NAMES, "loss")
    st.tol = r.f64()
    null = r.i32()
    if null:
        st.algo = None
    else:
        name = r.enum(ALGO_NAMES, "optimiser")
        if name == "NONE":
            raise ValueError("restart file: invalid optimiser NONE")
        adim = r.count("optimiser dimension")
        if r.enum(ALGO_NAMES, "optimiser") != name:
            raise ValueError("restart file: optimiser record of another algorithm")
        scal, vecs = ALGO_FIELDS[name]
        state = {ref: r.f64() for ref, _ in scal}
        if adim:
            for ref, _ in vecs:
                if ref in ("B", "BOld"):
                    state[ref] = r.f64s(adim * adim).reshape(adim, adim).T.copy()
                else:
                    state[ref] = r.f64s(adim)
        st.algo, st.algo_state = name, state
    st.decay = r.enum(DECAY_NAMES, "decay")
    st.decay_period = r.i32()
    st.decay_alpha = r.f64()
    st.norm = r.enum(NORM_NAMES, "norm")
    st.prescale = r.enum(PRESCALE_NAMES, "pre-scaling")
    st.prebias = r.enum(PREBIAS_NAMES, "pre-bias")
    st.reg = r.enum(REG_NAMES, "regularization")
    (st.inscale, st.inbias, st.delta, st.beta, st.betai, st.rmse_t_a, st.rmse_v_a, st.eta) = (r.f64() for _ in range(8))
    if r.p != len(data):
        raise ValueError("restart file: %i bytes after the trainer record" % (len(data) - r.p))
    return st


def write_restart(path, st: RestartState):
    """NNPTEFR::write_restart (nnptefr.cpp:509-537)."""
    data = pack_restart(st)
    with open(path, "wb") as f:
        f.write(data)


def read_restart(path) -> RestartState:
    """NNPTEFR::read_restart (nnptefr.cpp:539-574)."""
    with open(path, "rb") as f:
        return unpack_restart(f.read())
