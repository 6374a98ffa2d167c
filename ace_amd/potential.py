"""ace's text formats for a neural network potential and for per-species basis files.

Potential ("ann") file, NNP::write / NNP::read (src/nnp/nnp.cpp:197-307):
    ann
    ntypes <n>
    name <s> energy <f> mass <f> charge <f> radius <f> z <f> alpha <f> ie <f> rcov <f> rvdw <f> amp <f>   (x n,
        NNH::Type::write / read, src/nnp/nnh.cpp:58-108; tags case-insensitive, any order)
    rc <f>
    then for every central species I:  basis <I>
        for every J:        basis_radial <J>      + a BasisR block
        for every J <= K:   basis_angular <J> <K> + a BasisA block
    then for every species I:  nn <I>  + an ANN block
BasisR block (src/nnp/basis_radial.cpp:108-143):  "BasisR <cutoff> <rc> <form> <n>", then n lines "rs eta".
BasisA block (src/nnp/basis_angular.cpp:101-138): "BasisA <cutoff> <rc> <form> <n>", then n lines
    "eta zeta lambda rflag" (rflag: index of the run of equal eta, BasisA::init, basis_angular.cpp:178-205;
    written, ignored when read).
ANN block (src/ml/nn.cpp:1471-1615): "nn <nInp> <nodes of each layer>", "neuron <NAME> <c>", inpw, outw, inpb,
    outb lines, "bias[l] ..." per layer, "weight[l-1,l] ..." per layer with W (outputs x inputs) row by row.
Basis file, NNP::read_basis (src/nnp/nnp.cpp:96-156): "<tag> rc", "<tag> ntypes", "<tag> <central>", then the
    radial and angular blocks as above.

Numbers are written as the reference writes them: basis parameters, cutoff radii and type properties with
%f (six decimals -- a written potential carries its basis rounded to 1e-6, as the reference's does),
network parameters with %.15f.  Errors the reference throws as std::invalid_argument raise ValueError.
"""
from __future__ import annotations

import numpy as np

from .model import Angular, Model, Network, Radial, tri_index

CUTOFF_NAMES = ["NONE", "STEP", "COS", "TANH", "POLY3"]                                   # cutoff.hpp:23-29
RADIAL_NAMES = ["NONE", "GAUSSIAN", "SECH", "LOGISTIC", "TANH", "LOGCOSH", "LOGCOSH2"]      # basis_radial.cpp
ANGULAR_NAMES = ["NONE", "GAUSS", "GAUSS2", "SECH", "STUDENT3", "STUDENT4", "STUDENT5"]    # basis_angular.cpp
NEURON_NAMES = ["NONE", "LINEAR", "SIGMOID", "TANH", "ISRU", "ARCTAN", "RELU", "ELU", "TANHRE", "SQRE", "SWISH",
                "GELU", "MISH", "PFLU", "LOGISH", "SOFTPLUS", "SQPLUS", "ATISH", "IERF", "TEST"]   # nn.cpp:95-140
TYPE_FIELDS = ["energy", "mass", "charge", "radius", "z", "alpha", "ie", "rcov", "rvdw", "amp"]


def _name(table, v):
    return table[v] if 0 <= v < len(table) else "NONE"


def _index(table, s):
    return table.index(s) if s in table else 0


def _rflags(eta):
    """Index of the run of consecutive equal eta (|difference| < 1e-6) of every angular function."""
    f, out = 0, []
    for n in range(len(eta)):
        if n > 0 and abs(eta[n] - eta[n - 1]) >= 1.0e-6:
            f += 1
        out.append(f)
    return out


class _Lines:
    """Line reader with the reference's whitespace tokenising (string::WS = " \\r\\t\\n")."""

    def __init__(self, text):
        self.lines = text.splitlines()
        self.i = 0

    def raw(self):
        if self.i >= len(self.lines):
            raise ValueError("unexpected end of file")
        s = self.lines[self.i]
        self.i += 1
        return s

    def tokens(self):
        return self.raw().split()


def _type_line(s):
    """NNH::Type::read: '#' also separates tokens (string::WSC); tags are case-insensitive."""
    tok = s.replace("#", " ").split()
    t = {"name": "NONE", **{k: 0.0 for k in TYPE_FIELDS}}
    q = 0
    while q < len(tok):
        tag = tok[q].upper()
        val = tok[q + 1] if q + 1 < len(tok) else "0"
        if tag == "NAME":
            t["name"] = val
        elif tag.lower() in TYPE_FIELDS:
            t[tag.lower()] = float(val)
        q += 2
    return t


def _read_radial(L: _Lines) -> Radial:
    tok = L.tokens()
    cut, rc, form, n = _index(CUTOFF_NAMES, tok[1]), float(tok[2]), _index(RADIAL_NAMES, tok[3]), int(tok[4])
    rs, eta = np.zeros(n), np.zeros(n)
    for i in range(n):
        t = L.tokens()
        rs[i], eta[i] = float(t[0]), float(t[1])
    return Radial(form, cut, rc, rs, eta)


def _read_angular(L: _Lines) -> Angular:
    tok = L.tokens()
    cut, rc, form, n = _index(CUTOFF_NAMES, tok[1]), float(tok[2]), _index(ANGULAR_NAMES, tok[3]), int(tok[4])
    eta, zeta, lam = np.zeros(n), np.zeros(n), np.zeros(n, dtype=np.int32)
    for i in range(n):
        t = L.tokens()
        eta[i], zeta[i], lam[i] = float(t[0]), float(t[1]), int(t[2])
    return Angular(form, cut, rc, eta, zeta, lam)


def _write_radial(out, b: Radial):
    out.append("BasisR %s %f %s %i" % (_name(CUTOFF_NAMES, b.cut), b.rc, _name(RADIAL_NAMES, b.form), len(b.rs)))
    for i in range(len(b.rs)):
        out.append("\t%f %f" % (b.rs[i], b.eta[i]))


def _write_angular(out, b: Angular):
    out.append("BasisA %s %f %s %i" % (_name(CUTOFF_NAMES, b.cut), b.rc, _name(ANGULAR_NAMES, b.form), len(b.eta)))
    for i, f in enumerate(_rflags(b.eta)):
        out.append("\t%f %f %i %i" % (b.eta[i], b.zeta[i], int(b.lam[i]), f))


def _read_network(L: _Lines) -> Network:
    widths = [int(x) for x in L.tokens()[1:]]
    tok = L.tokens()
    act, c = _index(NEURON_NAMES, tok[1]), float(tok[2])
    if act == 0:
        raise ValueError("ANN::read: invalid neuron")
    K, nl = widths[0], len(widths) - 1
    inpw = np.array([float(x) for x in L.tokens()[1:1 + K]])
    outw = float(L.tokens()[1])
    inpb = np.array([float(x) for x in L.tokens()[1:1 + K]])
    outb = float(L.tokens()[1])
    biases = [np.array([float(x) for x in L.tokens()[1:1 + widths[l + 1]]]) for l in range(nl)]
    weights = []
    for l in range(nl):
        no, ni = widths[l + 1], widths[l]
        w = np.array([float(x) for x in L.tokens()[1:1 + no * ni]])
        if len(w) != no * ni:
            raise ValueError("ANN::read: weight line too short")
        weights.append(w.reshape(no, ni))                         # written row by row: W[o][k]
    # flat order of the parameter vector (nn.cpp:881-896): biases, then W column-major (index o + k no)
    params = np.concatenate(biases + [w.T.reshape(-1) for w in weights])
    return Network(widths, act, c, params, inpw, inpb, outw, outb)


def _write_network(out, n: Network):
    out.append("nn " + "".join("%i " % w for w in n.widths))
    out.append("neuron %s %f" % (_name(NEURON_NAMES, n.act), n.c))
    out.append("inpw " + "".join("%.15f " % v for v in np.asarray(n.inpw, dtype=float)))
    out.append("outw %.15f " % n.outw)
    out.append("inpb " + "".join("%.15f " % v for v in np.asarray(n.inpb, dtype=float)))
    out.append("outb %.15f " % n.outb)
    bs, ws = n.unpack()
    for l in range(n.nlayer):
        out.append("bias[%i] " % (l + 1) + "".join("%.15f " % v for v in bs[l]))
    for l in range(n.nlayer):
        out.append("weight[%i,%i] " % (l, l + 1) + "".join("%.15f " % v for v in ws[l].reshape(-1)))


def potential_text(model: Model) -> str:
    """NNP::write."""
    nt = model.ntypes
    names = model.names or ["X%i" % I for I in range(nt)]
    types = model.types or [{} for _ in range(nt)]
    e0 = np.zeros(nt) if model.E0 is None else np.asarray(model.E0, dtype=float)
    out = ["ann", "ntypes %i" % nt]
    for I in range(nt):
        t = {k: float(types[I].get(k, 0.0)) for k in TYPE_FIELDS}
        t["energy"] = float(e0[I])
        out.append("name %s " % names[I] + "".join("%s %f " % (k, t[k]) for k in TYPE_FIELDS))
    out.append("rc %f" % model.rc)
    for I in range(nt):
        out.append("basis %s" % names[I])
        for J in range(nt):
            out.append("basis_radial %s" % names[J])
            _write_radial(out, model.radial[I][J])
        for J in range(nt):
            for K in range(J, nt):
                out.append("basis_angular %s %s" % (names[J], names[K]))
                _write_angular(out, model.angular[I][tri_index(J, K)])
    for I in range(nt):
        out.append("nn %s" % names[I])
        _write_network(out, model.nets[I])
    return "\n".join(out) + "\n"


def read_potential_text(text: str) -> Model:
    """NNP::read."""
    L = _Lines(text)
    L.raw()                                            # header
    nt = int(L.tokens()[1])
    if nt <= 0:
        raise ValueError("NNP::read: invalid number of types")
    types = [_type_line(L.raw()) for _ in range(nt)]
    names = [t["name"] for t in types]

    def index(s):
        if s not in names:
            raise ValueError("NNP::read: unknown species '%s'" % s)
        return names.index(s)

    rc = float(L.tokens()[1])
    if rc <= 0:
        raise ValueError("NNP::read: invalid cutoff")
    nab = nt * (nt + 1) // 2
    radial = [[None] * nt for _ in range(nt)]
    angular = [[None] * nab for _ in range(nt)]
    for _ in range(nt):
        II = index(L.tokens()[1])
        for _ in range(nt):
            JJ = index(L.tokens()[1])
            radial[II][JJ] = _read_radial(L)
        for _ in range(nab):
            t = L.tokens()
            angular[II][tri_index(index(t[1]), index(t[2]))] = _read_angular(L)
    nets = [None] * nt
    for _ in range(nt):
        II = index(L.tokens()[1])
        nets[II] = _read_network(L)
    m = Model(nt, rc, radial, angular, nets, E0=np.array([t["energy"] for t in types]), names=names,
              types=[{k: t[k] for k in TYPE_FIELDS} for t in types])
    for I in range(nt):
        if nets[I].widths[0] != m.input_width(I):
            raise ValueError("NNP::read: network input width %i does not match the basis of species %s (%i)"
                             % (nets[I].widths[0], names[I], m.input_width(I)))
    return m


def write_potential(path, model: Model):
    with open(path, "w") as f:
        f.write(potential_text(model))


def read_potential(path) -> Model:
    with open(path) as f:
        return read_potential_text(f.read())


def basis_text(model: Model, I: int) -> str:
    """A basis file of central species I in the layout NNP::read_basis reads."""
    nt = model.ntypes
    names = model.names or ["X%i" % q for q in range(nt)]
    out = ["rc %f" % model.rc, "ntypes %i" % nt, "central %s" % names[I]]
    for J in range(nt):
        out.append("basis_radial %s" % names[J])
        _write_radial(out, model.radial[I][J])
    for J in range(nt):
        for K in range(J, nt):
            out.append("basis_angular %s %s" % (names[J], names[K]))
            _write_angular(out, model.angular[I][tri_index(J, K)])
    return "\n".join(out) + "\n"


def read_basis_text(text: str, model: Model, name: str):
    """NNP::read_basis: replaces the basis of species `name` in model (same rc and species count required)."""
    names = model.names or ["X%i" % q for q in range(model.ntypes)]

    def index(s):
        if s not in names:
            raise ValueError("NNP::read_basis: unknown species '%s'" % s)
        return names.index(s)

    L = _Lines(text)
    rc = float(L.tokens()[1])
    if rc != model.rc:
        raise ValueError("NNP::read_basis: invalid cutoff")
    nt = int(L.tokens()[1])
    if nt != model.ntypes:
        raise ValueError("NNP::read_basis: invalid number of types")
    II = index(L.tokens()[1])
    if names[II] != name:
        raise ValueError("NNP::read_basis: invalid central type")
    for _ in range(nt):
        JJ = index(L.tokens()[1])
        model.radial[II][JJ] = _read_radial(L)
    for _ in range(nt * (nt + 1) // 2):
        t = L.tokens()
        model.angular[II][tri_index(index(t[1]), index(t[2]))] = _read_angular(L)
    return model
