"""The reference's binary restart layout (NNPTEFR::write_restart/read_restart, src/nnp/nnptefr.cpp:509-574).

The reference ships no restart file, so the layout is pinned by (a) the species-name hash evaluated by hand
from string.cpp:18-33, (b) a tiny trainer whose expected bytes are assembled here field by field from the
reference's pack functions (nnptefr.cpp:121-157, nnp.cpp:525-537, map.cpp:50-64, nnh.cpp:260-298,
nn.cpp:2264-2301, basis_radial.cpp:402-413, basis_angular.cpp:590-602, cutoff.cpp:177-182,
serialize.cpp:63-69, batch.cpp:115-123, objective.cpp:85-105, iter.cpp:77-92, algo.cpp:896-925 and
990-1000, decay.cpp:118-125) and whose length is checked against the reference's nbytes sums, and
(c) round trips, resumption and refusals."""
import struct

import numpy as np
import pytest

from ace_amd import configs
from ace_amd import optimizers as opt
from ace_amd import restart as R
from ace_amd.model import ACT_TANH, ANG_GAUSS, CUT_COS, Angular, Model, Network, RAD_GAUSSIAN, Radial
from ace_amd.training import Trainer


class NoDevice:
    def set_targets(self, *a):
        pass


def test_species_name_hash():
    """sum_i (c_i - 'a' + 1) 31^i (string.cpp:18-33), by hand."""
    assert R.name_hash("a") == 1
    assert R.name_hash("ab") == 1 + 2 * 31
    assert R.name_hash("H") == ord("H") - ord("a") + 1 == -24          # C: -24 % m = -24
    assert R.name_hash("Na") == (78 - 97 + 1) + 1 * 31 == 13
    assert R.name_hash("Si") == (83 - 97 + 1) + (105 - 97 + 1) * 31 == 266
    assert R.name_hash("") == 0


def tiny_trainer():
    rad = Radial(RAD_GAUSSIAN, CUT_COS, 5.5, np.array([0.25]), np.array([1.5]))
    ang = Angular(ANG_GAUSS, CUT_COS, 5.0, np.array([3.0]), np.array([2.0]), np.array([-1], dtype=np.int32))
    net = Network([2, 1], ACT_TANH, 1.0, np.array([0.125, 0.5, -0.75]), np.array([2.0, 4.0]), np.array([-1.0, -3.0]),
                  0.5, 0.25)
    m = Model(1, 6.0, [[rad]], [[ang]], [net], E0=np.array([-1.5]), names=["Ar"],
              types=[{"mass": 39.948, "charge": 0.0, "z": 18.0, "radius": 1.0, "rcov": 1.06, "rvdw": 1.88, "amp": 0.0}])
    t = Trainer(NoDevice(), m, [], gamma=0.01, w_force=0.0, algo="ADAM", decay=opt.Decay("EXP", 0.5), reg="RIDGE",
                eta_reg=0.25, loss=2, norm=1, delta=0.75)
    t.nbatch = 2
    t.update(3.0, np.array([0.5, -1.0, 2.0]))
    return t


def test_tiny_trainer_bytes_follow_the_reference_pack_order(tmp_path):
    t = tiny_trainer()
    t.write_restart(tmp_path / "r.bin", files=("p.dat", "e.dat", "ann", "r"), n_print=10, n_write=100, max_iter=7)
    data = (tmp_path / "r.bin").read_bytes()
    o, a = t.obj, t.algo
    i = lambda *v: struct.pack("<%ii" % len(v), *v)
    d = lambda *v: struct.pack("<%id" % len(v), *v)
    s = lambda x: i(len(x)) + x.encode()
    exp = b"".join([
        i(1), i(1), i(R.name_hash("Ar")), i(0),                          # NNP: ntypes, Map<int,int>
        s("Ar"), d(-1.5, 39.948, 0.0, 18.0, 1.0, 1.06, 1.88, 0.0),        # NNH::Type
        i(3), d(1.0), i(1), i(2), i(1), d(*o.p), d(2.0, 4.0), d(-1.0, -3.0), d(0.5), d(0.25),   # NN::ANN
        i(1),                                                             # NNH ntypes
        i(1), i(1), i(2), d(5.5), d(0.25), d(1.5),                        # BasisR
        i(1), i(1), i(2), d(5.0), d(3.0), d(2.0), i(-1),                  # BasisA
        d(6.0),                                                           # NNP rc
        s("p.dat"), s("e.dat"), s("ann"), s("r"), b"\x00\x00\x00\x00",    # files, four bools
        i(0, 2, 2), i(0, 1), i(0, 1),                                     # training Batch
        i(0, 0, 0),                                                       # validation Batch
        d(o.gamma, 3.0, 3.0, t.dv, t.dp), i(3), d(*o.p), d(*o.p_old), d(*o.g_old), d(*o.g_old),   # Objective
        i(10, 100, 1, 1, 7), i(0), i(3), d(0.0),                          # Iterator: HUBER = 3
        i(0), i(7), i(3), i(7), d(a.eps), d(*a.m), d(*a.v),               # ADAM
        i(2), i(1), d(0.5),                                               # Decay EXP
        i(1), i(1), i(1), i(1),                                           # Norm LINEAR, DEV, MEAN, RIDGE
        d(1.0, 0.0, 0.75, 0.0, 0.0, 0.0, 0.0, 0.25),
    ])
    assert data == exp
    # nbytes of the reference: NNP 4 + (4 + 4 + 4) + NNH + 8; NNH = Type (4 + 2 + 64) + ANN (4 + 8 + 4 + 4 + 4 +
    # 3*8 + 2*16 + 2*8) + 4 + BasisR (4 + 4 + 12 + 16) + BasisA (4 + 4 + 12 + 20); files 4*4 + 14; bools 4;
    # batches 12 + 8 + 8 and 12; Objective 40 + 4 + 4*24; Iterator 28 + 8; ADAM 8 + 8 + 8 + 48; Decay 16;
    # enums 16; doubles 64
    nnh = (4 + 2 + 64) + (4 + 8 + 4 + 4 + 4 + 24 + 32 + 16) + 4 + 36 + 40
    assert len(data) == (4 + 12 + nnh + 8) + (16 + 14) + 4 + 28 + 12 + 140 + 36 + 72 + 16 + 16 + 64


@pytest.mark.parametrize("algo", sorted(R.ALGO_FIELDS))
def test_every_optimizer_round_trips(tmp_path, algo):
    """write -> read -> pack is byte-identical, and a trainer reading the file holds the written state."""
    rng = np.random.default_rng(7)
    mk = lambda: Trainer(NoDevice(), configs.config1()[0], [], gamma=0.01, w_force=0.0, algo=algo,
                         decay=opt.Decay("INV", 0.1), reg="HUBER", eta_reg=0.5)
    a = mk()
    dim = len(a.params)
    if algo == "BFGS":
        a.algo.B = np.diag(rng.uniform(1, 2, dim))          # kept symmetric positive definite
    for k in range(3):
        a.update(float(rng.uniform(1, 2)), rng.normal(size=dim) * (0.1 if algo == "BFGS" else 1.0))
    a.write_restart(tmp_path / "r")
    data = (tmp_path / "r").read_bytes()
    st = R.read_restart(tmp_path / "r")
    assert R.pack_restart(st) == data
    assert st.algo == algo and st.count == 3 and st.decay == "INV" and st.reg == "HUBER"
    b = mk()
    b.read_restart(tmp_path / "r")
    assert np.array_equal(b.obj.p, a.obj.p) and np.array_equal(b.obj.g_old, a.obj.g_old) and b.it == 3
    assert b.obj.gamma == a.obj.gamma
    for ref, attr in R.ALGO_FIELDS[algo][0] + R.ALGO_FIELDS[algo][1]:
        if hasattr(a.algo, attr):
            assert np.array_equal(np.asarray(getattr(b.algo, attr)), np.asarray(getattr(a.algo, attr))), (algo, attr)


@pytest.mark.parametrize("algo", ["SDM", "NAG", "RPROP", "ADAM", "AMSGRAD"])
def test_resuming_from_a_reference_restart(tmp_path, algo):
    """Five updates, the reference-layout restart, five more: the uninterrupted trajectory for optimisers
    without bias powers; for the moment methods the trajectory whose beta1^t, beta2^t restart at the
    restart point, as the reference's unpack leaves them (algo.cpp:1230-1244)."""
    rng = np.random.default_rng(3)
    grads = [rng.normal(size=861) for _ in range(10)]
    losses = list(rng.uniform(1, 2, 10))
    mk = lambda: Trainer(NoDevice(), configs.config1()[0], [], gamma=0.01, w_force=0.0, algo=algo,
                         decay=opt.Decay("SQRT", 0.1), reg="LASSO", eta_reg=0.5)
    a = mk()
    for k in range(10):
        if k == 5 and isinstance(a.algo, opt._Moments):
            a.algo.b1t, a.algo.b2t = a.algo.b1, a.algo.b2
        a.update(losses[k], grads[k])
    b = mk()
    for k in range(5):
        b.update(losses[k], grads[k])
    b.write_restart(tmp_path / "r")
    c = mk()
    c.read_restart(tmp_path / "r")
    for k in range(5, 10):
        c.update(losses[k], grads[k])
    assert np.array_equal(a.params, c.params) and a.obj.gamma == c.obj.gamma and c.it == 10


def test_reading_restores_the_potential(tmp_path):
    t = tiny_trainer()
    t.write_restart(tmp_path / "r")
    st = R.read_restart(tmp_path / "r")
    m = st.model
    assert m.names == ["Ar"] and m.rc == 6.0 and m.E0[0] == -1.5 and m.types[0]["rvdw"] == 1.88
    n = m.nets[0]
    assert n.widths == [2, 1] and n.act == ACT_TANH and np.array_equal(n.params, t.obj.p)
    assert n.outw == 0.5 and n.outb == 0.25 and np.array_equal(n.inpb, [-1.0, -3.0])
    assert np.array_equal(m.radial[0][0].rs, [0.25]) and m.angular[0][0].lam[0] == -1


def test_malformed_restart_files_raise(tmp_path):
    t = tiny_trainer()
    t.write_restart(tmp_path / "r")
    data = (tmp_path / "r").read_bytes()
    with pytest.raises(ValueError, match="truncated"):
        R.unpack_restart(data[:-3])
    with pytest.raises(ValueError, match="after the trainer record"):
        R.unpack_restart(data + b"\x00")
    bad = bytearray(data)
    bad[8:12] = struct.pack("<i", R.name_hash("Kr"))          # species map key of another name
    with pytest.raises(ValueError, match="species map"):
        R.unpack_restart(bytes(bad))
    other = Trainer(NoDevice(), configs.config1()[0], [], gamma=0.01, algo="ADAM")
    with pytest.raises(ValueError, match="does not match"):
        other.read_restart(tmp_path / "r")
