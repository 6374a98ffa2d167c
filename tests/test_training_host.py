"""Host-side training pieces: Taylor augmentation (nnptefr.cpp:2733-2840) and ADAM (algo.cpp:280-296)
against the oracle.  CPU only."""
import numpy as np

from ace_amd import configs
from ace_amd.training import Adam, taylor_augment
from oracle import oracle as O


def test_taylor_augmentation_first_order_energy():
    rng = np.random.default_rng(0)
    s1 = configs.Structure(np.eye(3) * 10, rng.uniform(0, 10, (7, 3)), np.zeros(7, dtype=np.int32),
                           energy=-3.2, forces=rng.normal(size=(7, 3)))
    s0 = configs.Structure(np.eye(3) * 10, rng.uniform(0, 10, (5, 3)), np.zeros(5, dtype=np.int32),
                           energy=1.0, forces=np.zeros((5, 3)))          # zero forces: no copies
    out = taylor_augment([s1, s0], nadd=3, rdelta=0.05, rng=rng)
    assert len(out) == 1 + 3 + 1
    assert out[0] is s1 and out[4] is s0
    for c in out[1:4]:
        dr = c.pos - s1.pos
        assert c.forces is None
        assert np.isclose(c.energy, s1.energy - np.sum(dr * s1.forces), rtol=0, atol=1e-13)
        assert np.all(np.linalg.norm(dr, axis=1) < 0.05 * 8)


def test_taylor_direction_is_isotropic():
    rng = np.random.default_rng(1)
    s = configs.Structure(np.eye(3) * 10, np.zeros((4000, 3)), np.zeros(4000, dtype=np.int32),
                          energy=0.0, forces=np.ones((4000, 3)))
    c = taylor_augment([s], nadd=1, rdelta=1.0, rng=rng, dist="uniform")[1]
    u = c.pos / np.linalg.norm(c.pos, axis=1)[:, None]
    assert np.all(np.abs(u.mean(axis=0)) < 0.05)          # uniform on the sphere
    assert np.allclose((u ** 2).mean(axis=0), 1 / 3, atol=0.03)


def test_adam_matches_oracle():
    rng = np.random.default_rng(2)
    dim = 17
    p = rng.normal(size=dim)
    q = p.copy()
    a = Adam(dim, gamma=0.01)
    m, v, b = np.zeros(dim), np.zeros(dim), np.array([0.9, 0.999])
    L = O.lib()
    for _ in range(5):
        g = rng.normal(size=dim)
        p = a.step(p, g)
        L.or_adam_step(dim, O._d(q), O._d(g), 0.01, 1e-16, O._d(m), O._d(v), O._d(b))
    assert np.allclose(p, q, rtol=0, atol=1e-15)


def test_output_scaling_scales_energies_and_forces():
    """configs.scale_output multiplies the output layer, so NN energies (outb = 0, E0 = 0) and forces
    scale by exactly the factor (oracle, FP64)."""
    model, [s] = configs.config1()
    orc = O.Oracle(model)
    E1, F1, _ = orc.compute(s)
    configs.scale_output(model, 0.37)
    orc = O.Oracle(model)
    E2, F2, _ = orc.compute(s)
    assert np.isclose(E2, 0.37 * E1, rtol=1e-12, atol=0)
    assert np.allclose(F2, 0.37 * F1, rtol=1e-10, atol=1e-14)
    f = configs.calibrate_forces(model, F2, rms=0.1)
    assert np.isclose(np.sqrt(np.mean((f * F2) ** 2)), 0.1, rtol=1e-12)


class _DeviceStub:
    """The device calls the trainer makes, recorded (no GPU): a device ADAM restated with the host ADAM."""

    def __init__(self, model):
        from ace_amd.optimizers import ADAM, Objective
        self.model = model
        self.calls = []
        self.obj = Objective(model.all_params(), 0.0)
        self.adam = None

    def train_graph(self, on):
        self.calls.append(("graph", on))

    def set_targets(self, E, F):
        self.calls.append(("targets",))

    def train_optimizer(self, algo, gamma, eps):
        from ace_amd.optimizers import ADAM
        self.calls.append(("optimizer", algo, gamma, eps))
        self.obj.gamma = gamma
        self.adam = ADAM(len(self.obj.p), eps)

    def train_step(self, loss, norm, delta, w_energy, w_force, want_grad=True):
        self.calls.append(("step", want_grad))
        g = np.linspace(-1.0, 1.0, len(self.obj.p))
        out = np.array([1.0, 0.5, 0.5, 0.0])
        if self.adam is not None:
            self.obj.g = g
            self.adam.step(self.obj)
        return out, (g if want_grad else None)

    def get_all_params(self):
        return self.obj.p.copy()

    def set_all_params(self, p):
        self.obj.p = np.array(p, dtype=np.float64)

    def train_opt_state(self, state=None):
        if state is None:
            return {"p": self.obj.p.copy(), "m": self.adam.m.copy(), "v": self.adam.v.copy(),
                    "bt": np.array([self.adam.b1t, self.adam.b2t])}
        self.obj.p = np.array(state["p"])
        self.adam.m, self.adam.v = np.array(state["m"]), np.array(state["v"])
        self.adam.b1t, self.adam.b2t = float(state["bt"][0]), float(state["bt"][1])


def _small():
    model, [s] = configs.config1()
    s.energy, s.forces = -1.0, np.zeros((s.natoms, 3))
    return model, [s]


def test_device_optimizer_refuses_what_it_does_not_restate():
    import pytest
    from ace_amd.optimizers import Decay
    from ace_amd.training import Trainer
    for kw in ({"algo": "SGD"}, {"decay": Decay("EXP", 0.1)}, {"reg": "RIDGE", "eta_reg": 0.1}):
        model, strucs = _small()
        with pytest.raises(ValueError):
            Trainer(_DeviceStub(model), model, strucs, device_opt=True, **kw)


def test_device_optimizer_loop_and_restart(tmp_path):
    """With device_opt the trainer asks for no gradient, leaves the update to the device and reads the
    parameters back on demand; a restart written from the device state and loaded into a new trainer
    continues the same trajectory as an uninterrupted run, and equals the host ADAM on the same
    gradients."""
    from ace_amd.training import Trainer
    model, strucs = _small()
    dev = _DeviceStub(model)
    tr = Trainer(dev, model, strucs, gamma=1e-2, device_opt=True)
    assert ("optimizer", "ADAM", 1e-2, 1e-16) in dev.calls
    for _ in range(3):
        tr.step()
    assert [c for c in dev.calls if c[0] == "step"] == [("step", False)] * 3
    tr.save_restart(str(tmp_path / "r.npz"))
    for _ in range(2):
        tr.step()
    p_cont = tr.params.copy()

    model2, strucs2 = _small()
    dev2 = _DeviceStub(model2)
    tr2 = Trainer(dev2, model2, strucs2, gamma=1e-2, device_opt=True)
    tr2.load_restart(str(tmp_path / "r.npz"))
    for _ in range(2):
        tr2.step()
    assert np.array_equal(tr2.params, p_cont)

    model3, strucs3 = _small()
    dev3 = _DeviceStub(model3)
    host = Trainer(dev3, model3, strucs3, gamma=1e-2)
    for _ in range(5):
        host.step()
    assert np.array_equal(host.params, p_cont)


def _stats_model():
    model, [s] = configs.config1()
    return model


def test_input_statistics_match_library_routines():
    """Per-species count, extremes, mean, sample deviation (N - 1), RMS and harmonic mean
    (nnptefr.cpp:689-786) against numpy / scipy on positive random inputs of two species."""
    from scipy import stats as sps
    from ace_amd.training import input_statistics
    rng = np.random.default_rng(5)
    G = rng.uniform(0.1, 3.0, (57, 6))
    types = rng.integers(0, 2, 57)
    st = input_statistics(types, G, 2)
    for I in range(2):
        g = G[types == I]
        assert st[I]["n"] == len(g)
        assert np.allclose(st[I]["avg"], g.mean(axis=0), rtol=1e-14, atol=0)
        assert np.allclose(st[I]["dev"], g.std(axis=0, ddof=1), rtol=1e-12, atol=0)
        assert np.allclose(st[I]["rms"], np.sqrt((g ** 2).mean(axis=0)), rtol=1e-14, atol=0)
        assert np.allclose(st[I]["har"], sps.hmean(g, axis=0), rtol=1e-13, atol=0)
        assert np.array_equal(st[I]["max"], g.max(axis=0)) and np.array_equal(st[I]["min"], g.min(axis=0))


def test_precondition_every_bias_and_scale():
    """Each PreBias / PreScale of nnptefr.cpp:788-880 with inbias and inscale, the guards for zero
    deviation, empty span and non-positive maximum, the output reset, and MEAN/DEV leaving the scaled
    inputs with zero mean and unit sample deviation."""
    import pytest
    from ace_amd.training import input_statistics, precondition
    model = _stats_model()
    K = model.input_width(0)
    rng = np.random.default_rng(6)
    G = rng.uniform(-1.0, 2.0, (40, K))
    G[:, 0] = 0.5                       # constant column (exact mean): deviation and span 0
    G[:, 1] = -np.abs(G[:, 1]) - 0.1    # negative column: maximum < 0
    types = np.zeros(40, dtype=int)
    st = input_statistics(types, G, 1)[0]
    expect_b = {"IDENTITY": np.zeros(K), "MEAN": -st["avg"], "HAR": -st["har"], "RMS": -st["rms"],
                "MID": -0.5 * (st["max"] + st["min"]), "MIN": -st["min"]}
    for pb, eb in expect_b.items():
        for ps in ("IDENTITY", "DEV", "MINMAX", "MAX"):
            model.nets[0].outb, model.nets[0].outw = 0.3, 2.0
            precondition(model, types, G, pb, ps, inbias=0.25, inscale=1.5)
            net = model.nets[0]
            assert np.allclose(net.inpb, eb + 0.25, rtol=0, atol=1e-15)
            if ps == "DEV":
                ew = np.r_[1.0, 1.0 / st["dev"][1:]]
            elif ps == "MINMAX":
                ew = np.r_[1.0, 1.0 / (st["max"] - st["min"])[1:]]
            elif ps == "MAX":
                ew = np.r_[2.0, 1.0, 1.0 / st["max"][2:]]     # column 0: max 0.5; column 1: max < 0
            else:
                ew = np.ones(K)
            assert np.allclose(net.inpw, 1.5 * ew, rtol=1e-15, atol=0), (pb, ps)
            assert net.outb == 0.0 and net.outw == 1.0
    precondition(model, types, G, "MEAN", "DEV")
    x = model.nets[0].inpw * (G + model.nets[0].inpb)
    assert np.allclose(x.mean(axis=0), 0.0, atol=1e-13)
    assert np.allclose(x[:, 1:].std(axis=0, ddof=1), 1.0, rtol=1e-12)
    with pytest.raises(ValueError):
        precondition(model, types, G, "MEDIAN", "DEV")
    with pytest.raises(ValueError):
        precondition(model, types, G, "MEAN", "RANGE")
    two, _ = configs.config2()
    with pytest.raises(ValueError):
        precondition(two, np.zeros(40, dtype=int), np.zeros((40, two.input_width(0))), "MEAN", "DEV")
