"""Host side of training around the GPU step: Taylor augmentation, ADAM, and the iteration loop.

``taylor_augment`` restates the perturbed-copy augmentation of NNPTEFR (src/nnp/nnptefr.cpp:2733-2840):
each structure whose mean |F| exceeds 1e-12 gets ``nadd`` copies; every atom of a copy moves by
dr = r (sin t cos p, sin t sin p, cos t) with r drawn from the chosen distribution (scale ``rdelta``),
cos t ~ U(-1,1), p ~ U(-pi,pi), and the copy's energy becomes E - sum_atoms dr . F (first-order Taylor
expansion); the copy's forces are dropped.
``Adam`` restates ADAM::step (src/opt/algo.cpp:280-296, beta1 0.9, beta2 0.999, eps 1e-16).
``Trainer.step`` is one iteration of the NNPTEFR::train loop (src/nnp/nnptefr.cpp:977-1081) on one rank:
loss and gradient from the GPU, then the optimiser step and the parameter upload.
"""
from __future__ import annotations

import numpy as np

from .configs import Structure


def taylor_augment(strucs, nadd, rdelta, rng, dist="normal"):
    """Originals first, then ``nadd`` perturbed copies after each structure with forces."""
    draw = {
        "uniform": lambda: rng.uniform(0.0, rdelta),
        "normal": lambda: rng.normal(0.0, rdelta),
        "laplace": lambda: rng.laplace(0.0, rdelta),
        "logistic": lambda: rng.logistic(0.0, rdelta),
    }
    if dist not in draw:
        raise ValueError("Invalid position distribution.")
    out = []
    for s in strucs:
        out.append(s)
        if s.forces is None:
            continue
        favg = np.mean(np.linalg.norm(s.forces, axis=1))
        if favg <= 1.0e-12:
            continue
        for _ in range(nadd):
            pos = s.pos.copy()
            e = float(s.energy)
            for j in range(s.natoms):
                vr = draw[dist]()
                ct = rng.uniform(-1.0, 1.0)
                st = np.sqrt(1.0 - ct * ct)
                ph = rng.uniform(-1.0, 1.0) * np.pi
                dr = np.array([vr * st * np.cos(ph), vr * st * np.sin(ph), vr * ct])
                pos[j] += dr
                e -= dr @ s.forces[j]
            out.append(Structure(s.cell, pos, s.type, energy=e, forces=None))
    return out


PREBIAS = ("IDENTITY", "MEAN", "HAR", "RMS", "MID", "MIN")
PRESCALE = ("IDENTITY", "DEV", "MINMAX", "MAX")


def input_statistics(types, G, ntypes):
    """Per species and input: count, max, min, harmonic mean, mean, sample deviation and RMS of the
    symmetry functions over the training atoms (NNPTEFR::train, src/nnp/nnptefr.cpp:689-786; the MPI
    reductions there are sums, minima and maxima over ranks, so one rank's full set gives the same)."""
    types = np.asarray(types)
    out = []
    for I in range(ntypes):
        g = np.asarray(G[types == I], dtype=np.float64)
        n = len(g)
        if n == 0:
            out.append(None)
            continue
        with np.errstate(divide="ignore"):
            har = n / np.sum(1.0 / g, axis=0)
        avg = np.sum(g, axis=0) / n
        dev = np.sqrt(np.sum((avg - g) * (avg - g), axis=0) / (n - 1.0)) if n > 1 else np.full(g.shape[1], np.nan)
        rms = np.sqrt(np.sum(g * g, axis=0) / n)
        out.append({"n": n, "max": g.max(axis=0), "min": g.min(axis=0), "har": har, "avg": avg, "dev": dev,
                    "rms": rms})
    return out


def precondition(model, types, G, prebias="MEAN", prescale="DEV", inbias=0.0, inscale=1.0):
    """The network input bias and weight of every species from the input statistics, and the output
    bias/weight reset to (0, 1) (src/nnp/nnptefr.cpp:788-880).  G: one row per training atom (the species'
    inputs first in each row), types: that atom's species.  prebias: IDENTITY, MEAN (-mean), HAR (-harmonic
    mean), RMS (-rms), MID (-(max + min)/2), MIN (-min), then + inbias; prescale: IDENTITY, DEV (1/dev),
    MINMAX (1/(max - min)), MAX (1/max), 1 where the statistic is not positive, then x inscale.  A species
    with no training atom raises ValueError (the reference divides by its zero count)."""
    pb, ps = prebias.upper(), prescale.upper()
    if pb not in PREBIAS:
        raise ValueError("Invalid input bias.")
    if ps not in PRESCALE:
        raise ValueError("Invalid input scaling.")
    stats = input_statistics(types, G, model.ntypes)
    for I in range(model.ntypes):
        K = model.input_width(I)
        st = stats[I]
        if st is None:
            raise ValueError("species %d has no atom in the training set" % I)
        st = {k: (v[:K] if isinstance(v, np.ndarray) else v) for k, v in st.items()}
        bias = {"IDENTITY": np.zeros(K), "MEAN": -st["avg"], "HAR": -st["har"], "RMS": -st["rms"],
                "MID": -0.5 * (st["max"] + st["min"]), "MIN": -st["min"]}[pb]
        if ps == "DEV":
            w = np.where(st["dev"] == 0, 1.0, 1.0 / np.where(st["dev"] == 0, 1.0, st["dev"]))
        elif ps == "MINMAX":
            span = st["max"] - st["min"]
            w = np.where(span > 0, 1.0 / np.where(span > 0, span, 1.0), 1.0)
        elif ps == "MAX":
            w = np.where(st["max"] > 0, 1.0 / np.where(st["max"] > 0, st["max"], 1.0), 1.0)
        else:
            w = np.ones(K)
        net = model.nets[I]
        net.inpb = bias + inbias
        net.inpw = w * inscale
        net.outb, net.outw = 0.0, 1.0


class Adam:
    """ADAM::step: m = b1 m + (1-b1) g, v = b2 v + (1-b2) g^2, p -= gamma m/(1-b1^t) / (sqrt(v/(1-b2^t)) + eps)."""

    def __init__(self, dim, gamma=1e-3, eps=1e-16):
        self.m = np.zeros(dim)
        self.v = np.zeros(dim)
        self.b1t, self.b2t = 0.9, 0.999
        self.gamma, self.eps = gamma, eps

    def step(self, p, g):
        b1, b2 = 0.9, 0.999
        b1ii = 1.0 / (1.0 - self.b1t)
        b2ii = 1.0 / (1.0 - self.b2t)
        self.m = self.m * b1 + (1.0 - b1) * g
        self.v = self.v * b2 + (1.0 - b2) * g * g
        p -= self.gamma * b1ii * self.m / (np.sqrt(b2ii * self.v) + self.eps)
        self.b1t *= b1
        self.b2t *= b2
        return p


class Trainer:
    """Full-batch training on one GPU: every uploaded structure is in the batch.  The update follows the
    training loop of src/nnp/nnptefr.cpp:1031-1066: gamma <- decay(gamma), the optimiser step (any of
    src/opt/algo.cpp, default ADAM), weight decay of the weights from the previous parameters, then the
    "old" value, parameters and gradient.

    device_opt=True runs the ADAM update on the GPU after each gradient (constant gamma, no weight decay:
    the configuration it restates exactly); the parameters then stay on the device between steps and
    `params`, `save_restart` read them back."""

    def __init__(self, dev, model, strucs, gamma=1e-3, w_force=0.1, loss=0, norm=0, delta=1.0, algo="ADAM",
                 decay=None, reg="NONE", eta_reg=0.0, graph=False, device_opt=False):
        from .optimizers import Decay, Objective, make
        self.dev, self.model = dev, model
        self.nat = sum(s.natoms for s in strucs)
        self.nbatch = len(strucs)
        self.obj = Objective(model.all_params(), gamma)
        self.algo = make(algo, len(self.obj.p))
        self.decay = decay if decay is not None else Decay()
        self.reg, self.eta_reg = reg, eta_reg
        # weight mask: per species the biases (0) then the weights (1), the parameter vector order
        self.mask = np.concatenate([np.r_[np.zeros(sum(n.widths[1:])), np.ones(n.nparams() - sum(n.widths[1:]))]
                                    for n in model.nets])
        self.it = 0
        self.w_force, self.loss, self.norm, self.delta = w_force, loss, norm, delta
        if graph:
            dev.train_graph(True)
        self.device_opt = device_opt
        if device_opt:
            if algo.upper() != "ADAM" or self.decay.name != "CONST" or reg.upper() != "NONE":
                raise ValueError("the device optimiser runs ADAM with constant gamma and no weight decay")
            dev.train_optimizer("ADAM", gamma, self.algo.eps)
        have_f = bool(strucs) and all(s.forces is not None for s in strucs)
        dev.set_targets([s.energy for s in strucs],
                        np.concatenate([s.forces for s in strucs]) if have_f else None)

    @property
    def params(self):
        if self.device_opt:
            self.obj.p = self.dev.get_all_params()
            self.model.set_all_params(self.obj.p)
        return self.obj.p

    def _pull_device_state(self):
        st = self.dev.train_opt_state()
        self.obj.p = st["p"]
        self.algo.m, self.algo.v = st["m"], st["v"]
        self.algo.b1t, self.algo.b2t = float(st["bt"][0]), float(st["bt"][1])
        self.model.set_all_params(self.obj.p)

    def update(self, loss, g):
        """One parameter update from the loss value and its gradient (host, FP64)."""
        from .optimizers import weight_decay
        o = self.obj
        o.g = np.array(g, dtype=np.float64)
        o.val = float(loss)
        o.gamma = self.decay.step(o.gamma, self.it)
        self.algo.step(o)
        weight_decay(o, self.mask, self.eta_reg, self.reg)
        self.dv, self.dp = abs(o.val - o.val_old), float(np.linalg.norm(o.p - o.p_old))   # nnptefr.cpp:1060-1061
        o.val_old, o.p_old, o.g_old = o.val, o.p.copy(), o.g.copy()
        self.it += 1
        return o.p

    def save_restart(self, path):
        """Training state for a restart (the content of NNPTEFR::write_restart, src/nnp/nnptefr.cpp:509-537:
        parameters, optimiser state, gamma, iteration, old value/parameters/gradient) as a portable .npz
        rather than the reference's memory image of the trainer object."""
        if self.device_opt:
            self._pull_device_state()
        o = self.obj
        state = {"p": o.p, "p_old": o.p_old, "g_old": o.g_old, "val_old": o.val_old, "gamma": o.gamma,
                 "it": self.it, "algo": type(self.algo).__name__}
        for k, v in vars(self.algo).items():
            state["algo_" + k] = np.asarray(v)
        np.savez(path, **state)

    def write_restart(self, path, files=("", "", "", ""), n_print=1, n_write=1, max_iter=0):
        """NNPTEFR::write_restart (src/nnp/nnptefr.cpp:509-537) in the reference's binary layout (restart.py):
        the potential, the full training batch in order, the objective (gamma, last loss, parameters, last
        gradient), the iteration count, the optimiser's state, decay, loss, norm and weight decay."""
        from . import restart as R
        if self.device_opt:
            self._pull_device_state()
        o = self.obj
        self.model.set_all_params(o.p)
        name = type(self.algo).__name__
        scal, vecs = R.ALGO_FIELDS[name]
        dim = len(o.p)
        state = {ref: float(getattr(self.algo, attr)) for ref, attr in scal}
        for ref, attr in vecs:
            v = getattr(self.algo, attr, None)
            state[ref] = np.zeros((dim, dim) if ref in ("B", "BOld") else dim) if v is None else np.asarray(v)
        order = np.arange(self.nbatch, dtype=np.int32)
        st = R.RestartState(
            model=self.model, files=tuple(files), batch_train=R.Batch(0, self.nbatch, self.nbatch, order, order),
            gamma=o.gamma, val=o.val_old, val_old=o.val_old, dv=getattr(self, "dv", 0.0), dp=getattr(self, "dp", 0.0),
            p=o.p, p_old=o.p_old, g=o.g_old, g_old=o.g_old, n_print=n_print, n_write=n_write, step=self.it,
            count=self.it, max_iter=max_iter, loss=R.LOSS_NAMES[self.loss + 1], algo=name, algo_state=state,
            decay=self.decay.name, decay_period=self.decay.period, decay_alpha=self.decay.alpha,
            norm=R.NORM_NAMES[self.norm], reg=self.reg.upper(), delta=self.delta, eta=self.eta_reg)
        R.write_restart(path, st)

    def read_restart(self, path):
        """NNPTEFR::read_restart (nnptefr.cpp:539-574) of a file in the reference's layout: parameters, the
        objective, the optimiser (as the reference unpacks it: a fresh optimiser holding the file's state,
        so ADAM-type bias-correction powers restart), decay, loss, norm, weight decay and the iteration
        count.  The file's potential must have this trainer's species, bases and network shapes."""
        from . import restart as R
        from .optimizers import Decay, make
        st = R.read_restart(path)
        m = self.model
        if st.model.ntypes != m.ntypes or any(a.widths != b.widths for a, b in zip(st.model.nets, m.nets)) or \
                any(st.model.input_width(I) != m.input_width(I) for I in range(m.ntypes)):
            raise ValueError("read_restart: the restart's potential does not match the trainer's model")
        if st.algo is None or st.p is None:
            raise ValueError("read_restart: the restart holds no optimiser state")
        o = self.obj
        o.p, o.p_old, o.g_old = st.p.copy(), st.p_old.copy(), st.g_old.copy()
        o.val_old, o.gamma, self.it = st.val_old, st.gamma, st.count
        self.dv, self.dp = st.dv, st.dp
        self.algo = make(st.algo, len(o.p))
        scal, vecs = R.ALGO_FIELDS[st.algo]
        for ref, attr in scal:
            setattr(self.algo, attr, st.algo_state[ref])
        for ref, attr in vecs:
            if hasattr(self.algo, attr) and ref in st.algo_state:
                setattr(self.algo, attr, np.array(st.algo_state[ref]))
        self.decay = Decay(st.decay, st.decay_alpha, st.decay_period)
        self.loss = R.LOSS_NAMES.index(st.loss) - 1
        self.norm = R.NORM_NAMES.index(st.norm)
        self.delta, self.reg, self.eta_reg = st.delta, st.reg, st.eta
        self.model.set_all_params(o.p)
        if self.device_opt:
            self.dev.train_opt_state({"p": o.p, "m": self.algo.m, "v": self.algo.v,
                                      "bt": np.array([self.algo.b1t, self.algo.b2t])})
        elif hasattr(self.dev, "set_all_params"):
            self.dev.set_all_params(o.p)

    def load_restart(self, path):
        """Continues from a state written by save_restart (NNPTEFR::read_restart, nnptefr.cpp:539-574)."""
        from .optimizers import make
        z = np.load(path, allow_pickle=False)
        o = self.obj
        o.p, o.p_old, o.g_old = z["p"].copy(), z["p_old"].copy(), z["g_old"].copy()
        o.val_old, o.gamma, self.it = float(z["val_old"]), float(z["gamma"]), int(z["it"])
        self.algo = make(str(z["algo"]), len(o.p))
        for k in vars(self.algo):
            v = z["algo_" + k]
            setattr(self.algo, k, v.copy() if v.ndim else v.item())
        self.model.set_all_params(o.p)
        if self.device_opt:
            self.dev.train_opt_state({"p": o.p, "m": self.algo.m, "v": self.algo.v,
                                      "bt": np.array([self.algo.b1t, self.algo.b2t])})
        elif hasattr(self.dev, "set_all_params"):
            self.dev.set_all_params(o.p)

    def step(self):
        """Loss terms of one iteration; the parameters are updated on the host and uploaded, or on the
        device (device_opt)."""
        wf = self.w_force / (3 * self.nat) if self.w_force else 0.0
        if self.device_opt:
            out, _ = self.dev.train_step(self.loss, self.norm, self.delta, 1.0 / self.nbatch, wf, want_grad=False)
            self.obj.val_old = self.obj.val = float(out[0])
            self.it += 1
            return out
        out, g = self.dev.train_step(self.loss, self.norm, self.delta, 1.0 / self.nbatch, wf)
        p = self.update(out[0], g)
        self.dev.set_all_params(p)
        self.model.set_all_params(p)
        return out
