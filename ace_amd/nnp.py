"""Host interface over the C-ABI, named after the reference's ``NNP`` / ``NNPTEFR`` methods.

``NNPDevice`` mirrors ``NNP::compute_symm``, ``compute_force``, ``compute_energy`` and ``compute``
(src/nnp/nnp.hpp:101-119) for a batch of structures held in HBM, and ``train_step`` mirrors one
iteration of ``NNPTEFR::error_cost_o1`` (src/nnp/nnptefr.cpp:1292-1425) with the force-matching term.
Errors are raised as ``RuntimeError`` carrying ``ace_last_error`` (the reference throws
``std::invalid_argument`` / ``std::runtime_error``).
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib
from .model import LossDesc, Model

PREC_FP32, PREC_BF16L1, PREC_FP64_BASIS = 0, 1, 2


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double)) if a is not None else None


def _ip(a):
    return a.ctypes.data_as(C.POINTER(C.c_int)) if a is not None else None


class NNPDevice:
    """One HIP context: model tables, one stream, the uploaded batch of structures."""

    def __init__(self, model: Model):
        self.lib = _lib.load()
        self.model = model
        self.ctx = C.c_void_p()
        self._desc = model.desc()
        rc = self.lib.ace_create(C.byref(self.ctx), C.cast(C.pointer(self._desc), C.c_void_p))
        if rc != 0:
            msg = self._err()
            self.close()
            raise RuntimeError(f"ace_create failed ({rc}): {msg}")
        self.structures = []
        self.N = 0

    # -- plumbing --------------------------------------------------------------------------------
    @classmethod
    def from_file(cls, path):
        """A context for the potential in an ace potential file (NNP::read, src/nnp/nnp.cpp:233-307)."""
        from .potential import read_potential
        return cls(read_potential(path))

    def _err(self):
        if not self.ctx:
            return "no context"
        m = self.lib.ace_last_error(self.ctx)
        return m.decode() if m else ""

    def _check(self, rc, what):
        if rc != 0:
            raise RuntimeError(f"{what} failed ({rc}): {self._err()}")

    def close(self):
        if self.ctx:
            self.lib.ace_destroy(self.ctx)
            self.ctx = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def sync(self):
        self._check(self.lib.ace_sync(self.ctx), "ace_sync")

    # -- model -----------------------------------------------------------------------------------
    def input_width(self, I):
        return self.lib.ace_input_width(self.ctx, I)

    def set_params(self, I, params):
        p = np.ascontiguousarray(params, dtype=np.float64)
        self._check(self.lib.ace_set_params(self.ctx, I, _dp(p), len(p)), "ace_set_params")

    def get_params(self, I):
        n = self.lib.ace_param_count(self.ctx, I)
        p = np.zeros(n)
        self._check(self.lib.ace_get_params(self.ctx, I, _dp(p), n), "ace_get_params")
        return p

    def set_all_params(self, x):
        off = 0
        for I in range(self.model.ntypes):
            n = self.lib.ace_param_count(self.ctx, I)
            self.set_params(I, x[off:off + n])
            off += n

    def set_input_scaling(self, I, inpw, inpb):
        w = np.ascontiguousarray(inpw, dtype=np.float64)
        b = np.ascontiguousarray(inpb, dtype=np.float64)
        self._check(self.lib.ace_set_input_scaling(self.ctx, I, _dp(w), _dp(b), len(w)), "ace_set_input_scaling")

    def sync_model_scaling(self):
        for I, n in enumerate(self.model.nets):
            self.set_input_scaling(I, n.inpw, n.inpb)

    # -- structures ------------------------------------------------------------------------------
    def upload(self, structures):
        self.structures = list(structures)
        nat = np.array([s.natoms for s in structures], dtype=np.int32)
        pos = np.ascontiguousarray(np.concatenate([s.pos for s in structures]), dtype=np.float64)
        typ = np.ascontiguousarray(np.concatenate([s.type for s in structures]), dtype=np.int32)
        cell = np.ascontiguousarray(np.concatenate([np.asarray(s.cell, dtype=np.float64).T.reshape(-1)
                                                    for s in structures]))
        self.N = int(nat.sum())
        self.natoms = nat
        self._check(self.lib.ace_upload(self.ctx, len(nat), _ip(nat), _dp(pos), _ip(typ), _dp(cell)), "ace_upload")

    def update_positions(self, pos):
        p = np.ascontiguousarray(pos, dtype=np.float64)
        assert p.shape == (self.N, 3)
        self._check(self.lib.ace_update_positions(self.ctx, _dp(p)), "ace_update_positions")

    def build_neighbors(self, skin=0.0):
        """NeighborList::build for every uploaded structure (src/struc/neighbor.cpp:46-87)."""
        self._check(self.lib.ace_build_neighbors(self.ctx, float(skin)), "ace_build_neighbors")

    def neighbors(self):
        """(offsets[N+1], j local index, img[P,3]) in the reference order."""
        n = C.c_longlong()
        self._check(self.lib.ace_neighbor_count(self.ctx, C.byref(n)), "ace_neighbor_count")
        off = np.zeros(self.N + 1, dtype=np.int64)
        j = np.zeros(max(n.value, 1), dtype=np.int32)
        img = np.zeros((max(n.value, 1), 3), dtype=np.int32)
        self._check(self.lib.ace_export_neighbors(self.ctx, off.ctypes.data_as(C.POINTER(C.c_longlong)), _ip(j),
                                                  _ip(img)), "ace_export_neighbors")
        return off, j[: n.value], img[: n.value]

    # -- NNP methods -----------------------------------------------------------------------------
    def set_precision(self, precision):
        """Precision of compute_symm and train_step (PREC_FP32 or PREC_FP64_BASIS)."""
        self._check(self.lib.ace_set_precision(self.ctx, precision), "ace_set_precision")

    def compute_symm(self, ldg=None):
        """NNP::compute_symm: G row per atom (FP32 on the GPU, exported as FP64)."""
        kmax = max(self.model.input_width(I) for I in range(self.model.ntypes))
        ldg = kmax if ldg is None else ldg
        G = np.zeros((self.N, ldg))
        self._check(self.lib.ace_symm(self.ctx, _dp(G), ldg), "ace_symm")
        return G

    def compute(self, precision=PREC_FP32, want_virial=True):
        """NNP::compute over the batch: per-structure energies (with E0), forces, virials."""
        E = np.zeros(len(self.natoms))
        F = np.zeros((self.N, 3))
        W = np.zeros((len(self.natoms), 9)) if want_virial else None
        self._check(self.lib.ace_energy_forces(self.ctx, precision, _dp(E), _dp(F), _dp(W)), "ace_energy_forces")
        return E, F, (W.reshape(-1, 3, 3) if want_virial else None)

    def compute_energy(self, precision=PREC_FP32):
        """NNP::compute_energy convention: sum over atoms of (NN + E0)."""
        return self.compute(precision, want_virial=False)[0]

    def compute_force(self, precision=PREC_FP32):
        """NNP::compute_force convention: fills forces, returns sum of NN outputs without E0."""
        E, F, _ = self.compute(precision, want_virial=False)
        e0 = np.asarray(self.model.E0 if self.model.E0 is not None else np.zeros(self.model.ntypes))
        off = np.concatenate([[0], np.cumsum(self.natoms)])
        typ = np.concatenate([s.type for s in self.structures])
        pe = np.array([E[s] - e0[typ[off[s]:off[s + 1]]].sum() for s in range(len(self.natoms))])
        return pe, F

    def compute_device(self, precision=PREC_FP32):
        """Energy/force pass whose results stay in HBM (timed loops)."""
        self._check(self.lib.ace_energy_forces(self.ctx, precision, None, None, None), "ace_energy_forces")

    def compute_step(self, precision=PREC_FP32, skin=0.0):
        """NNP::compute(struc) (nnp.cpp:489-494) for the whole batch: neighbour list, then energies and
        forces, in one call; results stay in HBM."""
        self._check(self.lib.ace_compute(self.ctx, precision, float(skin), None, None, None), "ace_compute")

    def autotune(self, precision=PREC_FP32):
        """Measures the energy/force paths on the uploaded batch and keeps the fastest; returns the
        milliseconds of every path of PATHS (fused, split with staged weights, split with weights in memory;
        four or two waves per atom) and the choice."""
        n = len(self.PATHS)
        ms = np.zeros(n + 1)
        self._check(self.lib.ace_autotune(self.ctx, precision, _dp(ms)), "ace_autotune")
        names = self.PATHS
        return {"ms": {names[i]: float(ms[i]) for i in range(n) if ms[i] >= 0}, "choice": names[int(ms[n])],
                "tile": self.tile_shape(precision)}

    PATHS = ("fused", "split_staged", "split", "fused_2w", "split_staged_2w", "split_2w")

    def tile_shape(self, precision=PREC_FP32):
        """The shapes the energy/force pass uses: path (None: size rule), waves per atom, network weights staged
        (None: not tuned), the angular chunk weight w and the milliseconds measured for w = 1, 3, 8, the level
        split, and the waves per network tile with the milliseconds measured for 1, 2, 4."""
        v = np.zeros(13)
        self._check(self.lib.ace_tile_shape(self.ctx, precision, _dp(v)), "ace_tile_shape")
        return {"path": None if v[0] < 0 else self.PATHS[int(v[0])], "waves": int(v[1]),
                "staged": None if v[2] < 0 else bool(v[2]), "chunk_w": int(v[3]),
                "chunk_ms": {w: float(v[5 + k]) for k, w in enumerate((1, 3, 8)) if v[5 + k] >= 0},
                "level_split": int(v[8]), "net_waves": int(v[9]),
                "net_waves_ms": {w: float(v[10 + k]) for k, w in enumerate((1, 2, 4)) if v[10 + k] >= 0}}

    def set_tile_shape(self, chunk_w=0, net_waves=0, precision=PREC_FP32):
        """Fixes the angular chunk weight (1, 3, 8) and the waves per network tile (1, 2, 4); 0: tuned/default."""
        self._check(self.lib.ace_set_tile_shape(self.ctx, precision, chunk_w, net_waves), "ace_set_tile_shape")

    def set_level_split(self, enable):
        """Level-split angular kernels (two eta groups over disjoint zeta levels) on or off."""
        self._check(self.lib.ace_set_level_split(self.ctx, 1 if enable else 0), "ace_set_level_split")

    def set_force_gather(self, enable):
        """Split energy/force passes: per-pair gradients summed by atom (True, default; from the next
        neighbour build) or fixed-point atomics per pair (False)."""
        self._check(self.lib.ace_set_force_gather(self.ctx, 1 if enable else 0), "ace_set_force_gather")

    def set_path(self, path, precision=PREC_FP32):
        """Fixes the energy/force path (a name of PATHS, or None for the size rule) for every batch."""
        k = -1 if path is None else self.PATHS.index(path)
        self._check(self.lib.ace_set_path(self.ctx, precision, k), "ace_set_path")

    def md_init(self, masses, vel=None):
        """Per-species masses (amu) and velocities (N x 3, A/ps; None = at rest) for md_run."""
        m = np.ascontiguousarray(masses, dtype=np.float64)
        v = None if vel is None else np.ascontiguousarray(vel, dtype=np.float64).reshape(self.N, 3)
        self._check(self.lib.ace_md_init(self.ctx, _dp(m), _dp(v)), "ace_md_init")

    def md_run(self, nsteps, dt=1e-3, rebuild_every=20, skin=1.0, precision=PREC_FP32, graph=True):
        """nsteps velocity-Verlet steps of dt ps (metal units); returns per-step (potential, kinetic) energy in
        eV, shape (nsteps, 2)."""
        e = np.zeros((max(nsteps, 1), 2))
        self._check(self.lib.ace_md_run(self.ctx, precision, nsteps, float(dt), rebuild_every, float(skin),
                                        1 if graph else 0, _dp(e)), "ace_md_run")
        return e[:nsteps]

    def md_state(self):
        """Current positions (A) and velocities (A/ps) on the device."""
        pos = np.zeros((self.N, 3))
        vel = np.zeros((self.N, 3))
        self._check(self.lib.ace_md_state(self.ctx, _dp(pos), _dp(vel)), "ace_md_state")
        return pos, vel

    def plain_compute(self):
        """The unbatched FP64 GPU port: per-structure energies and forces."""
        E = np.zeros(len(self.natoms))
        F = np.zeros((self.N, 3))
        self._check(self.lib.ace_plain_energy_forces(self.ctx, _dp(E), _dp(F)), "ace_plain_energy_forces")
        return E, F

    # -- training --------------------------------------------------------------------------------
    def set_targets(self, E_ref, F_ref=None):
        e = np.ascontiguousarray(E_ref, dtype=np.float64)
        f = None if F_ref is None else np.ascontiguousarray(F_ref, dtype=np.float64).reshape(-1)
        self._check(self.lib.ace_set_targets(self.ctx, _dp(e), _dp(f)), "ace_set_targets")

    def train_graph(self, on=True):
        """Replay the device part of train_step from a captured HIP graph."""
        self._check(self.lib.ace_train_graph(self.ctx, 1 if on else 0), "ace_train_graph")

    def train_step(self, loss=0, norm=0, delta=1.0, w_energy=1.0, w_force=0.0, want_grad=True):
        d = LossDesc(loss, norm, delta, w_energy, w_force)
        out = np.zeros(4)
        g = np.zeros(sum(self.model.param_counts())) if want_grad else None
        self._check(self.lib.ace_train_step(self.ctx, C.cast(C.pointer(d), C.c_void_p), _dp(out), _dp(g)),
                    "ace_train_step")
        return out, g

    def forward(self):
        """Autograd-style forward at the current parameters and positions (precision of set_precision):
        per-structure energies (with E0), forces, virials; keeps the network inputs for backward()."""
        E = np.zeros(len(self.natoms))
        F = np.zeros((self.N, 3))
        W = np.zeros((len(self.natoms), 9))
        self._check(self.lib.ace_forward(self.ctx, _dp(E), _dp(F), _dp(W)), "ace_forward")
        return E, F, W.reshape(-1, 3, 3)

    def backward(self, dL_dE=None, dL_dF=None, want_params=True, want_pos=False):
        """Backward of forward(): from upstream dL/dE (per structure) and dL/dF (per atom, 3) of a loss L(E, F),
        dL/dparams (species concatenated, get_all_params order) and dL/dpos (energy term only; needs
        dL_dF None)."""
        e = None if dL_dE is None else np.ascontiguousarray(dL_dE, dtype=np.float64)
        f = None if dL_dF is None else np.ascontiguousarray(dL_dF, dtype=np.float64).reshape(-1)
        gp = np.zeros(sum(self.model.param_counts())) if want_params else None
        gx = np.zeros((self.N, 3)) if want_pos else None
        self._check(self.lib.ace_backward(self.ctx, _dp(e), _dp(f), _dp(gp), _dp(gx)), "ace_backward")
        return gp, gx

    def train_optimizer(self, algo=None, gamma=1e-3, eps=1e-16):
        """algo "ADAM": every later train_step also applies one ADAM step on the GPU (FP64 master
        parameters, device layouts refreshed in place); None: the caller updates the parameters."""
        k = {None: 0, "ADAM": 1}.get(None if algo is None else algo.upper(), -1)
        if k < 0:
            raise ValueError("only ADAM runs on the device")
        self._check(self.lib.ace_train_optimizer(self.ctx, k, gamma, eps), "ace_train_optimizer")

    def train_opt_state(self, state=None):
        """The device ADAM's state as a dict (p, m, v, bt), or sets it from one."""
        n = sum(self.model.param_counts())
        if state is None:
            p, m, v, bt = np.zeros(n), np.zeros(n), np.zeros(n), np.zeros(2)
            self._check(self.lib.ace_train_opt_state(self.ctx, 0, _dp(p), _dp(m), _dp(v), _dp(bt)),
                        "ace_train_opt_state")
            return {"p": p, "m": m, "v": v, "bt": bt}
        arr = {k: np.ascontiguousarray(state[k], dtype=np.float64) for k in ("p", "m", "v", "bt")}
        self._check(self.lib.ace_train_opt_state(self.ctx, 1, _dp(arr["p"]), _dp(arr["m"]), _dp(arr["v"]),
                                                 _dp(arr["bt"])), "ace_train_opt_state")

    def get_all_params(self):
        return np.concatenate([self.get_params(I) for I in range(self.model.ntypes)])

    # -- timing ----------------------------------------------------------------------------------
    def profile(self, on=True):
        self._check(self.lib.ace_profile_enable(self.ctx, 1 if on else 0), "ace_profile_enable")

    def profile_read(self, kernel):
        ms = C.c_double()
        n = C.c_longlong()
        self._check(self.lib.ace_profile_read(self.ctx, kernel.encode(), C.byref(ms), C.byref(n)), "ace_profile_read")
        return ms.value, n.value
