"""Seeded synthetic structures and potentials for the five benchmark workloads.

Symmetry-function parameters are given in Behler units and mapped to the reference's conventions
(src/nnp/basis_radial.cpp:268-272: exp(-(eta (r-rs))^2), so eta_ref = sqrt(eta_B);
src/nnp/basis_angular.cpp:376-382: exp(-(rij^2+rik^2)/eta^2), so eta_ref = 1/sqrt(eta_B)).
All randomness comes from numpy's PCG64 seeded per workload (the reference's std::mt19937 streams
are not reproduced; oracle and GPU receive the same arrays).
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np

from .model import (ACT_TANH, ANG_GAUSS, CUT_COS, RAD_GAUSSIAN, Angular, Model, Network, Radial,
                    tri_index, xavier_params)


@dataclass
class Structure:
    """Structure: cell columns are lattice vectors (Structure::R()), pos N x 3 in Angstrom."""
    cell: np.ndarray
    pos: np.ndarray
    type: np.ndarray
    energy: float = 0.0
    forces: np.ndarray = None

    @property
    def natoms(self) -> int:
        return len(self.type)


def fcc(n, a):
    base = np.array([[0, 0, 0], [0.5, 0.5, 0], [0.5, 0, 0.5], [0, 0.5, 0.5]])
    g = np.array([[i, j, k] for i in range(n[0]) for j in range(n[1]) for k in range(n[2])], dtype=float)
    pos = (g[:, None, :] + base[None, :, :]).reshape(-1, 3) * a
    return np.diag([n[0] * a, n[1] * a, n[2] * a]).astype(float), pos


def bcc(n, a):
    base = np.array([[0, 0, 0], [0.5, 0.5, 0.5]])
    g = np.array([[i, j, k] for i in range(n[0]) for j in range(n[1]) for k in range(n[2])], dtype=float)
    pos = (g[:, None, :] + base[None, :, :]).reshape(-1, 3) * a
    return np.diag([n[0] * a, n[1] * a, n[2] * a]).astype(float), pos


def random_packed(n, rho, dmin, rng, max_tries=2000):
    """Random sequential addition of n points in a cubic periodic box at density rho with minimum
    image distance >= dmin."""
    L = (n / rho) ** (1.0 / 3.0)
    pos = np.zeros((n, 3))
    k = 0
    tries = 0
    while k < n:
        p = rng.uniform(0, L, size=3)
        if k:
            d = pos[:k] - p
            d -= L * np.round(d / L)
            if np.min(np.einsum("ij,ij->i", d, d)) < dmin * dmin:
                tries += 1
                if tries > max_tries * n:
                    raise RuntimeError("random_packed: density too high")
                continue
        pos[k] = p
        k += 1
    return np.diag([L, L, L]), pos


def wrap(cell, pos):
    f = np.linalg.solve(cell, pos.T).T
    f -= np.floor(f)
    return (cell @ f.T).T


def radial_block(eta_b, rc=6.0):
    eta_b = np.asarray(eta_b, dtype=float)
    return Radial(RAD_GAUSSIAN, CUT_COS, rc, np.zeros(len(eta_b)), np.sqrt(eta_b))


def angular_block(funcs, rc=6.0):
    """funcs: list of (eta_B, zeta, lambda); consecutive equal etas share one exp."""
    eta = np.array([1.0 / np.sqrt(f[0]) for f in funcs])
    zeta = np.array([float(f[1]) for f in funcs])
    lam = np.array([int(f[2]) for f in funcs], dtype=np.int32)
    return Angular(ANG_GAUSS, CUT_COS, rc, eta, zeta, lam)


def make_model(ntypes, rc, rad_eta_b, ang_funcs_per_block, hidden, rng, act=ACT_TANH):
    """Same basis for every centre species; ang_funcs_per_block[b] lists (eta_B, zeta, lambda)."""
    nab = ntypes * (ntypes + 1) // 2
    radial = [[radial_block(rad_eta_b, rc) for _ in range(ntypes)] for _ in range(ntypes)]
    angular = [[angular_block(ang_funcs_per_block[b], rc) for b in range(nab)] for _ in range(ntypes)]
    m = Model(ntypes, rc, radial, angular, [], E0=np.zeros(ntypes))
    for I in range(ntypes):
        K = m.input_width(I)
        widths = [K] + list(hidden) + [1]
        m.nets.append(Network(widths, act, 1.0, xavier_params(widths, rng), np.ones(K), np.zeros(K)))
    return m


def zeta_lambda(zetas):
    return [(z, l) for z in zetas for l in (1, -1)]


def set_input_scaling(model: Model, structures, G, dev_floor=1e-3):
    """inpb = -mean, inpw = 1/max(dev, dev_floor) over the atoms of each species (NNPTEFR::train
    preconditioning with PreBias::MEAN and PreScale::DEV, src/nnp/nnptefr.cpp:786-880).  The floor keeps a
    nearly constant input (e.g. a G2 with eta_B = 2 A^-2 that is ~1e-6 for every atom) from being
    multiplied by 1e5 or more, which would amplify rounding noise into the energies."""
    types = np.concatenate([s.type for s in structures])
    for I in range(model.ntypes):
        K = model.input_width(I)
        g = G[types == I][:, :K]
        if len(g) == 0:
            continue
        mu = g.mean(axis=0)
        dev = np.maximum(g.std(axis=0), dev_floor)
        w = 1.0 / dev
        model.nets[I].inpb = -mu
        model.nets[I].inpw = w


# ---- workloads -----------------------------------------------------------------------------------

def config1(seed=0xACE1):
    """1 element, 3x3x3 fcc a=4.05, jitter 0.1, rc 6, 12 G2 + 8 G4, 20-20-20-1 tanh."""
    rng = np.random.default_rng(seed)
    cell, pos = fcc((3, 3, 3), 4.05)
    pos = wrap(cell, pos + rng.normal(0, 0.1, pos.shape))
    s = Structure(cell, pos, np.zeros(len(pos), dtype=np.int32))
    rad = 0.01 * 200.0 ** (np.arange(12) / 11.0)
    ang = [[(0.005, z, l) for z, l in zeta_lambda([1, 2, 4, 8])]]
    m = make_model(1, 6.0, rad, ang, [20, 20], rng)
    return m, [s]


def c2_basis():
    rad = 0.01 * 200.0 ** (np.arange(16) / 15.0)
    funcs = [(eb, z, l) for eb in (0.005, 0.02) for z, l in zeta_lambda([1, 2, 4, 8])]
    return rad, funcs


def config2(seed=0xACE2):
    """2 elements 50/50, 1000 atoms bcc (5x10x10, rho 0.085), jitter 0.15, rc 6, 80-32-32-1."""
    rng = np.random.default_rng(seed)
    a = (2.0 / 0.085) ** (1.0 / 3.0)
    cell, pos = bcc((5, 10, 10), a)
    pos = wrap(cell, pos + rng.normal(0, 0.15, pos.shape))
    t = np.zeros(len(pos), dtype=np.int32)
    t[rng.permutation(len(pos))[: len(pos) // 2]] = 1
    rad, funcs = c2_basis()
    m = make_model(2, 6.0, rad, [funcs] * 3, [32, 32], rng)
    return m, [Structure(cell, pos, t)]


def c3_model(rng, ntypes=3):
    rad = 0.01 * 200.0 ** (np.arange(8) / 7.0)
    zl = zeta_lambda([1, 2, 4])
    sizes = [6, 6, 5, 5, 5, 5]
    blocks = []
    for n in sizes:
        f = [(0.005 if q < 4 else 0.02, zl[q][0], zl[q][1]) for q in range(n)]
        blocks.append(f)
    return make_model(ntypes, 6.0, rad, blocks, [64, 64], rng)


def config3(seed=0xACE3, nstruc=64, nmin=32, nmax=1024):
    """64 structures, N log-uniform in [32,1024], 3 elements, random packing rho 0.08, rc 6,
    24 G2 + 32 G4 (blocks 6,6,5,5,5,5), 56-64-64-1; teacher of the same architecture."""
    rng = np.random.default_rng(seed)
    strucs = []
    for _ in range(nstruc):
        n = int(round(np.exp(rng.uniform(np.log(nmin), np.log(nmax)))))
        cell, pos = random_packed(n, 0.08, 1.6, rng)
        pos = wrap(cell, pos + rng.normal(0, 0.05, pos.shape))
        comp = rng.dirichlet([1.0, 1.0, 1.0])
        t = rng.choice(3, size=n, p=comp).astype(np.int32)
        strucs.append(Structure(cell, pos, t))
    model = c3_model(rng)
    teacher = c3_model(np.random.default_rng(seed + 1000))
    return model, teacher, strucs


def config4(seed=0xACE4, ncell=30):
    """fcc 30^3 (108,000 atoms), rho 0.085, 70/30, masses 27/16, rc 6 + skin 1, 80-32-32-1."""
    rng = np.random.default_rng(seed)
    a = (4.0 / 0.085) ** (1.0 / 3.0)
    cell, pos = fcc((ncell, ncell, ncell), a)
    pos = wrap(cell, pos + rng.normal(0, 0.05, pos.shape))
    n = len(pos)
    t = np.zeros(n, dtype=np.int32)
    t[rng.permutation(n)[: int(round(0.3 * n))]] = 1
    rad, funcs = c2_basis()
    m = make_model(2, 6.0, rad, [funcs] * 3, [32, 32], rng)
    masses = np.array([27.0, 16.0])
    return m, Structure(cell, pos, t), masses


KB_EV = 8.617333262e-5     # Boltzmann constant, eV/K
MVV2E = 1.0364269e-4       # eV per amu (A/ps)^2 (metal units)


def maxwell_boltzmann(atom_masses, T, rng):
    """Velocities (A/ps) drawn from the Maxwell-Boltzmann distribution at T kelvin (per component
    sigma = sqrt(kB T / m)), with the centre-of-mass velocity removed."""
    m = np.asarray(atom_masses, dtype=float)[:, None]
    v = rng.normal(size=(len(m), 3)) * np.sqrt(KB_EV * T / (m * MVV2E))
    v -= (m * v).sum(axis=0) / m.sum()
    return v


def scale_output(model: Model, factor):
    """Multiplies every network's output layer (weights and bias, the [biases][weights col-major] layout of
    src/ml/nn.cpp:881-896) by factor, which scales NN energies (about outb) and all forces by factor."""
    for n in model.nets:
        w = n.widths
        nb = sum(w[1:])
        n.params = np.array(n.params, dtype=float)
        n.params[nb - 1] *= factor                 # bias of the width-1 output layer
        n.params[len(n.params) - w[-2]:] *= factor   # its 1 x w[-2] weights, stored last


def calibrate_forces(model: Model, F, rms=0.1):
    """Random-initialised networks on a nearly perfect crystal give forces of several eV/A (the input
    scaling divides by the small spread of G); molecular dynamics with 1 fs steps needs the force scale of a
    real solid.  Scales the output layers so that the force components of F have the given rms (eV/A);
    returns the factor."""
    f = rms / float(np.sqrt(np.mean(np.asarray(F) ** 2)))
    scale_output(model, f)
    return f


def c5_model(rng):
    rad = 0.01 * 200.0 ** (np.arange(8) / 7.0)
    zl = zeta_lambda([1, 2, 4, 8, 16])
    sizes = [10] * 6 + [9] * 4
    blocks = [[(0.005 if q < 6 else 0.02, zl[q][0], zl[q][1]) for q in range(n)] for n in sizes]
    return make_model(4, 6.5, rad, blocks, [128, 64], rng)


def config5(seed=0xACE5, nstruc=256, nmin=256, nmax=8192):
    """256 structures, N log-uniform in [256,8192], 4 elements, rho 0.085, rc 6.5,
    32 G2 + 96 G4, 128-128-64-1."""
    rng = np.random.default_rng(seed)
    strucs = []
    a = (4.0 / 0.085) ** (1.0 / 3.0)
    for _ in range(nstruc):
        n = int(round(np.exp(rng.uniform(np.log(nmin), np.log(nmax)))))
        # jittered fcc block trimmed to n sites keeps generation O(n)
        m = int(np.ceil((n / 4.0) ** (1.0 / 3.0)))
        cell, pos = fcc((m, m, m), a)
        keep = np.sort(rng.permutation(len(pos))[:n])
        pos = pos[keep]
        L = m * a * (n / (4.0 * m ** 3)) ** (1.0 / 3.0)
        pos = pos * (L / (m * a))
        cell = np.diag([L, L, L])
        pos = wrap(cell, pos + rng.normal(0, 0.1, pos.shape))
        t = rng.integers(0, 4, size=n).astype(np.int32)
        strucs.append(Structure(cell, pos, t))
    return c5_model(rng), strucs
