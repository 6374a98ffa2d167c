"""Velocity-Verlet molecular dynamics on the GPU (ace_md_*), through the C-ABI.

* Graph replay and plain launches give bit-identical trajectories (fixed-point force sums, same kernels).
* A 108-atom crystal follows a Verlet loop driven by the FP64 oracle's forces (the integrator template of
  src/sim/integrator.cpp:548-571 with metal units): after 24 steps positions agree to 1e-6 A, velocities
  to 1e-3 A/ps, and per-step potential/kinetic energies to 1e-6 eV/atom (FP64 symmetry functions) --
  the FP32 network error of ~2e-5 eV/A in the forces moves positions by ~1e-8 A in 24 fs.
* The 108,000-atom two-species fcc cell at 300 K (list radius 7 A, rebuilt every 20 steps; output layers
  scaled to an rms force component of 0.1 eV/A, a real solid's scale, see configs.calibrate_forces): total energy
  drift in FP32 and in FP64 symmetry functions, each below 1e-4 eV/atom over 200 steps, and the two
  runs' drifts within 2e-5 eV/atom of each other.
"""
import numpy as np
import pytest

from ace_amd import PREC_FP32, PREC_FP64_BASIS, NNPDevice, configs
from oracle import oracle as O

from conftest import log

pytestmark = pytest.mark.gpu

MVV2E = configs.MVV2E


def oracle_verlet(model, s, masses, v0, nsteps, dt):
    """Half kick, drift with the return to the cell, forces, half kick (integrator.cpp:548-571) with the
    oracle's FP64 forces; per-step (PE, KE)."""
    orc = O.Oracle(model)
    m = np.asarray(masses)[s.type][:, None]
    pos, vel = s.pos.copy(), v0.copy()
    _, F, _ = orc.compute(s)
    rec = []
    for _ in range(nsteps):
        vel = vel + 0.5 * dt * F / (m * MVV2E)
        pos = configs.wrap(s.cell, pos + dt * vel)
        E, F, _ = orc.compute(configs.Structure(s.cell, pos, s.type))
        vel = vel + 0.5 * dt * F / (m * MVV2E)
        rec.append((E, 0.5 * (m[:, 0] * MVV2E * (vel ** 2).sum(1)).sum()))
    return pos, vel, np.array(rec)


def periodic_dist(cell, a, b):
    """|a - b| under the minimum image (positions returned to the cell may sit on opposite faces)."""
    d = np.linalg.solve(cell, (a - b).T).T
    d -= np.round(d)
    return np.linalg.norm(d @ np.asarray(cell).T, axis=1)


def test_md_graph_matches_eager_and_oracle():
    model, [s] = configs.config1()
    orc = O.Oracle(model)
    configs.set_input_scaling(model, [s], orc.symm(s))
    masses = [27.0]
    v0 = configs.maxwell_boltzmann(np.full(s.natoms, 27.0), 600.0, np.random.default_rng(3))
    runs = {}
    for graph in (True, False):
        dev = NNPDevice(model)
        dev.upload([s])
        dev.md_init(masses, v0)
        e = dev.md_run(24, dt=1e-3, rebuild_every=8, skin=1.0, precision=PREC_FP64_BASIS, graph=graph)
        runs[graph] = (e,) + dev.md_state()
        dev.close()
    (eg, pg, vg), (ee, pe, ve) = runs[True], runs[False]
    assert np.array_equal(eg, ee) and np.array_equal(pg, pe) and np.array_equal(vg, ve), "graph != plain launches"
    pr, vr, er = oracle_verlet(model, s, masses, v0, 24, 1e-3)
    dp, dv = periodic_dist(s.cell, pg, pr).max(), np.abs(vg - vr).max()
    de = np.abs(eg - er).max() / s.natoms
    log(f"md c1 24 steps: max|dx| {dp:.1e} A, max|dv| {dv:.1e} A/ps, max|dE|/atom {de:.1e} eV")
    assert dp < 1e-6 and dv < 1e-3 and de < 1e-6


def test_md_c4_energy_drift():
    model, s, masses = configs.config4()
    rng = np.random.default_rng(0xACE4)
    v0 = configs.maxwell_boltzmann(masses[s.type], 300.0, rng)
    dev = NNPDevice(model)
    dev.upload([s])
    dev.build_neighbors()
    configs.set_input_scaling(model, [s], dev.compute_symm())
    dev.sync_model_scaling()
    _, F, _ = dev.compute()
    configs.calibrate_forces(model, F, rms=0.1)
    dev.close()
    drift = {}
    for prec in (PREC_FP32, PREC_FP64_BASIS):
        dev = NNPDevice(model)
        dev.upload([s])
        dev.md_init(masses, v0)
        e = dev.md_run(200, dt=1e-3, rebuild_every=20, skin=1.0, precision=prec, graph=True)
        tot = e.sum(axis=1)
        drift[prec] = (tot - tot[0]) / s.natoms
        T = 2 * e[-1, 1] / (3 * s.natoms * configs.KB_EV)
        log(f"md c4 prec {prec}: E0 {tot[0] / s.natoms:.6f} eV/atom, max|drift| {np.abs(drift[prec]).max():.2e} "
            f"eV/atom, end drift {drift[prec][-1]:.2e}, T_end {T:.1f} K")
        dev.close()
    for prec in drift:
        assert np.abs(drift[prec]).max() < 1e-4
    assert np.abs(drift[PREC_FP32] - drift[PREC_FP64_BASIS]).max() < 2e-5


def test_md_fast_atom_crosses_the_cell_many_times():
    """An atom fast enough to cross its 7.5 A cell about six times (more than the image range
    ceil(3 rc / |a|) = 3 the list search covers) in 60 steps with the list rebuilt every step: the positions
    returned to the cell at each rebuild keep every pair in the list, and the trajectory follows the
    oracle-driven Verlet (which returns atoms to the cell after every drift, integrator.cpp:553)."""
    rng = np.random.default_rng(12)
    m = configs.c3_model(rng)
    L_ = 7.5
    cell = np.eye(3) * L_
    pos = np.array([[1.0, 1.0, 1.0], [4.0, 4.5, 3.5], [2.0, 5.5, 6.0]])
    s = configs.Structure(cell, pos, np.array([0, 1, 2], dtype=np.int32))
    configs.set_input_scaling(m, [s], O.Oracle(m).symm(s))
    # weak forces (output layers x 0.01): at full strength the random potential swings by eV per step
    # along the fast atom's path, and the three-body trajectory amplifies the FP32 networks' 1e-7 errors
    configs.scale_output(m, 0.01)
    masses = [27.0, 16.0, 40.0]
    v0 = np.zeros((3, 3))
    v0[0] = [750.0, 310.0, -120.0]          # A/ps: 0.75 A per 1 fs step
    dev = NNPDevice(m)
    dev.upload([s])
    dev.md_init(masses, v0)
    e = dev.md_run(60, dt=1e-3, rebuild_every=1, skin=1.0, precision=PREC_FP64_BASIS, graph=False)
    pg, vg = dev.md_state()
    dev.close()
    pr, vr, er = oracle_verlet(m, s, masses, v0, 60, 1e-3)
    f = np.linalg.solve(cell, pg.T).T
    dp, dv = periodic_dist(cell, pg, pr).max(), np.abs(vg - vr).max()
    de = np.abs(e - er).max() / 3
    log(f"md fast atom 60 steps: max|dx| {dp:.1e} A, max|dv| {dv:.1e} A/ps, max|dE|/atom {de:.1e} eV")
    assert np.all((f >= 0) & (f < 1)), "positions returned to the cell"
    assert dp < 1e-6 and dv < 1e-3 and de < 1e-6
