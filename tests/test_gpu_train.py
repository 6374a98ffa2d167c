"""Batched training step (energy + force-matching loss, d(F)/dW) vs the FP64 oracle and vs FP64 central
finite differences of the oracle's loss on 50 random weights.  Tolerance: loss relative 1e-5, gradient
1e-5 of the gradient's largest entry (measured ~1e-7; FP32 symmetry-function terms, networks and adjoints
with FP64 geometry, sums and reductions)."""
import numpy as np
import pytest

from ace_amd import NNPDevice, configs
from ace_amd.model import LossDesc
from oracle import oracle as O

from conftest import log

pytestmark = pytest.mark.gpu


def non_gaussian_angular(model):
    """Every angular block of the model on one of GAUSS2, SECH, STUDENT3/4/5 (basis_angular.cpp:211-249)."""
    from ace_amd.model import Angular
    for I in range(model.ntypes):
        for b, a in enumerate(model.angular[I]):
            model.angular[I][b] = Angular(2 + (I + b) % 5, a.cut, a.rc, a.eta, a.zeta, a.lam)


def batch(seed=21, sizes=(24, 40, 57, 33), angular_forms=False):
    rng = np.random.default_rng(seed)
    model = configs.c3_model(rng)
    teacher = configs.c3_model(np.random.default_rng(seed + 1))
    if angular_forms:
        non_gaussian_angular(model)
        non_gaussian_angular(teacher)
    strucs = []
    for n in sizes:
        cell, pos = configs.random_packed(n, 0.08, 1.6, rng)
        strucs.append(configs.Structure(cell, pos, rng.integers(0, 3, n).astype(np.int32)))
    orc = O.Oracle(model)
    G = [orc.symm(s) for s in strucs]
    configs.set_input_scaling(model, strucs, np.concatenate(G))
    for I in range(3):
        teacher.nets[I].inpw, teacher.nets[I].inpb = model.nets[I].inpw, model.nets[I].inpb
    tor = O.Oracle(teacher)
    for s in strucs:
        E, F, _ = tor.compute(s)
        s.energy, s.forces = E, F
    return model, strucs, G


def oracle_loss(model, strucs, G, ld):
    orc = O.Oracle(model)
    grad = np.zeros(sum(model.param_counts()))
    tot = np.zeros(3)
    for s, g in zip(strucs, G):
        tot += orc.train_struct(ld, s, g, s.energy, s.forces, grad)
    return tot, grad


@pytest.mark.parametrize("loss,norm,wf,prec", [(0, 0, 0.0, 0), (0, 1, 0.1, 0), (2, 2, 0.1, 0), (0, 1, 0.1, 2),
                                               (1, 3, 0.1, 0), (3, 4, 0.1, 0), (3, 3, 0.0, 2)])
def test_train_step_matches_oracle(loss, norm, wf, prec):
    model, strucs, G = batch()
    nat = sum(s.natoms for s in strucs)
    ld = LossDesc(loss, norm, 0.5, 1.0 / len(strucs), wf / (3 * nat))
    ref, gref = oracle_loss(model, strucs, G, ld)
    dev = NNPDevice(model)
    dev.upload(strucs)
    dev.build_neighbors()
    dev.set_precision(prec)
    dev.set_targets([s.energy for s in strucs], np.concatenate([s.forces for s in strucs]))
    out, g = dev.train_step(loss, norm, 0.5, 1.0 / len(strucs), wf / (3 * nat))
    le = abs(out[1] - ref[0]) / max(abs(ref[0]), 1e-12)
    lf = abs(out[2] - ref[1]) / max(abs(ref[1]), 1e-12) if wf else 0.0
    ge = np.max(np.abs(g - gref)) / np.max(np.abs(gref))
    log(f"train loss={loss} norm={norm} wf={wf} prec={prec}: loss_E rel {le:.1e} loss_F rel {lf:.1e} grad {ge:.1e}")
    assert le < 1e-5 and lf < 1e-5 and ge < 1e-5
    # a second step with unchanged parameters reproduces the first (cached G, deterministic reductions)
    out2, g2 = dev.train_step(loss, norm, 0.5, 1.0 / len(strucs), wf / (3 * nat))
    assert np.array_equal(out, out2) and np.array_equal(g, g2)
    dev.close()


@pytest.mark.parametrize("split", [True, False])
def test_train_step_level_split(split):
    """The training basis (zeta 1-2 with one eta, 4 with the other) runs its symmetry-function, adjoint and
    tangent kernels level-split by default; with the split off the full kernels run.  Both against the
    oracle's loss and gradient."""
    model, strucs, G = batch(seed=45)
    nat = sum(s.natoms for s in strucs)
    ld = LossDesc(0, 1, 0.5, 1.0 / len(strucs), 0.1 / (3 * nat))
    ref, gref = oracle_loss(model, strucs, G, ld)
    dev = NNPDevice(model)
    dev.set_level_split(split)
    assert dev.tile_shape(0)["level_split"] == (2 if split else 0)
    dev.upload(strucs)
    dev.build_neighbors()
    dev.set_targets([s.energy for s in strucs], np.concatenate([s.forces for s in strucs]))
    out, g = dev.train_step(0, 1, 0.5, 1.0 / len(strucs), 0.1 / (3 * nat))
    le = abs(out[1] - ref[0]) / max(abs(ref[0]), 1e-12)
    lf = abs(out[2] - ref[1]) / max(abs(ref[1]), 1e-12)
    ge = np.max(np.abs(g - gref)) / np.max(np.abs(gref))
    log(f"train level split={split}: loss_E rel {le:.1e} loss_F rel {lf:.1e} grad {ge:.1e}")
    assert le < 1e-5 and lf < 1e-5 and ge < 1e-5
    dev.close()


@pytest.mark.parametrize("prec", [0, 2])
def test_train_step_every_angular_form(prec):
    """The force-matching gradient through the generic angular path's per-triplet distance factors (tangent and
    adjoint modes) for GAUSS2, SECH and STUDENT3/4/5 blocks."""
    model, strucs, G = batch(seed=33, angular_forms=True)
    nat = sum(s.natoms for s in strucs)
    ld = LossDesc(0, 1, 0.5, 1.0 / len(strucs), 0.1 / (3 * nat))
    ref, gref = oracle_loss(model, strucs, G, ld)
    dev = NNPDevice(model)
    dev.upload(strucs)
    dev.build_neighbors()
    dev.set_precision(prec)
    dev.set_targets([s.energy for s in strucs], np.concatenate([s.forces for s in strucs]))
    out, g = dev.train_step(0, 1, 0.5, 1.0 / len(strucs), 0.1 / (3 * nat))
    le = abs(out[1] - ref[0]) / max(abs(ref[0]), 1e-12)
    lf = abs(out[2] - ref[1]) / max(abs(ref[1]), 1e-12)
    ge = np.max(np.abs(g - gref)) / np.max(np.abs(gref))
    log(f"train angular forms prec={prec}: loss_E rel {le:.1e} loss_F rel {lf:.1e} grad {ge:.1e}")
    assert le < 1e-5 and lf < 1e-5 and ge < 1e-5
    dev.close()


def test_train_gradient_vs_fp64_fd_50_weights():
    model, strucs, G = batch(seed=33, sizes=(20, 31))
    nat = sum(s.natoms for s in strucs)
    ld = LossDesc(0, 0, 1.0, 0.5, 0.1 / (3 * nat))
    dev = NNPDevice(model)
    dev.upload(strucs)
    dev.build_neighbors()
    dev.set_targets([s.energy for s in strucs], np.concatenate([s.forces for s in strucs]))
    _, g = dev.train_step(0, 0, 1.0, 0.5, 0.1 / (3 * nat))
    dev.close()
    x0 = model.all_params()
    rng = np.random.default_rng(4)
    h = 1e-6
    worst = 0.0
    scale = np.max(np.abs(g))
    for k in rng.choice(len(x0), 50, replace=False):
        vals = []
        for sg in (1, -1):
            x = x0.copy()
            x[k] += sg * h
            model.set_all_params(x)
            vals.append(oracle_loss(model, strucs, G, ld)[0][:2].sum())
        fd = (vals[0] - vals[1]) / (2 * h)
        worst = max(worst, abs(fd - g[k]) / scale)
    model.set_all_params(x0)
    log(f"train FD (50 weights) worst rel {worst:.1e}")
    assert worst < 1e-5


def test_train_graph_replay_is_bit_exact():
    """Four ADAM iterations with the device part of each step replayed from captured HIP graphs give the same
    losses and gradients, bit for bit, as plain launches (same kernels; sums in fixed order)."""
    from ace_amd.training import Trainer
    runs = []
    for graph in (False, True):
        model, strucs, G = batch(seed=41)
        dev = NNPDevice(model)
        dev.upload(strucs)
        dev.build_neighbors()
        tr = Trainer(dev, model, strucs, gamma=1e-3, w_force=0.1, graph=graph)
        hist = []
        for _ in range(4):
            out = tr.step()
            hist.append((out.copy(), tr.params.copy()))
        runs.append(hist)
        dev.close()
    for (o1, p1), (o2, p2) in zip(*runs):
        assert np.array_equal(o1, o2) and np.array_equal(p1, p2)


def test_device_adam_matches_host_adam(tmp_path):
    """ADAM on the GPU (FP64 master parameters, device layouts refreshed in place) follows the host ADAM
    trajectory: losses and parameters agree to 1e-14 over five steps (same operation order, no
    contraction).  The device layouts it writes equal the ones uploaded from the host: energies and forces
    of the trained networks (FP32 and BF16 first layer) match a device given the parameters by
    ace_set_params, bit for bit.  A restart written after three steps and continued for two gives the
    uninterrupted run."""
    from ace_amd import PREC_BF16L1, PREC_FP32
    from ace_amd.training import Trainer
    runs = []
    for device_opt in (False, True):
        model, strucs, G = batch(seed=43)
        dev = NNPDevice(model)
        dev.upload(strucs)
        dev.build_neighbors()
        tr = Trainer(dev, model, strucs, gamma=1e-3, w_force=0.1, graph=True, device_opt=device_opt)
        hist = []
        for k in range(5):
            out = tr.step()
            hist.append((out.copy(), tr.params.copy()))
            if device_opt and k == 2:
                tr.save_restart(str(tmp_path / "restart.npz"))
        runs.append(hist)
        if device_opt:
            res = {prec: dev.compute(prec) for prec in (PREC_FP32, PREC_BF16L1)}
            p_final = tr.params.copy()
        dev.close()
    worst = 0.0
    for (o1, p1), (o2, p2) in zip(*runs):
        worst = max(worst, np.max(np.abs(p1 - p2)) / np.max(np.abs(p1)))
        assert np.allclose(o1, o2, rtol=1e-13, atol=0)
    log(f"device ADAM vs host ADAM, 5 steps: max relative parameter difference {worst:.1e}")
    assert worst <= 1e-14

    model, strucs, G = batch(seed=43)
    ref = NNPDevice(model)
    ref.upload(strucs)
    ref.build_neighbors()
    ref.set_all_params(p_final)
    for prec in (PREC_FP32, PREC_BF16L1):
        E1, F1, _ = res[prec]
        E2, F2, _ = ref.compute(prec)
        assert np.array_equal(E1, E2) and np.array_equal(F1, F2), prec
    ref.close()

    model, strucs, G = batch(seed=43)
    dev = NNPDevice(model)
    dev.upload(strucs)
    dev.build_neighbors()
    tr = Trainer(dev, model, strucs, gamma=1e-3, w_force=0.1, graph=True, device_opt=True)
    tr.load_restart(str(tmp_path / "restart.npz"))
    for k in range(3, 5):
        out = tr.step()
        assert np.array_equal(out, runs[1][k][0]) and np.array_equal(tr.params, runs[1][k][1])
    dev.close()


@pytest.mark.parametrize("force_only", [False, True])
def test_train_step_full_config3_batch(golden_dir, force_only):
    """One ace_train_step on the whole 64-structure, 20,245-atom force-matching batch (structures of 32 to
    1013 atoms, small cells with several images, the cell search and the multi-round adjoint and tangent
    kernels) against the oracle's loss and gradient in tests/golden/c3_train.npz (written by
    tests/golden/make_c3_train.py: the oracle needs minutes for this batch).  The force-only loss checks the
    d(F)/dW part, which the energy term would otherwise dominate."""
    import os
    g = np.load(os.path.join(golden_dir, "c3_train.npz"))
    model, _, strucs = configs.config3()
    o = 0
    for I in range(model.ntypes):
        K = model.input_width(I)
        model.nets[I].inpw, model.nets[I].inpb = g["inpw"][o:o + K], g["inpb"][o:o + K]
        o += K
    we = 0.0 if force_only else float(g["w_energy"])
    ref, gref = (g["loss_force"], g["grad_force"]) if force_only else (g["loss"], g["grad"])
    dev = NNPDevice(model)
    dev.upload(strucs)
    dev.build_neighbors()
    dev.set_targets(g["E_ref"], g["F_ref"].astype(np.float64))
    out, grad = dev.train_step(0, 0, 1.0, we, float(g["w_force"]))
    dev.close()
    le = abs(out[1] - ref[0]) / max(abs(ref[0]), 1e-300) if not force_only else abs(out[1])
    lf = abs(out[2] - ref[1]) / abs(ref[1])
    ge = np.max(np.abs(grad - gref)) / np.max(np.abs(gref))
    log(f"c3 full batch force_only={force_only}: loss_E rel {le:.1e} loss_F rel {lf:.1e} grad rel {ge:.1e}")
    assert le < 1e-5 and lf < 1e-5 and ge < 1e-5


def test_train_step_after_an_overflowing_async_rebuild():
    """Device ADAM after update_positions + an asynchronous list rebuild that outgrows its capacity: the
    step verifies the list first and rebuilds it with exact sizes, so its loss, gradient and ADAM update are
    those of a context given the new positions from the start."""
    from ace_amd.training import Trainer
    res = []
    for fresh in (False, True):
        model, strucs, G = batch(seed=47, sizes=(24, 40))
        dense = []
        for s in strucs:
            c = s.pos.mean(axis=0)
            dense.append(configs.Structure(s.cell, c + 0.75 * (s.pos - c), s.type, s.energy, s.forces))
        dev = NNPDevice(model)
        if fresh:
            dev.upload(dense)
            dev.build_neighbors()
        else:
            dev.upload(strucs)
            dev.build_neighbors()
            dev.build_neighbors()          # capacities known: the next rebuild runs without a host check
            dev.update_positions(np.concatenate([s.pos for s in dense]))
            dev.build_neighbors()          # asynchronous, outgrows the rows on the device
        tr = Trainer(dev, model, dense, gamma=1e-3, w_force=0.1, graph=False, device_opt=True)
        out = tr.step()
        res.append((out.copy(), tr.params.copy()))
        dev.close()
    (o1, p1), (o2, p2) = res
    assert np.array_equal(o1, o2) and np.array_equal(p1, p2)
