"""Autograd-style boundary (ace_forward / ace_backward): the backward seeded with a loss's dL/dE and dL/dF
reproduces ace_train_step's gradient bit for bit; random upstream vectors give dL/dparams equal to FP64 central
differences of the oracle's sum_s dL/dE_s E_s + sum_a dL/dF_a . F_a over the weights, and dL/dpos of the energy
term equal to central differences over the positions.  Reference: NN::D2ODZDI::grad (ml/nn.cpp:2047-2140)
seeded as nnptefr.cpp:1353-1378 seeds it."""
import numpy as np
import pytest

from ace_amd import NNPDevice
from oracle import oracle as O

from conftest import log
from test_gpu_train import batch

pytestmark = pytest.mark.gpu


def mse_seeds(E, F, strucs, w_energy, w_force):
    """The seeds of the MSE loss with the LINEAR norm, in k_loss's operation order (nnptefr.cpp:609-632, 1353)."""
    gE = np.zeros(len(strucs))
    for s, st in enumerate(strucs):
        nf = 1.0 / st.natoms
        dU = (E[s] - st.energy) * nf
        gE[s] = w_energy * (nf * dU)
    Fref = np.concatenate([st.forces for st in strucs])
    return gE, w_force * (F - Fref)


@pytest.mark.parametrize("prec", [0, 2])
def test_backward_with_mse_seeds_equals_train_step(prec):
    model, strucs, _ = batch(seed=41)
    nat = sum(s.natoms for s in strucs)
    we, wf = 1.0 / len(strucs), 0.1 / (3 * nat)
    dev = NNPDevice(model)
    dev.upload(strucs)
    dev.build_neighbors()
    dev.set_precision(prec)
    E, F, _ = dev.forward()
    gE, gF = mse_seeds(E, F, strucs, we, wf)
    gp, _ = dev.backward(gE, gF)
    gp2, _ = dev.backward(gE, gF)            # several backward calls per forward
    dev.set_targets([s.energy for s in strucs], np.concatenate([s.forces for s in strucs]))
    _, g = dev.train_step(0, 1, 0.5, we, wf)
    log(f"backward vs train_step prec={prec}: max |diff| {np.max(np.abs(gp - g)):.1e}")
    assert np.array_equal(gp, g) and np.array_equal(gp, gp2)
    # a changed parameter invalidates the forward
    dev.set_all_params(model.all_params())
    with pytest.raises(RuntimeError, match="ace_backward"):
        dev.backward(gE, gF)
    dev.close()


def test_backward_random_upstream_vs_fd():
    model, strucs, _ = batch(seed=43, sizes=(22, 29))
    rng = np.random.default_rng(8)
    dE = rng.normal(size=len(strucs))
    nat = sum(s.natoms for s in strucs)
    dF = rng.normal(size=(nat, 3))
    dev = NNPDevice(model)
    dev.upload(strucs)
    dev.build_neighbors()
    dev.set_precision(2)
    dev.forward()
    gp, _ = dev.backward(dE, dF)
    gE_only, gx = dev.backward(dE, None, want_pos=True)
    dev.close()
    orc = O.Oracle(model)

    def objective(m, with_forces=True):
        o = O.Oracle(m)
        tot, a0 = 0.0, 0
        for s, st in enumerate(strucs):
            E, F, _ = o.compute(st)
            tot += dE[s] * E + (np.sum(dF[a0:a0 + st.natoms] * F) if with_forces else 0.0)
            a0 += st.natoms
        return tot

    x0 = model.all_params()
    h = 1e-6
    worst = worst_e = 0.0
    scale, scale_e = np.max(np.abs(gp)), np.max(np.abs(gE_only))
    for k in rng.choice(len(x0), 30, replace=False):
        v = []
        for sg in (1, -1):
            x = x0.copy()
            x[k] += sg * h
            model.set_all_params(x)
            v.append((objective(model), objective(model, False)))
        worst = max(worst, abs((v[0][0] - v[1][0]) / (2 * h) - gp[k]) / scale)
        worst_e = max(worst_e, abs((v[0][1] - v[1][1]) / (2 * h) - gE_only[k]) / scale_e)
    model.set_all_params(x0)
    # dL/dpos of the energy term: -dL/dE_s F_a against central differences of sum_s dL/dE_s E_s
    hx, worst_x = 1e-5, 0.0
    st, a = strucs[1], 5
    for d in range(3):
        v = []
        for sg in (1, -1):
            p = st.pos.copy()
            st.pos = p.copy()
            st.pos[a, d] += sg * hx
            v.append(dE[1] * orc.compute(st)[0])
            st.pos = p
        worst_x = max(worst_x, abs((v[0] - v[1]) / (2 * hx) - gx[strucs[0].natoms + a, d]))
    log(f"backward FD: dparams worst rel {worst:.1e}, energy-only {worst_e:.1e}, dpos abs {worst_x:.1e}")
    assert worst < 1e-5 and worst_e < 1e-5 and worst_x < 1e-5
