"""ace_energy_forces_noalloc (static scratch, device pointers) against ace_energy_forces on the same batch.

The two entry points run the same kernels and sum forces, energies and virials in 64-bit fixed point, so
the results are compared bit for bit.  Two calls in a row share the device's scratch array: the first on a
1000-atom bcc structure, the second on a smaller batch (250 atoms), so the second must not see the first's
contents; then a network too large for the fused kernel (split path, network rows in scratch) with no
virial pointer."""
import ctypes as C

import numpy as np
import pytest

from ace_amd import PREC_FP32, PREC_FP64_BASIS, NNPDevice, configs
from oracle import oracle as O

from conftest import log

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")


def scaled(model, strucs):
    G = np.concatenate([O.Oracle(model).symm(s) for s in strucs])
    configs.set_input_scaling(model, strucs, G)


def noalloc(dev, prec, strucs, pos=True, virial=True):
    N = sum(s.natoms for s in strucs)
    cuda = torch.device("cuda:0")
    # row-major N x 3 (structure positions may be transposed views; the tensor must own C-order memory)
    p = torch.from_numpy(np.ascontiguousarray(np.concatenate([s.pos for s in strucs]))).to(cuda).contiguous() if pos else None
    assert p is None or (p.is_contiguous() and p.stride() == (3, 1))
    E = torch.full((len(strucs),), np.nan, dtype=torch.float64, device=cuda)
    F = torch.full((N, 3), np.nan, dtype=torch.float64, device=cuda)
    W = torch.full((len(strucs), 9), np.nan, dtype=torch.float64, device=cuda) if virial else None
    ptr = lambda t: C.c_void_p(t.data_ptr()) if t is not None else None
    rc = dev.lib.ace_energy_forces_noalloc(dev.ctx, prec, ptr(p), ptr(E), ptr(F), ptr(W))
    assert rc == 0, dev._err()
    dev.sync()
    return E.cpu().numpy(), F.cpu().numpy(), (W.cpu().numpy() if virial else None)


def check(name, model, strucs, precs=(PREC_FP32, PREC_FP64_BASIS), virial=True):
    dev = NNPDevice(model)
    dev.upload(strucs)
    dev.build_neighbors()
    for prec in precs:
        E0, F0, W0 = dev.compute(prec)
        for pos in (True, False):
            E, F, W = noalloc(dev, prec, strucs, pos=pos, virial=virial)
            log(f"{name} prec {prec} pos {pos}: |dE| {np.abs(E - E0).max():.1e} |dF| {np.abs(F - F0).max():.1e}")
            assert np.array_equal(E, E0) and np.array_equal(F, F0), (name, prec, pos)
            if virial:
                assert np.array_equal(W, W0.reshape(W.shape)), (name, prec, pos)
    dev.close()


def test_noalloc_matches_energy_forces_twice_then_split():
    m, strucs = configs.config2()
    scaled(m, strucs)
    check("bcc1000", m, strucs)
    # second call, smaller input: 250 atoms of the same lattice
    a = (2.0 / 0.085) ** (1.0 / 3.0)
    cell, pos = configs.bcc((5, 5, 5), a)
    rng = np.random.default_rng(3)
    pos = configs.wrap(cell, pos + rng.normal(0, 0.15, pos.shape))
    t = rng.integers(0, 2, len(pos)).astype(np.int32)
    check("bcc250", m, [configs.Structure(cell, pos, t)])
    # 128-128-64-1 networks: the split path, network rows in scratch; no virial pointer
    rng = np.random.default_rng(21)
    m5 = configs.c5_model(rng)
    s5 = []
    for n in (30, 111):
        c, p = configs.random_packed(n, 0.085, 1.5, rng)
        s5.append(configs.Structure(c, p, rng.integers(0, 4, n).astype(np.int32)))
    scaled(m5, s5)
    check("c5model", m5, s5, precs=(PREC_FP32,), virial=False)
