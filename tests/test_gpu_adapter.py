"""The integration guide's C++ adapters, compiled (tests/adapter: nnp_device.hpp = NNPDevice, the drop-in for
NNP::compute callers, src/nnp/nnp.cpp:489-494; calc_nnp.hpp = CalcNNP, an MD Calculator, src/sim/calc.hpp:17-108)
against stand-in ace types and linked with libace_hip.so.  The program's energies and forces equal the Python
binding's through the same C-ABI bit for bit, and the oracle's within the FP32 tolerances; the calculator adds its
forces onto the engine's zeroed ones and follows a moved atom."""
import os
import subprocess

import numpy as np
import pytest

from ace_amd import NNPDevice, PREC_FP32, configs
from oracle import oracle as O

from conftest import log
from test_gpu_parity import TOL, scaled_model, structure

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))
EXE = os.path.join(HERE, "adapter", "adapter_test")


def dump_model(m, path):
    nt = m.ntypes
    out = [f"{nt} {m.rc!r}"]
    for I in range(nt):
        for J in range(nt):
            b = m.radial[I][J]
            out.append(" ".join([str(b.form), str(b.cut), repr(float(b.rc)), str(len(b.rs))] +
                                [repr(float(x)) for x in b.rs] + [repr(float(x)) for x in b.eta]))
        for b in m.angular[I]:
            out.append(" ".join([str(b.form), str(b.cut), repr(float(b.rc)), str(len(b.eta))] +
                                [repr(float(x)) for x in b.eta] + [repr(float(x)) for x in b.zeta] +
                                [str(int(x)) for x in b.lam]))
        n = m.nets[I]
        K = n.widths[0]
        inpw = np.ones(K) if n.inpw is None else n.inpw
        inpb = np.zeros(K) if n.inpb is None else n.inpb
        e0 = 0.0 if m.E0 is None else float(m.E0[I])
        out.append(" ".join([str(n.nlayer)] + [str(w) for w in n.widths] + [str(n.act), repr(float(n.c))] +
                            [repr(float(x)) for x in inpw] + [repr(float(x)) for x in inpb] +
                            [repr(float(n.outw)), repr(float(n.outb)), repr(e0), str(len(n.params))] +
                            [repr(float(x)) for x in n.params]))
    with open(path, "w") as f:
        f.write("\n".join(out) + "\n")


def dump_structure(s, path):
    lines = [str(s.natoms), " ".join(repr(float(s.cell[r, c])) for c in range(3) for r in range(3))]
    lines += [f"{int(s.type[i])} " + " ".join(repr(float(x)) for x in s.pos[i]) for i in range(s.natoms)]
    with open(path, "w") as f:
        f.write("\n".join(lines) + "\n")


def parse(path, n):
    E, F = {}, {}
    for line in open(path):
        t = line.split()
        if t[0].startswith("E"):
            E[t[0][1:]] = float(t[1])
        else:
            F.setdefault(t[0][1:], np.zeros((n, 3)))[int(t[1])] = [float(x) for x in t[2:5]]
    return E, F


def test_compiled_adapter_and_calculator(tmp_path):
    assert os.path.exists(EXE), "tests/adapter/adapter_test is built by __graft_entry__.build() (make -C tests/adapter)"
    rng = np.random.default_rng(51)
    m = configs.c3_model(rng)
    cell, pos = configs.random_packed(45, 0.08, 1.6, rng)
    s = structure(cell, pos, rng.integers(0, 3, 45))
    scaled_model(m, [s])
    m.E0 = np.array([-1.5, -2.25, -0.75])
    dump_model(m, tmp_path / "model.txt")
    dump_structure(s, tmp_path / "struc.txt")
    r = subprocess.run([EXE, str(tmp_path / "model.txt"), str(tmp_path / "struc.txt"), str(tmp_path / "out.txt")],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    E, F = parse(tmp_path / "out.txt", s.natoms)
    # the Python binding, same C-ABI and path: identical numbers
    dev = NNPDevice(m)
    dev.upload([s])
    dev.build_neighbors()
    Ep, Fp, _ = dev.compute(PREC_FP32)
    moved = s.pos.copy()
    moved[0, 0] += 0.01
    dev.update_positions(moved)
    dev.build_neighbors()
    Em, Fm, _ = dev.compute(PREC_FP32)
    dev.close()
    assert E["1"] == Ep[0] and np.array_equal(F["1"], Fp)
    assert E["2"] == Ep[0] and np.array_equal(F["2"], Fp)
    assert E["3"] == Em[0] and np.array_equal(F["3"], Fm)
    # and the oracle within the FP32 tolerances
    orc = O.Oracle(m)
    Er, Fr, _ = orc.compute(s)
    s2 = configs.Structure(s.cell, moved, s.type)
    Er2, Fr2, _ = orc.compute(s2)
    de = max(abs(E["1"] - Er), abs(E["3"] - Er2)) / s.natoms
    df = max(np.max(np.abs(F["1"] - Fr)), np.max(np.abs(F["3"] - Fr2)))
    log(f"compiled adapter: dE/atom {de:.1e} dF {df:.1e}")
    assert de < TOL[PREC_FP32]["E"] and df < TOL[PREC_FP32]["F"]
