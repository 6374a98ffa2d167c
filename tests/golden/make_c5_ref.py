"""Writes tests/golden/c5_ref.npz: FP64 oracle energies and forces of three structures of the 256-structure
mixed-precision inference batch (configs.config5: 568,089 atoms, 4 species, 128-128-64-1 networks) -- the
smallest, the median and the largest (8,126 atoms) -- so the GPU test can check its errors across the
size range without running the oracle there (the largest takes about a minute and a half on one core).

Input scaling (PreBias MEAN / PreScale DEV) comes from the oracle's G of the 16 smallest structures; the
GPU test uses the same values.

Contents: inpw, inpb (concatenated over species), idx (structure indices), E (oracle energies, with E0),
F_<k> (oracle forces of structure idx[k]).

Run from the repository root: python tests/golden/make_c5_ref.py
"""
import os
import sys
from multiprocessing import Pool

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from ace_amd import configs  # noqa: E402
from oracle import oracle as O  # noqa: E402


def checked_structures(strucs):
    """Indices of the smallest, median and largest structure of the batch."""
    order = np.argsort([s.natoms for s in strucs], kind="stable")
    return [int(order[0]), int(order[len(order) // 2]), int(order[-1])]


def scaling_structures(strucs):
    return [int(k) for k in np.argsort([s.natoms for s in strucs], kind="stable")[:16]]


def _symm(k):
    model, strucs = configs.config5()
    return O.Oracle(model).symm(strucs[k])


def _compute(args):
    k, inpw, inpb = args
    model, strucs = configs.config5()
    apply_scaling(model, inpw, inpb)
    E, F, _ = O.Oracle(model).compute(strucs[k])
    return E, F


def apply_scaling(model, inpw, inpb):
    o = 0
    for I in range(model.ntypes):
        K = model.input_width(I)
        model.nets[I].inpw, model.nets[I].inpb = inpw[o:o + K], inpb[o:o + K]
        o += K


def main():
    model, strucs = configs.config5()
    sc = scaling_structures(strucs)
    idx = checked_structures(strucs)
    with Pool(min(8, os.cpu_count() or 1)) as pool:
        Gs = pool.map(_symm, sc)
        configs.set_input_scaling(model, [strucs[k] for k in sc], np.concatenate(Gs))
        inpw = np.concatenate([model.nets[I].inpw for I in range(model.ntypes)])
        inpb = np.concatenate([model.nets[I].inpb for I in range(model.ntypes)])
        res = pool.map(_compute, [(k, inpw, inpb) for k in idx])
    out = dict(inpw=inpw, inpb=inpb, idx=np.array(idx), E=np.array([e for e, _ in res]))
    for q, (_, F) in enumerate(res):
        out[f"F_{q}"] = F
    np.savez(os.path.join(HERE, "c5_ref.npz"), **out)
    print("c5_ref.npz: structures", idx, "atoms", [strucs[k].natoms for k in idx])


if __name__ == "__main__":
    main()
