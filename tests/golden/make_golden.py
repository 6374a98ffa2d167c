"""Writes the committed fixtures under tests/golden/.

nacl.json  -- the NaCl structures of the reference's periodic-invariance test: the 2-atom primitive
              cell and the 216-atom conventional supercell, lattice constants and fractional
              coordinates as listed in src/nnp/test_nnp_symm.cpp:25-290 (data only).  Needs
              /root/reference (this container); the JSON is committed.
c1_oracle.npz -- config-1 structure (108-atom jittered fcc) with the oracle's neighbour list, G,
              energy, forces and virial under the seeded model (regenerated from the oracle).
"""
import json
import os
import re
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)


def nacl_fixture(src="/root/reference/src/nnp/test_nnp_symm.cpp"):
    txt = open(src).read()
    small = txt[txt.index("Structure make_NaCl_small()"):txt.index("Structure make_NaCl_large()")]
    large = txt[txt.index("Structure make_NaCl_large()"):txt.index("void test_symm_super_crystal()")]
    a_small = float(re.search(r"const double a0=([0-9.eE+-]+);", small).group(1))
    a_large = float(re.search(r"const double a0=([0-9.eE+-]+);", large).group(1))
    lv = [float(x) for x in re.search(r"lv<<([^;]+);", small).group(1).split(",")]
    ps = {int(m.group(1)): [float(v) for v in m.group(2).split(",")]
          for m in re.finditer(r"struc\.posn\((\d+)\)<<([^;]+);", small)}
    pl = {int(m.group(1)): [float(v) for v in m.group(2).split(",")]
          for m in re.finditer(r"struc\.posn\((\d+)\)<<([^;]+);", large)}
    return {
        "small": {"a0": a_small, "lv_rowmajor": lv, "pos_over_a0": [ps[i] for i in sorted(ps)], "type": [0, 1]},
        "large": {"a0": a_large, "pos_over_a0": [pl[i] for i in sorted(pl)], "type": [0] * 108 + [1] * 108},
    }


def c1_fixture():
    from ace_amd import configs
    from oracle import oracle as O
    model, [s] = configs.config1()
    orc = O.Oracle(model)
    nl = orc.neighbors(s)
    G = orc.symm(s, nl)
    configs.set_input_scaling(model, [s], G)
    orc.refresh()
    pe, F, W = orc.force(s, G, nl)
    E = pe + 0.0
    return dict(cell=s.cell, pos=s.pos, type=s.type, off=nl.off, j=nl.j, img=nl.img, G=G, E=E, F=F, W=W,
                inpw=model.nets[0].inpw, inpb=model.nets[0].inpb, params=model.nets[0].params)


if __name__ == "__main__":
    if os.path.exists("/root/reference/src/nnp/test_nnp_symm.cpp"):
        with open(os.path.join(HERE, "nacl.json"), "w") as f:
            json.dump(nacl_fixture(), f)
    np.savez_compressed(os.path.join(HERE, "c1_oracle.npz"), **c1_fixture())
    print("fixtures written")
