"""The C-ABI library loads and exports every entry point include/ace_hip.h declares (no GPU calls),
and the host-side model packing follows the reference layout.  CPU only."""
import ctypes as C
import os
import re

import numpy as np

from ace_amd import _lib, configs
from ace_amd.model import tri_index

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    txt = open(os.path.join(ROOT, "include", "ace_hip.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(ace_[a-z_0-9]+)\s*\(", txt)))


def test_library_exports_every_declared_symbol():
    assert os.path.exists(_lib.LIB_PATH), "libace_hip.so not built"
    lib = C.CDLL(_lib.LIB_PATH)
    names = declared_symbols()
    assert len(names) >= 20
    for n in names:
        assert hasattr(lib, n), n
    assert set(names) == set(_lib.SIGNATURES), "ctypes binding and header disagree"


def test_tensortri_order():
    # TensorTri<2> (tensortri.hpp:99-111): (0,0)=0, (0,1)=1, (1,1)=2, (0,2)=3, (1,2)=4, (2,2)=5
    assert [tri_index(j, k) for j, k in [(0, 0), (0, 1), (1, 1), (0, 2), (1, 2), (2, 2)]] == [0, 1, 2, 3, 4, 5]
    assert tri_index(2, 1) == tri_index(1, 2)


def test_param_layout_column_major():
    """Parameter vector = all biases, then weights column-major (nn.cpp:881-896)."""
    from ace_amd.model import Network
    n = Network([3, 2, 1], params=np.arange(2 + 1 + 6 + 2, dtype=float))
    bs, ws = n.unpack()
    assert np.array_equal(bs[0], [0, 1]) and np.array_equal(bs[1], [2])
    assert np.array_equal(ws[0], np.array([[3, 5, 7], [4, 6, 8]]))   # W0(o,i) = p[3 + o + 2 i]
    assert np.array_equal(ws[1], [[9, 10]])


def test_config_shapes():
    m, [s] = configs.config1()
    assert s.natoms == 108 and m.input_width(0) == 20 and m.nets[0].nparams() == 861
    m, [s] = configs.config2()
    assert s.natoms == 1000 and m.input_width(0) == 80 and m.nets[0].nparams() == 3681
    assert sorted(np.bincount(s.type)) == [500, 500]


def test_invalid_model_rejected_on_host():
    """ace_create validates the model before touching the device (errors mirror the reference's
    std::invalid_argument checks, e.g. cutoff.cpp:51-52, nn.cpp:904)."""
    import ctypes as C
    from ace_amd import NNPDevice
    m, _ = configs.config1()
    m.nets[0].widths = [21, 20, 20, 1]            # input width differs from the basis size
    m.nets[0].params = np.zeros(m.nets[0].nparams())
    import pytest
    with pytest.raises(Exception):
        NNPDevice(m)
    m, _ = configs.config1()
    m.radial[0][0].cut = 9                        # invalid cutoff name
    lib = _lib.load()
    ctx = C.c_void_p()
    d = m.desc()
    rc = lib.ace_create(C.byref(ctx), C.cast(C.pointer(d), C.c_void_p))
    msg = lib.ace_last_error(ctx).decode()
    lib.ace_destroy(ctx)
    assert rc == -4 and "cutoff" in msg
