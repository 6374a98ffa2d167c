"""HIP path vs the FP64 oracle, through the C-ABI, in both precision modes.

Tolerances (written per mode):
  FP32 (symmetry-function terms, networks, adjoints in FP32; pair geometry, cutoffs, exponents, sums of the
  symmetry functions and the input scaling in FP64):
      neighbour lists bit-exact; G 2e-5 relative (floor 1); energy 1e-6 eV/atom; forces 1e-5 eV/A.
  FP64_BASIS (symmetry functions and adjoints in FP64, networks in FP32):
      G 1e-6 relative; energy 1e-6 eV/atom; forces 1e-5 eV/A.
The config-1 crystal (108 nearly identical fcc environments, specified in FP64) is ill-conditioned for
FP32 basis functions: its input preconditioning divides by a G spread of ~2% of G, so its FP32 force
bound is 1e-4 eV/A (energy 1e-6 eV/atom); the FP64_BASIS mode meets 1e-6 and 2e-5.  The plain FP64 port
must match to 1e-9.
  BF16L1 (FP32 basis; first network layer on BF16 matrix cores, hi/lo split operands, FP32 sums):
      energy 1e-4 eV/atom; forces 1e-3 eV/A."""
import os

import numpy as np
import pytest

from ace_amd import PREC_BF16L1, PREC_FP32, PREC_FP64_BASIS, NNPDevice, configs
from oracle import oracle as O

from conftest import log

pytestmark = pytest.mark.gpu

TOL = {PREC_FP32: dict(G=2e-5, E=1e-6, F=1e-5, W=1e-5),
       PREC_FP64_BASIS: dict(G=1e-6, E=1e-6, F=1e-5, W=1e-6)}
PRECS = [PREC_FP32, PREC_FP64_BASIS]
NAME = {PREC_FP32: "fp32", PREC_FP64_BASIS: "fp64b"}


def structure(cell, pos, typ):
    return configs.Structure(np.asarray(cell, float), np.asarray(pos, float), np.asarray(typ, dtype=np.int32))


def check_against_oracle(model, strucs, name, check_plain=True, paths=(None,), tol=None, setup=None):
    """Neighbour lists bit-exact, G, energies, forces and virials within TOL (or tol[(path, prec)]), for
    each energy/force path in `paths` (None: the size rule); setup(dev) runs before the upload."""
    orc = O.Oracle(model)
    dev = NNPDevice(model)
    if setup is not None:
        setup(dev)
    dev.upload(strucs)
    dev.build_neighbors()
    off, j, img = dev.neighbors()
    res = {}
    for path in paths:
        for prec in PRECS:
            dev.set_path(path, prec)
            dev.set_precision(prec)
            Gd = dev.compute_symm()
            E, F, W = dev.compute(prec)
            res[(path, prec)] = (Gd, E, F, W)
    if check_plain:
        Ep, Fp = dev.plain_compute()
    a0 = 0
    p0 = 0
    for k, s in enumerate(strucs):
        nl = orc.neighbors(s)
        n = s.natoms
        # neighbour list: same pairs, same order
        o = off[a0:a0 + n + 1] - off[a0]
        assert np.array_equal(o, nl.off), f"{name}[{k}] neighbour offsets"
        assert np.array_equal(j[p0:p0 + len(nl.j)], nl.j), f"{name}[{k}] neighbour indices"
        assert np.array_equal(img[p0:p0 + len(nl.j)], nl.img), f"{name}[{k}] image vectors"
        G = orc.symm(s, nl)
        kk = G.shape[1]
        pe, Fr, Wr = orc.force(s, G, nl)
        Er = pe + np.asarray(model.E0)[s.type].sum()
        for (path, prec), (Gd, E, F, W) in res.items():
            t = dict(TOL[prec], **((tol or {}).get((path, prec), {})))
            gerr = np.max(np.abs(Gd[a0:a0 + n, :kk] - G) / np.maximum(np.abs(G), 1.0))
            de = abs(E[k] - Er) / n
            df = np.max(np.abs(F[a0:a0 + n] - Fr))
            dw = np.max(np.abs(W[k] - Wr)) / max(1.0, np.max(np.abs(Wr)))
            tag = NAME[prec] + ("" if path is None else " " + path)
            log(f"{name}[{k}] {tag} n={n} pairs={len(nl.j)} G {gerr:.1e} dE/atom {de:.1e} dF {df:.1e} dW {dw:.1e}")
            assert gerr < t["G"] and de < t["E"] and df < t["F"] and dw < t["W"], (name, k, tag)
        if check_plain:
            assert abs(Ep[k] - Er) < 1e-9 * max(1, abs(Er)) and np.max(np.abs(Fp[a0:a0 + n] - Fr)) < 1e-9
        a0 += n
        p0 += len(nl.j)
    dev.close()


def scaled_model(model, strucs):
    orc = O.Oracle(model)
    G = np.concatenate([orc.symm(s) for s in strucs])
    configs.set_input_scaling(model, strucs, G)
    return model


@pytest.mark.parametrize("prec,tol_e,tol_f", [(PREC_FP32, 1e-6, 1e-4), (PREC_FP64_BASIS, 1e-6, 2e-5)])
def test_c1_golden(golden_dir, prec, tol_e, tol_f):
    g = np.load(os.path.join(golden_dir, "c1_oracle.npz"))
    model, [s] = configs.config1()
    model.nets[0].inpw, model.nets[0].inpb = g["inpw"], g["inpb"]
    dev = NNPDevice(model)
    dev.upload([s])
    dev.build_neighbors()
    off, j, img = dev.neighbors()
    assert np.array_equal(off, g["off"]) and np.array_equal(j, g["j"]) and np.array_equal(img, g["img"])
    dev.set_precision(prec)
    G = dev.compute_symm()
    assert np.max(np.abs(G - g["G"]) / np.maximum(np.abs(g["G"]), 1.0)) < TOL[prec]["G"]
    E, F, W = dev.compute(prec)
    de = abs(E[0] - float(g["E"])) / s.natoms
    df = np.max(np.abs(F - g["F"]))
    log(f"c1 {NAME[prec]}: dE/atom {de:.1e} dF {df:.1e}")
    assert de < tol_e and df < tol_f
    Ep, Fp = dev.plain_compute()
    assert abs(Ep[0] - float(g["E"])) < 1e-9 and np.max(np.abs(Fp - g["F"])) < 1e-9
    dev.close()


def test_small_and_triclinic_cells():
    """Cells shorter than 2 rc (several images of one neighbour), skewed cells, unwrapped atoms."""
    rng = np.random.default_rng(5)
    m = configs.c3_model(rng)
    strucs = []
    for L_, skew, n in [(7.4, 0.0, 32), (8.2, 0.2, 40), (12.0, 0.35, 60), (6.1, 0.1, 18)]:
        cell = np.array([[L_, skew * L_, 0.2 * skew * L_], [0, L_, 0.5 * skew * L_], [0, 0, L_]])
        f = rng.uniform(0, 1, (n, 3))
        f[:3] += np.array([[1, 0, 0], [-1, 2, 0], [0, 0, -3]])     # atoms outside the cell
        strucs.append(structure(cell, (cell @ f.T).T, rng.integers(0, 3, n)))
    scaled_model(m, strucs)
    check_against_oracle(m, strucs, "small")


def test_large_triclinic_cells_binning_free_search():
    """A batch of skewed cells whose heights all exceed twice the list radius, with atoms one or more cells
    outside their cell: the build tests every atom of each structure under the minimum image (no binning)
    and must still give the reference's pairs, images and order bit for bit."""
    rng = np.random.default_rng(31)
    m = configs.c3_model(rng)
    strucs = []
    for L_, skew, n in [(14.0, 0.25, 90), (15.5, 0.4, 120), (13.2, 0.1, 70)]:
        cell = np.array([[L_, skew * L_, 0.3 * skew * L_], [0, L_ * 1.05, 0.45 * skew * L_], [0, 0, L_ * 0.97]])
        heights = 1.0 / np.linalg.norm(np.linalg.inv(cell), axis=1)
        assert heights.min() > 2 * 6.0
        f = rng.uniform(0, 1, (n, 3))
        f[:6] += rng.integers(-2, 3, (6, 3))          # atoms one or two cells outside
        strucs.append(structure(cell, (cell @ f.T).T, rng.integers(0, 3, n)))
    scaled_model(m, strucs)
    check_against_oracle(m, strucs, "triclinic")


def test_neighbour_rows_of_every_length():
    """Binned search (a cell shorter than 2 rc in the batch) over densities whose rows hold 22 to 362
    neighbours: the 14 A cells (one image per neighbour) ranked from the j bitmap, the 7.4 A cell (several
    images of a neighbour) by the rank sort with one to four packed 32-bit keys per lane, or the full keys
    beyond 256; pairs, images and order bit for bit."""
    rng = np.random.default_rng(41)
    m = configs.c3_model(rng)
    L_ = 14.0
    strucs = [structure(np.eye(3) * 7.4, rng.uniform(0, 7.4, (30, 3)), rng.integers(0, 3, 30))]
    for n in (121, 303, 485, 697, 970):
        strucs.append(structure(np.eye(3) * L_, rng.uniform(0, L_, (n, 3)), rng.integers(0, 3, n)))
    orc = O.Oracle(m)
    dev = NNPDevice(m)
    dev.upload(strucs)
    dev.build_neighbors()
    off, j, img = dev.neighbors()
    dev.close()
    a0 = p0 = 0
    lengths = []
    for k, s in enumerate(strucs):
        nl = orc.neighbors(s)
        n = s.natoms
        assert np.array_equal(off[a0:a0 + n + 1] - off[a0], nl.off), f"rows[{k}] offsets"
        assert np.array_equal(j[p0:p0 + len(nl.j)], nl.j), f"rows[{k}] indices"
        assert np.array_equal(img[p0:p0 + len(nl.j)], nl.img), f"rows[{k}] images"
        lengths.append(np.diff(nl.off))
        a0 += n
        p0 += len(nl.j)
    lengths = np.concatenate(lengths)
    log(f"row lengths {lengths.min()}..{lengths.max()}")
    for lo, hi in ((1, 64), (65, 128), (129, 192), (193, 256), (257, 1024)):
        assert ((lengths >= lo) & (lengths <= hi)).any(), (lo, hi)


def test_isolated_and_single_atom():
    rng = np.random.default_rng(6)
    m = configs.c3_model(rng)
    strucs = [
        structure(np.eye(3) * 40.0, [[1.0, 1.0, 1.0], [20.0, 20.0, 20.0]], [0, 1]),      # no neighbours
        structure(np.eye(3) * 4.0, [[0.5, 0.5, 0.5]], [2]),                                # self images only
        structure(np.eye(3) * 9.0, rng.uniform(0, 9, (20, 3)), rng.integers(0, 3, 20)),
    ]
    for I in range(3):
        m.nets[I].inpw = np.full(m.input_width(I), 0.05)
    check_against_oracle(m, strucs, "edge")


def test_c2_full_size():
    m, strucs = configs.config2()
    scaled_model(m, strucs)
    # every path, both precisions, within 1e-6 eV/atom and 1e-5 eV/A
    check_against_oracle(m, strucs, "c2", paths=(None,) + NNPDevice.PATHS)


def test_every_radial_and_cutoff_form_and_generic_angular():
    """Each radial form (GAUSSIAN, SECH, LOGISTIC, TANH, LOGCOSH, LOGCOSH2) and cutoff (STEP, COS, TANH,
    POLY3) on one neighbour-species block each, and angular blocks outside the (lambda = +-1, zeta = 2^k)
    fast path: odd and fractional zeta, a repeated function, more than 16 functions in one block."""
    from ace_amd.model import Radial, CUT_STEP, CUT_COS, CUT_TANH, CUT_POLY3
    rng = np.random.default_rng(17)
    m = configs.c3_model(rng)
    cuts = [CUT_COS, CUT_TANH, CUT_POLY3, CUT_STEP]
    for I in range(3):
        for J in range(3):
            form = 1 + (3 * I + J) % 6
            eta = np.array([0.8, 1.3, 2.1, 0.5])
            rs = np.array([0.0, 1.5, 2.5, 3.5])
            m.radial[I][J] = Radial(form, cuts[(I + J) % 4], 6.0, rs, eta)
    generic = [(0.005, 1, 1), (0.005, 3, -1), (0.005, 1.5, 1), (0.02, 2, 1), (0.02, 2, 1), (0.02, 5, -1)]
    many = [(eb, z, l) for eb in (0.005, 0.02) for z in (1, 2, 3, 4, 6, 8, 16, 32, 64) for l in (1, -1)]
    for I in range(3):
        m.angular[I][0] = configs.angular_block(generic)
        m.angular[I][4] = configs.angular_block(many)
    for I in range(3):
        K = m.input_width(I)
        w = [K] + list(m.nets[I].widths[1:])
        m.nets[I] = configs.Network(w, m.nets[I].act, 1.0, configs.xavier_params(w, rng), np.ones(K), np.zeros(K))
    strucs = []
    for n in (30, 47):
        cell, pos = configs.random_packed(n, 0.08, 1.6, rng)
        strucs.append(structure(cell, pos, rng.integers(0, 3, n)))
    scaled_model(m, strucs)
    check_against_oracle(m, strucs, "forms", check_plain=False, paths=NNPDevice.PATHS)


def test_every_angular_form():
    """The five angular forms whose distance factor couples the two distances (GAUSS2, SECH, STUDENT3/4/5,
    basis_angular.cpp:211-249, 455-560): one form per block beside GAUSS blocks, through the generic angular
    path with its per-triplet FP64 factor, on every energy/force path and both precisions."""
    from ace_amd.model import Angular, CUT_COS
    rng = np.random.default_rng(29)
    m = configs.c3_model(rng)
    funcs = [(0.005, 1, 1), (0.005, 2, -1), (0.005, 4, 1), (0.02, 1, -1), (0.02, 3, 1), (0.02, 8, -1)]
    eta = np.array([1.0 / np.sqrt(f[0]) for f in funcs])
    zeta = np.array([float(f[1]) for f in funcs])
    lam = np.array([int(f[2]) for f in funcs], dtype=np.int32)
    for I in range(3):
        for b in range(6):
            form = 2 + (I + b) % 5 if b != 5 else 1
            m.angular[I][b] = Angular(form, CUT_COS, 6.0, eta, zeta, lam)
    for I in range(3):
        K = m.input_width(I)
        w = [K] + list(m.nets[I].widths[1:])
        m.nets[I] = configs.Network(w, m.nets[I].act, 1.0, configs.xavier_params(w, rng), np.ones(K), np.zeros(K))
    strucs = []
    for n in (33, 52):
        cell, pos = configs.random_packed(n, 0.08, 1.6, rng)
        strucs.append(structure(cell, pos, rng.integers(0, 3, n)))
    scaled_model(m, strucs)
    check_against_oracle(m, strucs, "angular forms", check_plain=False, paths=NNPDevice.PATHS)


@pytest.mark.parametrize("split", [True, False])
def test_level_split_angular_kernels(split):
    """Two eta groups over disjoint zeta levels (the headline basis: zeta 1-4 with one eta, 8-16 with the other,
    blocks of 10 and 9 functions; the training basis: zeta 1-2 and 4) take the level-split kernels, which
    form only those (level, group) products; with the split switched off the full kernels run.  Every
    path, both precisions, against the oracle; the FP64-basis mode never splits."""
    rng = np.random.default_rng(41)
    models = [("c5 basis", configs.c5_model(rng), 4, 3), ("c3 basis", configs.c3_model(rng), 3, 2)]
    for name, m, nt, spl in models:
        for I in range(nt):
            K = m.input_width(I)
            w = [K, 24, 16, 1]
            m.nets[I] = configs.Network(w, m.nets[I].act, 1.0, configs.xavier_params(w, rng), np.ones(K), np.zeros(K))
        strucs = []
        for n in (45, 70):
            cell, pos = configs.random_packed(n, 0.085, 1.6, rng)
            strucs.append(structure(cell, pos, rng.integers(0, nt, n)))
        scaled_model(m, strucs)
        seen = {}

        def setup(dev):
            dev.set_level_split(split)
            seen["fp32"] = dev.tile_shape(PREC_FP32)["level_split"]
            seen["fp64b"] = dev.tile_shape(PREC_FP64_BASIS)["level_split"]

        check_against_oracle(m, strucs, f"{name} split={split}", check_plain=False, paths=NNPDevice.PATHS, setup=setup)
        assert seen == {"fp32": spl if split else 0, "fp64b": 0}, (name, seen)


@pytest.mark.parametrize("hidden,act", [((128, 128), 3), ((40, 24, 17), 2), ((8,), 10), ((33, 31), 15)])
def test_network_shapes_and_activations(hidden, act):
    """Widths that are not multiples of the 16x16 matrix tiles, 1 to 3 hidden layers, each activation,
    and a network too large to stage its weights in LDS (128-128)."""
    rng = np.random.default_rng(23)
    m = configs.c3_model(rng)
    for I in range(3):
        K = m.input_width(I)
        w = [K] + list(hidden) + [1]
        m.nets[I] = configs.Network(w, act, 1.0, configs.xavier_params(w, rng), np.ones(K), np.zeros(K))
    strucs = []
    for n in (17, 40):
        cell, pos = configs.random_packed(n, 0.08, 1.6, rng)
        strucs.append(structure(cell, pos, rng.integers(0, 3, n)))
    scaled_model(m, strucs)
    check_against_oracle(m, strucs, f"net{hidden}", check_plain=False)


@pytest.mark.parametrize("prec,tol", [(PREC_FP32, 1e-5), (PREC_FP64_BASIS, 1e-5)])
def test_c2_forces_central_fd(prec, tol):
    """Adjoint forces vs central finite differences (h = 1e-4 A) of the FP64 energy."""
    m, [s] = configs.config2()
    scaled_model(m, [s])
    dev = NNPDevice(m)
    dev.upload([s])
    dev.build_neighbors()
    _, F, _ = dev.compute(prec)
    h = 1e-4
    rng = np.random.default_rng(1)
    worst = 0.0
    for a in rng.choice(s.natoms, 6, replace=False):
        for d in range(3):
            e = []
            for sg in (1, -1):
                p = s.pos.copy()
                p[a, d] += sg * h
                dev.update_positions(p)
                e.append(dev.plain_compute()[0][0])
            fd = -(e[0] - e[1]) / (2 * h)
            worst = max(worst, abs(fd - F[a, d]))
    log(f"c2 {NAME[prec]} FD force max err {worst:.2e}")
    assert worst < tol
    dev.close()


@pytest.mark.parametrize("prec", PRECS)
def test_batching_is_invariant_and_deterministic(prec):
    rng = np.random.default_rng(9)
    m = configs.c3_model(rng)
    strucs = []
    for n in (37, 120, 64):
        cell, pos = configs.random_packed(n, 0.08, 1.6, rng)
        strucs.append(structure(cell, pos, rng.integers(0, 3, n)))
    scaled_model(m, strucs)
    dev = NNPDevice(m)
    dev.upload(strucs)
    dev.build_neighbors()
    E, F, W = dev.compute(prec)
    E2, F2, W2 = dev.compute(prec)
    assert np.array_equal(E, E2) and np.array_equal(F, F2) and np.array_equal(W, W2)
    a0 = 0
    for k, s in enumerate(strucs):
        one = NNPDevice(m)
        one.upload([s])
        one.build_neighbors()
        e, f, w = one.compute(prec)
        assert e[0] == E[k] and np.array_equal(f, F[a0:a0 + s.natoms]) and np.array_equal(w[0], W[k])
        a0 += s.natoms
        one.close()
    dev.close()


@pytest.mark.parametrize("path", ["split", "split_2w"])
def test_force_gather_matches_atomics(path):
    """Split passes: forces gathered from per-pair gradients at the reverse entries (default) against the
    fixed-point atomics, on cells with several images of a neighbour, self images, skewed cells and atoms
    outside their cell, with and without a list skin (pairs past the cutoff in the rows; the image range
    stays the reference's, so a skin changes nothing): forces within 1e-9 eV/A, energies and virials bit
    for bit; the gathered forces are deterministic and batch-invariant."""
    rng = np.random.default_rng(17)
    m = configs.c3_model(rng)
    strucs = []
    for L_, skew, n in [(7.4, 0.0, 32), (8.2, 0.2, 40), (12.0, 0.35, 60), (6.1, 0.1, 18), (4.0, 0.0, 1)]:
        cell = np.array([[L_, skew * L_, 0.2 * skew * L_], [0, L_, 0.5 * skew * L_], [0, 0, L_]])
        f = rng.uniform(0, 1, (n, 3))
        if n > 3:
            f[:3] += np.array([[1, 0, 0], [-1, 2, 0], [0, 0, -3]])
        strucs.append(structure(cell, (cell @ f.T).T, rng.integers(0, 3, n)))
    scaled_model(m, strucs)
    out = {}
    for gather in (True, False):
        for skin in (0.0, 1.0):
            dev = NNPDevice(m)
            dev.set_path(path, PREC_FP32)
            dev.set_force_gather(gather)
            dev.upload(strucs)
            dev.build_neighbors(skin=skin)
            out[(gather, skin)] = dev.compute(PREC_FP32)
            if gather:
                again = dev.compute(PREC_FP32)
                assert all(np.array_equal(a, b) for a, b in zip(out[(gather, skin)], again))
            dev.close()
    E0, F0, W0 = out[(False, 0.0)]
    for key, (E, F, W) in out.items():
        log(f"force gather {path} {key}: dF {np.max(np.abs(F - F0)):.1e} dE {np.max(np.abs(E - E0)):.1e}")
    for key, (E, F, W) in out.items():
        assert np.max(np.abs(F - F0)) < 1e-9, key
        assert np.array_equal(E, E0) and np.array_equal(W, W0), key
        assert np.array_equal(F, out[(key[0], 0.0)][1]), key          # the skin changes no bit
    E, F, W = out[(True, 0.0)]
    a0 = 0
    for k, s in enumerate(strucs):
        one = NNPDevice(m)
        one.set_path(path, PREC_FP32)
        one.upload([s])
        one.build_neighbors()
        e, f, w = one.compute(PREC_FP32)
        assert e[0] == E[k] and np.array_equal(f, F[a0:a0 + s.natoms]) and np.array_equal(w[0], W[k])
        a0 += s.natoms
        one.close()


def test_positions_update_matches_rebuild():
    m, [s] = configs.config1()
    scaled_model(m, [s])
    dev = NNPDevice(m)
    dev.upload([s])
    dev.build_neighbors(skin=1.0)
    p = s.pos + np.random.default_rng(2).normal(0, 0.05, s.pos.shape)
    dev.update_positions(p)
    E, F, _ = dev.compute()
    ref = NNPDevice(m)
    ref.upload([structure(s.cell, p, s.type)])
    ref.build_neighbors()
    Er, Fr, _ = ref.compute()
    assert abs(E[0] - Er[0]) / s.natoms < 1e-9 and np.max(np.abs(F - Fr)) < 1e-6
    dev.close()
    ref.close()


def test_async_rebuild_and_overflow_recovery():
    """Rebuilds after the first reuse capacities without a host round trip; a list that outgrows them
    is detected and rebuilt when results are read."""
    m, [s] = configs.config1()
    scaled_model(m, [s])
    dev = NNPDevice(m)
    dev.upload([s])
    dev.build_neighbors()
    E0, F0, _ = dev.compute()
    for _ in range(3):
        dev.build_neighbors()                  # asynchronous: same layout
    E1, F1, _ = dev.compute()
    assert E1[0] == E0[0] and np.array_equal(F1, F0)
    # pull the atoms together (same cell): far more neighbours than the capacities held
    c = s.pos.mean(axis=0)
    dense = structure(s.cell, c + 0.7 * (s.pos - c), s.type)
    dev.update_positions(dense.pos)
    dev.build_neighbors()                      # asynchronous build overflows on the device
    E2, F2, _ = dev.compute()                  # detected here, rebuilt with exact sizes, recomputed
    ref = NNPDevice(m)
    ref.upload([dense])
    ref.build_neighbors()
    Er, Fr, _ = ref.compute()
    dev.close()
    ref.close()
    assert E2[0] == Er[0] and np.array_equal(F2, Fr)


def test_compute_entry_matches_build_then_energy():
    """ace_compute (NNP::compute with the list built inside) equals ace_build_neighbors + ace_energy_forces,
    bit for bit (fixed-point force sums, fixed-order reductions)."""
    m, [s] = configs.config2()
    scaled_model(m, [s])
    dev = NNPDevice(m)
    dev.upload([s])
    dev.build_neighbors()
    E0, F0, W0 = dev.compute()
    for _ in range(3):
        dev.compute_step()                     # asynchronous steps, results in HBM
    import ctypes as C
    E = np.zeros(1)
    F = np.zeros((s.natoms, 3))
    W = np.zeros((1, 9))
    dp = lambda a: a.ctypes.data_as(C.POINTER(C.c_double))
    assert dev.lib.ace_compute(dev.ctx, 0, 0.0, dp(E), dp(F), dp(W)) == 0
    dev.close()
    assert E[0] == E0[0] and np.array_equal(F, F0) and np.array_equal(W.reshape(W0.shape), W0)


BF16_TOL = dict(E=1e-4, F=1e-3)


def check_bf16(model, strucs, name, net_waves=0):
    """BF16L1 energies and forces vs the FP64 oracle, next to the FP32 mode on the same batch."""
    orc = O.Oracle(model)
    dev = NNPDevice(model)
    if net_waves:
        for prec in (PREC_BF16L1, PREC_FP32):
            dev.set_path("split", prec)
            dev.set_tile_shape(0, net_waves, prec)
    dev.upload(strucs)
    dev.build_neighbors()
    Eb, Fb, Wb = dev.compute(PREC_BF16L1)
    Ef, Ff, _ = dev.compute(PREC_FP32)
    a0 = 0
    for k, s in enumerate(strucs):
        n = s.natoms
        Er, Fr, Wr = orc.compute(s)
        de, df = abs(Eb[k] - Er) / n, np.max(np.abs(Fb[a0:a0 + n] - Fr))
        de32, df32 = abs(Ef[k] - Er) / n, np.max(np.abs(Ff[a0:a0 + n] - Fr))
        dw = np.max(np.abs(Wb[k] - Wr)) / max(1.0, np.max(np.abs(Wr)))
        log(f"{name}[{k}] bf16l1 n={n} dE/atom {de:.1e} dF {df:.1e} dW {dw:.1e} (fp32: {de32:.1e} {df32:.1e})")
        assert de < BF16_TOL["E"] and df < BF16_TOL["F"] and dw < 1e-3, (name, k)
        a0 += n
    dev.close()


def test_bf16_first_layer():
    """128-128-64-1 networks on 4 species (the mixed-precision inference model), random packings of
    30 to 300 atoms, and the 80-32-32-1 model, which otherwise takes the fused FP32 kernel."""
    rng = np.random.default_rng(21)
    m = configs.c5_model(rng)
    strucs = []
    for n in (30, 111, 300):
        cell, pos = configs.random_packed(n, 0.085, 1.5, rng)
        strucs.append(structure(cell, pos, rng.integers(0, 4, n)))
    scaled_model(m, strucs)
    check_bf16(m, strucs, "bf16-c5model")
    m2, strucs2 = configs.config2()
    scaled_model(m2, strucs2)
    check_bf16(m2, strucs2, "bf16-c2")


@pytest.mark.parametrize("net_waves,chunk_w", [(1, 1), (2, 8), (4, 3)])
def test_network_and_chunk_tile_shapes(net_waves, chunk_w):
    """Each count of waves sharing a 16-atom network tile (1, 2, 4) and each angular chunk weight (1, 3, 8),
    pinned, on the split paths (weights staged in LDS or read from memory, two or four waves per atom), with
    layer widths that are not multiples of 16 next to a 128-wide layer; then the BF16 first layer."""
    rng = np.random.default_rng(53)
    m = configs.c3_model(rng)
    for I in range(3):
        K = m.input_width(I)
        w = [K, 128, 40, 17, 1]
        m.nets[I] = configs.Network(w, 3, 1.0, configs.xavier_params(w, rng), np.ones(K), np.zeros(K))
    strucs = []
    for n in (21, 52, 37):
        cell, pos = configs.random_packed(n, 0.08, 1.6, rng)
        strucs.append(structure(cell, pos, rng.integers(0, 3, n)))
    scaled_model(m, strucs)

    def setup(dev):
        for prec in PRECS:
            dev.set_tile_shape(chunk_w, net_waves, prec)
            t = dev.tile_shape(prec)
            assert t["chunk_w"] == chunk_w, t

    check_against_oracle(m, strucs, f"tile w={chunk_w} net_waves={net_waves}", check_plain=False,
                         paths=("split", "split_staged", "split_2w", "split_staged_2w"), setup=setup)
    m5 = configs.c5_model(rng)
    strucs5 = []
    for n in (40, 95):
        cell, pos = configs.random_packed(n, 0.085, 1.5, rng)
        strucs5.append(structure(cell, pos, rng.integers(0, 4, n)))
    scaled_model(m5, strucs5)
    check_bf16(m5, strucs5, f"bf16 net_waves={net_waves}", net_waves=net_waves)


def test_c5_full_size(golden_dir):
    """The 256-structure, 568k-atom mixed-precision inference batch: FP32 and BF16L1 energies and forces of
    the smallest, the median and the largest (8,126 atoms) structure vs the FP64 oracle's values in
    tests/golden/c5_ref.npz (written by tests/golden/make_c5_ref.py, with the input scaling of that file),
    and the two smallest structures uploaded alone give the same bits as inside the batch on the same path
    (every per-atom result is independent of the batch; sums are exact).  The path is the one the size rules
    take for the full batch: split, two waves per atom."""
    g = np.load(os.path.join(golden_dir, "c5_ref.npz"))
    model, strucs = configs.config5()
    o = 0
    for I in range(model.ntypes):
        K = model.input_width(I)
        model.nets[I].inpw, model.nets[I].inpb = g["inpw"][o:o + K], g["inpb"][o:o + K]
        o += K
    N = sum(s.natoms for s in strucs)
    path = "split_2w"
    dev = NNPDevice(model)
    dev.upload(strucs)
    dev.build_neighbors()
    for p in (PREC_FP32, PREC_BF16L1):
        dev.set_path(path, p)
    res = {p: dev.compute(p) for p in (PREC_FP32, PREC_BF16L1)}
    dev.close()
    off = np.concatenate([[0], np.cumsum([s.natoms for s in strucs])])
    for q, k in enumerate(g["idx"]):
        n = strucs[k].natoms
        Er, Fr = float(g["E"][q]), g[f"F_{q}"]
        for p, te, tf in ((PREC_FP32, TOL[PREC_FP32]["E"], TOL[PREC_FP32]["F"]),
                          (PREC_BF16L1, BF16_TOL["E"], BF16_TOL["F"])):
            E, F, _ = res[p]
            de = abs(E[k] - Er) / n
            df = np.max(np.abs(F[off[k]:off[k] + n] - Fr))
            log(f"c5 ({N} atoms) struct {k} n={n} prec {p}: dE/atom {de:.1e} dF {df:.1e}")
            assert de < te and df < tf, (k, p)
    small = [int(k) for k in np.argsort([s.natoms for s in strucs])[:2]]
    alone = NNPDevice(model)
    alone.upload([strucs[k] for k in small])
    alone.build_neighbors()
    for p in (PREC_FP32, PREC_BF16L1):
        alone.set_path(path, p)
        e1, f1, w1 = alone.compute(p)
        E, F, W = res[p]
        a0 = 0
        for q, k in enumerate(small):
            n = strucs[k].natoms
            assert e1[q] == E[k] and np.array_equal(f1[a0:a0 + n], F[off[k]:off[k] + n]) and np.array_equal(w1[q], W[k])
            a0 += n
    alone.close()


def test_c4_fp32_forces_vs_plain_fp64():
    """The 108,000-atom two-species fcc cell of the MD workload (list radius 6 A, output layers scaled to an
    rms force of 0.1 eV/A as for the MD runs, configs.calibrate_forces): FP32 energies and forces against
    the unbatched FP64 port (which matches the oracle to 1e-9 on the smaller cells above) -- the oracle
    itself would need ~10^11 distance tests for this cell's reference neighbour list."""
    model, s, _ = configs.config4()
    dev = NNPDevice(model)
    dev.upload([s])
    dev.build_neighbors()
    dev.set_precision(PREC_FP64_BASIS)
    configs.set_input_scaling(model, [s], dev.compute_symm())
    dev.sync_model_scaling()
    _, F, _ = dev.compute(PREC_FP64_BASIS)
    configs.calibrate_forces(model, F, rms=0.1)
    for I in range(model.ntypes):
        dev.set_params(I, model.nets[I].params)
    E32, F32, _ = dev.compute(PREC_FP32)
    Ep, Fp = dev.plain_compute()
    dev.close()
    de = abs(E32[0] - Ep[0]) / s.natoms
    df = np.abs(F32 - Fp).max()
    log(f"c4 ({s.natoms} atoms) fp32 vs plain fp64: dE/atom {de:.1e} dF {df:.1e} (|F| rms {np.sqrt(np.mean(Fp ** 2)):.2f})")
    assert de < TOL[PREC_FP32]["E"] and df < TOL[PREC_FP32]["F"]


def test_potential_file_on_gpu(tmp_path):
    """A potential written in ace's text format and read back runs on the GPU and matches the oracle of
    the read-back model (whose basis carries the format's 6-decimal rounding)."""
    from ace_amd import potential
    model, strucs = configs.config2()
    scaled_model(model, strucs)
    model.names = ["Fe", "Cr"]
    path = tmp_path / "fecr.nn"
    potential.write_potential(path, model)
    dev = NNPDevice.from_file(path)
    m2 = dev.model
    assert m2.names == ["Fe", "Cr"]
    dev.upload(strucs)
    dev.build_neighbors()
    E, F, _ = dev.compute(PREC_FP32)
    Er, Fr, _ = O.Oracle(m2).compute(strucs[0])
    de, df = abs(E[0] - Er) / strucs[0].natoms, np.abs(F - Fr).max()
    log(f"potential file: dE/atom {de:.1e} dF {df:.1e}")
    assert de < TOL[PREC_FP32]["E"] and df < TOL[PREC_FP32]["F"]
    dev.close()


def test_autotune_keeps_results_within_tolerance():
    """Every autotuner candidate computes the same quantities: after tuning, energies and forces still
    match the oracle to the FP32 bounds, and each candidate that was timed is reported."""
    m, strucs = configs.config2()
    scaled_model(m, strucs)
    dev = NNPDevice(m)
    dev.upload(strucs)
    dev.build_neighbors()
    r = dev.autotune(PREC_FP32)
    log(f"autotune c2: {r}")
    assert r["choice"] in r["ms"] and all(v > 0 for v in r["ms"].values()) and len(r["ms"]) >= 2
    # the tile shape of the angular triplet loops was measured on the chosen path and the fastest kept
    t = r["tile"]
    assert t["path"] == r["choice"] and sorted(t["chunk_ms"]) == [1, 3, 8]
    assert t["chunk_w"] == min(t["chunk_ms"], key=t["chunk_ms"].get)
    E, F, _ = dev.compute(PREC_FP32)
    Er, Fr, _ = O.Oracle(m).compute(strucs[0])
    assert abs(E[0] - Er) / strucs[0].natoms < TOL[PREC_FP32]["E"] and np.abs(F - Fr).max() < TOL[PREC_FP32]["F"]
    dev.close()
