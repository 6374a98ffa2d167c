"""Pins the FP64 CPU oracle: closed forms, invariants, finite differences and the reference's own
fixture (the NaCl periodic-invariance structures of src/nnp/test_nnp_symm.cpp).  CPU only."""
import json
import os

import numpy as np
import pytest

from ace_amd import configs
from ace_amd.model import (ACT_LINEAR, ACT_SIGMOID, ACT_SOFTPLUS, ACT_SWISH, ACT_TANH, ANG_GAUSS, CUT_COS,
                           CUT_POLY3, CUT_STEP, CUT_TANH, RAD_GAUSSIAN, Angular, LossDesc, Model, Network,
                           Radial, tri_index, xavier_params)
from oracle import oracle as O

L = O.lib()


# ---- scalar math (special.cpp) -------------------------------------------------------------------

def test_fmexp_matches_exp():
    x = np.linspace(-700.0, 700.0, 20001)
    r = np.array([L.or_fmexp(v) for v in x]) / np.exp(x)
    assert np.max(np.abs(r - 1)) < 1e-13
    assert L.or_fmexp(-709.0) == 0.0          # below -1022/log2(e): flushed to 0 (special.cpp:271)
    assert L.or_fmexp(0.0) == 1.0


def test_coscut_sincut_taylor():
    x = np.linspace(0.0, np.pi, 2001)
    assert max(abs(L.or_coscut(v) - np.cos(v)) for v in x) < 1e-11
    assert max(abs(L.or_sincut(v) - np.sin(v)) for v in x) < 1e-11


def test_powint():
    for x in (0.3, 1.7, -2.0):
        for n in range(-4, 9):
            assert np.isclose(L.or_powint(x, n), x ** n, rtol=1e-15, atol=0)
    assert L.or_powint(0.0, 3) == 0.0 and L.or_powint(5.0, 0) == 1.0


# ---- cutoff (cutoff.cpp, test_nnp_cutoff.cpp:53-58) ----------------------------------------------

@pytest.mark.parametrize("name,integral", [(CUT_COS, 0.5), (CUT_POLY3, 0.5), (CUT_STEP, 1.0)])
def test_cutoff_integral_closed_form(name, integral):
    rc = 6.0
    r = np.linspace(0, rc, 200001)
    f = np.array([L.or_cutf(name, rc, v) for v in r])
    assert abs(np.trapezoid(f, r) - integral * rc) < 1e-4
    assert L.or_cutf(name, rc, 0.0) == pytest.approx(1.0, abs=1e-11)   # Taylor truncation of coscut
    assert L.or_cutf(name, rc, rc) == 0.0 and L.or_cutf(name, rc, rc + 1) == 0.0


@pytest.mark.parametrize("name", [CUT_COS, CUT_TANH, CUT_POLY3])
def test_cutoff_derivative_fd(name):
    rc, h = 6.0, 1e-6
    for r in np.linspace(0.1, 5.9, 37):
        fd = (L.or_cutf(name, rc, r + h) - L.or_cutf(name, rc, r - h)) / (2 * h)
        assert abs(fd - L.or_cutg(name, rc, r)) < 1e-8
    assert abs(L.or_cutg(CUT_COS, rc, 0.0)) < 1e-12 and abs(L.or_cutg(CUT_POLY3, rc, 0.0)) < 1e-15


# ---- basis functions vs finite differences (test_basis_radial.cpp, test_basis_angular.cpp) -------

@pytest.mark.parametrize("form", [1, 2, 3, 4, 5, 6])
def test_radial_derivative_fd(form):
    rng = np.random.default_rng(form)
    rc, h = 6.0, 1e-6
    for _ in range(40):
        eta, rs, r = rng.uniform(0.3, 2.0), rng.uniform(0, 3), rng.uniform(0.3, 5.9)
        fd = (L.or_radial_f(form, CUT_COS, rc, eta, rs, r + h) - L.or_radial_f(form, CUT_COS, rc, eta, rs, r - h)) / (2 * h)
        d = L.or_radial_d(form, CUT_COS, rc, eta, rs, r)
        assert abs(fd - d) < 1e-7 * max(1.0, abs(d))


def test_radial_gaussian_closed_form():
    rc, eta, rs, r = 6.0, 0.7, 1.2, 3.3
    ref = np.exp(-(eta * (r - rs)) ** 2) * 0.5 * (np.cos(np.pi * r / rc) + 1)
    assert L.or_radial_f(RAD_GAUSSIAN, CUT_COS, rc, eta, rs, r) == pytest.approx(ref, rel=1e-12)


@pytest.mark.parametrize("form", [1, 2, 3, 4, 5, 6])
def test_angular_derivative_fd(form):
    rng = np.random.default_rng(10 + form)
    rc, h = 6.0, 1e-6
    fphi, feta = (np.zeros(1), np.zeros(2))
    for _ in range(40):
        eta, zeta, lam = rng.uniform(1.0, 8.0), float(rng.integers(1, 9)), int(rng.choice([-1, 1]))
        c, rij, rik = rng.uniform(-0.95, 0.95), rng.uniform(0.5, 5.9), rng.uniform(0.5, 5.9)
        f = lambda c_, a_, b_: L.or_angular_f(form, CUT_COS, rc, eta, zeta, lam, c_, a_, b_)
        L.or_angular_d(form, CUT_COS, rc, eta, zeta, lam, c, rij, rik, O._d(fphi), O._d(feta))
        scale = max(1.0, abs(f(c, rij, rik)))
        assert abs((f(c + h, rij, rik) - f(c - h, rij, rik)) / (2 * h) - fphi[0]) < 1e-7 * scale
        assert abs((f(c, rij + h, rik) - f(c, rij - h, rik)) / (2 * h) - feta[0]) < 1e-7 * scale
        assert abs((f(c, rij, rik + h) - f(c, rij, rik - h)) / (2 * h) - feta[1]) < 1e-7 * scale


# ---- neighbour list (neighbor.cpp:46-87) -----------------------------------------------------------

def brute_neighbors(cell, pos, rc, ext=6):
    """Independent set: every (i, j, n) with |n_d| <= ext and 0 < |d| < rc, in reference order."""
    R = np.asarray(cell)
    n = len(pos)
    cf = [int(np.ceil(3 * rc / np.linalg.norm(R[:, d]))) for d in range(3)]
    rng_ = [range(-min(ext, cf[d]), min(ext, cf[d]) + 1) for d in range(3)]
    imgs = np.array([[a, b, c] for a in rng_[0] for b in rng_[1] for c in rng_[2]])
    out = []
    for i in range(n):
        d = pos[i][None, None, :] - pos[None, :, :] - (imgs @ R.T)[:, None, :]   # images x n x 3
        ds2 = np.einsum("abk,abk->ab", d, d)
        for a, j in zip(*np.nonzero(ds2 < rc * rc)):
            if j == i and ds2[a, j] <= 1e-12:
                continue
            out.append((i, j, *imgs[a]))
    return sorted(out)


def as_set(nl):
    out = []
    for i in range(len(nl.off) - 1):
        for p in range(nl.off[i], nl.off[i + 1]):
            out.append((i, int(nl.j[p]), *map(int, nl.img[p])))
    return out


def test_neighbors_fcc_shells():
    cell, pos = configs.fcc((3, 3, 3), 4.05)
    nl = O.neighbors(cell, pos, 6.0)
    assert set(np.diff(nl.off)) == {54}      # 12 + 6 + 24 + 12 within 6 A


@pytest.mark.parametrize("L_,skew", [(7.4, 0.0), (9.0, 0.15), (13.0, 0.3)])
def test_neighbors_small_cells_brute_force(L_, skew):
    rng = np.random.default_rng(int(L_ * 10))
    cell = np.array([[L_, skew * L_, 0.1 * skew * L_], [0, L_, skew * L_ * 0.5], [0, 0, L_]])
    pos = (cell @ rng.uniform(0, 1, (12, 3)).T).T
    nl = O.neighbors(cell, pos, 6.0)
    got = as_set(nl)
    assert got == brute_neighbors(cell, pos, 6.0)
    # list order = reference order: j ascending, then image lexicographic
    assert got == sorted(got)
    # pair symmetry (i, j, n) <=> (j, i, -n)
    s = set(got)
    assert all((j, i, -a, -b, -c) in s for (i, j, a, b, c) in got)
    # a cell shorter than 2 rc holds several images of the same j
    if L_ < 12:
        from collections import Counter
        assert max(Counter((g[0], g[1]) for g in got).values()) > 1


# ---- symmetry functions --------------------------------------------------------------------------

def two_species_model(rc=6.0, nR=3, angular=((1.5, 1.0, 1), (3.0, 2.0, -1), (3.0, 4.0, 1)), hidden=(4,), seed=0,
                      act=ACT_TANH):
    rng = np.random.default_rng(seed)
    rad = Radial(RAD_GAUSSIAN, CUT_COS, rc, np.linspace(0, 2, nR), np.linspace(0.5, 1.5, nR))
    ang = Angular(ANG_GAUSS, CUT_COS, rc, np.array([a[0] for a in angular]), np.array([a[1] for a in angular]),
                  np.array([a[2] for a in angular], dtype=np.int32))
    m = Model(2, rc, [[rad, rad], [rad, rad]], [[ang, ang, ang], [ang, ang, ang]], [], E0=np.array([-1.5, -2.5]))
    for I in range(2):
        K = m.input_width(I)
        w = [K, *hidden, 1]
        m.nets.append(Network(w, act, 1.0, xavier_params(w, rng), rng.uniform(0.5, 1.5, K), rng.uniform(-0.5, 0.5, K)))
    return m


def structure(cell, pos, typ):
    return configs.Structure(np.asarray(cell, float), np.asarray(pos, float), np.asarray(typ, dtype=np.int32))


def fc(r, rc=6.0):
    return 0.5 * (np.cos(np.pi * r / rc) + 1) if r < rc else 0.0


def test_symm_two_atom_closed_form():
    m = two_species_model()
    s = structure(np.eye(3) * 30.0, [[1, 1, 1], [3.1, 2.0, 1.5]], [0, 1])
    G = O.Oracle(m).symm(s)
    r = np.linalg.norm(s.pos[1] - s.pos[0])
    rad = m.radial[0][1]
    expect = np.exp(-(rad.eta * (r - rad.rs)) ** 2) * fc(r)
    assert np.allclose(G[0, 0:3], 0.0) and np.allclose(G[0, 3:6], expect, rtol=1e-12, atol=0)
    assert np.allclose(G[1, 0:3], expect, rtol=1e-12, atol=0) and np.allclose(G[1, 3:6], 0.0)
    assert np.allclose(G[:, 6:], 0.0)


def test_symm_triplet_closed_form():
    m = two_species_model()
    pos = np.array([[10.0, 10, 10], [12.0, 10.5, 10], [9.0, 12.0, 11.0]])
    s = structure(np.eye(3) * 40.0, pos, [0, 0, 1])
    G = O.Oracle(m).symm(s)
    a = pos[1] - pos[0]
    b = pos[2] - pos[0]
    rij, rik = np.linalg.norm(a), np.linalg.norm(b)
    cs = a @ b / (rij * rik)
    ang = m.angular[0][tri_index(0, 1)]
    expect = fc(rij) * fc(rik) * np.abs(0.5 + 0.5 * ang.lam * cs) ** ang.zeta * np.exp(-(rij ** 2 + rik ** 2) / ang.eta ** 2)
    o = m.angular_offset(0, tri_index(0, 1))
    assert np.allclose(G[0, o:o + 3], expect, rtol=1e-12, atol=0)
    o00, o11 = m.angular_offset(0, tri_index(0, 0)), m.angular_offset(0, tri_index(1, 1))
    assert np.allclose(G[0, o00:o00 + 3], 0) and np.allclose(G[0, o11:o11 + 3], 0)


def nacl_structures(exact):
    d = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "nacl.json")))
    a_l = d["large"]["a0"]
    a_s = a_l / 3.0 if exact else d["small"]["a0"]
    lv = np.array(d["small"]["lv_rowmajor"]).reshape(3, 3) * a_s
    small = structure(lv.T, np.array(d["small"]["pos_over_a0"]) * a_s, d["small"]["type"])
    fl = np.array(d["large"]["pos_over_a0"])
    if exact:   # the listed coordinates are float32 roundings of multiples of 1/6
        fl = np.round(fl * 6.0) / 6.0
    large = structure(np.eye(3) * a_l, fl * a_l, d["large"]["type"])
    return small, large


@pytest.mark.parametrize("exact", [True, False])
def test_symm_periodic_invariance_nacl(exact):
    """test_nnp_symm.cpp:292-381: G of the 2-atom primitive cell equal the G of the 216-atom supercell."""
    m = two_species_model(rc=8.0)
    orc = O.Oracle(m)
    small, large = nacl_structures(exact)
    Gs, Gl = orc.symm(small), orc.symm(large)
    tol = 1e-10 if exact else 1e-5
    for i in range(216):
        ref = Gs[0] if large.type[i] == 0 else Gs[1]
        assert np.max(np.abs(Gl[i] - ref) / np.maximum(np.abs(ref), 1.0)) < tol


def test_symm_invariances():
    m = two_species_model()
    rng = np.random.default_rng(3)
    cell = np.diag([9.5, 10.0, 10.5])
    pos = rng.uniform(0, 9.5, (16, 3))
    typ = rng.integers(0, 2, 16)
    orc = O.Oracle(m)
    G = orc.symm(structure(cell, pos, typ))
    # translation
    Gt = orc.symm(structure(cell, pos + np.array([0.3, -1.1, 2.4]), typ))
    assert np.allclose(G, Gt, rtol=1e-10, atol=1e-12)
    # rotation of cell and atoms
    q, _ = np.linalg.qr(rng.normal(size=(3, 3)))
    Gr = orc.symm(structure(q @ cell, pos @ q.T, typ))
    assert np.allclose(G, Gr, rtol=1e-10, atol=1e-12)
    # permutation of atoms
    perm = rng.permutation(16)
    Gp = orc.symm(structure(cell, pos[perm], typ[perm]))
    assert np.allclose(G[perm], Gp, rtol=1e-10, atol=1e-12)


# ---- network (nn.cpp) ------------------------------------------------------------------------------

@pytest.mark.parametrize("act", [ACT_TANH, ACT_SIGMOID, ACT_SWISH, ACT_SOFTPLUS, ACT_LINEAR])
def test_network_dEdG_fd(act):
    m = two_species_model(hidden=(5, 4), act=act)
    orc = O.Oracle(m)
    rng = np.random.default_rng(act)
    g = rng.normal(size=m.input_width(0))
    out, d = orc.network(0, g)
    h = 1e-6
    for k in range(len(g)):
        gp, gm = g.copy(), g.copy()
        gp[k] += h
        gm[k] -= h
        fd = (orc.network(0, gp)[0] - orc.network(0, gm)[0]) / (2 * h)
        assert abs(fd - d[k]) < 1e-8


# ---- forces, virial --------------------------------------------------------------------------------

def random_struct(seed, n=14, L_=9.0):
    rng = np.random.default_rng(seed)
    cell = np.array([[L_, 0.4, 0.0], [0.0, L_ * 1.05, 0.3], [0.0, 0.0, L_ * 0.95]])
    f = rng.uniform(0, 1, (n, 3))
    return structure(cell, (cell @ f.T).T, rng.integers(0, 2, n))


def test_forces_fd_momentum_rotation():
    m = two_species_model(hidden=(6, 5))
    orc = O.Oracle(m)
    s = random_struct(1)
    E, F, W = orc.compute(s)
    h = 1e-5
    for a in range(0, s.natoms, 3):
        for d in range(3):
            pp, pm = s.pos.copy(), s.pos.copy()
            pp[a, d] += h
            pm[a, d] -= h
            fd = -(orc.compute(structure(s.cell, pp, s.type))[0] - orc.compute(structure(s.cell, pm, s.type))[0]) / (2 * h)
            assert abs(fd - F[a, d]) < 1e-7 * max(1.0, abs(F[a, d]))
    assert np.allclose(F.sum(axis=0), 0.0, atol=1e-11)
    q, _ = np.linalg.qr(np.random.default_rng(2).normal(size=(3, 3)))
    Er, Fr, _ = orc.compute(structure(q @ s.cell, s.pos @ q.T, s.type))
    assert abs(Er - E) < 1e-10 and np.allclose(Fr, F @ q.T, atol=1e-10)


def test_virial_strain_fd():
    """W_ab = -dE/d(eps_ab) under the homogeneous strain r -> (1+eps) r, cell -> (1+eps) cell."""
    m = two_species_model(hidden=(6,))
    orc = O.Oracle(m)
    s = random_struct(4, n=12)
    _, _, W = orc.compute(s)
    h = 1e-6
    for a in range(3):
        for b in range(3):
            e = np.zeros((3, 3))
            e[a, b] = h
            sp = structure((np.eye(3) + e) @ s.cell, s.pos @ (np.eye(3) + e).T, s.type)
            sm = structure((np.eye(3) - e) @ s.cell, s.pos @ (np.eye(3) - e).T, s.type)
            fd = -(orc.compute(sp)[0] - orc.compute(sm)[0]) / (2 * h)
            assert abs(fd - W[a, b]) < 1e-6 * max(1.0, abs(W[a, b]))


def test_c1_golden_reproduces(golden_dir):
    g = np.load(os.path.join(golden_dir, "c1_oracle.npz"))
    model, [s] = configs.config1()
    assert np.array_equal(s.pos, g["pos"])
    orc = O.Oracle(model)
    nl = orc.neighbors(s)
    assert np.array_equal(nl.off, g["off"]) and np.array_equal(nl.j, g["j"]) and np.array_equal(nl.img, g["img"])
    G = orc.symm(s, nl)
    assert np.array_equal(G, g["G"])


# ---- training (nnptefr.cpp:1292-1425 + force term) ----------------------------------------------------

def total_loss(orc, model, strucs, Gs, Eref, Fref, ld):
    tot = 0.0
    grad = np.zeros(sum(model.param_counts()))
    for s, G, e, f in zip(strucs, Gs, Eref, Fref):
        out = orc.train_struct(ld, s, G, e, f, grad)
        tot += out[0] + out[1]
    return tot, grad


@pytest.mark.parametrize("loss,wf", [(0, 0.0), (0, 0.3), (2, 0.3)])
def test_training_gradient_fd(loss, wf):
    m = two_species_model(hidden=(5, 4), seed=7)
    orc = O.Oracle(m)
    strucs = [random_struct(10 + k, n=9 + 3 * k) for k in range(2)]
    Gs = [orc.symm(s) for s in strucs]
    rng = np.random.default_rng(11)
    Eref = [orc.compute(s)[0] + rng.normal() for s in strucs]
    Fref = [orc.compute(s)[1] + rng.normal(scale=0.3, size=(s.natoms, 3)) for s in strucs]
    ld = LossDesc(loss, 1, 0.7, 0.5, wf)
    x0 = m.all_params()
    _, grad = total_loss(orc, m, strucs, Gs, Eref, Fref, ld)
    h = 1e-6
    for k in rng.choice(len(x0), 25, replace=False):
        for sgn, store in ((1, "p"), (-1, "m")):
            x = x0.copy()
            x[k] += sgn * h
            m.set_all_params(x)
            orc.refresh()
            val = total_loss(orc, m, strucs, Gs, Eref, Fref, ld)[0]
            if store == "p":
                lp = val
            else:
                lm = val
        fd = (lp - lm) / (2 * h)
        assert abs(fd - grad[k]) < 1e-6 * max(1.0, abs(grad[k])), (k, fd, grad[k])
    m.set_all_params(x0)


def test_adam_first_step_closed_form():
    g = np.array([0.3, -2.0, 1e-3])
    p = np.zeros(3)
    mm, v, b = np.zeros(3), np.zeros(3), np.array([0.9, 0.999])
    L.or_adam_step(3, O._d(p), O._d(g), 0.01, 1e-16, O._d(mm), O._d(v), O._d(b))
    assert np.allclose(p, -0.01 * np.sign(g), rtol=1e-9)
    assert np.allclose(b, [0.81, 0.999 ** 2])
