"""Independent pins for the oracle functions that closed forms and finite differences of their own
derivatives leave open: the training-loss values (every loss and normalisation of
src/nnp/nnptefr.cpp:609-632, 1353-1378, plus the force term), the six radial forms and four cutoffs
(src/nnp/basis_radial.cpp:175-257, src/nnp/cutoff.cpp:62-117) as two-atom values and structure-level
finite-difference forces, and the activation values of src/ml/nn.cpp:149-240.  Expected values are
written from the mathematical definitions (numpy's cosh, tanh, arcsinh, log1p, expm1), not from the
oracle's code.  CPU only."""
import numpy as np
import pytest

from ace_amd import configs
from ace_amd.model import (ACT_LINEAR, ACT_SIGMOID, ACT_SOFTPLUS, ACT_SWISH, ACT_TANH, ANG_GAUSS, CUT_COS,
                           CUT_POLY3, CUT_STEP, CUT_TANH, LOSS_ASINH, LOSS_HUBER, LOSS_MAE, LOSS_MSE,
                           NORM_CBRT, NORM_IDENTITY, NORM_LINEAR, NORM_LOG, NORM_SQRT, Angular, LossDesc,
                           Model, Network, Radial, xavier_params)
from oracle import oracle as O

RAD_FORMS = (1, 2, 3, 4, 5, 6)   # GAUSSIAN, SECH, LOGISTIC, TANH, LOGCOSH, LOGCOSH2
CUTS = (CUT_COS, CUT_TANH, CUT_POLY3, CUT_STEP)


def structure(cell, pos, typ):
    return configs.Structure(np.asarray(cell, float), np.asarray(pos, float), np.asarray(typ, dtype=np.int32))


# ---- textbook definitions ------------------------------------------------------------------------

def cutoff_value(cut, r, rc):
    """COS 1/2 (cos(pi r/rc) + 1), TANH tanh^3(1 - r/rc), POLY3 1 - 3x^2 + 2x^3 (x = r/rc), STEP 1;
    zero at and beyond rc."""
    if r >= rc:
        return 0.0
    x = r / rc
    return {CUT_COS: 0.5 * (np.cos(np.pi * x) + 1.0), CUT_TANH: np.tanh(1.0 - x) ** 3,
            CUT_POLY3: 1.0 - 3.0 * x * x + 2.0 * x ** 3, CUT_STEP: 1.0}[cut]


def radial_value(form, eta, rs, r):
    """GAUSSIAN exp(-(eta (r - rs))^2); SECH sech(eta (r - rs)); LOGISTIC sech^2(eta (r - rs));
    TANH (1 + tanh(-eta (r - rs))) / 2; LOGCOSH ln(1 + exp(-2 eta (r - rs))) / 2;
    LOGCOSH2 a / (1 - exp(-2 sqrt(2) a)), a = -eta (r - rs), with the limit 1 / (2 sqrt 2) at a = 0."""
    x = eta * (r - rs)
    if form == 1:
        return np.exp(-x * x)
    if form == 2:
        return 1.0 / np.cosh(x)
    if form == 3:
        return 1.0 / np.cosh(x) ** 2
    if form == 4:
        return 0.5 * (1.0 + np.tanh(-x))
    if form == 5:
        return 0.5 * np.log1p(np.exp(-2.0 * x))
    a = -x
    return 1.0 / (2.0 * np.sqrt(2.0)) if a == 0 else -a / np.expm1(-2.0 * np.sqrt(2.0) * a)


def activation(act, c, z):
    """linear z; sigmoid 1/(1 + e^(-cz)); tanh tanh(cz); swish z sigmoid(cz);
    softplus (ln(1 + e^(cz)) - ln 2) / c (zero at z = 0)."""
    sig = 1.0 / (1.0 + np.exp(-c * z))
    return {ACT_LINEAR: z, ACT_SIGMOID: sig, ACT_TANH: np.tanh(c * z), ACT_SWISH: z * sig,
            ACT_SOFTPLUS: (np.logaddexp(0.0, c * z) - np.log(2.0)) / c}[act]


def norm_value(norm, n):
    return {NORM_IDENTITY: 1.0, NORM_LINEAR: 1.0 / n, NORM_SQRT: 1.0 / np.sqrt(n), NORM_CBRT: n ** (-1.0 / 3.0),
            NORM_LOG: 1.0 / (np.log(n) + 1.0)}[norm]


def loss_value(loss, u, delta):
    """MSE u^2/2; MAE |u|; HUBER (pseudo-Huber) delta^2 (sqrt(1 + (u/delta)^2) - 1);
    ASINH delta^2 (1 - sqrt(1 + a^2) + a asinh(a)), a = u/delta (the integral of delta asinh(u/delta))."""
    a = u / delta
    return {LOSS_MSE: 0.5 * u * u, LOSS_MAE: abs(u), LOSS_HUBER: delta * delta * (np.sqrt(1 + a * a) - 1),
            LOSS_ASINH: delta * delta * (1 - np.sqrt(1 + a * a) + a * np.arcsinh(a))}[loss]


# ---- models ---------------------------------------------------------------------------------------

def radial_model(form, cut, rc=6.0, nR=4, hidden=(5,), seed=0, act=ACT_TANH, ntypes=2):
    rng = np.random.default_rng(seed)
    rad = Radial(form, cut, rc, np.linspace(0.0, 2.5, nR), np.linspace(0.6, 1.6, nR))
    ang = Angular(ANG_GAUSS, cut, rc, np.array([2.0, 3.0]), np.array([1.0, 2.0]), np.array([1, -1], dtype=np.int32))
    nab = ntypes * (ntypes + 1) // 2
    m = Model(ntypes, rc, [[rad] * ntypes for _ in range(ntypes)], [[ang] * nab for _ in range(ntypes)], [],
              E0=np.linspace(-1.0, -2.0, ntypes))
    for I in range(ntypes):
        K = m.input_width(I)
        w = [K, *hidden, 1]
        m.nets.append(Network(w, act, 1.0, xavier_params(w, rng), rng.uniform(0.5, 1.5, K), rng.uniform(-0.5, 0.5, K)))
    return m


# ---- radial forms and cutoffs: two-atom values -------------------------------------------------------

@pytest.mark.parametrize("cut", CUTS)
@pytest.mark.parametrize("form", RAD_FORMS)
def test_radial_two_atom_closed_form(form, cut):
    """One pair in a large box: each radial input of each atom is the form times the cutoff."""
    m = radial_model(form, cut)
    rad = m.radial[0][1]
    for r in (0.9, 2.2, 3.7, 5.6):
        s = structure(np.eye(3) * 30.0, [[5.0, 5.0, 5.0], [5.0 + r, 5.0, 5.0]], [0, 1])
        G = O.Oracle(m).symm(s)
        expect = np.array([radial_value(form, e, q, r) for e, q in zip(rad.eta, rad.rs)]) * cutoff_value(cut, r, 6.0)
        nR = len(rad.rs)
        # atom 0 sees one neighbour of species 1 (inputs nR..2nR), atom 1 one of species 0 (inputs 0..nR)
        assert np.allclose(G[0, nR:2 * nR], expect, rtol=1e-10, atol=1e-14), (form, cut, r)
        assert np.allclose(G[1, 0:nR], expect, rtol=1e-10, atol=1e-14), (form, cut, r)
        assert np.allclose(G[0, 0:nR], 0.0) and np.allclose(G[1, nR:2 * nR], 0.0)


def test_logcosh2_at_its_removable_point():
    """|d| = rs exactly (rs = 2.5 and two positions 2.5 apart): a = 0, where the reference takes the limit
    1/(2 sqrt 2) times the cutoff.  (Within ~1e-8 of rs, but not at it, the reference's a e / (e - 1)
    cancels in FP64, and the oracle restates that.)"""
    m = radial_model(6, CUT_COS)
    rad = m.radial[0][1]
    r = float(rad.rs[3])
    assert r == 2.5
    s = structure(np.eye(3) * 30.0, [[5.0, 5.0, 5.0], [7.5, 5.0, 5.0]], [0, 1])
    G = O.Oracle(m).symm(s)
    assert G[0, len(rad.rs) + 3] == pytest.approx(cutoff_value(CUT_COS, r, 6.0) / (2 * np.sqrt(2)), rel=1e-12)


# ---- radial forms and cutoffs: structure-level forces vs central differences ------------------------

@pytest.mark.parametrize("cut", CUTS)
@pytest.mark.parametrize("form", RAD_FORMS)
def test_forces_fd_every_radial_form_and_cutoff(form, cut):
    """The radial branches of compute_force (nnp.cpp:434-446) and the angular cutoff derivatives,
    against central differences of the energy, on a small periodic cell with several images."""
    m = radial_model(form, cut, hidden=(6, 4), seed=form + 10 * cut)
    orc = O.Oracle(m)
    rng = np.random.default_rng(100 * form + cut)
    cell = np.array([[8.5, 0.3, 0.0], [0.0, 8.9, 0.4], [0.0, 0.0, 8.2]])
    f = rng.uniform(0, 1, (10, 3))
    s = structure(cell, (cell @ f.T).T, rng.integers(0, 2, 10))
    E, F, _ = orc.compute(s)
    h = 1e-5
    for a in (0, 3, 7):
        for d in range(3):
            pp, pm = s.pos.copy(), s.pos.copy()
            pp[a, d] += h
            pm[a, d] -= h
            fd = -(orc.compute(structure(cell, pp, s.type))[0] - orc.compute(structure(cell, pm, s.type))[0]) / (2 * h)
            assert abs(fd - F[a, d]) < 1e-7 * max(1.0, abs(F[a, d])), (form, cut, a, d, fd, F[a, d])
    assert np.allclose(F.sum(axis=0), 0.0, atol=1e-10)


# ---- activations --------------------------------------------------------------------------------

@pytest.mark.parametrize("c", [1.0, 1.7])
@pytest.mark.parametrize("act", [ACT_TANH, ACT_SIGMOID, ACT_SWISH, ACT_SOFTPLUS, ACT_LINEAR])
def test_network_output_textbook_activations(act, c):
    """ANN::fp (nn.cpp:1378-1396): ins = inpw (G + inpb), z = b + W a, hidden activations with sharpness
    c, linear last layer, out = outb + outw z_L; weights in the reference's [biases][W col-major] order."""
    rng = np.random.default_rng(act)
    m = radial_model(1, CUT_COS, hidden=(3, 4), act=act, seed=act)
    net = m.nets[0]
    net.c = c
    net.outw, net.outb = 1.3, -0.4
    orc = O.Oracle(m)
    bs, ws = net.unpack()
    for _ in range(5):
        g = rng.normal(scale=2.0, size=m.input_width(0))
        a = net.inpw * (g + net.inpb)
        for l in range(len(ws)):
            z = ws[l] @ a + bs[l]
            a = activation(act, c, z) if l < len(ws) - 1 else z
        expect = net.outb + net.outw * a[0]
        assert orc.network(0, g)[0] == pytest.approx(expect, rel=1e-13, abs=1e-13)


def test_softplus_and_sigmoid_known_values():
    """Fixed points of the definitions: softplus(0) = 0, sigmoid(0) = 1/2, swish(0) = 0, and the
    large-|z| limits softplus(z) -> z - ln2/c, sigmoid -> 1."""
    assert activation(ACT_SOFTPLUS, 1.0, 0.0) == 0.0 and activation(ACT_SIGMOID, 2.0, 0.0) == 0.5
    assert activation(ACT_SOFTPLUS, 2.0, 40.0) == pytest.approx(40.0 - np.log(2) / 2, rel=1e-15)
    # the same points through a one-neuron network in the oracle
    for act, z, val in ((ACT_SOFTPLUS, 0.0, 0.0), (ACT_SIGMOID, 0.0, 0.5), (ACT_SWISH, 0.0, 0.0),
                        (ACT_SOFTPLUS, 30.0, 30.0 - np.log(2)), (ACT_SIGMOID, 40.0, 1.0)):
        m = radial_model(1, CUT_COS, hidden=(1,), act=act)
        K = m.input_width(0)
        net = m.nets[0]
        net.inpw, net.inpb = np.ones(K), np.zeros(K)
        net.outw, net.outb = 1.0, 0.0
        # b_0 = z, W_0 = 0, b_1 = 0, W_1 = 1: out = act(z)
        net.params = np.concatenate([[z], [0.0], np.zeros(K), [1.0]])
        assert O.Oracle(m).network(0, np.zeros(K))[0] == pytest.approx(val, rel=1e-14, abs=1e-14), (act, z)


# ---- training loss values ----------------------------------------------------------------------------

LOSSES = (LOSS_MSE, LOSS_MAE, LOSS_HUBER, LOSS_ASINH)
NORMS = (NORM_IDENTITY, NORM_LINEAR, NORM_SQRT, NORM_CBRT, NORM_LOG)


def loss_case(seed=5):
    m = radial_model(1, CUT_COS, hidden=(5, 4), seed=seed)
    orc = O.Oracle(m)
    rng = np.random.default_rng(seed)
    strucs = []
    for n in (7, 13):
        cell = np.diag([8.0, 8.5, 9.0])
        strucs.append(structure(cell, (cell @ rng.uniform(0, 1, (n, 3)).T).T, rng.integers(0, 2, n)))
    Gs = [orc.symm(s) for s in strucs]
    EF = [orc.compute(s) for s in strucs]
    Eref = [e + rng.normal(scale=0.8) for e, _, _ in EF]
    Fref = [f + rng.normal(scale=0.3, size=f.shape) for _, f, _ in EF]
    return m, orc, strucs, Gs, EF, Eref, Fref


@pytest.mark.parametrize("norm", NORMS)
@pytest.mark.parametrize("loss", LOSSES)
def test_loss_value_every_loss_and_norm(loss, norm):
    """error_cost_o1's loss (nnptefr.cpp:1353-1378: MSE carries 1/2) on dU = (E - E_ref) norm(N), times
    w_energy, plus w_force sum 1/2 (F - F_ref)^2, from the oracle's energies and forces (pinned by the
    closed forms and finite differences above) and the textbook formulas."""
    m, orc, strucs, Gs, EF, Eref, Fref = loss_case()
    delta, we, wf = 0.7, 0.5, 0.3
    ld = LossDesc(loss, norm, delta, we, wf)
    for s, G, (E, F, _), er, fr in zip(strucs, Gs, EF, Eref, Fref):
        grad = np.zeros(sum(m.param_counts()))
        out = orc.train_struct(ld, s, G, er, fr, grad)
        u = (E - er) * norm_value(norm, s.natoms)
        assert out[0] == pytest.approx(we * loss_value(loss, u, delta), rel=1e-12)
        assert out[1] == pytest.approx(wf * 0.5 * np.sum((F - fr) ** 2), rel=1e-12)
        assert out[2] == pytest.approx(((E - er) / s.natoms) ** 2, rel=1e-12)


@pytest.mark.parametrize("loss,norm", [(LOSS_MAE, NORM_CBRT), (LOSS_ASINH, NORM_LOG), (LOSS_MSE, NORM_SQRT),
                                       (LOSS_HUBER, NORM_IDENTITY)])
def test_loss_gradient_fd_other_losses_and_norms(loss, norm):
    """The parameter gradient of the pinned loss (energy and force terms) against central differences."""
    m, orc, strucs, Gs, EF, Eref, Fref = loss_case(seed=9)
    ld = LossDesc(loss, norm, 0.7, 0.5, 0.3)

    def total():
        g = np.zeros(sum(m.param_counts()))
        v = 0.0
        for s, G, er, fr in zip(strucs, Gs, Eref, Fref):
            o = orc.train_struct(ld, s, G, er, fr, g)
            v += o[0] + o[1]
        return v, g

    _, grad = total()
    x0 = m.all_params()
    rng = np.random.default_rng(3)
    h = 1e-6
    for k in rng.choice(len(x0), 20, replace=False):
        vals = []
        for sg in (1, -1):
            x = x0.copy()
            x[k] += sg * h
            m.set_all_params(x)
            orc.refresh()
            vals.append(total()[0])
        fd = (vals[0] - vals[1]) / (2 * h)
        assert abs(fd - grad[k]) < 1e-6 * max(1.0, abs(grad[k])), (k, fd, grad[k])
    m.set_all_params(x0)
    orc.refresh()


# ---- angular forms: three-atom values and structure-level forces ----------------------------------------

ANG_FORMS = (1, 2, 3, 4, 5, 6)   # GAUSS, GAUSS2, SECH, STUDENT3, STUDENT4, STUDENT5


def angular_radial_factor(form, r2s, eta):
    """Distance factor of each angular form (basis_angular.cpp:212-249), s = (rij^2 + rik^2) / eta^2:
    GAUSS exp(-s); GAUSS2 2 / (1 + exp(pi s / 2)); SECH sech(pi s / 2); STUDENT3/4/5 (1 + s)^(-2, -5/2, -3)."""
    s = r2s / eta ** 2
    x = 0.5 * np.pi * s
    return {1: np.exp(-s), 2: 2.0 / (1.0 + np.exp(x)), 3: 1.0 / np.cosh(x), 4: (1.0 + s) ** -2.0,
            5: (1.0 + s) ** -2.5, 6: (1.0 + s) ** -3.0}[form]


def angular_model(form, cut=CUT_COS, rc=6.0, seed=0, hidden=(5,)):
    rng = np.random.default_rng(seed)
    rad = Radial(1, cut, rc, np.array([0.0, 1.0]), np.array([0.8, 1.2]))
    ang = Angular(form, cut, rc, np.array([2.0, 3.0, 3.0, 4.5]), np.array([1.0, 2.5, 4.0, 1.0]),
                  np.array([1, -1, 1, -1], dtype=np.int32))
    m = Model(2, rc, [[rad, rad], [rad, rad]], [[ang] * 3 for _ in range(2)], [], E0=np.array([-1.0, -2.0]))
    for I in range(2):
        K = m.input_width(I)
        w = [K, *hidden, 1]
        m.nets.append(Network(w, ACT_TANH, 1.0, xavier_params(w, rng), rng.uniform(0.5, 1.5, K), rng.uniform(-0.5, 0.5, K)))
    return m


@pytest.mark.parametrize("form", ANG_FORMS)
def test_angular_three_atom_closed_form(form):
    """One triplet (i at the vertex, j and k of species 0 and 1) in a large box: fc(rij) fc(rik)
    |1/2 + lambda cos / 2|^zeta times the form's distance factor (no r_jk term, each pair once)."""
    m = angular_model(form)
    pos = np.array([[10.0, 10, 10], [12.1, 10.4, 10.2], [9.1, 12.3, 11.0]])
    s = structure(np.eye(3) * 40.0, pos, [1, 0, 1])
    G = O.Oracle(m).symm(s)
    a, b = pos[1] - pos[0], pos[2] - pos[0]
    rij, rik = np.linalg.norm(a), np.linalg.norm(b)
    cs = a @ b / (rij * rik)
    ang = m.angular[1][1]     # centre species 1, pair (0, 1): TensorTri index 1
    expect = (cutoff_value(CUT_COS, rij, 6.0) * cutoff_value(CUT_COS, rik, 6.0) * np.abs(0.5 + 0.5 * ang.lam * cs) ** ang.zeta
              * angular_radial_factor(form, rij ** 2 + rik ** 2, ang.eta))
    o = m.angular_offset(1, 1)
    assert np.allclose(G[0, o:o + 4], expect, rtol=1e-12, atol=0), form
    for b_ in (0, 2):
        ob = m.angular_offset(1, b_)
        assert np.allclose(G[0, ob:ob + 4], 0.0)


@pytest.mark.parametrize("form", ANG_FORMS)
def test_forces_fd_every_angular_form(form):
    """The angular branches of compute_force (nnp.cpp:448-471) for every form, against central differences."""
    m = angular_model(form, hidden=(6, 4), seed=form)
    orc = O.Oracle(m)
    rng = np.random.default_rng(200 + form)
    cell = np.array([[8.0, 0.2, 0.0], [0.0, 8.4, 0.3], [0.0, 0.0, 7.9]])
    s = structure(cell, (cell @ rng.uniform(0, 1, (11, 3)).T).T, rng.integers(0, 2, 11))
    E, F, _ = orc.compute(s)
    h = 1e-5
    for a in (1, 5, 9):
        for d in range(3):
            pp, pm = s.pos.copy(), s.pos.copy()
            pp[a, d] += h
            pm[a, d] -= h
            fd = -(orc.compute(structure(cell, pp, s.type))[0] - orc.compute(structure(cell, pm, s.type))[0]) / (2 * h)
            assert abs(fd - F[a, d]) < 1e-7 * max(1.0, abs(F[a, d])), (form, a, d, fd, F[a, d])
