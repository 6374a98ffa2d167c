"""ace's text potential and basis files (NNP::read/write, src/nnp/nnp.cpp:96-307; BasisR/BasisA/ANN read and
write).  CPU only.

tests/golden/tiny_potential.txt is written by hand in the documented layout (mixed-case type tags, a '#'
comment, W row by row) so that the parameter order is pinned independently of the writer."""
import numpy as np
import pytest

from ace_amd import configs
from ace_amd import potential as P
from oracle import oracle as O


def test_hand_written_file(golden_dir):
    m = P.read_potential(f"{golden_dir}/tiny_potential.txt")
    assert m.ntypes == 1 and m.names == ["Ar"] and m.rc == 6.0
    assert m.E0[0] == -0.5 and m.types[0]["mass"] == 39.948 and m.types[0]["z"] == 18
    r, a = m.radial[0][0], m.angular[0][0]
    assert r.cut == 2 and r.form == 1 and list(r.rs) == [0.0, 1.5] and list(r.eta) == [0.5, 1.25]
    assert a.cut == 2 and a.form == 1 and list(a.eta) == [2.0, 2.0] and list(a.zeta) == [1.0, 4.0]
    assert list(a.lam) == [1, -1]
    n = m.nets[0]
    assert n.widths == [4, 3, 1] and n.act == 3 and n.c == 1.0 and n.outw == 1.5 and n.outb == 0.01
    assert list(n.inpw) == [1.0, 2.0, 0.5, 0.25] and list(n.inpb) == [0.1, -0.2, 0.3, -0.4]
    bs, ws = n.unpack()
    assert np.array_equal(bs[0], [0.01, 0.02, 0.03]) and np.array_equal(bs[1], [-0.05])
    # the file lists W (outputs x inputs) row by row; W[o][k] = 0.1 (o + 1) + 0.01 (k + 1)
    assert np.allclose(ws[0], [[0.1 * (o + 1) + 0.01 * (k + 1) for k in range(4)] for o in range(3)], rtol=0, atol=1e-15)
    assert np.array_equal(ws[1], [[0.7, -0.8, 0.9]])
    # flat vector: biases, then W column-major (nn.cpp:881-896)
    assert np.allclose(n.params[4:8], [0.11, 0.21, 0.31, 0.12], rtol=0, atol=1e-15)
    # atomic energy of a known input against a hand evaluation of the textbook network
    x = np.array([0.3, -0.1, 0.7, 0.2])
    h = np.tanh(ws[0] @ ((x + n.inpb) * n.inpw) + bs[0])
    e_hand = n.outb + n.outw * (ws[1] @ h + bs[1])[0]
    from ace_amd.model import ACT_TANH
    assert n.act == ACT_TANH
    orc = O.Oracle(m)
    e, _ = orc.network(0, x)
    assert np.isclose(e, e_hand, rtol=0, atol=1e-14)


def test_write_read_is_idempotent_and_rounds_the_basis_like_the_reference():
    for model in (configs.config2()[0], configs.c5_model(np.random.default_rng(1))):
        model.names = ["A%i" % I for I in range(model.ntypes)]
        t1 = P.potential_text(model)
        m2 = P.read_potential_text(t1)
        assert P.potential_text(m2) == t1
        for I in range(model.ntypes):
            for J in range(model.ntypes):
                assert np.array_equal(m2.radial[I][J].eta, np.round(model.radial[I][J].eta, 6))
            for b in range(len(model.angular[I])):
                assert np.array_equal(m2.angular[I][b].eta, np.round(model.angular[I][b].eta, 6))
                assert np.array_equal(m2.angular[I][b].lam, model.angular[I][b].lam)
            assert np.allclose(m2.nets[I].params, model.nets[I].params, rtol=0, atol=6e-16)   # %.15f + parse
            assert m2.nets[I].widths == model.nets[I].widths


def test_read_potential_then_evaluate_matches_the_written_model():
    """Energy and forces of the read-back potential equal those of the model rounded the same way."""
    model, [s] = configs.config1()
    G = O.Oracle(model).symm(s)
    configs.set_input_scaling(model, [s], G)
    m2 = P.read_potential_text(P.potential_text(model))
    E1, F1, _ = O.Oracle(model).compute(s)
    E2, F2, _ = O.Oracle(m2).compute(s)
    # only the 6-decimal basis rounding separates them (eta_ref 0.1...1.4 to 1e-6)
    assert abs(E1 - E2) / s.natoms < 1e-5 and np.abs(F1 - F2).max() < 1e-3


def test_basis_file_round_trip():
    model, _ = configs.config2()
    model.names = ["Fe", "Cr"]
    text = P.basis_text(model, 1)
    other = configs.config2()[0]
    other.names = ["Fe", "Cr"]
    for J in range(2):
        other.radial[1][J].eta = other.radial[1][J].eta * 2.0
    P.read_basis_text(text, other, "Cr")
    for J in range(2):
        assert np.array_equal(other.radial[1][J].eta, np.round(model.radial[1][J].eta, 6))
    with pytest.raises(ValueError):
        P.read_basis_text(text, other, "Fe")                           # central species differs
    with pytest.raises(ValueError):
        P.read_basis_text(text.replace("rc 6.000000", "rc 5.000000"), other, "Cr")


@pytest.mark.parametrize("edit", [("ntypes 1", "ntypes 0"), ("rc 6.000000", "rc 0.000000"),
                                  ("neuron TANH", "neuron FOO"), ("basis_radial Ar", "basis_radial Kr"),
                                  ("nn 4 3 1", "nn 5 3 1")])
def test_invalid_files_raise(golden_dir, edit):
    text = open(f"{golden_dir}/tiny_potential.txt").read()
    with pytest.raises(ValueError):
        P.read_potential_text(text.replace(*edit, 1))
