"""Structure files (extended XYZ, VASP POSCAR) with the reference's parsing and unit rules
(src/format/xyz_struc.cpp, src/format/vasp_struc.cpp, src/chem/units.hpp).  CPU only.  The fixtures are
hand-written: tests/golden/mixed4.xyz (triclinic lattice, interleaved species, an atom outside the cell,
mixed-case header), tests/golden/nacl8.poscar (direct coordinates)."""
import numpy as np
import pytest

from ace_amd import configs
from ace_amd import structures as S


def test_xyz_fixture(golden_dir):
    s, names, per_atom = S.read_xyz(f"{golden_dir}/mixed4.xyz")
    assert names == ["Al", "Cu"] and per_atom == ["Al", "Cu", "Al", "Cu"] and list(s.type) == [0, 1, 0, 1]
    assert s.energy == -12.5
    # Lattice lists the lattice vectors a, b, c: the columns of R
    assert np.array_equal(s.cell[:, 1], [0.5, 4.0, 0.0]) and np.array_equal(s.cell[:, 2], [0.2, 0.3, 4.0])
    assert np.array_equal(s.forces[1], [-0.1, 0.2, -0.3])
    # positions returned to the cell: the same point modulo lattice vectors, fractional in [0, 1)
    raw = np.array([[0.1, 0.2, 0.3], [2.0, 2.0, 0.1], [-0.5, 1.0, 1.0], [1.0, 1.0, 4.5]])
    f = np.linalg.solve(s.cell, s.pos.T).T
    assert np.all((f >= 0) & (f < 1))
    d = np.linalg.solve(s.cell, (s.pos - raw).T).T
    assert np.allclose(d, np.round(d), atol=1e-12)
    assert np.allclose(s.pos[0], raw[0]) and not np.allclose(s.pos[2], raw[2])


def test_xyz_units_and_round_trip(golden_dir):
    s, names, _ = S.read_xyz(f"{golden_dir}/mixed4.xyz")
    a, _, _ = S.read_xyz(f"{golden_dir}/mixed4.xyz", units="au")
    assert np.allclose(a.pos, s.pos * S.ANG2BOHR) and np.isclose(a.energy, -12.5 * S.EV2EH)
    assert np.allclose(a.forces, s.forces * S.EV2EH / S.ANG2BOHR)
    s2, names2, _ = S.read_xyz_text(S.xyz_text(s, names))
    assert names2 == names and np.allclose(s2.pos, s.pos, atol=1e-9) and np.allclose(s2.cell, s.cell, atol=1e-6)
    assert np.allclose(s2.forces, s.forces, atol=1e-9) and s2.energy == s.energy
    a2, _, _ = S.read_xyz_text(S.xyz_text(a, names, units="au"), units="au")
    assert np.allclose(a2.pos, a.pos, atol=1e-8)


def test_poscar_fixture_and_round_trip(golden_dir):
    s, names, per_atom = S.read_poscar(f"{golden_dir}/nacl8.poscar")
    assert names == ["Na", "Cl"] and list(s.type) == [0] * 4 + [1] * 4 and per_atom[4] == "Cl"
    assert np.allclose(s.pos[1], [2.82, 2.82, 0.0]) and np.allclose(s.cell, np.eye(3) * 5.64)
    # a structure with interleaved species is written grouped by species
    t = configs.Structure(s.cell, s.pos[[0, 4, 1, 5, 2, 6, 3, 7]], np.array([0, 1] * 4, dtype=np.int32))
    r, names2, _ = S.read_poscar_text(S.poscar_text(t, names))
    assert names2 == names and np.allclose(r.pos, s.pos, atol=1e-7) and list(r.type) == list(s.type)
    r2, _, _ = S.read_poscar_text(S.poscar_text(t, names, direct=False))
    assert np.allclose(r2.pos, s.pos, atol=1e-7)


def test_poscar_cartesian_ignores_the_scale_factor_like_the_reference(golden_dir):
    text = open(f"{golden_dir}/nacl8.poscar").read().replace("   1.0000", "   2.0000", 1)
    text = text.replace("Direct", "Cartesian").replace("0.5", "1.5")
    s, _, _ = S.read_poscar_text(text)
    assert np.allclose(s.cell, np.eye(3) * 11.28) and np.allclose(s.pos[1], [1.5, 1.5, 0.0])


def test_model_type_mapping(golden_dir):
    s, names, _ = S.read_xyz(f"{golden_dir}/mixed4.xyz")
    m = S.to_model_types(s, names, ["Cu", "Al"])
    assert list(m.type) == [1, 0, 1, 0]
    with pytest.raises(ValueError):
        S.to_model_types(s, names, ["Cu"])


@pytest.mark.parametrize("text", ["0\nProperties=species:S:1:pos:R:3\n",
                                  "1\nProperties=species:S:2:pos:R:3\nAl 0 0 0\n",
                                  "1\nProperties=species:S:1:pos:I:3\nAl 0 0 0\n"])
def test_invalid_xyz_raises(text):
    with pytest.raises(ValueError):
        S.read_xyz_text(text)


def test_read_write_struc_by_format_name(golden_dir, tmp_path):
    """read_struc / write_struc (file_struc.cpp:13-50) dispatch on the format name, case-insensitively;
    known formats without a reader here and unknown names raise."""
    s1, n1, _ = S.read_struc(f"{golden_dir}/mixed4.xyz", "xyz")
    s2, n2, _ = S.read_xyz(f"{golden_dir}/mixed4.xyz")
    assert n1 == n2 and np.array_equal(s1.pos, s2.pos) and np.array_equal(s1.cell, s2.cell)
    p, names, _ = S.read_struc(f"{golden_dir}/nacl8.poscar", "POSCAR")
    S.write_struc(str(tmp_path / "out.poscar"), "POSCAR", p, names)
    q, names2, _ = S.read_struc(str(tmp_path / "out.poscar"), "POSCAR")
    assert names2 == names and np.allclose(q.pos, p.pos, atol=1e-7)
    for fmt in ("XDATCAR", "CUBE", "PDB"):
        with pytest.raises(ValueError):
            S.read_struc(f"{golden_dir}/nacl8.poscar", fmt)
    for fmt in ("QE", "CP2K", "VASP_XML"):       # readers present, but the file is not in their format
        with pytest.raises((ValueError, IndexError)):
            S.read_struc(f"{golden_dir}/nacl8.poscar", fmt)
    with pytest.raises(ValueError):
        S.write_struc(str(tmp_path / "x.cube"), "CUBE", p, names)


# ---- VASP_XML, CP2K and QE outputs: hand-written excerpts in each program's layout (tests/golden/formats) ----
EH2EV, BOHR2ANG = 27.211386245988, 0.529177210903     # src/chem/units.hpp:24,30


def test_vasp_xml(golden_dir):
    """VASP::XML::read (vasp_struc.cpp:303-470): types from <atominfo>, the first basis and positions, the last
    e_fr_energy, the first forces block."""
    s, names, per = S.read_struc(f"{golden_dir}/formats/vasprun_si2.xml", "VASP_XML")
    cell = np.array([[5.1, 0.2, 0.0], [0.0, 5.3, 0.1], [0.0, 0.0, 5.5]])
    frac = np.array([[0, 0, 0], [0.25, 0.5, 0.1], [0.5, 0.25, 0.75]])
    assert names == ["Si", "O"] and per == ["Si", "O", "Si"] and list(s.type) == [0, 1, 0]
    assert np.array_equal(s.cell, cell)
    assert np.allclose(s.pos, frac @ cell.T, atol=1e-14)
    assert s.energy == -12.3456789
    assert np.array_equal(s.forces, np.array([[0.1, -0.2, 0.3], [-0.4, 0.5, -0.6], [0.3, -0.3, 0.3]]))


def test_cp2k(golden_dir):
    """CP2K::read (cp2k_struc.cpp:15-168): names and positions (A) from ATOMIC COORDINATES, forces in
    Hartree/Bohr -> eV/A, the last ENERGY| line in Hartree -> eV, the cell from the three lines after the first
    CELL| line (the line after the CELL| block is skipped, as the reference's scan does)."""
    s, names, per = S.read_struc(f"{golden_dir}/formats/cp2k_h2o.out", "CP2K")
    assert names == ["O", "H"] and per == ["O", "H", "H"] and list(s.type) == [0, 1, 1]
    assert np.array_equal(s.cell, np.array([[10.0, 0.0, 0.0], [0.0, 10.0, 0.5], [0.0, 0.0, 10.0]]))
    assert np.array_equal(s.pos, np.array([[5.0, 5.0, 5.12], [5.76, 5.0, 4.52], [4.24, 5.0, 4.52]]))
    assert s.energy == -17.15 * EH2EV
    f = np.array([[0.01, -0.002, 0.003], [-0.005, 0.001, -0.0015], [-0.005, 0.001, -0.0015]])
    assert np.allclose(s.forces, f * (EH2EV / BOHR2ANG), rtol=1e-15, atol=0)
    au, _, _ = S.read_cp2k(f"{golden_dir}/formats/cp2k_h2o.out", units="au")
    assert au.energy == -17.15 and np.array_equal(au.forces, f) and np.array_equal(au.pos, s.pos)


def test_qe(golden_dir):
    """QE::OUT::read (qe_struc.cpp:40-306): alat (Bohr) x crystal axes, tau (alat units) returned to the cell,
    the total energy in Ry and the forces in Ry/Bohr converted to eV and eV/A, types in species order."""
    s, names, per = S.read_struc(f"{golden_dir}/formats/qe_si2.out", "QE")
    alat = 10.2 * BOHR2ANG
    lv = np.array([[1.0, 0.05, 0.0], [0.0, 0.95, 0.1], [0.0, 0.0, 1.05]])
    cell = alat * lv
    assert names == ["Si", "O"] and per == ["Si", "O", "Si"] and list(s.type) == [0, 1, 0]
    assert np.allclose(s.cell, cell, rtol=1e-15, atol=0)
    tau = np.array([[0, 0, 0], [0.25, 0.25, 0.25], [1.1, 0.5, 0.5]]) * alat
    f = np.linalg.solve(cell, tau.T).T
    assert np.allclose(s.pos, (cell @ (f - np.floor(f)).T).T, atol=1e-12)
    assert s.pos[2, 0] < tau[2, 0]                       # atom 3 came back into the cell
    assert s.energy == 0.5 * EH2EV * -93.45678
    fr = np.array([[0.001, 0.002, -0.003], [-0.0004, 0.0005, 0.0006], [-0.0006, -0.0025, 0.0024]])
    assert np.allclose(s.forces, fr * (0.5 * EH2EV / BOHR2ANG), rtol=1e-15, atol=0)


def test_dft_readers_feed_the_model_type_mapping(golden_dir):
    """Species names of any reader map onto a potential's species (nnptefr.cpp:2489-2496)."""
    s, names, per = S.read_struc(f"{golden_dir}/formats/qe_si2.out", "QE")
    t = S.to_model_types(s, names, ["O", "Si"])
    assert list(t.type) == [1, 0, 1]
