"""Structure files for real datasets: extended XYZ and VASP POSCAR, with the reference's unit handling.

Extended XYZ, XYZ::read (src/format/xyz_struc.cpp:39-236): line 1 the atom count; line 2 (case-insensitive)
    Properties=species:S:1:pos:R:3[:forces:R:3]... (also charge/mass R:1, dipole R:3; other properties are
    skipped by their count), potential_energy=<E>, Lattice="a_x a_y a_z b_x b_y b_z c_x c_y c_z" (lattice
    vectors, i.e. the columns of Structure::R); then one line per atom.  Positions are returned to the cell
    (Cell::returnToCell) when a lattice is given; species get type indices in order of first appearance.
VASP POSCAR, VASP::POSCAR::read (src/format/vasp_struc.cpp:41-205): comment, scale, three lattice vectors,
    species names, counts, a mode line ('D...' = direct, anything else Cartesian), positions.  As in the
    reference, Cartesian coordinates are not multiplied by the scale factor (only the lattice is).

Units (xyz_struc.cpp:51-62, vasp_struc.cpp:56-72; constants of src/chem/units.hpp): "metal" (A, eV) keeps the
numbers; "au" converts lengths to Bohr, energies to Hartree, forces to Hartree/Bohr.

Writers: write_xyz emits the Lattice field in the order its reader expects (the reference's writer puts
R(1,2) in the sixth slot, xyz_struc.cpp:257); write_poscar groups atoms by species in order of first
appearance so that the counts line describes the positions that follow (the reference writes the counts
of grouped species but the positions in atom order, vasp_struc.cpp:222-270).
"""
from __future__ import annotations

import re

import numpy as np

from .configs import Structure

BOHR2ANG = 0.529177210903
ANG2BOHR = 1.0 / BOHR2ANG
EH2EV = 27.211386245988
EV2EH = 1.0 / EH2EV


def _scales(units):
    u = units.lower()
    if u in ("metal", "lj"):
        return 1.0, 1.0
    if u == "au":
        return ANG2BOHR, EV2EH
    raise ValueError("invalid units '%s'" % units)


def _wrap(cell, pos):
    """Cell::returnToCell: fractional coordinates into [0, 1)."""
    f = np.linalg.solve(cell, pos.T).T
    f -= np.floor(f)
    return (cell @ f.T).T


def _types(names):
    order = []
    for s in names:
        if s not in order:
            order.append(s)
    return np.array([order.index(s) for s in names], dtype=np.int32), order


def read_xyz_text(text: str, units: str = "metal"):
    """One structure; returns (Structure, species names in type order, per-atom names)."""
    s_len, s_e = _scales(units)
    lines = text.splitlines()
    try:
        n = int(lines[0].split()[0])
    except (IndexError, ValueError):
        n = 0
    if n <= 0:
        raise ValueError("XYZ::read: found zero atoms")
    head = lines[1].upper()
    cols = {}
    nd = 0
    m = re.search(r"PROPERTIES\s*=\s*(\S+)", head)
    if m:
        f = m.group(1).split(":")
        q = 0
        while q + 2 < len(f):
            tag, typ, cnt = f[q], f[q + 1], int(f[q + 2])
            q += 3
            want = {"SPECIES": ("S", 1, "species"), "CHARGE": ("R", 1, "charge"), "CHG": ("R", 1, "charge"),
                    "Q": ("R", 1, "charge"), "MASS": ("R", 1, "mass"), "POS": ("R", 3, "pos"),
                    "POSN": ("R", 3, "pos"), "POSITION": ("R", 3, "pos"), "FORCES": ("R", 3, "forces"),
                    "FORCE": ("R", 3, "forces"), "DIPOLE": ("R", 3, "dipole"), "DIPOLES": ("R", 3, "dipole")}
            if tag in want:
                t, c, key = want[tag]
                if typ != t:
                    raise ValueError("XYZ::read: invalid %s data type" % key)
                if cnt != c:
                    raise ValueError("XYZ::read: invalid %s length" % key)
                cols[key] = nd
            nd += cnt
    m = re.search(r"POTENTIAL_ENERGY\s*=\s*(\S+)", head)
    pe = float(m.group(1)) if m else 0.0
    cell = None
    m = re.search(r'LATTICE\s*=\s*"([^"]*)"', head)
    if m:
        v = [float(x) for x in m.group(1).split()[:9]]
        cell = np.array(v).reshape(3, 3).T * s_len        # columns = lattice vectors
    names, pos, frc = [], np.zeros((n, 3)), None
    if "forces" in cols:
        frc = np.zeros((n, 3))
    for i in range(n):
        tok = lines[2 + i].split()
        if "species" in cols:
            names.append(tok[cols["species"]])
        if "pos" in cols:
            pos[i] = [float(tok[cols["pos"] + d]) * s_len for d in range(3)]
        if frc is not None:
            frc[i] = [float(tok[cols["forces"] + d]) * s_e / s_len for d in range(3)]
    if cell is not None and np.linalg.norm(cell) > 0:
        pos = _wrap(cell, pos)
    typ, order = _types(names) if names else (np.zeros(n, dtype=np.int32), [])
    return Structure(cell if cell is not None else np.zeros((3, 3)), pos, typ, energy=pe * s_e, forces=frc), order, names


def xyz_text(s: Structure, names, units: str = "metal") -> str:
    """Extended XYZ of one structure (names: species name of each type index)."""
    s_len, s_e = _scales(units)
    R = np.asarray(s.cell) / s_len
    lat = " ".join("%f" % R[r, c] for c in range(3) for r in range(3))
    fr = s.forces is not None
    props = "species:S:1:pos:R:3" + (":forces:R:3" if fr else "")
    out = ["%i" % s.natoms, 'Properties=%s potential_energy=%f pbc="T T T" Lattice="%s"' % (props, s.energy / s_e, lat)]
    for i in range(s.natoms):
        p = s.pos[i] / s_len
        line = "%-2s %19.10f %19.10f %19.10f" % (names[s.type[i]], p[0], p[1], p[2])
        if fr:
            f = s.forces[i] * s_len / s_e
            line += " %19.10f %19.10f %19.10f" % (f[0], f[1], f[2])
        out.append(line)
    return "\n".join(out) + "\n"


def read_poscar_text(text: str, units: str = "metal"):
    """Returns (Structure, species names in type order, per-atom names)."""
    s_len, _ = _scales(units)
    lines = text.splitlines()
    scale = float(lines[1].split()[0])
    R = np.array([[float(x) for x in lines[2 + c].split()[:3]] for c in range(3)]).T * (s_len * scale)
    names = lines[5].split()
    counts = [int(float(x)) for x in lines[6].split()]
    if not names or not counts or len(names) != len(counts):
        raise ValueError("VASP::POSCAR::read: invalid number of species")
    n = sum(counts)
    direct = lines[7][:1] == "D"
    p = np.array([[float(x) for x in lines[8 + i].split()[:3]] for i in range(n)])
    pos = (R @ p.T).T if direct else p * s_len
    per_atom = [names[q] for q in range(len(names)) for _ in range(counts[q])]
    typ = np.array([q for q in range(len(names)) for _ in range(counts[q])], dtype=np.int32)
    return Structure(R, pos, typ), list(names), per_atom


def poscar_text(s: Structure, names, direct: bool = True, units: str = "metal") -> str:
    """POSCAR with the atoms grouped by species in order of first appearance; returns the text.  The
    grouping permutation is stable, so atoms of one species keep their order."""
    s_len, _ = _scales(units)
    R = np.asarray(s.cell) / s_len
    order = []
    for t in s.type:
        if t not in order:
            order.append(int(t))
    perm = [i for t in order for i in range(s.natoms) if s.type[i] == t]
    out = ["system", "%10.4f" % 1.0]
    for c in range(3):
        out.append("".join("%11.6f " % R[r, c] for r in range(3)))
    out.append("".join("%s " % names[t] for t in order))
    out.append("".join("%i " % int(np.sum(s.type == t)) for t in order))
    out.append("Direct" if direct else "Cart")
    Rinv = np.linalg.inv(np.asarray(s.cell))
    for i in perm:
        v = Rinv @ s.pos[i] if direct else s.pos[i] / s_len
        out.append("%.8f %.8f %.8f" % (v[0], v[1], v[2]))
    return "\n".join(out) + "\n"


def read_xyz(path, units="metal"):
    with open(path) as f:
        return read_xyz_text(f.read(), units)


def read_poscar(path, units="metal"):
    with open(path) as f:
        return read_poscar_text(f.read(), units)


FILE_FORMATS = ("POSCAR", "XDATCAR", "VASP_XML", "LAMMPS", "XYZ", "CUBE", "QE", "CP2K")   # FILE_FORMAT::read


def read_struc(path, fmt, units="metal"):
    """read_struc (src/format/file_struc.cpp:13-35) by format name: POSCAR and XYZ are read here; the
    reference's other structure readers (VASP_XML, CP2K, QE) are not restated and raise ValueError, as an
    unknown name does."""
    f = fmt.upper()
    if f == "POSCAR":
        return read_poscar(path, units)
    if f == "XYZ":
        return read_xyz(path, units)
    if f in FILE_FORMATS:
        raise ValueError("read_struc: format %s is not supported by this build" % f)
    raise ValueError("read_struc: invalid file format")


def write_struc(path, fmt, s: Structure, names, units="metal"):
    """write_struc (src/format/file_struc.cpp:37-50): POSCAR or XYZ."""
    f = fmt.upper()
    if f == "POSCAR":
        return write_poscar(path, s, names, units=units)
    if f == "XYZ":
        return write_xyz(path, s, names, units)
    raise ValueError("write_struc: invalid file format")


def to_model_types(s: Structure, file_names, model_names):
    """Type indices of the model's species for a structure read from a file (the NNP index of every
    atom's name, src/nnp/nnptefr.cpp:2489-2496); unknown species raise ValueError."""
    idx = []
    for q in file_names:
        if q not in model_names:
            raise ValueError("species '%s' is not in the potential" % q)
        idx.append(model_names.index(q))
    return Structure(s.cell, s.pos, np.array([idx[t] for t in s.type], dtype=np.int32), s.energy, s.forces)


def write_xyz(path, s: Structure, names, units="metal"):
    with open(path, "w") as f:
        f.write(xyz_text(s, names, units))


def write_poscar(path, s: Structure, names, direct=True, units="metal"):
    with open(path, "w") as f:
        f.write(poscar_text(s, names, direct, units))
