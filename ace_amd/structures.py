"""Structure files for real datasets: extended XYZ, VASP POSCAR and vasprun.xml, CP2K and Quantum ESPRESSO
outputs, with the reference's unit handling.

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

VASP_XML, CP2K and QE outputs: read_vasp_xml_text, read_cp2k_text, read_qe_text (their docstrings give the
lines each one reads).

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


def _num(tok):
    """std::atof of one token: the longest numeric prefix, 0 when there is none."""
    m = re.match(r"\s*[-+]?(\d+\.?\d*|\.\d+)([eEdD][-+]?\d+)?", tok)
    return float(m.group(0).replace("d", "e").replace("D", "e")) if m else 0.0


def _toks(line, delims):
    """Token::read(line, delims): the non-empty pieces between any of the delimiter characters."""
    return [t for t in re.split("[" + re.escape(delims) + "]", line) if t]


def read_vasp_xml_text(text: str, units: str = "metal"):
    """VASP::XML::read (src/format/vasp_struc.cpp:303-470), line-oriented as the reference reads vasprun.xml:
    <atominfo> (the <atoms> and <types> counts, then after five lines one '<rc><c>name</c><c>type</c></rc>'
    row per atom, types 1-based), the first 'basis' block (three <v> lattice vectors), the last e_fr_energy,
    the first 'positions' block (fractional) and the first 'forces' block.  Units: metal keeps A and eV; "au"
    applies the reference's factors (lengths x Ang2Bohr, energies and forces x Eh2Ev (x 1/Ang2Bohr) as
    written there).  Returns (Structure, species names in type order, per-atom names)."""
    u = units.lower()
    if u in ("metal", "lj"):
        s_len, s_e = 1.0, 1.0
    elif u == "au":
        s_len, s_e = ANG2BOHR, EH2EV
    else:
        raise ValueError("invalid units '%s'" % units)
    lines = text.splitlines()
    n = nt = 0
    names, types = [], []
    for q, line in enumerate(lines):
        if "atominfo" in line:
            n = int(_num(_toks(lines[q + 1], "<> \r\t\n")[1]))
            nt = int(_num(_toks(lines[q + 2], "<> \r\t\n")[1]))
            names = [""] * nt
            for i in range(n):
                t = _toks(lines[q + 8 + i], "<> \r\t\n")
                idx = int(_num(t[5])) - 1
                names[idx] = t[2]
                types.append(idx)
            break
    if n <= 0:
        raise ValueError("VASP::XML::read: no atominfo block")
    cell = None
    for q, line in enumerate(lines):
        if "basis" in line:
            cell = np.array([[_num(x) for x in lines[q + 1 + c].split()[1:4]] for c in range(3)]).T * s_len
            break
    if cell is None:
        raise ValueError("VASP::XML::read: no basis block")
    pe = 0.0
    for line in lines:
        if "e_fr_energy" in line:
            pe = _num(_toks(line, "<> \r\t\n")[2]) * s_e
    pos = frc = None
    for q, line in enumerate(lines):
        if "positions" in line:
            fr = np.array([[_num(x) for x in lines[q + 1 + i].split()[1:4]] for i in range(n)]) * s_len
            pos = (cell @ fr.T).T
            break
    for q, line in enumerate(lines):
        if "forces" in line:
            frc = np.array([[_num(x) for x in lines[q + 1 + i].split()[1:4]] for i in range(n)]) * (s_e / s_len)
            break
    if pos is None:
        raise ValueError("VASP::XML::read: no positions block")
    typ = np.array(types, dtype=np.int32)
    return Structure(cell, pos, typ, energy=pe, forces=frc), list(names), [names[t] for t in types]


def read_cp2k_text(text: str, units: str = "metal"):
    """CP2K::read (src/format/cp2k_struc.cpp:15-168), an output file scanned line by line: 'version string'
    (the last number on it), 'Atoms:' (the count), 'ATOMIC COORDINATES' (after two header lines, three before
    version 1000: per atom the name in column 3 and x y z in columns 5-7, A), 'ATOMIC FORCES in' (after two
    lines, x y z from column 4, Hartree/Bohr) or 'Atomic forces' (after one line, from column 3), 'ENERGY|'
    (the last number of the line, Hartree; the last such line wins) and 'CELL|' (the three following 'CELL|'
    lines: vectors a, b, c from column 5, A).  Units: metal gives eV and eV/A; "au" keeps Hartree and
    Hartree/Bohr (positions and cell stay in A there, as in the reference).  Species get type indices in order
    of first appearance.  Returns (Structure, species names in type order, per-atom names)."""
    u = units.lower()
    if u in ("metal",):
        s_len, s_e = BOHR2ANG, EH2EV
    elif u in ("au", "lj"):
        s_len, s_e = 1.0, 1.0
    else:
        raise ValueError("invalid units '%s'" % units)
    lines = text.splitlines()
    version, n = -1.0, 0
    names, pos, frc, cell, pe = [], None, None, None, 0.0
    q = 0
    while q < len(lines):
        line = lines[q]
        if "version string" in line:
            for t in line.split():
                if re.fullmatch(r"[-+]?(\d+\.?\d*|\.\d+)([eE][-+]?\d+)?", t):
                    version = float(t)
            if version < 0:
                raise ValueError("CP2K::read: invalid version")
        elif "Atoms:" in line:
            n = int(_num(line[line.index("Atoms:") + 6:]))
            if n == 0:
                raise ValueError("CP2K::read: found zero atoms")
        elif "ATOMIC COORDINATES" in line:
            q += 3 if version < 1000.0 else 2
            names, rows = [], []
            for i in range(n):
                t = lines[q + 1 + i].split()
                names.append(t[2])
                rows.append([_num(t[4]), _num(t[5]), _num(t[6])])
            pos = np.array(rows)
            q += n
        elif "ATOMIC FORCES in" in line:
            q += 2
            frc = np.array([[_num(x) for x in lines[q + 1 + i].split()[3:6]] for i in range(n)]) * (s_e / s_len)
            q += n
        elif "Atomic forces" in line:
            q += 1
            frc = np.array([[_num(x) for x in lines[q + 1 + i].split()[2:5]] for i in range(n)]) * (s_e / s_len)
            q += n
        elif "ENERGY|" in line:
            pe = _num(line.split()[-1]) * s_e
        elif "CELL|" in line:
            cell = np.array([[_num(x) for x in lines[q + 1 + c].split()[4:7]] for c in range(3)]).T
            q += 4
            while q < len(lines) and "CELL|" in lines[q]:
                q += 1
            # the reference's scan also consumes the first line after the CELL| block
        q += 1
    if pos is None or n <= 0:
        raise ValueError("CP2K::read: no atomic coordinates")
    typ, order = _types(names)
    return Structure(cell if cell is not None else np.zeros((3, 3)), pos, typ, energy=pe, forces=frc), order, names


def read_qe_text(text: str, units: str = "metal"):
    """QE::OUT::read (src/format/qe_struc.cpp:40-306), a pw.x output: 'lattice parameter' (alat, Bohr),
    'crystal axes' (three a(i) = ( x y z ) lines, alat units), 'number of atoms/cell', 'number of atomic types',
    'atomic species' (one name per type), '!    total energy' (Ry; the last wins), 'site n.' blocks (name in
    column 2, tau = ( x y z ) in alat units; the last block wins; positions returned to the cell) and 'Forces'
    blocks (after one line, x y z from column 7, Ry/Bohr; the last wins).  Units: metal gives A, eV, eV/A;
    "au" keeps Bohr and Ry as the reference does.  Types follow the species order of the file.  Returns
    (Structure, species names in type order, per-atom names)."""
    u = units.lower()
    if u == "metal":
        s_pos, s_e = BOHR2ANG, 0.5 * EH2EV
    elif u in ("au", "lj"):
        s_pos, s_e = 1.0, 1.0
    else:
        raise ValueError("invalid units '%s'" % units)
    lines = text.splitlines()
    n = ns = 0
    species, alat, pe = [], 0.0, 0.0
    lv = np.zeros((3, 3))
    q = 0
    while q < len(lines):
        line = lines[q]
        if "!    total energy" in line:
            pe = s_e * _num(line.split("=", 1)[1].split("Ry")[0] if "=" in line else "")
        elif "crystal axes" in line:
            for c in range(3):
                t = lines[q + 1 + c].replace("(", " ").replace(")", " ").split()
                lv[:, c] = [_num(t[3]), _num(t[4]), _num(t[5])]
            q += 3
        elif "lattice parameter" in line:
            alat = s_pos * _num(line.split("=", 1)[1].split("a")[0] if "=" in line else "")
        elif "number of atoms/cell" in line:
            n = int(_num(line.split("=", 1)[1]))
        elif "number of atomic types" in line:
            ns = int(_num(line.split("=", 1)[1]))
        elif "atomic species  " in line:
            species = [lines[q + 1 + i].split()[0] for i in range(ns)]
            q += ns
        q += 1
    if ns <= 0:
        raise ValueError("QE::OUT::read: ERROR reading number of species.")
    if n <= 0:
        raise ValueError("QE::OUT::read: ERROR reading number of atoms.")
    if len(species) != ns:
        raise ValueError("QE::OUT::read: ERROR reading species names.")
    if alat <= 0:
        raise ValueError("QE::OUT::read: ERROR reading alat.")
    if abs(np.linalg.det(lv)) < 1e-12:
        raise ValueError("QE::OUT::read: Invalid lattice vector matrix.")
    cell = alat * lv
    pos, typ, names, frc = None, None, None, None
    for q, line in enumerate(lines):
        if "site n." in line:
            rows, tt, nm = [], [], []
            for i in range(n):
                t = lines[q + 1 + i].replace("(", " ").replace(")", " ").split()
                if t[1] not in species:
                    raise ValueError("QE::OUT::read: Invalid species name in atomic positions")
                nm.append(t[1])
                tt.append(species.index(t[1]))
                rows.append([_num(t[5]), _num(t[6]), _num(t[7])])
            pos, typ, names = np.array(rows) * alat, np.array(tt, dtype=np.int32), nm
        elif "Forces" in line:
            frc = np.array([[_num(x) for x in lines[q + 2 + i].split()[6:9]] for i in range(n)]) * (s_e / s_pos)
    if pos is None:
        raise ValueError("QE::OUT::read: no atomic positions")
    return Structure(cell, _wrap(cell, pos), typ, energy=pe, forces=frc), list(species), names


def read_vasp_xml(path, units="metal"):
    with open(path) as f:
        return read_vasp_xml_text(f.read(), units)


def read_cp2k(path, units="metal"):
    with open(path) as f:
        return read_cp2k_text(f.read(), units)


def read_qe(path, units="metal"):
    with open(path) as f:
        return read_qe_text(f.read(), units)


FILE_FORMATS = ("POSCAR", "XDATCAR", "VASP_XML", "LAMMPS", "XYZ", "CUBE", "QE", "CP2K")   # FILE_FORMAT::read


def read_struc(path, fmt, units="metal"):
    """read_struc (src/format/file_struc.cpp:13-35) by format name: POSCAR, VASP_XML, XYZ, CP2K and QE; the
    other names of FILE_FORMAT (trajectory and volumetric formats) have no structure reader in the reference
    either and raise ValueError, as an unknown name does."""
    f = fmt.upper()
    readers = {"POSCAR": read_poscar, "XYZ": read_xyz, "VASP_XML": read_vasp_xml, "CP2K": read_cp2k,
               "QE": read_qe}
    if f in readers:
        return readers[f](path, units)
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
