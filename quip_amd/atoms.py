"""Atoms container and extended-XYZ reader/writer.

Mirrors the parts of QUIP's ``type(Atoms)`` that the GAP evaluation path reads
(src/libAtoms/Atoms_types.f95:224-347): positions, atomic numbers, lattice
vectors, periodicity, per-config info and per-atom arrays.  The lattice is
stored with ROWS = lattice vectors (the extxyz ``Lattice=`` order); the C-ABI
takes the same memory as QUIP's column-major ``lattice(3,3)`` whose columns are
the vectors, so ``lattice`` passes through unchanged.
"""
import re
import numpy as np

_SYMBOLS = ("X H He Li Be B C N O F Ne Na Mg Al Si P S Cl Ar K Ca Sc Ti V Cr Mn Fe Co Ni Cu "
            "Zn Ga Ge As Se Br Kr Rb Sr Y Zr Nb Mo Tc Ru Rh Pd Ag Cd In Sn Sb Te I Xe Cs Ba La Ce "
            "Pr Nd Pm Sm Eu Gd Tb Dy Ho Er Tm Yb Lu Hf Ta W Re Os Ir Pt Au Hg Tl Pb Bi Po At Rn Fr "
            "Ra Ac Th Pa U Np Pu Am Cm Bk Cf Es Fm Md No Lr Rf Db Sg Bh Hs Mt Ds Rg Cn Nh Fl Mc Lv "
            "Ts Og").split()
SYMBOL_TO_Z = {s: i for i, s in enumerate(_SYMBOLS)}


class Atoms:
    """Positions (N,3) in Angstrom, numbers (N,), cell (3,3) rows = vectors, pbc (3,)."""

    def __init__(self, numbers, positions, cell, pbc=(True, True, True), info=None, arrays=None):
        self.numbers = np.ascontiguousarray(numbers, dtype=np.int32).reshape(-1)
        self.positions = np.ascontiguousarray(positions, dtype=np.float64).reshape(-1, 3)
        self.cell = np.ascontiguousarray(cell, dtype=np.float64).reshape(3, 3)
        self.pbc = np.array(pbc, dtype=bool).reshape(3)
        self.info = dict(info or {})
        self.arrays = dict(arrays or {})
        if len(self.numbers) != len(self.positions):
            raise ValueError("numbers and positions differ in length")

    def __len__(self):
        return len(self.numbers)

    def copy(self):
        return Atoms(self.numbers.copy(), self.positions.copy(), self.cell.copy(), self.pbc.copy(),
                     dict(self.info), {k: np.copy(v) for k, v in self.arrays.items()})

    def get_volume(self):
        return abs(float(np.linalg.det(self.cell)))

    # ASE-like accessors backed by an attached calculator
    calc = None

    def get_potential_energy(self):
        return self.calc.get_property("energy", self)

    def get_forces(self):
        return self.calc.get_property("forces", self)

    def get_stress(self):
        return self.calc.get_property("stress", self)


def _parse_kv(line):
    out = {}
    for m in re.finditer(r'(\w+)=("([^"]*)"|\S+)|(\w+)', line):
        if m.group(1):
            out[m.group(1)] = m.group(3) if m.group(3) is not None else m.group(2)
        else:
            out[m.group(4)] = "T"
    return out


def read_xyz(path, index=None):
    """Read an extended-XYZ file (the format QUIP's xyz.c / CInOutput.f95 read).

    ``path`` is a file name or an open text stream.  Returns a list of Atoms
    (or one Atoms when ``index`` is an int)."""
    frames = []
    if hasattr(path, "read"):
        lines = path.read().splitlines()
    else:
        with open(path) as fh:
            lines = fh.read().splitlines()
    i = 0
    while i < len(lines):
        if not lines[i].strip():
            i += 1
            continue
        n = int(lines[i].split()[0])
        kv = _parse_kv(lines[i + 1])
        lat = np.array([float(x) for x in kv["Lattice"].split()]).reshape(3, 3) if "Lattice" in kv else np.eye(3)
        pbc = [p == "T" for p in kv.get("pbc", "T T T").split()]
        props = kv.get("Properties", "species:S:1:pos:R:3").split(":")
        cols = []
        for k in range(0, len(props), 3):
            cols.append((props[k], props[k + 1], int(props[k + 2])))
        rows = [lines[i + 2 + a].split() for a in range(n)]
        arrays = {}
        c0 = 0
        for name, typ, ncol in cols:
            raw = [r[c0:c0 + ncol] for r in rows]
            if typ == "R":
                arr = np.array(raw, dtype=float)
            elif typ == "I":
                arr = np.array(raw, dtype=np.int64)
            elif typ == "L":
                arr = np.array([[v in ("T", "True", "1") for v in r] for r in raw])
            else:
                arr = np.array(raw, dtype=object)
            arrays[name] = arr[:, 0] if ncol == 1 else arr
            c0 += ncol
        if "Z" in arrays:
            numbers = arrays["Z"].astype(np.int32)
        else:
            numbers = np.array([SYMBOL_TO_Z[s] for s in arrays["species"]], dtype=np.int32)
        info = {}
        for k, v in kv.items():
            if k in ("Lattice", "Properties", "pbc"):
                continue
            try:
                info[k] = float(v)
            except ValueError:
                info[k] = v
        frames.append(Atoms(numbers, arrays["pos"], lat, pbc, info,
                            {k: v for k, v in arrays.items() if k not in ("pos", "species")}))
        i += 2 + n
    if isinstance(index, int):
        return frames[index]
    return frames


def _info_value(v):
    """extxyz value: arrays as quoted numbers in column-major order (QUIP's reshape(m, (/n/)))."""
    a = np.asarray(v)
    if a.ndim == 0:
        return str(v)
    if a.dtype.kind == "b":
        return '"%s"' % " ".join("T" if x else "F" for x in a.ravel(order="F"))
    return '"%s"' % " ".join("%.16g" % x for x in a.ravel(order="F"))


def write_xyz(path, frames, mode="w", real_format="%18.10f"):
    """Write frames as extended XYZ with species, pos and any float per-atom arrays.
    `path` is a file name or an open text stream; `real_format` formats per-atom reals."""
    if isinstance(frames, Atoms):
        frames = [frames]
    if hasattr(path, "write"):
        _write_frames(path, frames, real_format)
        return
    with open(path, mode) as fh:
        _write_frames(fh, frames, real_format)


def _write_frames(fh, frames, rf):
    for at in frames:
        extra = [(k, v) for k, v in at.arrays.items() if np.asarray(v).dtype.kind == "f"]
        props = "species:S:1:pos:R:3" + "".join(
            ":%s:R:%d" % (k, 1 if np.ndim(v) == 1 else np.shape(v)[1]) for k, v in extra)
        lat = " ".join("%.10f" % x for x in at.cell.ravel())
        info = " ".join("%s=%s" % (k, _info_value(v)) for k, v in at.info.items()
                        if not isinstance(v, str) or " " not in v)
        fh.write("%d\n" % len(at))
        fh.write('Lattice="%s" Properties=%s %s pbc="%s"\n' % (
            lat, props, info, " ".join("T" if p else "F" for p in at.pbc)))
        for a in range(len(at)):
            row = "%-2s" % _SYMBOLS[at.numbers[a]] + "".join(" " + rf % x for x in at.positions[a])
            for k, v in extra:
                vv = np.atleast_1d(np.asarray(v)[a])
                row += "".join(" " + rf % x for x in vv)
            fh.write(row + "\n")