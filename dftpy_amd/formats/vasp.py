"""VASP POSCAR/CONTCAR structures (formats/vasp.py:4-103 of the reference);
lengths in Angstrom on disk, bohr in memory."""
import numpy as np

from ..constants import Units
from ..ions import Ions
from ._util import open_text


def read_poscar(path, names=None):
    """POSCAR with one or three scale factors, species names on their own
    line (VASP 5) or given by ``names``, Direct or Cartesian positions."""
    with open_text(path) as f:
        f.readline()
        scale = np.array(f.readline().split(), dtype=np.float64)
        scale = np.ones(3) * scale[0] if scale.size == 1 else scale
        lat = np.array([[float(x) for x in f.readline().split()] for _ in range(3)]) * scale[:, None]
        ls = f.readline().split()
        if ls[0].isdigit():
            counts = [int(x) for x in ls]
        else:
            names = ls
            counts = [int(x) for x in f.readline().split()]
        if names is None:
            raise ValueError("the POSCAR has no species line: pass names")
        kind = f.readline().strip()[0].lower()
        if kind == "s":  # selective dynamics
            kind = f.readline().strip()[0].lower()
        pos = np.array([[float(x) for x in f.readline().split()[:3]] for _ in range(sum(counts))])
    syms = [n for n, c in zip(names, counts) for _ in range(c)]
    cart = pos if kind in "ck" else pos @ lat
    return Ions(symbols=syms, positions=cart, cell=lat, units="ase")


read_vasp = read_poscar


def write_vasp(path, ions, direct=False, fmt="%22.15f", header="DFTpy"):
    """POSCAR (VASP 5) with species sorted by name, as formats/vasp.py:60-103
    writes it: scale 1.0, Angstrom cell, Cartesian or Direct positions."""
    cell = ions.cell * Units.Bohr
    pos = ions.get_scaled_positions() if direct else ions.positions * Units.Bohr
    names, idx, counts = np.unique(np.asarray(ions.symbols), return_inverse=True, return_counts=True)
    order = np.argsort(idx, kind="stable")
    with open_text(path, "w") as f:
        f.write(header + "\n1.0\n")
        for a in cell:
            f.write("".join(fmt % x for x in a) + "\n")
        f.write("".join(" %-3s" % n for n in names) + "\n")
        f.write("".join(" %-3d" % c for c in counts) + "\n")
        f.write("Direct\n" if direct else "Cartesian\n")
        for p in pos[order]:
            f.write("%s %s %s\n" % (fmt % p[0], fmt % p[1], fmt % p[2]))
