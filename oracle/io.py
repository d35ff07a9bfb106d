"""File readers for the fixtures the reference's tests use.

* QE pp density (``formats/qepp.py:17-117``): grid shape, ibrav=0 lattice in
  alat units, atoms, then the data in Fortran order.
* POSCAR (``formats/vasp.py:4-57``): lattice/positions in Angstrom.
* XSF (``formats/xsf.py:15-141``): Angstrom lattice/positions, a general
  periodic datagrid whose last point per axis is dropped, density scaled by
  Bohr^3.
* recpot (``functional/pseudo/recpot.py:11-62``): v(q) in eV*A^3 on a
  uniform q grid in 1/A, converted to Ha*bohr^3 and 1/bohr.
"""
import gzip

import numpy as np

from . import units


def _open(path):
    path = str(path)
    return gzip.open(path, "rt") if path.endswith(".gz") else open(path)


class Atoms:
    """Minimal ions container in bohr (``ions.py:5-156``)."""

    def __init__(self, symbols, positions, cell, charges=None):
        self.symbols = list(symbols)
        self.positions = np.asarray(positions, dtype=np.float64).reshape(-1, 3)
        self.cell = np.asarray(cell, dtype=np.float64).reshape(3, 3)
        self.charges = None if charges is None else np.asarray(charges, dtype=np.float64)

    @property
    def nat(self):
        return len(self.symbols)

    def scaled_positions(self):
        # The reference's Ions are built without pbc, so ase's
        # get_scaled_positions() returns the unwrapped solve of cell^T s = r.
        return np.linalg.solve(self.cell.T, self.positions.T).T

    def set_charges(self, zval):
        self.charges = np.array([zval[s] for s in self.symbols], dtype=np.float64)

    def ncharge(self):
        return float(self.charges.sum())

    @property
    def volume(self):
        return abs(np.linalg.det(self.cell))


def read_qepp(path):
    """Returns (atoms, rho (nx,ny,nz) C-order array, lattice)."""
    with _open(path) as f:
        f.readline()
        vals = f.readline().split()
        nrx = [int(v) for v in vals[:3]]
        nat, ntyp = int(vals[6]), int(vals[7])
        line = f.readline().split()
        ibrav = int(line[0])
        celldm = np.asarray(line[1:], dtype=float)
        if ibrav != 0:
            raise NotImplementedError("only ibrav=0 pp files are read")
        at = np.array([np.asarray(f.readline().split(), dtype=float) for _ in range(3)]) * celldm[0]
        f.readline()
        zval, atm = {}, []
        for _ in range(ntyp):
            ls = f.readline().split()
            atm.append(ls[1])
            zval[ls[1]] = float(ls[2])
        symbols, pos = [], []
        for _ in range(nat):
            ls = f.readline().split()
            pos.append(np.asarray(ls[1:4], dtype=float) * celldm[0])
            symbols.append(atm[int(ls[4]) - 1])
        data = np.array(f.read().split(), dtype=float)
    rho = np.reshape(data, nrx, order="F").copy(order="C")
    atoms = Atoms(symbols, pos, at)
    atoms.set_charges(zval)
    return atoms, rho, at


def read_poscar(path, names=None):
    with _open(path) as f:
        f.readline()
        scale = np.asarray(f.readline().split(), dtype=float)
        if scale.size == 1:
            scale = np.ones(3) * scale[0]
        lat = np.array([np.asarray(f.readline().split(), dtype=float) for _ in range(3)])
        for i in range(3):
            lat[i] *= scale[i]
        ls = f.readline().split()
        if ls[0].isdigit():
            typ = [int(v) for v in ls]
        else:
            names = ls
            typ = [int(v) for v in f.readline().split()]
        fmt = f.readline().strip()[0].lower()
        nat = sum(typ)
        pos = np.array([np.asarray(f.readline().split()[:3], dtype=float) for _ in range(nat)])
    symbols = []
    for n, t in zip(names, typ):
        symbols += [n] * t
    if fmt in "ck":
        cart = pos
    else:
        cart = pos @ lat
    return Atoms(symbols, cart / units.Bohr, lat / units.Bohr)


def read_xsf(path):
    """Periodic CRYSTAL xsf with PRIMVEC/PRIMCOORD and one 3D datagrid."""
    with _open(path) as f:
        lines = [l.strip() for l in f if l.strip() and not l.strip().startswith("#")]
    it = iter(lines)
    assert next(it).upper() == "CRYSTAL"
    lattice = None
    symbols, pos = [], []
    line = next(it).upper()
    if line.startswith("PRIMVEC"):
        lattice = np.array([np.asarray(next(it).split(), dtype=float) for _ in range(3)])
        line = next(it).upper()
    if line.startswith("PRIMCOORD"):
        nat = int(next(it).split()[0])
        for _ in range(nat):
            ls = next(it).split()
            symbols.append(ls[0])
            pos.append(np.asarray(ls[1:4], dtype=float))
        line = next(it).upper()
    while not line.startswith("BEGIN_DATAGRID_3D"):
        line = next(it).upper()
    nrx = [int(v) for v in next(it).split()]
    next(it)  # origin
    vlat = np.array([np.asarray(next(it).split(), dtype=float) for _ in range(3)])
    data = []
    for line in it:
        if line[0] in "Ee":
            break
        data.extend(float(v) for v in line.split())
    data = np.asarray(data)[: np.prod(nrx)]
    data = np.reshape(data, nrx, order="F")
    data = data[: nrx[0] - 1, : nrx[1] - 1, : nrx[2] - 1] * units.Bohr ** 3
    atoms = Atoms(symbols, np.array(pos) / units.Bohr, lattice / units.Bohr)
    return atoms, np.ascontiguousarray(data), vlat / units.Bohr


def read_recpot(path):
    """Returns (q grid in 1/bohr, v(q) in Ha*bohr^3, zval)."""
    with _open(path) as f:
        lines = f.readlines()
    ibegin = 0
    iend = None
    for i, line in enumerate(lines):
        if "END COMMENT" in line:
            ibegin = i + 3
        elif ibegin > 1 and (line.strip() == "1000" or len(line.strip()) == 1):
            iend = i
            break
    gmax = float(lines[ibegin - 1].strip()) * units.Bohr
    v = np.array(" ".join(l.strip() for l in lines[ibegin:iend]).split(), dtype=np.float64)
    v = v / units.Ha / units.Bohr ** 3
    q = np.linspace(0, gmax, num=len(v))
    zval = round((v[0] - v[1]) * (q[1] ** 2) / (4.0 * np.pi))
    return q, v, zval


def write_recpot(path, q, v, header="DFTpy"):
    """Writer following ``functional/pseudo/recpot.py:44-62``."""
    gmax = q[-1] / units.Bohr
    vv = v * units.Ha * units.Bohr ** 3
    with open(path, "w") as fw:
        fw.write("START COMMENT\n" + header + "\nEND COMMENT\n")
        fw.write("3     5\n")
        fw.write(str(gmax) + "\n")
        nl = vv.size // 3
        for line in vv[: nl * 3].reshape(-1, 3):
            fw.write("     {0[0]:22.15E}     {0[1]:22.15E}     {0[2]:22.15E}\n".format(line))
        for x in vv[nl * 3:]:
            fw.write("     {0:22.15E}".format(x))
        if len(vv[nl * 3:]) > 0:
            fw.write("\n")
        fw.write("1000\n")
