"""Readers for the inputs the hot path consumes (formats/qepp.py:17-117,
formats/vasp.py:4-57, formats/xsf.py:15-141, functional/pseudo/recpot.py
of the reference).  Files may be gzip-compressed (``.gz``)."""
import gzip

import numpy as np

from .constants import Units
from .ions import Ions


def _open(path):
    path = str(path)
    return gzip.open(path, "rt") if path.endswith(".gz") else open(path)


def read_qepp(path):
    """QE pp density file (ibrav = 0).  Returns (ions, rho array C-order, lattice)."""
    with _open(path) as f:
        f.readline()
        v = f.readline().split()
        nrx = [int(x) for x in v[:3]]
        nat, ntyp = int(v[6]), int(v[7])
        line = f.readline().split()
        if int(line[0]) != 0:
            raise NotImplementedError("only ibrav = 0 pp files are supported")
        alat = float(line[1])
        at = np.array([[float(x) for x in f.readline().split()] for _ in range(3)]) * alat
        f.readline()
        names, zval = [], {}
        for _ in range(ntyp):
            ls = f.readline().split()
            names.append(ls[1])
            zval[ls[1]] = float(ls[2])
        syms, pos = [], []
        for _ in range(nat):
            ls = f.readline().split()
            pos.append([float(x) * alat for x in ls[1:4]])
            syms.append(names[int(ls[4]) - 1])
        data = np.array(f.read().split(), dtype=np.float64)
    rho = np.ascontiguousarray(data.reshape(nrx, order="F"))
    ions = Ions(symbols=syms, positions=pos, cell=at)
    ions.set_charges(zval)
    return ions, rho, at


def read_poscar(path, names=None):
    with _open(path) as f:
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
        kind = f.readline().strip()[0].lower()
        pos = np.array([[float(x) for x in f.readline().split()[:3]] for _ in range(sum(counts))])
    syms = [n for n, c in zip(names, counts) for _ in range(c)]
    cart = pos if kind in "ck" else pos @ lat
    return Ions(symbols=syms, positions=cart, cell=lat, units="ase")


def read_xsf(path):
    """CRYSTAL xsf with one 3D datagrid (periodic end points dropped,
    1/A^3 -> 1/bohr^3).  Returns (ions, data, lattice)."""
    with _open(path) as f:
        lines = [l.strip() for l in f if l.strip() and not l.strip().startswith("#")]
    it = iter(lines)
    if next(it).upper() != "CRYSTAL":
        raise ValueError("only CRYSTAL xsf is supported")
    line = next(it).upper()
    lattice = None
    syms, pos = [], []
    if line.startswith("PRIMVEC"):
        lattice = np.array([[float(x) for x in next(it).split()] for _ in range(3)])
        line = next(it).upper()
    if line.startswith("PRIMCOORD"):
        nat = int(next(it).split()[0])
        for _ in range(nat):
            ls = next(it).split()
            syms.append(ls[0])
            pos.append([float(x) for x in ls[1:4]])
        line = next(it).upper()
    while not line.startswith("BEGIN_DATAGRID_3D"):
        line = next(it).upper()
    nrx = [int(x) for x in next(it).split()]
    next(it)
    vlat = np.array([[float(x) for x in next(it).split()] for _ in range(3)])
    vals = []
    for line in it:
        if line[0] in "Ee":
            break
        vals.extend(float(x) for x in line.split())
    data = np.asarray(vals)[: int(np.prod(nrx))].reshape(nrx, order="F")
    data = data[: nrx[0] - 1, : nrx[1] - 1, : nrx[2] - 1] * Units.Bohr ** 3
    ions = Ions(symbols=syms, positions=pos, cell=lattice, units="ase")
    return ions, np.ascontiguousarray(data), vlat / Units.Bohr


def read_recpot(path):
    """CASTEP recpot: returns (q in 1/bohr, v(q) in Ha bohr^3, zval)
    (functional/pseudo/recpot.py:11-42)."""
    with _open(path) as f:
        lines = f.readlines()
    ibegin, iend = 0, None
    for i, line in enumerate(lines):
        if "END COMMENT" in line:
            ibegin = i + 3
        elif ibegin > 1 and (line.strip() == "1000" or len(line.strip()) == 1):
            iend = i
            break
    if iend is None:
        raise ValueError("recpot file %s has no terminating '1000' line" % path)
    gmax = float(lines[ibegin - 1].strip()) * Units.Bohr
    v = np.array(" ".join(l.strip() for l in lines[ibegin:iend]).split(), dtype=np.float64)
    v = v / Units.Ha / Units.Bohr ** 3
    q = np.linspace(0, gmax, num=len(v))
    zval = round((v[0] - v[1]) * (q[1] ** 2) / (4.0 * np.pi))
    return q, v, zval


def write_recpot(path, q, v, header="dftpy_amd"):
    """recpot writer in the layout of functional/pseudo/recpot.py:44-62."""
    gmax = q[-1] / Units.Bohr
    vv = np.asarray(v) * Units.Ha * Units.Bohr ** 3
    with open(path, "w") as fw:
        fw.write("START COMMENT\n" + header + "\nEND COMMENT\n3     5\n" + str(gmax) + "\n")
        n3 = vv.size // 3
        for row in vv[: n3 * 3].reshape(-1, 3):
            fw.write("     {0[0]:22.15E}     {0[1]:22.15E}     {0[2]:22.15E}\n".format(row))
        rest = vv[n3 * 3:]
        for x in rest:
            # separated: the reference's writer fuses negative leftovers
            fw.write("     {0:22.15E}".format(x))
        if rest.size:
            fw.write("\n")
        fw.write("1000\n")
