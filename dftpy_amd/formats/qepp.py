"""Quantum ESPRESSO post-processing (``pp.x``) volumetric file, the layout
read and written by formats/qepp.py:17-209 of the reference.

Header: title; ``nr1x nr2x nr3x nr1 nr2 nr3 nat ntyp``; ``ibrav celldm(1:6)``;
for ibrav = 0 the three lattice vectors in units of celldm(1);
``gcutm dual ecut plot_num``; one line per species ``ityp symbol zval``; one
line per atom ``iat tau(1:3) ityp`` with tau in units of celldm(1).  Then the
values in Fortran order, five per line.  Densities are e/bohr^3; potentials
are stored in Ry (read as Ha).
"""
import numpy as np

from ..ions import Ions
from ._util import field_values, open_text

DEFAULT_CUTOFFS = {"gcutm": 1.0, "dual": 4.0, "ecut": 1.0}


def celldm_lattice(ibrav, celldm):
    """Lattice (rows a1, a2, a3 in bohr) of the Bravais lattices the reader
    accepts (formats/qepp.py:120-135): ibrav 1 (simple cubic) and 2 (fcc,
    kept in the reference's column convention)."""
    a = float(celldm[0])
    if ibrav == 1:
        return a * np.eye(3)
    if ibrav == 2:
        cols = 0.5 * a * np.array([[-1.0, 0.0, 1.0], [0.0, 1.0, 1.0], [-1.0, 1.0, 0.0]])
        return cols.T.copy()
    raise NotImplementedError("ibrav = %d is not supported (0, 1 and 2 are)" % ibrav)


def read_qepp_all(path, data_type="density"):
    """(ions, values C-order (nx, ny, nz), lattice, info) where ``info`` holds
    ibrav, celldm, gcutm, dual, ecut, plot_num and the title."""
    with open_text(path) as f:
        title = f.readline().rstrip("\n")
        v = f.readline().split()
        nrx = [int(x) for x in v[:3]]
        nat, ntyp = int(v[6]), int(v[7])
        line = f.readline().split()
        ibrav = int(line[0])
        celldm = np.array([float(x) for x in line[1:]], dtype=np.float64)
        if ibrav == 0:
            lattice = np.array([[float(x) for x in f.readline().split()] for _ in range(3)]) * celldm[0]
        else:
            lattice = celldm_lattice(ibrav, celldm)
        gcutm, dual, ecut, plot_num = (float(x) for x in f.readline().split())
        names, zval = [], {}
        for _ in range(ntyp):
            ls = f.readline().split()
            names.append(ls[1])
            zval[ls[1]] = float(ls[2])
        syms, tau = [], []
        for _ in range(nat):
            ls = f.readline().split()
            tau.append([float(x) for x in ls[1:4]])
            syms.append(names[int(ls[4]) - 1])
        tau = np.asarray(tau, dtype=np.float64).reshape(-1, 3)
        values = np.array(f.read().split(), dtype=np.float64)
    n = int(np.prod(nrx))
    if values.size < n:
        raise ValueError("%s holds %d values, the grid %s needs %d" % (path, values.size, nrx, n))
    rho = np.ascontiguousarray(values[:n].reshape(nrx, order="F"))
    if data_type == "potential":
        rho *= 0.5
    if ibrav == 0:
        ions = Ions(symbols=syms, positions=tau * celldm[0], cell=lattice)
    else:
        ions = Ions(symbols=syms, scaled_positions=tau, cell=lattice)
    ions.set_charges(zval)
    info = dict(title=title, ibrav=ibrav, celldm=celldm, gcutm=gcutm, dual=dual, ecut=ecut, plot_num=int(plot_num))
    return ions, rho, lattice, info


def read_qepp(path, data_type="density"):
    """QE pp file.  Returns (ions, values C-order (nx, ny, nz), lattice)."""
    return read_qepp_all(path, data_type)[:3]


def _zval(ions):
    try:
        return ions.zval
    except AttributeError:
        return {}


def write_qepp(path, ions, data, data_type="density", header=None, information=None):
    """Write ions and one 3D field (DirectField or array) as a QE pp file
    (formats/qepp.py:137-209).  Potentials (Ha) are stored in Ry."""
    values, _ = field_values(data)
    if values.ndim != 3:
        raise ValueError("a pp file holds one 3D field")
    info = dict(DEFAULT_CUTOFFS)
    if information:
        info.update(information)
    ibrav = int(info.get("ibrav", 0))
    celldm = info.get("celldm")
    if celldm is None:
        celldm = np.zeros(6)
        celldm[0] = np.linalg.norm(ions.cell[0])
    celldm = np.asarray(celldm, dtype=np.float64)
    plot_num = info.get("plot_num", 1 if data_type == "potential" else 0)
    species = list(ions.symbols_uniq)
    zv = _zval(ions)
    nr = values.shape
    out = [header if header is not None else info.get("title", "DFTpy"),
           " ".join(str(x) for x in (*nr, *nr, ions.nat, len(species))),
           " ".join([str(ibrav)] + [repr(float(x)) for x in celldm])]
    if ibrav == 0:
        for a in ions.cell / celldm[0]:
            out.append(" ".join(repr(float(x)) for x in a))
    out.append(" ".join([repr(float(info["gcutm"])), repr(float(info["dual"])), repr(float(info["ecut"])),
                         str(int(plot_num))]))
    for i, s in enumerate(species):
        out.append("%d %s %s" % (i + 1, s, repr(float(zv.get(s, 0.0)))))
    tau = ions.positions / celldm[0]
    for i in range(ions.nat):
        out.append("%d %s %s %s %d" % (i + 1, repr(float(tau[i, 0])), repr(float(tau[i, 1])),
                                       repr(float(tau[i, 2])), species.index(ions.symbols[i]) + 1))
    flat = values.ravel(order="F")
    if data_type == "potential":
        flat = flat * 2.0
    nfull = flat.size // 5
    with open_text(path, "w") as f:
        f.write("\n".join(out) + "\n")
        if nfull:
            np.savetxt(f, flat[: nfull * 5].reshape(nfull, 5), fmt="%22.15e", delimiter=" ")
        if flat.size > nfull * 5:
            f.write(" ".join("%22.15e" % x for x in flat[nfull * 5:]) + "\n")
