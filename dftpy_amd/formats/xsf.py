"""XCrySDen structure file with 3D datagrids (``.xsf``), the subset read and
written by formats/xsf.py:15-236 of the reference: a CRYSTAL with PRIMVEC,
PRIMCOORD and one BEGIN_BLOCK_DATAGRID_3D holding one datagrid per
component.  Datagrids are periodic (the end points repeat the first plane);
lengths are Angstrom, densities 1/A^3 and potentials eV unless
``units='bohr'``.
"""
import numpy as np

from ..constants import Units
from ..ions import Ions
from ._util import components, field_values, open_text, stack


def _lines(f):
    for line in f:
        line = line.strip()
        if line and not line.startswith("#"):
            yield line


def _floats(line):
    return [float(x) for x in line.split()]


def read_xsf_all(path, units="angstrom", data_type="density", pbc=True, kind="all"):
    """(ions, values, lattice): values C-order (nx, ny, nz), or (rank, nx, ny,
    nz) for several datagrids; lattice (bohr) spans the datagrid."""
    ang = units.lower() == "angstrom"
    with open_text(path) as f:
        it = _lines(f)
        if next(it).upper() != "CRYSTAL":
            raise ValueError("only CRYSTAL xsf files are supported")
        line = next(it).upper()
        cell = None
        syms, pos = [], []
        if line.startswith("PRIMVEC"):
            cell = np.array([_floats(next(it)) for _ in range(3)])
            line = next(it).upper()
        if line.startswith("CONVVEC"):
            for _ in range(3):
                next(it)
            line = next(it).upper()
        if line.startswith("PRIMCOORD"):
            nat = int(next(it).split()[0])
            for _ in range(nat):
                ls = next(it).split()
                syms.append(ls[0])
                pos.append([float(x) for x in ls[1:4]])
            line = next(it, "").upper()
        ions = Ions(symbols=syms, positions=np.asarray(pos).reshape(-1, 3), cell=cell,
                    units="ase" if ang else "au")
        if kind == "ions":
            return ions, None, None
        blocks, vlat = [], None
        while line and not line.startswith("END_BLOCK_DATAGRID"):
            if line.startswith("BEGIN_DATAGRID_3D"):
                nrx = [int(x) for x in next(it).split()]
                next(it)  # origin
                v = np.array([_floats(next(it)) for _ in range(3)])
                if vlat is None:
                    vlat = v
                n = int(np.prod(nrx))
                vals = []
                for line in it:
                    if line.upper().startswith("END_DATAGRID"):
                        break
                    vals.extend(_floats(line))
                if len(vals) < n:
                    raise ValueError("datagrid %d of %s is short: %d of %d values" % (len(blocks), path, len(vals), n))
                blocks.append(np.asarray(vals[:n]).reshape(nrx, order="F"))
            elif line.startswith("BEGIN_DATAGRID_"):
                raise NotImplementedError("only 3D datagrids are supported")
            line = next(it, "").upper()
    if not blocks:
        raise ValueError("%s has no 3D datagrid" % path)
    if any(b.shape != blocks[0].shape for b in blocks):
        raise ValueError("all datagrids must have the same shape")
    out = []
    for b in blocks:
        if pbc:
            b = b[tuple(slice(0, max(1, s - 1)) for s in b.shape)]
        b = np.ascontiguousarray(b)
        if ang:
            b = b * Units.Bohr ** 3
        if data_type == "potential":
            b = b / Units.Ha
        out.append(b)
    lattice = vlat / Units.Bohr if ang else vlat
    return ions, stack(out), lattice


def read_xsf(path, units="angstrom", data_type="density"):
    """Returns (ions, values C-order, lattice in bohr)."""
    return read_xsf_all(path, units=units, data_type=data_type)


def write_xsf(path, ions, data, data_type="density", units="angstrom", title="DFTpy"):
    """Write ions and a field (DirectField or array, 3D or (rank, 3D)) as xsf
    (formats/xsf.py:150-236)."""
    values, lattice = field_values(data)
    if lattice is None:
        lattice = ions.cell
    ang = units.lower() == "angstrom"
    lscale = Units.Bohr if ang else 1.0
    head = ["# " + title, "CRYSTAL", "PRIMVEC"]
    head += [" ".join(repr(float(x)) for x in a) for a in ions.cell * lscale]
    head += ["PRIMCOORD", "%d 1" % ions.nat]
    for s, p in zip(ions.symbols, ions.positions * lscale):
        head.append("%s %s %s %s" % (s, repr(float(p[0])), repr(float(p[1])), repr(float(p[2]))))
    head += ["BEGIN_BLOCK_DATAGRID_3D", "3d_datagrid_" + data_type]
    with open_text(path, "w") as f:
        f.write("\n".join(head) + "\n")
        for i, comp in enumerate(components(values)):
            v = np.pad(comp, [(0, 1)] * 3, mode="wrap")
            if ang:
                v = v / Units.Bohr ** 3
            if data_type == "potential":
                v = v * Units.Ha
            f.write("BEGIN_DATAGRID_3D#%d\n" % i)
            f.write("%d %d %d\n" % v.shape)
            f.write("0.0 0.0 0.0\n")
            for a in np.asarray(lattice) * lscale:
                f.write(" ".join(repr(float(x)) for x in a) + "\n")
            flat = v.ravel(order="F")
            nfull = flat.size // 5
            if nfull:
                np.savetxt(f, flat[: nfull * 5].reshape(nfull, 5), fmt="%.16e", delimiter=" ")
            if flat.size > nfull * 5:
                f.write(" ".join("%.16e" % x for x in flat[nfull * 5:]) + "\n")
            f.write("END_DATAGRID_3D\n")
        f.write("END_BLOCK_DATAGRID_3D\n")
