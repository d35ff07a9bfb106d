"""Format-dispatching read/write of structures and volumetric data, the
interface of formats/io.py:9-197 of the reference (``read``, ``read_all``,
``read_density``, ``read_potential``, ``write``, ``write_density`` ...).

Volumetric data come back as a DirectField on ``device`` (the GPU by
default) in ``dtype`` (FP64 by default); the files themselves are parsed on
the host.  Formats: snpy, xsf, qepp (alias pp), vasp (alias poscar,
contcar).  Structure-only formats return the ions for ``kind='all'``.
"""
import os
from collections import namedtuple

import numpy as np
import torch

from ..field import DirectField
from ..grid import DirectGrid
from ..ions import Ions
from . import qepp, snpy, vasp, xsf

IOFormat = namedtuple("IOFormat", ["format", "read", "write", "kind"])


def _read_snpy(path, kind="all", **kw):
    return snpy.read_snpy(path, kind=kind)


def _read_xsf(path, kind="all", units="angstrom", data_type="density", pbc=True, **kw):
    ions, values, lattice = xsf.read_xsf_all(path, units=units, data_type=data_type, pbc=pbc, kind=kind)
    return ions, values, lattice, None


def _read_qepp(path, kind="all", data_type="density", **kw):
    return qepp.read_qepp_all(path, data_type=data_type)


def _read_vasp(path, names=None, **kw):
    return vasp.read_poscar(path, names=names), None, None, None


def _write_snpy(path, ions, data, information=None, **kw):
    snpy.write_snpy(path, ions, data, information=information)


def _write_xsf(path, ions, data, data_type="density", units="angstrom", title="DFTpy", **kw):
    xsf.write_xsf(path, ions, data, data_type=data_type, units=units, title=title)


def _write_qepp(path, ions, data, data_type="density", header=None, information=None, **kw):
    qepp.write_qepp(path, ions, data, data_type=data_type, header=header, information=information)


def _write_vasp(path, ions, data=None, direct=False, header="DFTpy", **kw):
    vasp.write_vasp(path, ions, direct=direct, header=header)


IOFormats = {
    "snpy": IOFormat("snpy", _read_snpy, _write_snpy, ("ions", "data", "all")),
    "xsf": IOFormat("xsf", _read_xsf, _write_xsf, ("ions", "data", "all")),
    "qepp": IOFormat("qepp", _read_qepp, _write_qepp, ("ions", "data", "all")),
    "vasp": IOFormat("vasp", _read_vasp, _write_vasp, ("ions",)),
}
IOFormats["pp"] = IOFormats["qepp"]
IOFormats["poscar"] = IOFormats["vasp"]
IOFormats["contcar"] = IOFormats["vasp"]


def guess_format(infile):
    """Format from the file name (``POSCAR``, ``CONTCAR``) or its extension;
    a trailing ``.gz`` is ignored (formats/io.py:35-46)."""
    name = os.path.basename(str(infile)).lower()
    if name.endswith(".gz"):
        name = name[:-3]
    if name in IOFormats:
        return IOFormats[name].format
    ext = os.path.splitext(name)[1][1:]
    if ext in IOFormats:
        return IOFormats[ext].format
    return ext or name


guessType = guess_format


def get_io_driver(infile, format=None):
    fmt = format or guess_format(infile)
    if fmt not in IOFormats:
        raise ValueError("%s format is not supported (%s)" % (fmt, ", ".join(sorted(IOFormats))))
    return IOFormats[fmt]


def _to_field(values, lattice, grid, full, dtype, device, ecut):
    if grid is None:
        nr = values.shape[-3:]
        grid = DirectGrid(lattice, nr, full=full, ecut=ecut, dtype=dtype, device=device)
    elif tuple(grid.nrR) != tuple(values.shape[-3:]):
        raise ValueError("data of shape %s do not fit the grid %s" % (values.shape, tuple(grid.nrR)))
    t = torch.as_tensor(values, dtype=grid.dtype, device=grid.device)
    return DirectField(grid, data=t)


def read(infile, format=None, kind="ions", ecut=None, grid=None, full=False, dtype=torch.float64, device=None,
         **kwargs):
    """kind 'ions' -> Ions; 'data' -> DirectField; 'all' -> (ions, field,
    information).  ``data_type='potential'`` converts stored potentials to
    Ha; xsf takes ``units`` ('angstrom' or 'bohr'), vasp ``names``."""
    drv = get_io_driver(infile, format)
    if kind == "data" and "data" not in drv.kind:
        raise ValueError("%s files hold no volumetric data" % drv.format)
    out = drv.read(infile, kind=kind, **kwargs)
    if drv.format == "snpy":
        ions, values, info = out
        lattice = ions.cell if ions is not None else None
    else:
        ions, values, lattice, info = out
    if kind == "ions":
        return ions
    field = None
    if values is not None:
        field = _to_field(values, lattice, grid, full, dtype, device, ecut)
    if kind == "data":
        return field
    return ions, field, info


def read_data(infile, format=None, **kwargs):
    return read(infile, format=format, kind="data", **kwargs)


def read_all(infile, format=None, **kwargs):
    return read(infile, format=format, kind="all", **kwargs)


def read_density(infile, format=None, **kwargs):
    kwargs.pop("data_type", None)
    return read(infile, format=format, kind="data", data_type="density", **kwargs)


def read_potential(infile, format=None, **kwargs):
    kwargs.pop("data_type", None)
    return read(infile, format=format, kind="data", data_type="potential", **kwargs)


def write(outfile, ions=None, data=None, information=None, format=None, **kwargs):
    """Write ions and/or a field; the arguments may come in either order
    (formats/io.py:138-172)."""
    if isinstance(data, Ions) and (ions is None or hasattr(ions, "grid") or isinstance(ions, np.ndarray)):
        ions, data = data, ions
    drv = get_io_driver(outfile, format)
    if "data" in drv.kind and data is None and drv.format != "snpy":
        raise ValueError("%s files need volumetric data" % drv.format)
    drv.write(outfile, ions, data, information=information, **kwargs)


def write_all(outfile, ions=None, data=None, information=None, format=None, **kwargs):
    return write(outfile, ions=ions, data=data, information=information, format=format, **kwargs)


def write_density(outfile, data=None, ions=None, format=None, **kwargs):
    kwargs.pop("data_type", None)
    return write(outfile, ions=ions, data=data, format=format, data_type="density", **kwargs)


def write_potential(outfile, data=None, ions=None, format=None, **kwargs):
    kwargs.pop("data_type", None)
    return write(outfile, ions=ions, data=data, format=format, data_type="potential", **kwargs)
