"""File formats (formats/ of the reference): QE pp, xsf, POSCAR, snpy
volumetric/structure files and CASTEP recpot pseudopotentials, plus the
format-dispatching ``io`` interface.  Files may be gzip-compressed (``.gz``).

The per-format readers return host arrays: ``(ions, values, lattice)`` with
values in C order (nx, ny, nz) and lengths in bohr; ``io.read_*`` wraps them
as DirectFields."""
from .qepp import read_qepp, write_qepp
from .recpot import read_recpot, write_recpot
from .snpy import read_snpy, write_snpy
from .vasp import read_poscar, read_vasp, write_vasp
from .xsf import read_xsf, write_xsf
from . import io  # noqa: E402  (after the format modules it dispatches to)

__all__ = ["read_qepp", "write_qepp", "read_recpot", "write_recpot", "read_snpy", "write_snpy", "read_poscar",
           "read_vasp", "write_vasp", "read_xsf", "write_xsf", "io"]
