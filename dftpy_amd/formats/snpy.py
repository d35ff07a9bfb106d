"""``.snpy``: a sequence of NumPy ``.npy`` records, the format of
formats/snpy.py:1-144 of the reference.  The first record is the description
(integers naming the records that follow, in order): 1 the cell (3x3, bohr),
2 the atomic numbers, 3 the positions (N x 3, bohr), 4 the volumetric data
((nx, ny, nz) or (rank, nx, ny, nz), C order), 5 extra information (an
array).  Records are read with pickling disabled, so object arrays are
rejected.
"""
import numpy as np

from ..ions import Ions
from ._util import field_values, open_binary


def write_snpy(path, ions=None, data=None, information=None, kind="all", desc=None):
    if desc is None:
        if kind == "ions" or data is None:
            desc = [1, 2, 3]
        elif information is not None:
            desc = [1, 2, 3, 4, 5]
        else:
            desc = [1, 2, 3, 4]
    desc = np.asarray(desc, dtype=np.int64)
    with open_binary(path, "wb") as f:
        np.save(f, desc, allow_pickle=False)
        for key in desc:
            if key == 1:
                np.save(f, np.asarray(ions.cell, dtype=np.float64), allow_pickle=False)
            elif key == 2:
                np.save(f, ions.numbers, allow_pickle=False)
            elif key == 3:
                np.save(f, np.asarray(ions.positions, dtype=np.float64), allow_pickle=False)
            elif key == 4:
                np.save(f, np.ascontiguousarray(field_values(data)[0]), allow_pickle=False)
            elif key == 5:
                np.save(f, np.asarray(information), allow_pickle=False)
            else:
                raise ValueError("unknown snpy record %d" % key)


def read_snpy(path, kind="all"):
    """Returns (ions, values, information); values and information are None
    when the file holds no such record."""
    ions = values = information = None
    cell = numbers = None
    with open_binary(path, "rb") as f:
        desc = np.load(f, allow_pickle=False)
        for key in desc:
            if key == 1:
                cell = np.load(f, allow_pickle=False)
            elif key == 2:
                numbers = np.load(f, allow_pickle=False)
            elif key == 3:
                ions = Ions(numbers=numbers, positions=np.load(f, allow_pickle=False), cell=cell)
                if kind == "ions":
                    break
            elif key == 4:
                values = np.ascontiguousarray(np.load(f, allow_pickle=False))
            elif key == 5:
                information = np.load(f, allow_pickle=False)
            else:
                raise ValueError("unknown snpy record %d" % key)
    return ions, values, information
