"""Shared helpers of the file formats: gzip-transparent open and the
conversion of a field (DirectField or array) to host arrays."""
import gzip

import numpy as np


def open_text(path, mode="r"):
    """Open a text file; a ``.gz`` name is read or written through gzip."""
    path = str(path)
    if path.endswith(".gz"):
        return gzip.open(path, mode + "t")
    return open(path, mode)


def open_binary(path, mode="rb"):
    path = str(path)
    if path.endswith(".gz"):
        return gzip.open(path, mode)
    return open(path, mode)


def field_values(data):
    """Host FP64 copy of a field: (values, lattice or None).  ``values`` is
    (nx, ny, nz) for one component or (rank, nx, ny, nz) for several."""
    lattice = None
    if hasattr(data, "grid"):
        lattice = np.asarray(data.grid.lattice, dtype=np.float64)
        data = data.numpy() if hasattr(data, "numpy") else np.asarray(data)
    values = np.asarray(data, dtype=np.float64)
    if values.ndim not in (3, 4):
        raise ValueError("volumetric data must be 3D (or 4D with a leading component axis), got %s" %
                         (values.shape,))
    return values, lattice


def components(values):
    """List of 3D component arrays of a 3D or 4D array."""
    return [values] if values.ndim == 3 else list(values)


def stack(blocks):
    return blocks[0] if len(blocks) == 1 else np.stack(blocks)
