"""Synthetic workloads: fcc Al supercells and the Ashcroft empty-core
local pseudopotential written as a recpot file.

* empty core: v(q) = -4 pi Z cos(q rc) / q^2 on q in [0, 40] 1/bohr, step
  0.01; v(0) holds the regular limit 2 pi Z rc^2 of v + 4 pi Z / q^2, so the
  reader's valence estimate round((v0 - v1) dq^2 / 4 pi) gives Z.
* fcc Al: conventional cubic cell a = 7.60 bohr, 4 atoms; n1 x n2 x n3
  supercells, isotropic strain and Gaussian displacements from a seed.
"""
import os
import tempfile

import numpy as np

from .formats import write_recpot
from .ions import Ions

FCC_BASIS = np.array([[0.0, 0.0, 0.0], [0.5, 0.5, 0.0], [0.5, 0.0, 0.5], [0.0, 0.5, 0.5]])
A_AL = 7.60


def empty_core_table(Z=3.0, rc=1.12, qmax=40.0, dq=0.01):
    n = int(round(qmax / dq)) + 1
    q = np.linspace(0.0, qmax, n)
    v = np.empty(n)
    v[1:] = -4.0 * np.pi * Z * np.cos(q[1:] * rc) / q[1:] ** 2
    v[0] = 2.0 * np.pi * Z * rc ** 2
    return q, v


def empty_core_recpot(path=None, Z=3.0, rc=1.12):
    if path is None:
        d = os.path.join(tempfile.gettempdir(), "dftpy_amd")
        os.makedirs(d, exist_ok=True)
        path = os.path.join(d, "al_emptycore_Z%g_rc%g.recpot" % (Z, rc))
    if not os.path.exists(path):
        q, v = empty_core_table(Z, rc)
        tmp = path + ".%d.tmp" % os.getpid()
        write_recpot(tmp, q, v, header="Ashcroft empty core Z=%g rc=%g bohr" % (Z, rc))
        os.replace(tmp, path)
    return path


def fcc_al(reps=(1, 1, 1), a=A_AL, sigma=0.0, strain=0.0, seed=None):
    """fcc Al supercell; positions/cell in bohr."""
    reps = tuple(int(r) for r in reps)
    cells = np.array([[i, j, k] for i in range(reps[0]) for j in range(reps[1]) for k in range(reps[2])], float)
    frac = (cells[:, None, :] + FCC_BASIS[None, :, :]).reshape(-1, 3) / np.array(reps, float)
    L = a * (1.0 + strain)
    cell = np.diag([L * reps[0], L * reps[1], L * reps[2]])
    pos = frac @ cell
    if sigma > 0:
        rng = np.random.default_rng(seed)
        pos = pos + rng.normal(0.0, sigma, pos.shape)
    return Ions(symbols=["Al"] * len(pos), positions=pos, cell=cell)


def c1_system():
    """4-atom fcc Al, a = 7.60 bohr, 32^3 grid."""
    return fcc_al((1, 1, 1)), (32, 32, 32)


def c2_system(seed=2024):
    """500-atom fcc Al 5x5x5 (38.0 bohr), sigma = 0.1 bohr displacements, 160^3."""
    return fcc_al((5, 5, 5), sigma=0.1, seed=seed), (160, 160, 160)
