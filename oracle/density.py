"""Start densities from atomic radial functions, FP64 numpy restatement of
``density/density.py`` (test infrastructure: used only by tests as the
checker).

* ``place_atoms``: the real-space placement of ``AtomicDensity.step``
  (density.py:18-67) as used by ``guess_rho_atom`` (:209-249) and
  ``build_pseudo_density`` (:365-406): for each atom, the box of grid points
  within ``border = min(int(rcut / spacing) + 1, nr // 2)`` of the point below
  it, displacements ``((i + px) / nr)`` mapped through the cell matrix as the
  reference's einsum writes it (``lattice[k, j] * s_j``), values added with
  numpy's fancy-index ``+=`` (a grid point that the box reaches twice keeps the
  later value).
* ``gen_gaussian_density`` (:353-363) and ``build_pseudo_density``: normalised
  3D Gaussians (``math_utils.gaussian``, :470-478), sigma at least
  ``2 max(spacing) / 2.354820``.
* ``guess_rho_atom``: cubic spline (FITPACK splrep) of r, rho(r) inside
  ``(r[0], rc)``, rc = min(rcut or max(r), L_min / 2), then scaled to the
  electron count.
* ``get_3d_value_recipe`` (:252-317): rho(G) = sum_s spline_s(|G|) S_s(G)
  (exact or PME), zero for |G| >= r[-1]; inverse FFT, values below dtol set to
  dtol, scaled to the electron count.

Checks (tests/test_oracle.py): a resolved Gaussian integrates to its scale and
peaks at its closed form; the recipe of the Fourier transform of a Gaussian is
the periodic sum of that Gaussian; PME and exact recipes agree.
"""
import numpy as np
from scipy.interpolate import splev, splrep

from .pseudo import barray, pme_qarray, strf

FWHM = 2.354820


def spacings(grid):
    return np.linalg.norm(grid.lattice, axis=1) / grid.nr


def gaussian(x, sigma):
    return 1.0 / (np.sqrt(2.0 * np.pi) * sigma) ** 3 * np.exp(-0.5 * (x / sigma) ** 2)


def box_values(grid, scaled_position, rcut, rtol, fn):
    """(l123A, values) of one atom's box (``AtomicDensity.step``)."""
    border = (rcut / spacings(grid)).astype(np.int32) + 1
    border = np.minimum(border, grid.nr // 2)
    ixyz = np.mgrid[-border[0]:border[0] + 1, -border[1]:border[1] + 1, -border[2]:border[2] + 1].reshape((3, -1))
    atomp = (np.asarray(scaled_position) * grid.nr).reshape((3, 1))
    ipoint = np.floor(atomp + 1e-8)
    px = atomp - ipoint
    l123 = np.mod(ipoint.astype(np.int32) - ixyz, grid.nr[:, None])
    pos = (ixyz + px) / grid.nr[:, None]
    pos = np.einsum("j...,kj->k...", pos, grid.lattice)
    dists = np.linalg.norm(pos, axis=0)
    index = np.logical_and(dists < rcut, dists > rtol) if rtol else dists < rcut
    vals = np.zeros(dists.shape)
    vals[index] = fn(dists[index])
    return l123, vals


def place(grid, rho, scaled_position, rcut, rtol, fn):
    l123, vals = box_values(grid, scaled_position, rcut, rtol, fn)
    rho[l123[0], l123[1], l123[2]] += vals
    return rho


def build_pseudo_density(grid, scaled_position, scale=0.0, sigma=0.3, rcut=5.0, density=None):
    if density is None:
        density = np.zeros(tuple(grid.nr))
    sigma = max(sigma, np.max(spacings(grid)) * 2 / FWHM)
    return place(grid, density, scaled_position, rcut, None, lambda d: gaussian(d, sigma) * scale)


def gen_gaussian_density(grid, symbols, scaled_positions, options, density=None):
    if density is None:
        density = np.zeros(tuple(grid.nr))
    ncharge = 0.0
    for key, o in options.items():
        scale = o.get("scale", 0.0)
        if scale is None or abs(scale) < 1e-10:
            continue
        for s, sp in zip(symbols, scaled_positions):
            if s != key:
                continue
            density = build_pseudo_density(grid, sp, scale, o.get("sigma", 0.3), o.get("rcut", 5.0), density)
            ncharge += scale
    return density, ncharge


def guess_rho_atom(grid, symbols, scaled_positions, r, arho, ncharge, rcut=None, dtol=1e-30):
    rho = np.zeros(tuple(grid.nr)) + dtol
    latp = np.linalg.norm(grid.lattice, axis=1)
    for key in r:
        if arho[key] is None:
            continue
        rk = np.asarray(r[key])
        rc = rcut.get(key, np.max(rk)) if isinstance(rcut, dict) else (rcut or np.max(rk))
        rc = min(rc, 0.5 * np.min(latp))
        tck = splrep(rk, arho[key])
        for s, sp in zip(symbols, scaled_positions):
            if s == key:
                rho = place(grid, rho, sp, rc, rk[0], lambda d: splev(d, tck, der=0))
    nc = rho.sum() * grid.dV
    if ncharge > 1e-6:
        rho *= ncharge / nc
    return rho


def recipe_g(grid, symbols, positions, scaled_positions, r, arho, pme=True, order=10):
    """rho(G) of ``get_3d_value_recipe`` before the inverse transform."""
    q = grid.q
    rho_g = np.zeros(grid.nrG, dtype=np.complex128)
    for key in r:
        if arho[key] is None:
            continue
        rk = np.asarray(r[key])
        line = np.zeros_like(q)
        m = q < rk[-1]
        line[m] = splev(q[m], splrep(rk, arho[key]), der=0, ext=1)
        idx = [i for i, s in enumerate(symbols) if s == key]
        if pme:
            Q = pme_qarray(grid, np.asarray(scaled_positions)[idx], np.ones(len(idx)), order)
            rho_g = rho_g + line * grid.fft(Q)
        else:
            for i in idx:
                rho_g = rho_g + line * strf(grid, positions[i])
    if pme:
        rho_g = rho_g * barray(grid, order) * grid.nnr / grid.volume
    return rho_g


def get_3d_value_recipe(grid, symbols, positions, scaled_positions, r, arho, ncharge, dtol=0.0, pme=True, order=10):
    rho = grid.ifft(recipe_g(grid, symbols, positions, scaled_positions, r, arho, pme, order))
    rho[rho < dtol] = dtol
    nc = rho.sum() * grid.dV
    if ncharge > 1e-6:
        rho *= ncharge / nc
    return rho
