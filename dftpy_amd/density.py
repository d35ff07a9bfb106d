"""Start densities (density/density.py of the reference): the uniform gas,
superpositions of atomic radial densities placed in real space or through
structure factors, and Gaussian atoms.

The placement runs on the GPU: ``ofdft_atomic_density`` (real-space boxes
around each atom, density.py:18-67,209-249,353-406), ``ofdft_local_pp`` with
the density's radial table (the reciprocal recipe, density.py:252-317, whose
PME and exact structure-factor sums have the same form as the local
pseudopotential's) and ``ofdft_clamp_normalize``.  Host work is limited to the
radial tables.  Not built: spin-resolved and core densities, densities from
pseudopotential files (the recpot reader carries none, so a generator built
from one falls back to the uniform gas as the reference's does) and the
derivative form of ``build_pseudo_density``.
"""
import os

import numpy as np
import torch
from scipy.interpolate import PPoly, splev, splrep

from . import _lib
from .field import DirectField
from .math_utils import cubic_table

FWHM = 2.354820


def _frac(grid, ions_or_pos):
    if hasattr(ions_or_pos, "get_scaled_positions"):
        return np.ascontiguousarray(ions_or_pos.get_scaled_positions(), dtype=np.float64)
    pos = np.asarray(ions_or_pos, dtype=np.float64).reshape(-1, 3)
    return np.ascontiguousarray(np.linalg.solve(np.asarray(grid.lattice).T, pos.T).T)


def _fp64_out(grid, rho, fill):
    """FP64 device buffer for the placement (rho's values when given)."""
    if rho is None:
        return torch.full(tuple(int(n) for n in grid.nrR), fill, dtype=torch.float64, device=grid.device)
    return rho.data.to(torch.float64).clone() if rho.data.dtype != torch.float64 else rho.data


def _finish(grid, rho, buf):
    if rho is None:
        return DirectField(grid, data=buf.to(grid.dtype))
    if buf.data_ptr() != rho.data.data_ptr():
        rho.data.copy_(buf)
    return rho


def _spline_ppoly(r, v):
    """Breakpoints and c0..c3 coefficients of FITPACK's interpolating cubic
    spline through (r, v) (splrep, k = 3, s = 0), zero-length pieces dropped."""
    pp = PPoly.from_spline(splrep(np.asarray(r, dtype=np.float64), np.asarray(v, dtype=np.float64)))
    x, c = pp.x, pp.c
    keep = np.nonzero(np.diff(x) > 0)[0]
    br = np.append(x[keep], x[keep[-1] + 1])
    coef = np.stack([c[3, keep], c[2, keep], c[1, keep], c[0, keep]], axis=1)
    return np.ascontiguousarray(br), np.ascontiguousarray(coef.reshape(-1))


def place_atoms(grid, ions, keys, kind, rcut, rtol=None, sigma=None, scale=None, tables=None, out=None):
    """Add the per-species radial functions of ``keys`` around every atom of
    those species into ``out`` (FP64 device tensor); see ofdft_atomic_density."""
    ns = len(keys)
    idx = [i for i, s in enumerate(ions.symbols) if s in keys]
    if not idx:
        return out
    spec = np.ascontiguousarray([keys.index(ions.symbols[i]) for i in idx], dtype=np.int32)
    frac = np.ascontiguousarray(_frac(grid, ions)[idx])
    f64 = lambda a: np.ascontiguousarray(a, dtype=np.float64)
    rcut, rtol = f64(rcut), f64(rtol if rtol is not None else np.zeros(ns))
    sigma, scale = f64(sigma if sigma is not None else np.ones(ns)), f64(scale if scale is not None else np.ones(ns))
    keep = [rcut, rtol, sigma, scale, frac, spec]
    nb = br = co = None
    if kind == 0:
        brs = [t[0] for t in tables]
        cos = [t[1] for t in tables]
        keep += [brs, cos]
        nb = np.ascontiguousarray([len(b) for b in brs], dtype=np.int32)
        br = (_lib._DP * ns)(*[_lib.dptr(b) for b in brs])
        co = (_lib._DP * ns)(*[_lib.dptr(c) for c in cos])
        keep.append(nb)
    plan = grid.plan(torch.float64)
    _lib.check(_lib.lib.ofdft_atomic_density(plan.handle, int(kind), ns, _lib.iptr(nb) if nb is not None else None,
                                             br, co, _lib.dptr(rcut), _lib.dptr(rtol), _lib.dptr(sigma),
                                             _lib.dptr(scale), len(idx), _lib.dptr(frac), _lib.iptr(spec),
                                             out.data_ptr(), _lib.stream()))
    del keep
    return out


def clamp_normalize(grid, buf, dtol, ncharge, clamp=True):
    plan = grid.plan(torch.float64)
    _lib.check(_lib.lib.ofdft_clamp_normalize(plan.handle, buf.data_ptr(), int(bool(clamp)), float(dtol),
                                              float(ncharge), _lib.stream()))
    return buf


def guess_rho_heg(ions, grid, ncharge=None, rho=None, dtol=1e-30, **kwargs):
    """Uniform gas N_e / V (density.py:199-207)."""
    if ncharge is None:
        ncharge = ions.get_ncharges()
    if rho is None:
        rho = DirectField(grid)
    rho.data.fill_(ncharge / grid.volume)
    return rho


def _recipe_table(r, arho):
    """(q0, dq, coef, qmax) of the radial table on a uniform q grid: the
    FITPACK spline itself when r is uniform, else that spline resampled on
    8 len(r) uniform points."""
    r = np.asarray(r, dtype=np.float64)
    v = np.asarray(arho, dtype=np.float64)
    d = np.diff(r)
    if not np.allclose(d, d[0], rtol=1e-10, atol=0.0):
        x = np.linspace(r[0], r[-1], 8 * len(r))
        v = splev(x, splrep(r, v), der=0)
        r = x
    return float(r[0]), float((r[-1] - r[0]) / (len(r) - 1)), np.ascontiguousarray(cubic_table(r, v).reshape(-1)), \
        float(r[-1])


def get_3d_value_recipe(r, arho, ions, grid, ncharge=None, rho=None, dtol=0.0, direct=True, pme=True, order=10,
                        **kwargs):
    """Superposition through structure factors (density.py:252-317):
    rho(G) = sum_s spline_s(|G|) S_s(G), zero at |G| >= r_s[-1] and below
    r_s[0]; inverse transform, values below dtol set to dtol, scaled to the
    electron count.  direct=False returns rho(G) as a ReciprocalField."""
    grid = grid.get_direct() if hasattr(grid, "get_direct") else grid
    keys = sorted(k for k in r if arho.get(k) is not None)
    buf = torch.zeros(tuple(int(n) for n in grid.nrR), dtype=torch.float64, device=grid.device)
    if keys:
        tabs = [_recipe_table(r[k], arho[k]) for k in keys]
        ns = len(keys)
        q0 = np.array([t[0] for t in tabs])
        dq = np.array([t[1] for t in tabs])
        coefs = [t[2] for t in tabs]
        nco = np.array([len(c) // 4 for c in coefs], dtype=np.int32)
        qmax = np.array([t[3] for t in tabs])
        # below r[0] the table is zero (splev ext=1): the small-q quartic through zeros
        vsmall = np.zeros(3 * ns)
        dqs = q0.copy()
        cptr = (_lib._DP * ns)(*[_lib.dptr(c) for c in coefs])
        sel = [i for i, s in enumerate(ions.symbols) if s in keys]
        pos = np.ascontiguousarray(np.asarray(ions.positions, dtype=np.float64)[sel])
        frac = np.ascontiguousarray(_frac(grid, ions)[sel])
        spec = np.ascontiguousarray([keys.index(ions.symbols[i]) for i in sel], dtype=np.int32)
        plan = grid.plan(torch.float64)
        _lib.check(_lib.lib.ofdft_local_pp(plan.handle, ns, _lib.dptr(q0), _lib.dptr(dq), _lib.iptr(nco), cptr,
                                           _lib.dptr(qmax), _lib.dptr(vsmall), _lib.dptr(dqs), len(sel),
                                           _lib.dptr(pos), _lib.dptr(frac), _lib.iptr(spec), int(bool(pme)),
                                           int(order), buf.data_ptr(), _lib.stream()))
        del coefs, cptr
    if not direct:
        return DirectField(grid, data=buf).fft()
    if ncharge is None:
        ncharge = ions.get_ncharges()
    clamp_normalize(grid, buf, dtol, ncharge)
    return _finish(grid, rho, buf)


def gen_gaussian_density(ions, grid, options={}, density=None, deriv=0, **kwargs):
    """Normalised Gaussians of charge ``scale`` on the atoms of each species
    in ``options`` ({symbol: {'rcut', 'sigma', 'scale'}}, density.py:353-363);
    returns (density, total charge placed)."""
    if deriv != 0:
        raise NotImplementedError("the derivative form of the Gaussian density is not implemented")
    keys, rc, sg, sc = [], [], [], []
    ncharge = 0.0
    smin = np.max(grid.spacings) * 2 / FWHM
    for key, o in options.items():
        scale = o.get("scale", 0.0)
        if scale is None or abs(scale) < 1e-10:
            continue
        keys.append(key)
        rc.append(o.get("rcut", 5.0))
        sg.append(max(o.get("sigma", 0.3), smin))
        sc.append(scale)
        ncharge += scale * sum(1 for s in ions.symbols if s == key)
    buf = _fp64_out(grid, density, 0.0)
    if keys:
        place_atoms(grid, ions, keys, 1, rc, sigma=sg, scale=sc, out=buf)
    return _finish(grid, density, buf), ncharge


class _OnePosition:
    def __init__(self, grid, pos):
        self.symbols = ["X"]
        self._frac = _frac(grid, pos)

    def get_scaled_positions(self):
        return self._frac


def build_pseudo_density(pos, grid, scale=0.0, sigma=0.3, rcut=5.0, density=None, add=True, deriv=0,
                         generator=None):
    """One normalised Gaussian of charge ``scale`` at cartesian ``pos``
    (density.py:365-406), sigma at least 2 max(spacing) / 2.354820."""
    if deriv != 0 or np.asarray(scale).size > 1:
        raise NotImplementedError("the derivative form of the Gaussian density is not implemented")
    if not add:
        raise NotImplementedError("build_pseudo_density(add=False) is not implemented")
    buf = _fp64_out(grid, density, 0.0)
    if scale is None or abs(float(scale)) < 1e-10:
        return _finish(grid, density, buf)
    sigma = max(sigma, np.max(grid.spacings) * 2 / FWHM)
    place_atoms(grid, _OnePosition(grid, pos), ["X"], 1, [rcut], sigma=[sigma], scale=[float(scale)], out=buf)
    return _finish(grid, density, buf)


class DensityGenerator:
    """Start density from atomic radial densities (density.py:70-317):
    tables given as ``r``/``arho`` dicts or ``.list`` files (two columns),
    placed by structure factors (``direct=False``, the default) or in real
    space (``direct=True``); without tables the uniform gas."""

    def __init__(self, files=None, pseudo=None, is_core=False, direct=False, r=None, arho=None, **kwargs):
        self._r = dict(r or {})
        self._arho = dict(arho or {})
        self.direct = direct
        self.is_core = is_core
        self.files = files
        self.pseudo = pseudo
        if is_core:
            raise NotImplementedError("core densities are not implemented")
        if files:
            for key, infile in files.items():
                if not os.path.isfile(infile):
                    raise Exception("Density file " + infile + " for atom type " + str(key) + " not found")
                if infile[-4:].lower() != "list":
                    raise NotImplementedError("atomic densities from pseudopotential files are not implemented")
                self._r[key], self._arho[key] = self.read_density_list(infile)
        elif pseudo is not None:
            # recpot files carry no atomic density
            readpp = getattr(pseudo, "readpp", pseudo)
            for key in getattr(readpp, "PP_list", {}):
                self._r.setdefault(key, None)
                self._arho.setdefault(key, None)

    @property
    def r(self):
        return self._r

    @r.setter
    def r(self, value):
        self._r = value

    @property
    def arho(self):
        return self._arho

    @arho.setter
    def arho(self, value):
        self._arho = value

    @staticmethod
    def read_density_list(infile):
        data = np.loadtxt(infile, usecols=(0, 1), ndmin=2)
        return data[:, 0].copy(), data[:, 1].copy()

    def guess_rho(self, ions, grid=None, ncharge=None, rho=None, dtol=1e-30, nspin=1, **kwargs):
        if nspin != 1 or (rho is not None and getattr(rho, "rank", 1) > 1):
            raise NotImplementedError("spin-polarised start densities are not implemented")
        if grid is None:
            if rho is None:
                raise AttributeError("Please give one grid.")
            grid = rho.grid
        return self.guess_rho_unspin(ions, grid=grid, ncharge=ncharge, rho=rho, dtol=dtol, **kwargs)

    def guess_rho_unspin(self, ions, grid=None, ncharge=None, rho=None, dtol=1e-30, **kwargs):
        if len(self._arho) == 0:
            return self.guess_rho_heg(ions, grid, ncharge, rho, dtol=dtol, **kwargs)
        if self.direct:
            return self.guess_rho_atom(ions, grid, ncharge, rho, dtol=dtol, **kwargs)
        return self.get_3d_value_recipe(ions, grid, ncharge, rho, dtol=dtol, **kwargs)

    def guess_rho_heg(self, ions, grid, ncharge=None, rho=None, dtol=1e-30, **kwargs):
        return guess_rho_heg(ions, grid, ncharge, rho, dtol)

    def guess_rho_atom(self, ions, grid, ncharge=None, rho=None, dtol=1e-30, rcut=None, **kwargs):
        """Real-space superposition (density.py:209-249): splines of r, rho(r)
        within (r[0], rc), rc = min(rcut or max(r), L_min / 2), then scaled to
        the electron count."""
        buf = _fp64_out(grid, None, dtol)
        latp = np.linalg.norm(np.asarray(grid.lattice), axis=1)
        keys = [k for k in self._r if self._arho.get(k) is not None]
        if keys:
            rcs, rts, tabs = [], [], []
            for key in keys:
                rk = np.asarray(self._r[key], dtype=np.float64)
                rc = rcut.get(key, np.max(rk)) if isinstance(rcut, dict) else (rcut or np.max(rk))
                rcs.append(min(rc, 0.5 * np.min(latp)))
                rts.append(rk[0])
                tabs.append(_spline_ppoly(rk, self._arho[key]))
            place_atoms(grid, ions, keys, 0, rcs, rtol=rts, tables=tabs, out=buf)
        if ncharge is None:
            ncharge = ions.get_ncharges()
        clamp_normalize(grid, buf, 0.0, ncharge, clamp=False)
        return _finish(grid, rho, buf)

    def get_3d_value_recipe(self, ions, grid, ncharge=None, rho=None, dtol=1e-30, direct=True, **kwargs):
        return get_3d_value_recipe(self._r, self._arho, ions, grid, ncharge, rho, dtol, direct, **kwargs)
