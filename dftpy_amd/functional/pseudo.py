"""LocalPseudo and the Ewald ion-ion energy (functional/pseudo/__init__.py:
28-246,587-784 and ewald.py:197-506 of the reference).

Setup (once per geometry) runs on the GPU in FP64: radial tables are the
interpolating cubic spline of the recpot v(q) (host, 1D), then the
structure factor / PME spreading, the G-space product and the inverse FFT
are HIP kernels (``ofdft_local_pp``), and the Ewald sums likewise
(``ofdft_ewald``).  An FP32 grid receives the FP64 result rounded once.
"""
import os

import numpy as np
import torch

from .. import _lib
from ..field import DirectField
from ..formats import read_recpot
from ..math_utils import cubic_table
from . import AbstractFunctional, evaluate


class RadialTable:
    """v(q) of one species as the device-side piecewise cubic."""

    def __init__(self, gp, vp):
        gp = np.asarray(gp, dtype=np.float64)
        vp = np.asarray(vp, dtype=np.float64)
        self.coef = cubic_table(gp[1:], vp[1:])
        self.q0 = float(gp[1])
        self.dq = float((gp[-1] - gp[1]) / (len(gp) - 2))
        self.qmax = float(gp[-1])
        self.vsmall = (float(vp[0]), float(vp[1]), float(vp[2]))
        self.dqs = float(gp[1] - gp[0])


class ReadPseudo:
    """PP_list {symbol: file} -> radial tables and valence charges
    (pseudo/__init__.py:587-658).  Only recpot files are read."""

    def __init__(self, PP_list):
        if isinstance(PP_list, (list, tuple)):
            PP_list = {os.path.basename(str(p)).split(".")[0].split("_")[0].capitalize(): p for p in PP_list}
        self.PP_list = dict(PP_list)
        self.gp, self.vp, self.zval, self.tables = {}, {}, {}, {}
        for key, fname in self.PP_list.items():
            fname = str(fname)
            base = fname[:-3] if fname.endswith(".gz") else fname
            if not base.lower().endswith(".recpot"):
                raise NotImplementedError("only recpot pseudopotentials are supported (%s)" % fname)
            if not os.path.isfile(fname):
                raise FileNotFoundError("'{}' PP file for atom type {} not found".format(fname, key))
            gp, vp, z = read_recpot(fname)
            self.gp[key], self.vp[key], self.zval[key] = gp, vp, z
            self.tables[key] = RadialTable(gp, vp)


class Ewald:
    """Ewald ion-ion energy (ewald.py:197-506); PME reciprocal part by
    default as in LocalPseudo.get_ewald."""

    def __init__(self, ions, grid, PME=True, precision=1.0e-8, BsplineOrder=10):
        self.ions, self.grid, self.PME = ions, grid, PME
        self.precision, self.order = precision, BsplineOrder
        self._parts = None

    def _compute(self):
        plan = self.grid.plan(torch.float64)
        pos = np.ascontiguousarray(self.ions.positions, dtype=np.float64)
        frac = np.ascontiguousarray(self.ions.get_scaled_positions(), dtype=np.float64)
        q = np.ascontiguousarray(self.ions.charges, dtype=np.float64)
        out = np.zeros(5)
        _lib.check(_lib.lib.ofdft_ewald(plan.handle, self.ions.nat, _lib.dptr(pos), _lib.dptr(frac), _lib.dptr(q),
                                        int(bool(self.PME)), int(self.order), float(self.precision), _lib.dptr(out),
                                        _lib.stream()))
        self._parts = dict(energy=out[0], corr=out[1], real=out[2], rec=out[3], eta=out[4])

    @property
    def energy(self):
        if self._parts is None:
            self._compute()
        return self._parts["energy"]

    @property
    def eta(self):
        if self._parts is None:
            self._compute()
        return self._parts["eta"]

    def Energy_corr(self):
        self.energy
        return self._parts["corr"]

    def Energy_real(self):
        self.energy
        return self._parts["real"]

    def Energy_rec(self):
        self.energy
        return self._parts["rec"]

    def _forces_stress(self):
        plan = self.grid.plan(torch.float64)
        pos = np.ascontiguousarray(self.ions.positions, dtype=np.float64)
        frac = np.ascontiguousarray(self.ions.get_scaled_positions(), dtype=np.float64)
        q = np.ascontiguousarray(self.ions.charges, dtype=np.float64)
        f = np.zeros((self.ions.nat, 3))
        st = np.zeros((3, 3))
        if self.PME:
            _lib.check(_lib.lib.ofdft_ewald_forces_stress_pme(plan.handle, self.ions.nat, _lib.dptr(pos),
                                                              _lib.dptr(frac), _lib.dptr(q), float(self.precision),
                                                              int(self.order), _lib.dptr(f), _lib.dptr(st),
                                                              _lib.stream()))
        else:
            _lib.check(_lib.lib.ofdft_ewald_forces_stress(plan.handle, self.ions.nat, _lib.dptr(pos), _lib.dptr(frac),
                                                          _lib.dptr(q), float(self.precision), _lib.dptr(f),
                                                          _lib.dptr(st), _lib.stream()))
        self._fs = (f, st)

    @property
    def forces(self):
        """Ewald forces (ewald.py:510-521): real space plus the reciprocal
        part, PME (Forces_rec_PME) or exact (Forces_rec) as the Ewald was
        built, nat x 3 in Ha/bohr."""
        if getattr(self, "_fs", None) is None:
            self._forces_stress()
        return self._fs[0].copy()

    @property
    def stress(self):
        """Ewald stress (ewald.py:523-539; Stress_rec_PME or Stress_rec),
        3 x 3 in Ha/bohr^3."""
        if getattr(self, "_fs", None) is None:
            self._forces_stress()
        return self._fs[1].copy()


class LocalPseudo(AbstractFunctional):
    """``LocalPseudo(grid, ions, PP_list, PME=True)`` (pseudo/__init__.py:80-246)."""

    def __init__(self, grid=None, ions=None, PP_list=None, PME=True, readpp=None, BsplineOrder=10, **kwargs):
        self.type = "PSEUDO"
        self.name = "PSEUDO"
        self.terms = _lib.TERM_PP
        if PP_list is not None:
            self.readpp = ReadPseudo(PP_list)
        elif readpp is not None:
            self.readpp = readpp
        else:
            raise AttributeError("Must specify PP_list for Pseudopotentials")
        self.PME = PME
        self.BsplineOrder = BsplineOrder
        self.restart(grid, ions)

    def restart(self, grid=None, ions=None, duplicate=False):
        if duplicate:
            return LocalPseudo(grid=grid, ions=ions, readpp=self.readpp, PME=self.PME, BsplineOrder=self.BsplineOrder)
        self._vreal = {}
        self.ewald = None
        self._grid = grid
        self._ions = None
        if ions is not None:
            self.ions = ions

    @property
    def grid(self):
        if self._grid is None:
            raise AttributeError("Must specify grid for Pseudopotentials")
        return self._grid

    @property
    def ions(self):
        if self._ions is None:
            raise AttributeError("Must specify ions for Pseudopotentials")
        return self._ions

    @ions.setter
    def ions(self, value):
        for key in value.symbols_uniq:
            if key not in self.readpp.PP_list:
                raise ValueError("There is no pseudopotential for {:4s} atom".format(key))
        self._ions = value
        self._ions.set_charges(self.readpp.zval)

    @property
    def zval(self):
        return self.readpp.zval

    def get_ewald(self, PME=None):
        if self.ewald is None:
            self.ewald = Ewald(self.ions, self.grid, PME=self.PME if PME is None else PME,
                               BsplineOrder=self.BsplineOrder)
        return self.ewald

    def _table_args(self):
        """(keep-alive list, ctypes arguments) of the radial tables and the
        ions, in the order of ofdft_local_pp's nspecies .. species."""
        keys = sorted(self.readpp.tables)
        tabs = [self.readpp.tables[k] for k in keys]
        ns = len(keys)
        q0 = np.array([t.q0 for t in tabs])
        dq = np.array([t.dq for t in tabs])
        nco = np.array([t.coef.shape[0] for t in tabs], dtype=np.int32)
        qmax = np.array([t.qmax for t in tabs])
        vsmall = np.array([v for t in tabs for v in t.vsmall])
        dqs = np.array([t.dqs for t in tabs])
        coefs = [np.ascontiguousarray(t.coef.reshape(-1)) for t in tabs]
        cptr = (_lib._DP * ns)(*[_lib.dptr(c) for c in coefs])
        ions = self.ions
        pos = np.ascontiguousarray(ions.positions, dtype=np.float64)
        frac = np.ascontiguousarray(ions.get_scaled_positions(), dtype=np.float64)
        spec = ions.species_index(keys)
        keep = [q0, dq, nco, qmax, vsmall, dqs, coefs, cptr, pos, frac, spec]
        args = (ns, _lib.dptr(q0), _lib.dptr(dq), _lib.iptr(nco), cptr, _lib.dptr(qmax), _lib.dptr(vsmall),
                _lib.dptr(dqs), ions.nat, _lib.dptr(pos), _lib.dptr(frac), _lib.iptr(spec))
        return keep, args

    def _compute_vreal64(self):
        grid = self.grid
        plan = grid.plan(torch.float64)
        keep, args = self._table_args()
        nr = tuple(int(n) for n in grid.nrR)
        out = torch.empty(nr, dtype=torch.float64, device=grid.device)
        _lib.check(_lib.lib.ofdft_local_pp(plan.handle, *args, int(bool(self.PME)), int(self.BsplineOrder),
                                           out.data_ptr(), _lib.stream()))
        del keep
        return out

    def forces(self, rho):
        """LocalPseudo.forces (pseudo/__init__.py:260-270): nat x 3 in
        Ha/bohr, the PME form (_ForcePME, :475-520) when PME is set, as in
        the reference, else exact structure factors (_Force, :456-470)."""
        from . import _fp64
        if not isinstance(rho, DirectField):
            raise TypeError("rho must be DirectField")
        r = _fp64(rho)
        plan = r.grid.plan(torch.float64)
        keep, args = self._table_args()
        f = np.zeros((self.ions.nat, 3))
        if self.PME:
            _lib.check(_lib.lib.ofdft_local_pp_forces_pme(plan.handle, r.data.contiguous().data_ptr(), *args,
                                                          int(self.BsplineOrder), _lib.dptr(f), _lib.stream()))
        else:
            _lib.check(_lib.lib.ofdft_local_pp_forces(plan.handle, r.data.contiguous().data_ptr(), *args,
                                                      _lib.dptr(f), _lib.stream()))
        del keep
        return f

    def stress(self, rho, energy=None):
        """LocalPseudo.stress (pseudo/__init__.py:248-258), 3 x 3 in
        Ha/bohr^3: the PME form (_StressPME, :522-563) when PME is set, else
        the exact one (_Stress, :414-454); energy: the PP energy at rho
        (computed if None)."""
        from . import _fp64
        if not isinstance(rho, DirectField):
            raise TypeError("rho must be DirectField")
        r = _fp64(rho)
        if energy is None:
            energy = float((self.vreal_for(r.grid).data * r.data).sum().item() * r.grid.dV)
        plan = r.grid.plan(torch.float64)
        keep, args = self._table_args()
        st = np.zeros((3, 3))
        if self.PME:
            _lib.check(_lib.lib.ofdft_local_pp_stress_pme(plan.handle, r.data.contiguous().data_ptr(), *args,
                                                          int(self.BsplineOrder), float(energy), _lib.dptr(st),
                                                          _lib.stream()))
        else:
            _lib.check(_lib.lib.ofdft_local_pp_stress(plan.handle, r.data.contiguous().data_ptr(), *args,
                                                      float(energy), _lib.dptr(st), _lib.stream()))
        del keep
        return st

    def vreal_for(self, grid):
        """vreal as a DirectField in the precision of ``grid``."""
        key = grid.dtype
        if key not in self._vreal:
            if torch.float64 not in self._vreal:
                self._vreal[torch.float64] = DirectField(self.grid.with_dtype(torch.float64)
                                                         if self.grid.dtype != torch.float64 else self.grid,
                                                         data=self._compute_vreal64())
            if key != torch.float64:
                self._vreal[key] = DirectField(grid, data=self._vreal[torch.float64].data.to(key))
        return self._vreal[key]

    @property
    def vreal(self):
        return self.vreal_for(self.grid)

    def local_PP(self):
        return self.vreal

    @property
    def v(self):
        return self.vreal.fft()

    def compute(self, density, calcType={"E", "V"}, **kwargs):
        return evaluate([self], density, calcType, name="eN")
