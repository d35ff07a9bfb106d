"""DirectGrid / ReciprocalGrid (grid.py:26-57,291-734 of the reference).

A DirectGrid owns the HIP plans (FFT plans + workspace) of libofdft_hip for
its shape and lattice, one per precision, created on first use.  Grid
coordinate arrays (r, g, gg, q, invgg) are built on the GPU with torch on
demand for user code; the hot path computes G vectors inside its kernels.
"""
import ctypes

import numpy as np
import torch

from . import _lib
from .math_utils import ecut2nr, spacing2ecut


class Plan:
    """Owning handle of one ofdft_plan."""

    def __init__(self, nr, lattice, precision, full, max_batch=4, autotune=True):
        _lib.require_gpu()
        h = ctypes.c_void_p()
        nr_c = np.ascontiguousarray(nr, dtype=np.int32)
        lat = np.ascontiguousarray(lattice, dtype=np.float64).reshape(9)
        _lib.check(_lib.lib.ofdft_plan_create(_lib.iptr(nr_c), _lib.dptr(lat), int(precision), int(bool(full)),
                                              int(max_batch), ctypes.byref(h)))
        self.handle = h
        self.precision = precision
        self.full = full
        self.variants = None
        if autotune:
            self.autotune()

    def autotune(self):
        """Pick the fastest variant of each FFT pass (ofdft_plan_autotune): tiles
        of the row passes; tiles and radix orders of the column passes."""
        v = np.zeros(4, dtype=np.int32)
        t = np.zeros(24, dtype=np.float32)
        _lib.check(_lib.lib.ofdft_plan_autotune(self.handle, _lib.iptr(v), t.ctypes.data_as(ctypes.POINTER(ctypes.c_float)),
                                                _lib.stream()))
        self.variants = v.tolist()
        self.variant_ms = t.reshape(4, 6)
        return self.variants

    def set_variants(self, variants):
        v = np.ascontiguousarray(variants, dtype=np.int32)
        _lib.check(_lib.lib.ofdft_plan_set_variants(self.handle, _lib.iptr(v)))
        self.variants = v.tolist()

    @property
    def workspace_bytes(self):
        return int(_lib.lib.ofdft_plan_workspace_bytes(self.handle))

    def set_lattice(self, lattice):
        lat = np.ascontiguousarray(lattice, dtype=np.float64).reshape(9)
        _lib.check(_lib.lib.ofdft_plan_set_lattice(self.handle, _lib.dptr(lat)))

    def __del__(self):
        h = getattr(self, "handle", None)
        if h is None or not h.value:
            return
        try:
            torch.cuda.synchronize()
            _lib.lib.ofdft_plan_destroy(h)
        except Exception:  # interpreter shutdown: module globals may be gone
            pass
        self.handle = None


def _dtype_precision(dtype):
    if dtype == torch.float64:
        return _lib.FP64
    if dtype == torch.float32:
        return _lib.FP32
    raise TypeError("grid dtype must be torch.float64 or torch.float32")


class BaseGrid:
    def __init__(self, lattice, nr=None, origin=(0.0, 0.0, 0.0), full=True, ecut=None, spacing=None,
                 dtype=torch.float64, device=None, **kwargs):
        lattice = np.asarray(getattr(lattice, "array", lattice), dtype=np.float64).reshape(3, 3)
        if nr is None:
            if spacing is None:
                if ecut is None:
                    raise ValueError("give nr, ecut or spacing")
                spacing = np.sqrt(np.pi ** 2 / ecut * 0.5)
            nr = ecut2nr(lattice=lattice, spacing=spacing)
        self._lattice = lattice
        self._nrR = np.array(nr, dtype=np.int64).reshape(3)
        self._full = bool(full)
        self._origin = np.asarray(origin, dtype=np.float64)
        self._ecut = ecut
        self.dtype = dtype
        self.device = torch.device(device) if device is not None else torch.device("cuda")

    @property
    def lattice(self):
        return self._lattice

    @property
    def cell(self):
        return self._lattice

    @property
    def nrR(self):
        return self._nrR

    @property
    def nr(self):
        return self._nrR

    @property
    def nnrR(self):
        return int(np.prod(self._nrR))

    @property
    def nnr(self):
        return self.nnrR

    @property
    def volume(self):
        return abs(float(np.linalg.det(self._lattice)))

    @property
    def dV(self):
        return self.volume / self.nnrR

    @property
    def full(self):
        return self._full

    @property
    def origin(self):
        return self._origin

    @property
    def nrG(self):
        g = self._nrR.copy()
        if not self._full:
            g[-1] = g[-1] // 2 + 1
        return g

    @property
    def nnrG(self):
        return int(np.prod(self.nrG))

    @property
    def spacings(self):
        return np.sqrt((self._lattice ** 2).sum(axis=1)) / self._nrR

    def __eq__(self, other):
        return (isinstance(other, BaseGrid) and np.allclose(self.lattice, other.lattice)
                and np.allclose(self.nrR, other.nrR))

    __hash__ = object.__hash__


class DirectGrid(BaseGrid):
    """``DirectGrid(lattice, nr, full=True)`` (grid.py:291-465).  ``dtype``
    selects the precision of fields and kernels on this grid (float64 =
    FP64 mode, float32 = FP32-bulk mode with FP64 reductions)."""

    def __init__(self, lattice, nr=None, origin=(0.0, 0.0, 0.0), full=True, **kwargs):
        super().__init__(lattice, nr, origin, full, **kwargs)
        self._plans = {}
        self._rgrid = None
        self._r = None

    def plan(self, dtype=None, max_batch=4):
        """The HIP plan for this grid in the given precision."""
        prec = _dtype_precision(dtype or self.dtype)
        key = (prec, max_batch)
        if key not in self._plans:
            self._plans[key] = Plan(self._nrR, self._lattice, prec, self._full, max_batch)
        return self._plans[key]

    def set_lattice(self, lattice):
        """Move this grid (and its plans) to a new cell of the same shape, e.g.
        a strained copy; cached coordinate arrays are dropped."""
        lattice = np.asarray(getattr(lattice, "array", lattice), dtype=np.float64).reshape(3, 3)
        self._lattice = lattice
        for p in self._plans.values():
            p.set_lattice(lattice)
        self._rgrid = None
        self._r = None

    def with_dtype(self, dtype):
        """A grid of the same geometry with another precision (shares nothing)."""
        return DirectGrid(self._lattice, self._nrR, self._origin, self._full, dtype=dtype, device=self.device,
                          ecut=self._ecut)

    def get_reciprocal(self, scale=None, convention="physics"):
        if self._rgrid is None:
            self._rgrid = ReciprocalGrid(2.0 * np.pi * np.linalg.inv(self._lattice).T, self._nrR, full=self._full,
                                         dtype=self.dtype, device=self.device, direct=self)
        return self._rgrid

    @property
    def s(self):
        ax = [torch.arange(n, dtype=torch.float64, device=self.device) / n for n in self._nrR]
        return torch.stack(torch.meshgrid(*ax, indexing="ij"))

    @property
    def r(self):
        """Cartesian coordinates (3, nx, ny, nz) (grid.py:355-372)."""
        if self._r is None:
            lat = torch.as_tensor(self._lattice, device=self.device)
            self._r = torch.einsum("jxyz,jk->kxyz", self.s, lat).to(self.dtype)
        return self._r

    def guess_ecut(self):
        lengths = np.sqrt((self._lattice ** 2).sum(axis=1))
        sp2 = lengths / (self._nrR - 1)
        sp = 0.5 * (self.spacings + sp2)
        return spacing2ecut(sp.max())

    @property
    def ecut(self):
        return self._ecut if self._ecut is not None else self.guess_ecut()


class ReciprocalGrid(BaseGrid):
    """Reciprocal grid (grid.py:467-734): g = sum_j k_j b_j with fftfreq
    ordering, half grid on the last axis unless full."""

    def __init__(self, lattice, nr, full=True, direct=None, **kwargs):
        super().__init__(lattice, nr, full=full, **kwargs)
        self._direct = direct
        self._g = None

    def get_direct(self):
        return self._direct

    @property
    def nr(self):
        return self.nrG

    @property
    def g(self):
        if self._g is None:
            n = self._nrR
            ax = [torch.fft.fftfreq(int(n[0]), d=1.0 / int(n[0]), dtype=torch.float64, device=self.device),
                  torch.fft.fftfreq(int(n[1]), d=1.0 / int(n[1]), dtype=torch.float64, device=self.device)]
            if self._full:
                ax.append(torch.fft.fftfreq(int(n[2]), d=1.0 / int(n[2]), dtype=torch.float64, device=self.device))
            else:
                ax.append(torch.fft.rfftfreq(int(n[2]), d=1.0 / int(n[2]), dtype=torch.float64, device=self.device))
            S = torch.stack(torch.meshgrid(*ax, indexing="ij"))
            B = torch.as_tensor(self._lattice, device=self.device)
            self._g = torch.einsum("jxyz,jk->kxyz", S, B)
        return self._g

    @property
    def gg(self):
        g = self.g
        return (g * g).sum(0)

    @property
    def q(self):
        return torch.sqrt(self.gg)

    @property
    def invgg(self):
        gg = self.gg.clone()
        gg[0, 0, 0] = 1.0
        inv = 1.0 / gg
        inv[0, 0, 0] = 0.0
        return inv
