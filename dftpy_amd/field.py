"""DirectField / ReciprocalField (field.py:11-785 of the reference).

A field is a torch tensor on the GPU plus the grid it lives on.  Arithmetic
between fields and scalars is plain torch (user-level glue); the DFTpy
methods on the hot path -- fft/ifft, integral -- call libofdft_hip.
"""
import numbers

import numpy as np
import torch

from . import _lib
from .grid import DirectGrid, ReciprocalGrid


def _tensor_of(x):
    return x.data if isinstance(x, BaseField) else x


class BaseField:
    __array_priority__ = 1000

    def __init__(self, grid, data):
        self.grid = grid
        self.data = data

    # ---- container protocol
    @property
    def shape(self):
        return tuple(self.data.shape)

    @property
    def dtype(self):
        return self.data.dtype

    @property
    def ndim(self):
        return self.data.dim()

    @property
    def rank(self):
        return self.data.shape[0] if self.data.dim() == 4 else 1

    def __len__(self):
        return self.data.shape[0]

    def numpy(self):
        return self.data.detach().cpu().numpy()

    def __array__(self, dtype=None, copy=None):
        a = self.numpy()
        return a if dtype is None else a.astype(dtype)

    def __getitem__(self, idx):
        return self.data[idx]

    def __setitem__(self, idx, value):
        v = _tensor_of(value)
        if isinstance(v, np.ndarray):
            v = torch.as_tensor(v, device=self.data.device)
        self.data[idx] = v

    def _wrap(self, t):
        return type(self)(self.grid, data=t)

    # ---- arithmetic (torch)
    def __add__(self, o): return self._wrap(self.data + _tensor_of(o))
    def __radd__(self, o): return self._wrap(_tensor_of(o) + self.data)
    def __sub__(self, o): return self._wrap(self.data - _tensor_of(o))
    def __rsub__(self, o): return self._wrap(_tensor_of(o) - self.data)
    def __mul__(self, o): return self._wrap(self.data * _tensor_of(o))
    def __rmul__(self, o): return self._wrap(_tensor_of(o) * self.data)
    def __truediv__(self, o): return self._wrap(self.data / _tensor_of(o))
    def __rtruediv__(self, o): return self._wrap(_tensor_of(o) / self.data)
    def __pow__(self, o): return self._wrap(self.data ** _tensor_of(o))
    def __neg__(self): return self._wrap(-self.data)

    def __iadd__(self, o):
        self.data += _tensor_of(o)
        return self

    def __isub__(self, o):
        self.data -= _tensor_of(o)
        return self

    def __imul__(self, o):
        self.data *= _tensor_of(o)
        return self

    def __itruediv__(self, o):
        self.data /= _tensor_of(o)
        return self

    def __lt__(self, o): return self.data < _tensor_of(o)
    def __le__(self, o): return self.data <= _tensor_of(o)
    def __gt__(self, o): return self.data > _tensor_of(o)
    def __ge__(self, o): return self.data >= _tensor_of(o)

    def copy(self):
        return self._wrap(self.data.clone())

    def sum(self):
        return self.data.sum().item()

    def amax(self):
        return self.data.max().item()

    def amin(self):
        return self.data.min().item()

    def asum(self):
        return self.data.sum().item()

    def amean(self):
        return self.data.mean().item()

    def mean(self):
        return self.data.mean().item()

    @property
    def real(self):
        return self._wrap(self.data.real) if self.data.is_complex() else self

    @property
    def imag(self):
        return self._wrap(self.data.imag)

    def conj(self):
        return self._wrap(self.data.conj().resolve_conj())


class DirectField(BaseField):
    """``DirectField(grid, data=..., rank=1, cplx=False)`` (field.py:141-173).
    ``order='F'`` reads flat data in Fortran order as the reference does."""

    def __init__(self, grid, memo="", rank=1, data=None, order="C", cplx=False, griddata_F=None, griddata_C=None,
                 griddata_3d=None, **kwargs):
        if hasattr(grid, "get_direct") and not isinstance(grid, DirectGrid):
            grid = grid.get_direct()
        if not isinstance(grid, DirectGrid):
            raise TypeError("the grid argument is not an instance of DirectGrid")
        if griddata_F is not None:
            data, order = griddata_F, "F"
        elif griddata_C is not None:
            data = griddata_C
        elif griddata_3d is not None:
            data = griddata_3d
        nr = tuple(int(n) for n in grid.nrR)
        shape = nr if rank == 1 else (rank,) + nr
        dtype = grid.dtype if not cplx else (torch.complex128 if grid.dtype == torch.float64 else torch.complex64)
        if data is None:
            t = torch.zeros(shape, dtype=dtype, device=grid.device)
        elif isinstance(data, torch.Tensor):
            t = data.to(device=grid.device)
            if not t.is_complex() and not cplx:
                t = t.to(grid.dtype)
            t = t.reshape(shape)
        else:
            a = np.asarray(getattr(data, "data", data) if isinstance(data, BaseField) else data)
            a = np.reshape(a, shape, order=order)
            t = torch.as_tensor(np.ascontiguousarray(a), device=grid.device)
            t = t.to(dtype if (cplx or not t.is_complex()) else t.dtype)
        self.grid = grid
        self.data = t.contiguous()
        self.memo = memo
        self._N = None

    @property
    def cplx(self):
        return self.data.is_complex()

    def _wrap(self, t):
        return DirectField(self.grid, data=t, rank=t.shape[0] if t.dim() == 4 else 1, cplx=t.is_complex())

    def plan(self):
        return self.grid.plan(self.grid.dtype)

    def integral(self):
        """sum * dV with an FP64-accumulated HIP reduction (field.py:186-192)."""
        if self.data.is_complex() or self.rank > 1 or self.data.dtype != self.grid.dtype:
            return (self.data.sum(dim=(-3, -2, -1)) * self.grid.dV).cpu().numpy() if self.rank > 1 else \
                (self.data.sum() * self.grid.dV).item()
        out = torch.empty(1, dtype=torch.float64, device=self.data.device)
        _lib.check(_lib.lib.ofdft_dot(self.plan().handle, self.data.data_ptr(), None, out.data_ptr(), _lib.stream()))
        return out.item() * self.grid.dV

    @property
    def N(self):
        if self._N is None:
            self._N = self.integral()
        return self._N

    def fft(self):
        """rfftn * dV on a half grid, or fftn * dV on a full grid (field.py:373-390)."""
        rg = self.grid.get_reciprocal()
        plan = self.plan()
        nr = tuple(int(n) for n in self.grid.nrR)
        ctype = torch.complex128 if self.grid.dtype == torch.float64 else torch.complex64
        if self.grid.full or self.data.is_complex():
            src = self.data.to(ctype).contiguous()
            out = torch.empty(nr, dtype=ctype, device=src.device)
            _lib.check(_lib.lib.ofdft_fft_c2c(plan.handle, src.data_ptr(), out.data_ptr(), 0, 1, _lib.stream()))
            return ReciprocalField(rg, data=out, cplx=self.data.is_complex())
        out = torch.empty(nr[:2] + (nr[2] // 2 + 1,), dtype=ctype, device=self.data.device)
        src = self.data.contiguous()
        _lib.check(_lib.lib.ofdft_fft_r2c(plan.handle, src.data_ptr(), out.data_ptr(), 1, _lib.stream()))
        return ReciprocalField(rg, data=out)

    def laplacian(self, check_real=False, force_real=False, sigma=0.025):
        """IFFT(-G^2 exp(-G^2 sigma^2/2) FFT(f)) (field.py:343-350)."""
        F = self.fft()
        gg = F.grid.gg
        k = -gg if not sigma else -gg * torch.exp(-gg * sigma ** 2 / 2.0)
        return ReciprocalField(F.grid, data=F.data * k.to(F.data.real.dtype), cplx=F.cplx).ifft(
            force_real=force_real)


class ReciprocalField(BaseField):
    """Spectrum on the reciprocal grid (field.py:606-785)."""

    def __init__(self, grid, memo="", rank=1, data=None, cplx=False, griddata_3d=None, **kwargs):
        if hasattr(grid, "get_reciprocal"):
            grid = grid.get_reciprocal()
        if not isinstance(grid, ReciprocalGrid):
            raise TypeError("the grid argument is not an instance of ReciprocalGrid")
        if griddata_3d is not None:
            data = griddata_3d
        shape = tuple(int(n) for n in grid.nrG)
        ctype = torch.complex128 if grid.dtype == torch.float64 else torch.complex64
        if data is None:
            t = torch.zeros(shape, dtype=ctype, device=grid.device)
        elif isinstance(data, torch.Tensor):
            t = data.to(device=grid.device)
        else:
            t = torch.as_tensor(np.asarray(data), device=grid.device)
        self.grid = grid
        self.data = t.to(ctype).reshape(shape).contiguous()
        self._cplx = cplx

    @property
    def cplx(self):
        return self._cplx

    def _wrap(self, t):
        return ReciprocalField(self.grid, data=t, cplx=self._cplx)

    def ifft(self, check_real=False, force_real=False):
        """irfftn(s=nr) / dV, or ifftn / dV on a full grid (field.py:739-764)."""
        dg = self.grid.get_direct()
        plan = dg.plan(dg.dtype)
        nr = tuple(int(n) for n in dg.nrR)
        src = self.data.contiguous()
        if dg.full or self._cplx:
            out = torch.empty(nr, dtype=src.dtype, device=src.device)
            _lib.check(_lib.lib.ofdft_fft_c2c(plan.handle, src.data_ptr(), out.data_ptr(), 1, 1, _lib.stream()))
            if force_real or (not self._cplx):
                out = out.real.contiguous()
            return DirectField(dg, data=out, cplx=out.is_complex())
        out = torch.empty(nr, dtype=dg.dtype, device=src.device)
        _lib.check(_lib.lib.ofdft_fft_c2r(plan.handle, src.data_ptr(), out.data_ptr(), 1, _lib.stream()))
        return DirectField(dg, data=out)

    def integral(self):
        return (self.data.sum() * self.grid.get_direct().dV * self.grid.nnrG / (2.0 * np.pi) ** 3).item()


def is_number(x):
    return isinstance(x, numbers.Number)
