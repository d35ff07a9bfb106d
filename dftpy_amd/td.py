"""TD-OFDFT: Hamiltonian and the Crank-Nicolson propagator
(td/hamiltonian.py:12-179, td/propagator/crank_nicholson.py:59-188,
td/utils.py:10-16, utils/utils.py:40-41 of the reference).  The CN solve
with its Hamiltonian applications runs in ``ofdft_cn_step``."""
import numpy as np
import torch

from . import _lib
from .field import DirectField


class Hamiltonian:
    """H = -1/2 lap_sigma + v for a real potential v (no vector potential)."""

    def __init__(self, v=None, A=None, full=True):
        if A is not None:
            raise NotImplementedError("vector potentials are not implemented")
        self.v = v
        self.full = full

    @property
    def grid(self):
        return self.v.grid if self.v is not None else None

    def __call__(self, psi, force_real=None, sigma=0.025):
        return -0.5 * psi.laplacian(force_real=False, sigma=sigma) + self.v * psi

    def energy(self, psi, sigma=0.025):
        hp = self(psi, sigma=sigma)
        return (torch.real(psi.data.conj() * hp.data).sum() * psi.grid.dV).item()


class CrankNicholson:
    """(1 + i H dt/2) psi1 = (1 - i H dt/2) psi0 with the complex-symmetric CG
    (linear_solver.py:24-47); ``linearsolver='cg'`` only."""

    def __init__(self, hamiltonian, interval, linearsolver="cg", tol=1e-8, maxiter=100, atol=None, sigma=0.025,
                 **kwargs):
        if linearsolver != "cg":
            raise NotImplementedError("only the 'cg' linear solver is implemented")
        if atol is not None:
            raise NotImplementedError("atol is not supported")
        self.hamiltonian = hamiltonian
        self.interval = interval
        self.tol = tol
        self.maxiter = maxiter
        self.sigma = sigma
        self.last_info = None

    def __call__(self, psi0):
        grid = psi0.grid
        plan = grid.plan(grid.dtype)
        ctype = torch.complex128 if grid.dtype == torch.float64 else torch.complex64
        src = psi0.data.to(ctype).contiguous()
        v = self.hamiltonian.v.data.to(grid.dtype).contiguous()
        out = torch.empty_like(src)
        info = np.zeros(4)
        _lib.check(_lib.lib.ofdft_cn_step(plan.handle, src.data_ptr(), v.data_ptr(), float(self.interval),
                                          float(self.sigma), float(self.tol), int(self.maxiter), out.data_ptr(),
                                          _lib.dptr(info), _lib.stream()))
        self.last_info = dict(status=int(info[0]), iterations=int(info[1]), residual=info[2], atol=info[3])
        return DirectField(grid, data=out, cplx=True), int(info[0])


class CrankNicholsonBatch:
    """Crank-Nicolson steps of several independent runs at once
    (``ofdft_cn_step_batch``): each run has its own orbital, potential and
    CG scalars; every Hamiltonian application is one batched transform pair.
    The reference propagates one run per process
    (td/propagator/crank_nicholson.py:147-188)."""

    def __init__(self, interval, tol=1e-8, maxiter=100, sigma=0.025):
        self.interval = interval
        self.tol = tol
        self.maxiter = maxiter
        self.sigma = sigma
        self._work = {}
        self.last_info = None

    def __call__(self, grid, psi0, v):
        """psi0: (nb, nx, ny, nz) complex tensor, v: (nb, nx, ny, nz) real.
        Returns psi1 and the per-run status list."""
        plan = grid.plan(grid.dtype)
        nb = int(psi0.shape[0])
        ctype = torch.complex128 if grid.dtype == torch.float64 else torch.complex64
        src = psi0.to(ctype).contiguous()
        vv = v.to(grid.dtype).contiguous()
        out = torch.empty_like(src)
        nbytes = int(_lib.lib.ofdft_cn_batch_workspace_bytes(plan.handle, nb))
        key = (id(plan), nb)
        if key not in self._work:
            self._work = {key: torch.empty(nbytes, dtype=torch.uint8, device=src.device)}
        work = self._work[key]
        info = np.zeros(4 * nb)
        _lib.check(_lib.lib.ofdft_cn_step_batch(plan.handle, nb, src.data_ptr(), vv.data_ptr(), float(self.interval),
                                                float(self.sigma), float(self.tol), int(self.maxiter),
                                                out.data_ptr(), _lib.dptr(info), work.data_ptr(), nbytes,
                                                _lib.stream()))
        self.last_info = [dict(status=int(info[4 * r]), iterations=int(info[4 * r + 1]), residual=info[4 * r + 2],
                               atol=info[4 * r + 3]) for r in range(nb)]
        return out, [d["status"] for d in self.last_info]


class Propagator:
    """``Propagator(hamiltonian, interval, name='crank-nicholson')``."""

    def __new__(cls, *args, name=None, **kwargs):
        if name in ("crank-nicholson", "CrankNicholson", "cn"):
            return CrankNicholson(*args, **kwargs)
        raise AttributeError("{0} is not a supported propagator type".format(name))


def initial_kick(k, direction, psi, mask=None):
    """psi * exp(i k x) (td/utils.py:10-16)."""
    x = psi.grid.r[direction]
    if mask is not None:
        x = x * mask.data
    ph = torch.exp(1j * k * x.to(torch.float64))
    ctype = torch.complex128 if psi.grid.dtype == torch.float64 else torch.complex64
    return DirectField(psi.grid, data=(psi.data.to(ctype) * ph.to(ctype)), cplx=True)


def calc_rho(psi, N=1):
    """|psi|^2 (utils/utils.py:40-41)."""
    return DirectField(psi.grid, data=(torch.real(psi.data * psi.data.conj()) * N).contiguous())


def dipole(rho, rho0):
    """int (rho - rho0) r dV (real_time_runner.py:251-253)."""
    d = (rho.data - rho0.data).to(torch.float64)
    r = rho.grid.r.to(torch.float64)
    return (torch.einsum("xyz,kxyz->k", d, r) * rho.grid.dV).cpu().numpy()
