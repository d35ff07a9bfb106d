"""TD-OFDFT: Hamiltonian and the Crank-Nicolson propagator
(td/hamiltonian.py:12-179, td/propagator/crank_nicholson.py:59-188,
td/utils.py:10-16, utils/utils.py:40-41 of the reference).  The CN solve
with its Hamiltonian applications runs in ``ofdft_cn_step``."""
import numpy as np
import torch

from . import _lib
from .field import DirectField, ReciprocalField


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


# linear solvers of crank_nicholson.py:65-77 available here -> ofdft_cn_step_solver code
LINEAR_SOLVERS = {"cg": 0, "bicg": 1, "bicgstab": 2, "gmres": 3, "gmres_scipy": 3}


class CrankNicholson:
    """(1 + i H dt/2) psi1 = (1 - i H dt/2) psi0 (crank_nicholson.py:59-188)
    with the linear solver ``linearsolver``: 'cg' (complex-symmetric CG,
    linear_solver.py:24-47), 'bicg', 'bicgstab' (linear_solver.py:49-110) or
    'gmres' / 'gmres_scipy' (restarted GMRES, restart 20 as scipy's)."""

    def __init__(self, hamiltonian, interval, linearsolver="cg", tol=1e-8, maxiter=100, atol=None, sigma=0.025,
                 restart=20, **kwargs):
        if linearsolver not in LINEAR_SOLVERS:
            raise NotImplementedError("linear solver %r is not implemented (choose from %s)" %
                                      (linearsolver, sorted(LINEAR_SOLVERS)))
        if atol is not None:
            raise NotImplementedError("atol is not supported")
        self.hamiltonian = hamiltonian
        self.interval = interval
        self.linear_solver = linearsolver
        self.tol = tol
        self.maxiter = maxiter
        self.sigma = sigma
        self.restart = int(restart)
        self.last_info = None
        self._ws = None

    def __call__(self, psi0):
        grid = psi0.grid
        plan = grid.plan(grid.dtype)
        ctype = torch.complex128 if grid.dtype == torch.float64 else torch.complex64
        src = psi0.data.to(ctype).contiguous()
        v = self.hamiltonian.v.data.to(grid.dtype).contiguous()
        out = torch.empty_like(src)
        info = np.zeros(4)
        code = LINEAR_SOLVERS[self.linear_solver]
        if code == 0:
            _lib.check(_lib.lib.ofdft_cn_step(plan.handle, src.data_ptr(), v.data_ptr(), float(self.interval),
                                              float(self.sigma), float(self.tol), int(self.maxiter), out.data_ptr(),
                                              _lib.dptr(info), _lib.stream()))
        else:
            nbytes = int(_lib.lib.ofdft_cn_solver_workspace_bytes(plan.handle, code, self.restart))
            if self._ws is None or self._ws.numel() < nbytes or self._ws.device != src.device:
                self._ws = torch.empty(nbytes, dtype=torch.uint8, device=src.device)
            _lib.check(_lib.lib.ofdft_cn_step_solver(plan.handle, code, self.restart, src.data_ptr(), v.data_ptr(),
                                                     float(self.interval), float(self.sigma), float(self.tol),
                                                     int(self.maxiter), out.data_ptr(), _lib.dptr(info),
                                                     self._ws.data_ptr(), nbytes, _lib.stream()))
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


def calc_j(psi, N=1):
    """Current density (utils/utils.py:44-47): Re(-i/2 (conj(psi) grad psi -
    psi grad conj(psi))) with the spectral gradients IFFT(i G FFT(.)) of
    field.py:267-281 on the full grid (HIP C2C transforms).  grad conj(psi)
    is transformed on its own, as the reference does: on an even grid the
    Nyquist plane makes it differ from conj(grad psi).  Returns a
    (3, nx, ny, nz) real tensor."""
    # conj_physical: a lazy conj() view would hand the library the unconjugated memory
    pc = DirectField(psi.grid, data=psi.data.conj_physical().contiguous(), cplx=True)
    f, fc = psi.fft(), pc.fft()
    G = f.grid.g
    out = []
    for k in range(3):
        ig = (1j * torch.as_tensor(G[k], dtype=torch.float64, device=f.data.device)).to(f.data.dtype)
        gp = ReciprocalField(f.grid, data=(ig * f.data).contiguous(), cplx=True).ifft(force_real=False).data
        gc = ReciprocalField(f.grid, data=(ig * fc.data).contiguous(), cplx=True).ifft(force_real=False).data
        out.append(torch.real(-0.5j * (pc.data * gp - psi.data * gc)) * N)
    return torch.stack(out)


def _norm_rho(rho, dV):
    return float(torch.sqrt((rho.to(torch.float64) ** 2).sum() * dV))


def _norm_j(j, dV):
    return float(torch.sqrt((j.to(torch.float64) ** 2).sum(dim=(1, 2, 3)) * dV).max())


class PredictorCorrector:
    """Predictor-corrector time step (td/predictor_corrector.py:10-150):
    repeat up to ``max_steps`` times: psi_pred = CN(psi) with the current
    Hamiltonian potential, then the potential of ``functionals`` at
    (rho + rho_pred) / 2 becomes the Hamiltonian's; stop when rho_pred and
    j_pred change by less than tol * |rho| and tol * max|j_k| (or atol)
    between two predictions.  ``run()`` returns the convergence flag; the new
    state is ``psi_pred``, ``rho_pred``, ``j_pred``.  Vector potentials are
    not implemented."""

    def __init__(self, psi, propagator=None, tol=1.0e-8, atol=None, max_steps=2, propagate_vector_potential=False,
                 functionals=None, Omega=None, A_t=None, A_tm1=None, N0=None):
        if propagate_vector_potential or A_t is not None:
            raise NotImplementedError("vector potentials are not implemented")
        self.propagator, self.functionals = propagator, functionals
        self.max_steps, self.tol, self.atol = max_steps, tol, atol
        self.nsteps = 0
        self.psi = psi
        self.dV = psi.grid.dV
        self.rho = calc_rho(psi).data
        self.j = calc_j(psi)
        self.atol_rho = _get_atol(tol, atol, _norm_rho(self.rho, self.dV))
        self.atol_j = _get_atol(tol, atol, _norm_j(self.j, self.dV))
        self.old_rho_pred = torch.zeros_like(self.rho)
        self.rho_pred = self.rho
        self.old_j_pred = torch.zeros_like(self.j)
        self.j_pred = self.j
        self.psi_pred = None
        self.diff_rho = self.diff_j = None

    def step(self):
        if self.nsteps > 0:
            self.old_rho_pred = self.rho_pred
            self.old_j_pred = self.j_pred
        self.psi_pred, info = self.propagator(self.psi)
        self.rho_pred = calc_rho(self.psi_pred).data
        self.j_pred = calc_j(self.psi_pred)
        rho_corr = DirectField(self.psi.grid, data=((self.rho + self.rho_pred) * 0.5).contiguous())
        self.propagator.hamiltonian.v = self.functionals(rho_corr, calcType={"V"}).potential

    def converged(self):
        self.diff_rho = _norm_rho(self.old_rho_pred - self.rho_pred, self.dV)
        self.diff_j = _norm_j(self.old_j_pred - self.j_pred, self.dV)
        return self.diff_rho < self.atol_rho and self.diff_j < self.atol_j

    def run(self):
        # Dynamics.irun (optimize.py:85-128)
        while not self.converged() and self.nsteps < self.max_steps:
            self.step()
            self.nsteps += 1
        return self.converged()


def _get_atol(tol, atol, bnrm2):
    """linear_solver.py:8-17."""
    tol = float(tol)
    if atol is None:
        return tol if bnrm2 == 0 else tol * float(bnrm2)
    return max(float(atol), tol * float(bnrm2))
