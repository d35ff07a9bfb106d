"""TD-OFDFT Crank-Nicolson propagation, complex128, restating
``td/hamiltonian.py:70-114``, ``field.py:343-350`` (smoothed Laplacian),
``td/propagator/crank_nicholson.py:44-56,147-188``, ``linear_solver.py:8-47``
(CG without conjugation), ``utils/utils.py:40-41`` (calc_rho) and
``td/utils.py:10-16`` (delta kick), ``linear_solver.py:49-110`` (bicg,
bicgstab) and scipy's gmres (the ``gmres_scipy`` solver), ``utils/utils.py:44-47`` (current
density) and ``td/predictor_corrector.py:84-150`` (predictor-corrector step)."""
import numpy as np


def hamiltonian_apply(psi, v, grid, sigma=0.025):
    """H psi = -1/2 lap_sigma(psi) + v psi, lap_sigma = IFFT(-G^2 e^{-G^2 s^2/2} FFT)."""
    gg = grid.ggF
    lap = np.fft.ifftn(-gg * np.fft.fftn(psi) * np.exp(-gg * sigma ** 2 / 2.0))
    return -0.5 * lap + v * psi


def _norm(x):
    return np.sqrt(np.real(np.sum(np.conj(x) * x)))


def cg(A, b, x0, tol=1e-8, maxiter=100, atol=None):
    """``linear_solver.py:24-47``: complex-symmetric CG, sums without conj."""
    bn = _norm(b)
    atol = (tol * bn if bn != 0 else tol) if atol is None else max(atol, tol * bn)
    res = b - A(x0)
    if _norm(res) < atol:
        return x0, 0, 0
    p = res
    x = x0
    k = 0
    while k < maxiter:
        v = A(p)
        alpha = np.sum(res * res) / np.sum(p * v)
        x = x + alpha * p
        old = res
        res = old - alpha * v
        if _norm(res) < atol:
            return x, 0, k + 1
        beta = np.sum(res * res) / np.sum(old * old)
        p = res + beta * p
        k += 1
    return x, 1, k


def bicg(A, b, x0, tol=1e-8, maxiter=100):
    """``linear_solver.py:49-75``: CG with conjugated sums."""
    bn = _norm(b)
    atol = tol * bn if bn != 0 else tol
    res = b - A(x0)
    if _norm(res) < atol:
        return x0, 0, 0
    p, x, k = res, x0, 0
    while k < maxiter:
        v = A(p)
        rr = np.sum(np.conj(res) * res)
        alpha = rr / np.sum(np.conj(p) * v)
        x = x + alpha * p
        new = res - alpha * v
        if _norm(new) < atol:
            return x, 0, k + 1
        beta = np.sum(np.conj(new) * new) / rr
        p = new + beta * p
        res = new
        k += 1
    return x, 1, k


def bicgstab(A, b, x0, tol=1e-8, maxiter=100):
    """``linear_solver.py:78-110`` (sums without conjugation)."""
    bn = _norm(b)
    atol = tol * bn if bn != 0 else tol
    r0 = b - A(x0)
    if _norm(r0) < atol:
        return x0, 0, 0
    rho_old, omega, alpha = 1.0, 1.0, 1.0
    v = 0 * b
    p = v
    x, res, r, k = x0, r0, r0, 0
    while k < maxiter:
        rho = np.sum(r0 * r)
        beta = rho / rho_old * alpha / omega
        p = r + beta * (p - omega * v)
        v = A(p)
        alpha = rho / np.sum(r0 * v)
        h = x + alpha * p
        res = res - alpha * v
        if _norm(res) < atol:
            return h, 0, k + 1
        s = r - alpha * v
        t = A(s)
        omega = np.sum(t * s) / np.sum(t * t)
        x = h + omega * s
        res = res - omega * t
        if _norm(res) < atol:
            return x, 0, k + 1
        r = s - omega * t
        rho_old = rho
        k += 1
    return x, 1, k


def gmres(A, b, x0, tol=1e-8, maxiter=100, restart=20):
    """The ``gmres_scipy`` choice of ``crank_nicholson.py:65-77``: scipy's
    restarted GMRES on the operator, relative tolerance tol (maxiter in
    restart cycles, as scipy counts them)."""
    import scipy.sparse.linalg as sla
    shape = b.shape
    n = b.size
    op = sla.LinearOperator((n, n), matvec=lambda y: A(y.reshape(shape)).ravel(), dtype=np.complex128)
    x, info = sla.gmres(op, b.ravel(), x0=x0.ravel(), rtol=tol, atol=0.0, restart=restart, maxiter=maxiter)
    return x.reshape(shape), info, None


SOLVERS = {"cg": cg, "bicg": bicg, "bicgstab": bicgstab, "gmres": gmres}


def cn_step(psi0, v, grid, dt, tol=1e-8, maxiter=100, sigma=0.025, solver="cg"):
    """``crank_nicholson.py:147-188``: (1 + i H dt/2) psi1 = (1 - i H dt/2) psi0."""
    H = lambda x: hamiltonian_apply(x, v, grid, sigma)
    A = lambda x: x + 1j * H(x) * dt / 2.0
    b = psi0 - 1j * H(psi0) * dt / 2.0
    return SOLVERS[solver](A, b, psi0, tol, maxiter)


def initial_kick(k, direction, psi, grid):
    return psi * np.exp(1j * k * grid.r[direction])


def energy(psi, v, grid, sigma=0.025):
    return np.real(np.conj(psi) * hamiltonian_apply(psi, v, grid, sigma)).sum() * grid.dV


def dipole(rho, rho0, grid):
    """delta mu = int (rho - rho0) r dV (``real_time_runner.py:251-253``)."""
    d = rho - rho0
    return np.array([(d * grid.r[i]).sum() * grid.dV for i in range(3)])


def calc_j(psi, grid):
    """``utils/utils.py:44-47``: Re(-i/2 (conj(psi) grad psi - psi grad conj(psi)))
    with the spectral gradients of ``field.py:267-281`` on the full grid."""
    k = np.meshgrid(*[np.fft.fftfreq(int(n), 1.0 / int(n)) for n in grid.nr], indexing="ij")
    G = np.einsum("jxyz,jk->kxyz", np.array(k), grid.B)
    f, fc = np.fft.fftn(psi), np.fft.fftn(np.conj(psi))
    out = []
    for i in range(3):
        gp, gc = np.fft.ifftn(1j * G[i] * f), np.fft.ifftn(1j * G[i] * fc)
        out.append(np.real(-0.5j * (np.conj(psi) * gp - psi * gc)))
    return np.array(out)


def predictor_corrector(psi, v, grid, dt, potential, tol=1e-8, atol=None, max_steps=2):
    """``td/predictor_corrector.py:84-150`` driven as ``optimize.py:85-128``
    does: returns (psi_pred, v, nsteps, converged), v being the potential of
    ``potential`` at (rho + rho_pred) / 2 after the last prediction."""
    def nrm(x):
        return np.sqrt((x ** 2).sum() * grid.dV)

    def nrm_j(j):
        return np.max(np.sqrt((j ** 2).sum(axis=(1, 2, 3)) * grid.dV))

    def get_atol(b):
        return (tol * b if b != 0 else tol) if atol is None else max(atol, tol * b)

    rho, j = np.abs(psi) ** 2, calc_j(psi, grid)
    at_rho, at_j = get_atol(nrm(rho)), get_atol(nrm_j(j))
    old_rho, old_j, rho_pred, j_pred = np.zeros_like(rho), np.zeros_like(j), rho, j
    psi_pred, n = None, 0

    def conv():
        return nrm(old_rho - rho_pred) < at_rho and nrm_j(old_j - j_pred) < at_j

    while not conv() and n < max_steps:
        if n > 0:
            old_rho, old_j = rho_pred, j_pred
        psi_pred, info, _ = cn_step(psi, v, grid, dt, tol)
        rho_pred, j_pred = np.abs(psi_pred) ** 2, calc_j(psi_pred, grid)
        v = potential((rho + rho_pred) * 0.5)
        n += 1
    return psi_pred, v, n, conv()
