"""Direct minimisation of the energy over phi = sqrt(rho), restating
``optimization/optimization.py:281-539`` for one spin channel, with the
truncated-Newton direction (``:91-147``), the CG, steepest-descent and
L-BFGS directions of ``math_utils.py:374-455``.

``evaluator(rho, need_e)`` returns (energy, potential).

Checks (tests/test_oracle.py): the GaAs TN golden of test_density_opt.py:46,
the CG betas on a quadratic with an exact line search (A-conjugacy; the
reference's CG-LS and CG-CD forms), and the L-BFGS recursion against the dense
inverse update.
"""
import numpy as np

from .dcsrch import line_search


def get_direction_TN(evaluator, res0, phi, mu, maxfun=50):
    """``optimization.py:91-147``: CG on the Hessian with a forward-difference
    Hessian-vector product, eps = 1e-9."""
    direction = np.zeros_like(res0)
    epsi = 1.0e-9
    res = -res0.copy()
    p = res.copy()
    r0Norm = np.einsum("ijk,ijk->", res, res)
    r1Norm = r0Norm
    rConv = r0Norm * 0.1
    best = direction
    rLists = [1e10]
    it = 0
    for it in range(maxfun):
        phi1 = phi + epsi * p
        rho1 = phi1 * phi1
        _, v = evaluator(rho1, False)
        Ap = ((v - mu) * phi1 - res0) / epsi
        pAp = np.einsum("ijk,ijk->", p, Ap)
        if pAp < 0.0:
            if it == 0:
                direction = r0Norm / pAp * p
            break
        alpha = r0Norm / pAp
        direction = direction + alpha * p
        res = res - alpha * Ap
        r1Norm = np.einsum("ijk,ijk->", res, res)
        if r1Norm < min(rLists):
            best = direction.copy()
        rLists.append(r1Norm)
        if r1Norm < rConv:
            break
        elif r1Norm > 1000 * min(rLists[:-1]):
            direction = best
            break
        elif it > 10 and abs(r0Norm - r1Norm) < 0.1 * r0Norm:
            direction = best
            break
        beta = r1Norm / r0Norm
        r0Norm = r1Norm
        p = res + beta * p
    return direction, it + 1


def get_direction_CG(resA, dirA, method="CG-HS"):
    """``math_utils.py:395-423``.  The reference's CG-LS divides by
    sum(d r0) (no minus sign) and its CG-CD numerator -(r1 ** 2) is not
    summed, so CG-CD's beta is a field; both are kept as written there."""
    if len(resA) == 1:
        beta = 0.0
    else:
        r1, r0 = resA[-1], resA[-2]
        if method in ("CG-HS", "CG") and len(dirA) > 0:
            beta = np.sum(r1 * (r1 - r0)) / np.sum(dirA[-1] * (r1 - r0))
        elif method == "CG-PR":
            beta = max(np.sum(r1 * (r1 - r0)) / np.sum(r0 ** 2), 0.0)
        elif method == "CG-DY" and len(dirA) > 0:
            beta = np.sum(r1 ** 2) / np.sum(dirA[-1] * (r1 - r0))
        elif method == "CG-CD" and len(dirA) > 0:
            beta = -(r1 ** 2) / np.sum(dirA[-1] * r0)
        elif method == "CG-LS" and len(dirA) > 0:
            beta = np.sum(r1 * (r1 - r0)) / np.sum(dirA[-1] * r0)
        else:
            beta = np.sum(r1 ** 2) / np.sum(r0 ** 2)
    if len(dirA) == 0:
        return -resA[-1]
    return -resA[-1] + beta * dirA[-1]


class LBFGS:
    """``math_utils.py:374-392``: at most bound + 1 (s, y) pairs, oldest
    dropped first, rho_i = 1 / sum(y_i s_i)."""

    def __init__(self, h0=1.0, bound=5):
        self.bound, self.h0 = bound, h0
        self.s, self.y, self.rho = [], [], []

    def update(self, dx, dg):
        if len(self.s) > self.bound:
            self.s.pop(0)
            self.y.pop(0)
            self.rho.pop(0)
        self.s.append(dx)
        self.y.append(dg)
        self.rho.append(1.0 / np.sum(dg * dx))


def get_direction_LBFGS(res, lbfgs):
    """``math_utils.py:431-455``: the two-loop recursion as the reference
    writes it (its first loop stops before the oldest pair, whose alpha
    stays 0); the initial matrix is h0 I, or gamma I with
    gamma = s.y / y.y of the newest pair when h0 is falsy."""
    m = len(lbfgs.s)
    q = -res
    alpha = np.zeros(m)
    for i in range(m - 1, 0, -1):
        alpha[i] = lbfgs.rho[i] * np.sum(lbfgs.s[i] * q)
        q = q - alpha[i] * lbfgs.y[i]
    if not lbfgs.h0:
        gamma = 1.0 if m < 1 else np.sum(lbfgs.s[-1] * lbfgs.y[-1]) / np.sum(lbfgs.y[-1] * lbfgs.y[-1])
        d = gamma * q
    else:
        d = lbfgs.h0 * q
    for i in range(m):
        beta = lbfgs.rho[i] * np.sum(lbfgs.y[i] * d)
        d = d + lbfgs.s[i] * (alpha[i] - beta)
    return d


def optimize(evaluator, rho, dV, method="TN", econv=1e-5, maxiter=200, maxfun=50, maxls=30,
             c1=1e-4, c2=0.2, xtol=1e-12, ncheck=2, maxcor=6, h0=1.0, log=None, callback=None):
    """``optimization.py:281-520``.  Returns dict with rho, energy, mu,
    converged flag (0 ok, 1 line-search failure, 2 maxiter), history and the
    per-iteration (Nd, Nls) counts."""
    integral = lambda f: f.sum() * dV
    phi = rho.copy()
    m = rho > 0
    phi[m] = np.sqrt(rho[m])
    phi[~m] = 1e-300
    phi *= np.sqrt(rho.sum() / (phi * phi).sum())
    rho = phi * phi
    energy, pot = evaluator(rho, True)
    N = integral(rho)
    mu = integral(pot * rho) / N
    residual = (pot - mu) * phi
    resA = [residual]
    dirA = []
    history = [energy]
    counts = []
    theta = 0.1
    converged = 2
    lbfgs = LBFGS(h0, maxcor) if method == "LBFGS" else None
    for it in range(1, maxiter):
        if method == "TN":
            p, nd = get_direction_TN(evaluator, resA[-1], phi, mu, maxfun)
        elif method == "LBFGS":
            p, nd = get_direction_LBFGS(resA[-1], lbfgs), 1
        elif method == "GD":
            p, nd = -resA[-1], 1
        else:
            p, nd = get_direction_CG(resA, dirA, method), 1
        # OrthogonalNormalization (optimization.py:204-225)
        Ne = integral(phi * phi)
        p = p - (integral(p * phi) / Ne) * phi
        pNorm = integral(p * p)
        theta0 = np.sqrt(pNorm / Ne)
        p = p * np.sqrt(Ne / pNorm)
        d0 = np.einsum("ijk,ijk,ijk->", pot, phi, p) * 2.0
        if d0 > 0:
            p = -resA[-1]
            Ne = integral(phi * phi)
            p = p - (integral(p * phi) / Ne) * phi
            pNorm = integral(p * p)
            theta0 = np.sqrt(pNorm / Ne)
            p = p * np.sqrt(Ne / pNorm)
        theta = min(theta0, theta)

        def fvd(th, cur=None):
            # ValueAndDerivative (optimization.py:227-279), EMM
            newphi = phi * np.cos(th) + p * np.sin(th)
            newrho = newphi * newphi
            if cur is None:
                e, v = evaluator(newrho, True)
            else:
                e, v = cur
            p2 = p * np.cos(th) - phi * np.sin(th)
            grad = 2.0 * np.einsum("ijk,ijk,ijk->", v, newphi, p2) * dV
            return [e, grad, newphi, (e, v)]

        func0 = fvd(0.0, (energy, pot))
        theta, task, nls, vd = line_search(fvd, theta, func0, c1=c1, c2=c2, amax=np.pi, amin=0.0,
                                           xtol=xtol, maxiter=maxls)
        if theta is None:
            converged = 1
            break
        old_phi, phi = phi, vd[2]
        energy, pot = vd[3]
        rho = phi * phi
        mu = integral(pot * rho) / integral(rho)
        residual = (pot - mu) * phi
        resA.append(residual)
        if lbfgs is not None:
            lbfgs.update(phi - old_phi, resA[-1] - resA[-2])
        history.append(energy)
        counts.append((nd, nls))
        if log is not None:
            log(it, energy, history[-1] - history[-2], integral(residual * residual), nd, nls)
        if callback is not None:
            callback(it, energy)
        # check_converge (optimization.py:522-535)
        if len(history) - 1 >= ncheck and all(abs(history[-2 - i] - energy) <= econv for i in range(ncheck)):
            converged = 0
            break
        dirA.append(p)
        if len(resA) > 2:
            resA.pop(0)
        if len(dirA) > 2:
            dirA.pop(0)
    return dict(rho=rho, phi=phi, energy=energy, mu=mu, converged=converged, history=history,
                counts=counts, potential=pot)
