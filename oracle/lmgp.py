"""LMGP nonlocal kinetic functional (density-dependent kernel), FP64,
restating ``functional/kedf/lwt.py:36-82,85-341,501-578`` (LWT with
kerneltype 'MGP' and the reference defaults: linear interpolation in k_F,
kdd = 3, ratio 1.15, etamax 50, neta 50000, maxpoints 1000, no lumpfactor)
and ``functional/kedf/kernel.py:266-322,403-431`` (MGPKernelTable,
LWTKernelKf).  Test infrastructure only.

Pinned by the reference golden 2.9145872498710492 (TF + vW + LMGP on
Al_fde_rho.pp, ``examples/test/test_functional.py:29-34``, atol 1e-3) and the
homogeneous-gas limit (zero energy and potential).
"""
import numpy as np
from scipy.interpolate import splev, splrep

from .functionals import CTF, lindhard

ETAMAX = 50.0
NETA = 50000
MAXPOINTS = 1000
RATIO = 1.15


def mgp_kernel_table(eta, maxpoints=MAXPOINTS):
    """``MGPKernelTable`` (kernel.py:266-322), no symmetrization: the t-integral
    coe * sum_i G(eta / t_i^(1/3)) / t_i^(1/6), t_i = i / maxpoints, with G the
    inverse Lindhard function (lambda = mu = 1) splined on a log grid."""
    dt = 1.0 / maxpoints
    coe = 4.0 / 5.0 * CTF * 5.0 / 6.0 * dt
    gp = np.logspace(np.log10(eta[1] / 10), np.log10(eta[-1] + 2), num=eta.size)
    tck = splrep(gp, lindhard(gp, 1.0, 1.0))
    kernel = np.zeros_like(eta)
    for i in range(1, maxpoints + 1):
        t = i * dt
        kernel += splev(eta / np.cbrt(t), tck, ext=3) / t ** (1.0 / 6.0)
    return kernel * coe


def kernel_tables(etamax=ETAMAX, neta=NETA, maxpoints=MAXPOINTS):
    """Splines of the MGP kernel and of -eta dK/deta (lwt.py:541-557, order 3)."""
    eta = np.linspace(0, etamax, neta)
    k = mgp_kernel_table(eta, maxpoints)
    tck = splrep(eta, k, k=3)
    deriv = splev(eta, tck, der=1) * (-1.0 * eta)
    return eta, k, tck, splrep(eta, deriv, k=3)


def kernel_kf(q, kf, tck, etamax=ETAMAX):
    """``LWTKernelKf`` (kernel.py:403-431) for a spline table."""
    eta = q / kf
    out = np.empty_like(q)
    c0 = eta < etamax
    out[~c0] = splev(etamax, tck)
    if c0.any():
        out[c0] = splev(eta[c0], tck)
    if q.flat[0] < 1e-30:
        out.flat[0] = 0.0
    return out


def guess_kf_bound(kf, state, kftol=1e-3):
    """``guess_kf_bound`` (lwt.py:36-82) with the KEDF's saved bounds."""
    kfmin_prev, kfmax_prev = state.get("kfmin"), state.get("kfmax")
    kf_l, kf_r = kf.min(), kf.max()
    kfmin = kfmax = None
    if kfmin_prev is None:
        kfmin_prev = 10
        kfmin = kf_l
    if kfmax_prev is None:
        kfmax_prev = 1.0
        kfmax = kf_r
    if kfmin is None or kfmin > kfmin_prev:
        kfmin = kfmin_prev
    if kfmax is None or kfmax < kfmax_prev:
        kfmax = kfmax_prev
    if kfmin > kf_l:
        kfl = [1e-5, 1e-4, 5e-3, 1e-3, 5e-3, 1e-2, 5e-2, 0.1, 0.5, 1.0]
        ratio = kf_l / kfmin
        for i in range(len(kfl) - 1, 0, -1):
            if ratio > kfl[i]:
                kfmin *= kfl[i]
                break
    if kfmin < kftol:
        kfmin = kftol
    if kfmax < kf_r:
        kfmax += (np.round((kf_r - kfmax) / 0.5) + 1) * 0.5
    state["kfmin"], state["kfmax"] = kfmin, kfmax
    return kfmin, kfmax


def ladder(kfmin, kfmax, ratio=RATIO):
    nsp = int(np.ceil(np.log(kfmax / kfmin) / np.log(ratio))) + 1
    kfl = kfmin * ratio ** np.arange(nsp)
    kfl[0] -= 1e-16
    kfl[-1] += 1e-16
    return kfl


def lmgp_nl(rho, grid, tables, state=None, alpha=5.0 / 6.0, beta=5.0 / 6.0, ldw=1.0 / 6.0):
    """``LWTPotentialEnergy`` (lwt.py:85-341), linear interpolation, kdd = 3.
    Returns (energy, potential)."""
    state = {} if state is None else state
    _, _, tck, tckd = tables
    q = grid.q
    kf = 2.0 * np.cbrt(3.0 * np.pi ** 2 * rho)
    if abs(kf.max() - kf.min()) < 1e-8:
        return 0.0, np.zeros_like(rho)
    kfmin, kfmax = guess_kf_bound(kf, state)
    kf = np.minimum(kf, kfmax)
    kfl = ladder(kfmin, kfmax)
    nsp = kfl.size
    r = rho.copy()
    r[~(rho > 0)] = 1e-30
    ra = r ** alpha
    ra1 = ra / r
    rb, rb1 = (ra, ra1) if abs(alpha - beta) < 1e-8 else (r ** beta, r ** beta / r)
    rbG = grid.fft(rb)
    pot1 = np.zeros_like(rho)
    pot3 = np.zeros_like(rho)
    pot2G = None

    def samp(j):
        k = kernel_kf(q, kfl[j], tck)
        kd = kernel_kf(q, kfl[j], tckd) / kfl[j]
        return k, grid.ifft(k * rbG), grid.ifft(kd * rbG)

    k0, p0, m0 = samp(0)
    for i in range(nsp - 1):
        k1, p1, m1 = samp(i + 1)
        mask = (kf > kfl[i]) & (kf < kfl[i + 1] + 1e-18)
        if i == 0:
            small = kf < kfl[0] + 1e-18
            t = (kf - kfl[0]) / kfl[0]
            pot1[small] = p0[small] * t[small]
            pot2G = k0 * grid.fft(ra * small * t)
        t = (kf - kfl[i]) / (kfl[i + 1] - kfl[i])
        t0 = 1 - t
        pot1[mask] = p0[mask] * t0[mask] + p1[mask] * t[mask]
        pot2G = pot2G + k0 * grid.fft(ra * mask * t0) + k1 * grid.fft(ra * mask * t)
        pot3[mask] = m0[mask] * t0[mask] + m1[mask] * t[mask]
        k0, p0, m0 = k1, p1, m1
    pot2 = grid.ifft(pot2G)
    factor = np.ones_like(rho)
    low = rho < 1e-6
    ld = max(0.1, ldw)
    factor[low] = np.abs(rho[low]) ** ld / rho.max() ** ld
    factor[rho < 0] = 0.0
    pot1 = pot1 * factor
    energy = float((ra * pot1).sum() * grid.dV)
    pot = pot1 * alpha * ra1 + pot2 * factor * beta * rb1 + pot3 * factor * (kf / 3.0) * ra1
    return energy, pot
