"""Host-side tables of the LMGP nonlocal kinetic functional
(functional/kedf/lwt.py:501-578 with kerneltype 'MGP').

The MGP kernel K(eta) on eta = linspace(0, 50, 50000) is computed on the GPU
(``ofdft_mgp_kernel_table``: the t-integral of kernel.py:266-322 with 1000
terms); it and -eta dK/deta are then made interpolating cubic splines with
not-a-knot ends (the reference's splrep(k=3), lwt.py:551-553) as
per-interval coefficients for the device.  The tables depend on no grid and
are built once per process.
"""
import numpy as np

from .. import _lib
from ..math_utils import cubic_table, lindhard

ETAMAX = 50.0
NETA = 50000
MAXPOINTS = 1000
_TABLES = None


def _spline_coefs(x, y):
    """Not-a-knot interpolating cubic through (x, y) on any increasing grid
    (FITPACK splrep(k=3, s=0)), per interval: c0 + c1 t + c2 t^2 + c3 t^3."""
    from scipy.interpolate import CubicSpline
    c = CubicSpline(x, y, bc_type="not-a-knot").c
    return np.ascontiguousarray(np.stack([c[3], c[2], c[1], c[0]], axis=1).reshape(-1))


def lmgp_tables(etamax=ETAMAX, neta=NETA, maxpoints=MAXPOINTS):
    """(eta, coefficients of K, coefficients of -eta K') for ofdft_set_lmgp."""
    global _TABLES
    key = (etamax, neta, maxpoints)
    if _TABLES is not None and _TABLES[0] == key:
        return _TABLES[1]
    eta = np.linspace(0.0, etamax, neta)
    # the reference evaluates G through a cubic spline on a log grid from
    # eta[1] / 10 to eta[-1] + 2 (kernel.py:282-291); the same interpolant goes
    # to the device
    gp = np.logspace(np.log10(eta[1] / 10), np.log10(eta[-1] + 2), num=neta)
    gc = _spline_coefs(gp, lindhard(gp, 1.0, 1.0))
    k = np.zeros(neta)
    _lib.check(_lib.lib.ofdft_mgp_kernel_table(int(neta), float(etamax), int(maxpoints), int(gp.size),
                                               _lib.dptr(gp), _lib.dptr(gc), _lib.dptr(k)))
    kc = cubic_table(eta, k)
    h = eta[1] - eta[0]
    # derivative of the spline at the knots: c1 of each piece, and at the last
    # knot the end slope of the last piece
    dk = np.empty(neta)
    dk[:-1] = kc[:, 1]
    last = kc[-1]
    dk[-1] = last[1] + 2.0 * last[2] * h + 3.0 * last[3] * h * h
    dc = cubic_table(eta, dk * (-1.0 * eta))
    out = (eta, np.ascontiguousarray(kc.reshape(-1)), np.ascontiguousarray(dc.reshape(-1)))
    _TABLES = (key, out)
    return out
