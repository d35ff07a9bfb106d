"""Local pseudopotential on the grid and the Ewald ion-ion energy, FP64.

Restates ``functional/pseudo/__init__.py`` (vloc lines, exact and PME
structure factors), ``ewald.py`` (cardinal B-splines, Ewald sums) and
``ions.py:121-124`` (structure factor).
"""
import numpy as np
from scipy import special as sp
from scipy.interpolate import splrep, splev
from scipy.spatial.distance import cdist


def quartic(f, dx):
    """``math_utils.py:457-467``."""
    t0 = f[2]
    t1 = 2.0 * f[0] - 16.0 * f[1] + 16.0 * f[3] - 2.0 * f[4]
    t2 = -f[0] + 16.0 * f[1] - 30.0 * f[2] + 16.0 * f[3] - f[4]
    t3 = -2.0 * f[0] + 4.0 * f[1] - 4.0 * f[3] + 2.0 * f[4]
    t4 = f[0] - 4.0 * f[1] + 6.0 * f[2] - 4.0 * f[3] + f[4]
    return t0 + dx * (t1 + dx * (t2 + dx * (t3 + dx * t4))) / 24.0


def vloc_line(q, gp, vp):
    """vloc(|G|): cubic spline (splrep on gp[1:], ``pseudo/__init__.py:641-658``)
    for q < gp[-1], quartic patch below gp[1] (``:350-375``), zero beyond."""
    tck = splrep(gp[1:], vp[1:], k=3)
    out = np.zeros_like(q)
    m = q < gp[-1]
    if m.any():
        out[m] = splev(q[m], tck, der=0)
    m = q < gp[1]
    dp = gp[1] - gp[0]
    f = [vp[2], vp[1], vp[0], vp[1], vp[2]]
    out[m] = quartic(f, q[m] / dp)
    return out


def strf(grid, pos):
    """exp(-i G.R) on the half grid (``ions.py:121-124``)."""
    return np.exp(-1j * np.einsum("lijk,l->ijk", grid.g, pos))


# ---------------------------------------------------------------- B-splines
def calc_Mn(x, order=10):
    """Cardinal B-spline values M_n(x + k), k = 0..order-1, stored at index
    k+1 (``ewald.py:66-86``)."""
    Mn = np.zeros(order + 1)
    Mn[1] = x
    Mn[2] = 1.0 - x
    for i in range(3, order + 1):
        for j in range(0, i):
            n = i - j
            Mn[n] = ((x + n - 1) * Mn[n] + (j + 1 - x) * Mn[n - 1]) / (i - 1)
    return Mn


def bspline_bm(grid, order=10):
    """Per-axis Euler exponential-spline factors (``ewald.py:88-104``)."""
    Mn = calc_Mn(1.0, order)
    bm = []
    nrG = grid.nrG
    for i in range(3):
        q = 2.0 * np.pi * np.arange(nrG[i]) / grid.nr[i]
        tmp = np.exp(-1j * (order - 1.0) * q)
        fac = np.zeros_like(tmp)
        for k in range(1, order):
            fac += Mn[k] * np.exp(-1j * k * q)
        bm.append(tmp / fac)
    return bm


def barray(grid, order=10):
    bm = bspline_bm(grid, order)
    return np.einsum("i,j,k->ijk", bm[0], bm[1], bm[2])


def pme_qarray(grid, scaled_positions, weights, order=10):
    """Spread point weights with order-n cardinal B-splines
    (``ewald.py:151-194``): Q[floor(u)+1-i] += M(frac(u)+i) per axis."""
    nr = grid.nr
    Q = np.zeros(nr)
    ix = np.arange(order)
    for s, w in zip(scaled_positions, weights):
        Up = s * nr
        fl = np.floor(Up).astype(np.int64)
        M = [calc_Mn(Up[j] - np.floor(Up[j]), order)[1:] for j in range(3)]
        lx = np.mod(1 + fl[0] - ix, nr[0])
        ly = np.mod(1 + fl[1] - ix, nr[1])
        lz = np.mod(1 + fl[2] - ix, nr[2])
        blk = np.einsum("i,j,k->ijk", w * M[0], M[1], M[2])
        np.add.at(Q, np.ix_(lx, ly, lz), blk)
    return Q


# ---------------------------------------------------------------- local PP
def local_pp_v(grid, atoms, pp, pme=True, order=10):
    """Reciprocal local potential v(G) (``pseudo/__init__.py:377-404``).
    ``pp`` maps species -> (gp, vp)."""
    q = grid.q
    v = np.zeros(grid.nrG, dtype=np.complex128)
    spos = atoms.scaled_positions()
    for key in sorted(pp):
        gp, vp = pp[key]
        line = vloc_line(q, gp, vp)
        idx = [i for i, s in enumerate(atoms.symbols) if s == key]
        if not idx:
            continue
        if pme:
            Q = pme_qarray(grid, spos[idx], np.ones(len(idx)), order)
            v = v + line * grid.fft(Q)
        else:
            for i in idx:
                v = v + line * strf(grid, atoms.positions[i])
    if pme:
        v = v * barray(grid, order) * grid.nnr / grid.volume
    return v


def local_pp(grid, atoms, pp, pme=True, order=10):
    """Returns vreal = IFFT(v(G)) (``pseudo/__init__.py:236-246``)."""
    return grid.ifft(local_pp_v(grid, atoms, pp, pme, order))


# ---------------------------------------------------------------- Ewald
def ewald_eta(grid, atoms, precision=1e-8):
    """``ewald.py:256-279``."""
    gmax = np.sqrt(2.0 * grid.ecut())
    chargeSquare = np.sum(atoms.charges * atoms.charges)
    eta = 1.6
    while True:
        ub = 4.0 * np.pi * atoms.nat * chargeSquare * np.sqrt(eta / np.pi) * sp.erfc(gmax / 2.0 * np.sqrt(1.0 / eta))
        if ub < precision:
            return eta
        eta = eta - 0.01


def ewald_real(grid, atoms, eta, precision=1e-8):
    """``ewald.py:280-320``."""
    L = np.sqrt(np.einsum("ij->i", grid.lattice ** 2))
    rmax = sp.erfcinv(precision / 3.0) / np.sqrt(eta)
    N = np.ceil(rmax / L)
    charges, positions = [], []
    for ix in np.arange(-N[0], N[0] + 1):
        for iy in np.arange(-N[1], N[1] + 1):
            for iz in np.arange(-N[2], N[2] + 1):
                R = np.array([ix, iy, iz]) @ grid.lattice
                for i in range(atoms.nat):
                    charges.append(atoms.charges[i])
                    positions.append(atoms.positions[i] - R)
    charges = np.asarray(charges)
    positions = np.asarray(positions)
    etaSqrt = np.sqrt(eta)
    Esum = 0.0
    for i in range(atoms.nat):
        d = cdist(positions, atoms.positions[i].reshape((1, 3))).ravel()
        m = np.logical_and(d < rmax, d > 0.001)
        Esum += atoms.charges[i] * np.sum(charges[m] * sp.erfc(etaSqrt * d[m]) / d[m])
    return Esum / 2.0


def ewald_rec(grid, atoms, eta):
    """Exact structure factor form (``ewald.py:448-465``)."""
    s = np.zeros(grid.nrG, dtype=np.complex128)
    for i in range(atoms.nat):
        s += strf(grid, atoms.positions[i]) * atoms.charges[i]
    s2 = np.conjugate(s) * s
    m = grid.mask
    e = np.sum(s2[m] * np.exp(-grid.gg[m] / (4.0 * eta)) * grid.invgg[m])
    return 4.0 * np.pi * e.real / grid.volume


def ewald_rec_pme(grid, atoms, eta, order=10):
    """PME form (``ewald.py:786-808``)."""
    Q = pme_qarray(grid, atoms.scaled_positions(), atoms.charges, order)
    s = grid.fft(Q) * barray(grid, order)
    s2 = np.conjugate(s) * s
    m = grid.mask
    e = np.sum(s2[m] * np.exp(-grid.gg[m] / (4.0 * eta)) * grid.invgg[m])
    return 4.0 * np.pi * e.real / grid.volume / grid.dV ** 2


def ewald_corr(grid, atoms, eta):
    """``ewald.py:467-482``."""
    dc = -np.sqrt(eta / np.pi) * np.sum(atoms.charges ** 2)
    gz = -4.0 * np.pi * (1.0 / (4.0 * eta * grid.volume) / 2.0) * atoms.ncharge() ** 2
    return dc + gz


def ewald_energy(grid, atoms, pme=True, precision=1e-8, order=10):
    eta = ewald_eta(grid, atoms, precision)
    e_rec = ewald_rec_pme(grid, atoms, eta, order) if pme else ewald_rec(grid, atoms, eta)
    return ewald_corr(grid, atoms, eta) + ewald_real(grid, atoms, eta, precision) + e_rec
