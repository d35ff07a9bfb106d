"""Forces and stress, FP64: Ewald sum, local pseudopotential, TF, vW, WT,
Hartree and PZ-LDA.  Test infrastructure only (tests/, never the product).

Restates ``ewald.py:541-760`` (Forces_real, Forces_rec, Stress_real,
Stress_rec with the G = 0 term), ``ewald.py:810-960`` (Forces_rec_PME,
Stress_rec_PME), ``functional/pseudo/__init__.py:406-563`` (_PP_Derivative,
_Stress, _Force, _ForcePME, _StressPME), ``functional/kedf/tf.py:52-62``,
``functional/kedf/vw.py:97-111``, ``functional/kedf/wt.py:75-105`` with
``functional/kedf/kernel.py:197-211`` (LindhardDerivative),
``functional/hartree.py:67-91`` and ``functional/xc/semilocal_xc.py:308-322``
(LDAStress).  Sums over the half space use the grid's deduplication mask with
the reference's factors.

Pinned by the goldens of ``examples/test/test_ewald.py:38-43`` (Stress_rec
[0, 0] = 8.04967853e-03 and zero forces on the symmetric Al cell) and, for
everything those leave open, by the mathematics: each force against central
differences of the oracle's energy under a displacement of one atom, each
stress term against central differences of its energy under homogeneous
strain (density carried along, electron count fixed), in
tests/test_oracle_forces.py.
"""
import numpy as np
from scipy import special as sp
from scipy.interpolate import splev, splrep
from scipy.spatial.distance import cdist

from .functionals import CTF
from .pseudo import barray, calc_Mn, ewald_eta, pme_qarray, strf, vloc_line


def _images(grid, atoms, rmax):
    L = np.sqrt(np.einsum("ij->i", grid.lattice ** 2))
    N = np.ceil(rmax / L)
    charges, positions = [], []
    for ix in np.arange(-N[0], N[0] + 1):
        for iy in np.arange(-N[1], N[1] + 1):
            for iz in np.arange(-N[2], N[2] + 1):
                R = np.array([ix, iy, iz]) @ grid.lattice
                for i in range(atoms.nat):
                    charges.append(atoms.charges[i])
                    positions.append(atoms.positions[i] - R)
    return np.asarray(charges), np.asarray(positions)


def _sym(v6):
    s = np.zeros((3, 3))
    k = 0
    for i in range(3):
        for j in range(i, 3):
            s[i, j] = s[j, i] = v6[k]
            k += 1
    return s


# ---------------------------------------------------------------- Ewald
def ewald_forces_real(grid, atoms, eta, precision=1e-8):
    """``ewald.py:541-578``."""
    rmax = sp.erfcinv(precision / 3.0) / np.sqrt(eta)
    charges, positions = _images(grid, atoms, rmax)
    es = np.sqrt(eta)
    F = np.zeros((atoms.nat, 3))
    for i in range(atoms.nat):
        d = cdist(positions, atoms.positions[i].reshape((1, 3))).ravel()
        m = np.logical_and(d < rmax, d > 0.001)
        x = d[m] * es
        F[i] = atoms.charges[i] * np.einsum(
            "ij,i->j", (atoms.positions[i] - positions[m]) * charges[m][:, None],
            sp.erfc(x) / x ** 3 + 2.0 / np.sqrt(np.pi) * np.exp(-x ** 2) / x ** 2)
    return F * es ** 3


def _sz(grid, atoms):
    s = np.zeros(grid.nrG, dtype=np.complex128)
    for i in range(atoms.nat):
        s += strf(grid, atoms.positions[i]) * atoms.charges[i]
    return s


def ewald_forces_rec(grid, atoms, eta):
    """``ewald.py:580-607`` (exact structure factors)."""
    s = _sz(grid, atoms)
    m = grid.mask
    w = np.exp(-grid.gg[m] / (4.0 * eta)) * grid.invgg[m]
    F = np.empty((atoms.nat, 3))
    for i in range(atoms.nat):
        si = strf(grid, atoms.positions[i])[m] * atoms.charges[i]
        F[i] = np.einsum("ij,j->i", grid.g[:, m], (si.real * s[m].imag - si.imag * s[m].real) * w)
    return F * 8.0 * np.pi / grid.volume


def ewald_forces(grid, atoms, precision=1e-8):
    eta = ewald_eta(grid, atoms, precision)
    return ewald_forces_real(grid, atoms, eta, precision) + ewald_forces_rec(grid, atoms, eta)


def ewald_stress_real(grid, atoms, eta, precision=1e-8):
    """``ewald.py:609-662``."""
    rmax = sp.erfcinv(precision / 3.0) / np.sqrt(eta)
    charges, positions = _images(grid, atoms, rmax)
    es = np.sqrt(eta)
    st = np.zeros(6)
    for a in range(atoms.nat):
        d = cdist(positions, atoms.positions[a].reshape((1, 3))).ravel()
        m = np.logical_and(d < rmax, d > 0.001)
        R = atoms.positions[a] - positions[m]
        Rv = np.stack([R[:, i] * R[:, j] / d[m] ** 2 for i in range(3) for j in range(i, 3)], axis=1)
        st += atoms.charges[a] * np.einsum(
            "i,ij->j", charges[m] * (2 * es / np.sqrt(np.pi) * np.exp(-eta * d[m] ** 2) + sp.erfc(es * d[m]) / d[m]), Rv)
    return _sym(st * (-0.5 / grid.volume))


def ewald_stress_rec(grid, atoms, eta):
    """``ewald.py:719-760`` (exact structure factors, G = 0 term on the diagonal)."""
    s = _sz(grid, atoms)
    s2 = (np.conjugate(s) * s).real
    m = grid.mask
    gg, invgg, g = grid.gg[m], grid.invgg[m], grid.g[:, m]
    w = s2[m] * np.exp(-gg / (4.0 * eta)) * invgg
    st = []
    for i in range(3):
        for j in range(i, 3):
            f = g[i] * g[j] * 2.0 * invgg * (1 + gg / (4.0 * eta))
            if i == j:
                f = f - 1.0
            st.append(np.sum(w * f))
    S = _sym(np.array(st) * 4.0 * np.pi / grid.volume ** 2)
    S += np.eye(3) * atoms.ncharge() ** 2 * 4.0 * np.pi * (1.0 / (4.0 * eta * grid.volume ** 2) / 2.0)
    return S


def ewald_stress(grid, atoms, precision=1e-8):
    eta = ewald_eta(grid, atoms, precision)
    return ewald_stress_real(grid, atoms, eta, precision) + ewald_stress_rec(grid, atoms, eta)


# ---------------------------------------------------------------- local PP
def vloc_deriv_line(q, gp, vp):
    """d vloc / dq: derivative of the cubic spline of ``vloc_line`` for
    q < gp[-1], zero beyond (``pseudo/__init__.py:406-412``)."""
    tck = splrep(gp[1:], vp[1:], k=3)
    out = np.zeros_like(q)
    m = q < np.max(gp)
    out[m] = splev(q[m], tck, der=1)
    return out


def pp_forces(rho, grid, atoms, pp):
    """``pseudo/__init__.py:456-470``: F_i = (2/V) sum_mask g vloc(q) Im(rho(G) e^{iG.R_i})."""
    rhoG = grid.fft(rho)
    m = grid.mask
    lines = {k: vloc_line(grid.q, *pp[k])[m] for k in pp}
    F = np.zeros((atoms.nat, 3))
    for i in range(atoms.nat):
        istrf = np.conjugate(strf(grid, atoms.positions[i]))[m]
        den = lines[atoms.symbols[i]] * (rhoG[m] * istrf).imag
        F[i] = grid.g[:, m] @ den
    return F * 2.0 / grid.volume


def pp_stress(rho, grid, atoms, pp, energy):
    """``pseudo/__init__.py:414-454`` (exact structure factors)."""
    rhoG = grid.fft(rho)
    v = np.zeros(grid.nrG, dtype=np.complex128)
    for key in sorted(pp):
        d = vloc_deriv_line(grid.q, *pp[key])
        for i in range(atoms.nat):
            if atoms.symbols[i] == key:
                v += d * np.conjugate(strf(grid, atoms.positions[i]))
    qq = grid.q.copy()
    qq[0, 0, 0] = 1.0
    invq = 1.0 / qq
    invq[0, 0, 0] = 0.0
    m = grid.mask
    x = (rhoG * v * invq)[m]
    g = grid.g[:, m]
    S = np.zeros((3, 3))
    for i in range(3):
        for j in range(i, 3):
            S[i, j] = S[j, i] = -(np.sum(g[i] * g[j] * x)).real / grid.volume * 2.0
    S -= np.eye(3) * energy
    return S / grid.volume


# ---------------------------------------------------------------- PME forms
def _pme_stencil_forces(field, grid, frac, order):
    """-sum over each atom's B-spline stencil of grad_R Q(l) field(l) (the
    Q_derivativeA contraction of ``pseudo/__init__.py:500-520`` and
    ``ewald.py:833-862``), cartesian via the inverse lattice."""
    nr = grid.nr
    cell_inv = np.linalg.inv(grid.lattice)
    ix = np.arange(order)
    F = np.zeros((len(frac), 3))
    for a, s in enumerate(frac):
        Up = s * nr
        fl = np.floor(Up).astype(np.int64)
        M = [calc_Mn(Up[j] - fl[j], order) for j in range(3)]
        # order - 1 splines in an array of order + 1 entries, as calc_Mn
        # (ewald.py:66-86) returns them
        M2 = [np.append(calc_Mn(Up[j] - fl[j], order - 1), 0.0) for j in range(3)]
        d = [M2[j][1:order + 1] - M2[j][0:order] for j in range(3)]
        m = [M[j][1:order + 1] for j in range(3)]
        Qd = np.stack([nr[0] * np.einsum("i,j,k->ijk", d[0], m[1], m[2]),
                       nr[1] * np.einsum("i,j,k->ijk", m[0], d[1], m[2]),
                       nr[2] * np.einsum("i,j,k->ijk", m[0], m[1], d[2])])
        sub = field[np.ix_(np.mod(1 + fl[0] - ix, nr[0]), np.mod(1 + fl[1] - ix, nr[1]),
                           np.mod(1 + fl[2] - ix, nr[2]))]
        F[a] = -np.einsum("dijk,ijk->d", Qd, sub) @ cell_inv
    return F


def pp_forces_pme(rho, grid, atoms, pp, order=10):
    """``pseudo/__init__.py:475-520`` (_ForcePME): per species the field
    IFFT(conj(B) vloc rho(G)) (G = 0 dropped) contracted with the derivative
    of each atom's B-spline stencil."""
    rhoG = grid.fft(rho) * np.conjugate(barray(grid, order))
    frac = atoms.scaled_positions()
    F = np.zeros((atoms.nat, 3))
    for key in sorted(pp):
        idx = [i for i, s in enumerate(atoms.symbols) if s == key]
        if not idx:
            continue
        x = rhoG * vloc_line(grid.q, *pp[key])
        x[0, 0, 0] = 0.0
        F[idx] = _pme_stencil_forces(grid.ifft(x), grid, frac[idx], order)
    return F


def pp_stress_pme(rho, grid, atoms, pp, energy, order=10):
    """``pseudo/__init__.py:522-563`` (_StressPME): the exact form with each
    species' structure factor replaced by B rfftn(Q_s)."""
    rhoG = grid.fft(rho)
    B = barray(grid, order)
    frac = atoms.scaled_positions()
    v = np.zeros(grid.nrG, dtype=np.complex128)
    for key in sorted(pp):
        idx = [i for i, s in enumerate(atoms.symbols) if s == key]
        if not idx:
            continue
        S = B * grid.fft(pme_qarray(grid, frac[idx], np.ones(len(idx)), order)) / grid.dV
        v += vloc_deriv_line(grid.q, *pp[key]) * np.conjugate(S)
    qq = grid.q.copy()
    qq[0, 0, 0] = 1.0
    invq = 1.0 / qq
    invq[0, 0, 0] = 0.0
    m = grid.mask
    x = (rhoG * v * invq)[m]
    g = grid.g[:, m]
    S = np.zeros((3, 3))
    for i in range(3):
        for j in range(i, 3):
            S[i, j] = S[j, i] = -(np.sum(g[i] * g[j] * x)).real / grid.volume * 2.0
    S -= np.eye(3) * energy
    return S / grid.volume


def ewald_forces_rec_pme(grid, atoms, eta, order=10):
    """``ewald.py:810-864`` (Forces_rec_PME)."""
    Q = pme_qarray(grid, atoms.scaled_positions(), atoms.charges, order)
    B = barray(grid, order)
    x = grid.fft(Q) * (B * np.conjugate(B)) * np.exp(-grid.gg / (4.0 * eta)) * grid.invgg
    F = _pme_stencil_forces(grid.ifft(x), grid, atoms.scaled_positions(), order)
    return F * atoms.charges[:, None] * 4.0 * np.pi / grid.dV


def ewald_stress_rec_pme(grid, atoms, eta, order=10):
    """``ewald.py:915-960`` (Stress_rec_PME): the exact form with S_Z(G)
    replaced by B rfftn(Q_Z)."""
    Q = pme_qarray(grid, atoms.scaled_positions(), atoms.charges, order)
    s = grid.fft(Q) * barray(grid, order) / grid.dV
    s2 = (np.conjugate(s) * s).real
    m = grid.mask
    gg, invgg, g = grid.gg[m], grid.invgg[m], grid.g[:, m]
    w = s2[m] * np.exp(-gg / (4.0 * eta)) * invgg
    st = []
    for i in range(3):
        for j in range(i, 3):
            f = g[i] * g[j] * 2.0 * invgg * (1 + gg / (4.0 * eta))
            if i == j:
                f = f - 1.0
            st.append(np.sum(w * f))
    S = _sym(np.array(st) * 4.0 * np.pi / grid.volume ** 2)
    S += np.eye(3) * atoms.ncharge() ** 2 * 4.0 * np.pi * (1.0 / (4.0 * eta * grid.volume ** 2) / 2.0)
    return S


def ewald_forces_pme(grid, atoms, precision=1e-8, order=10):
    eta = ewald_eta(grid, atoms, precision)
    return ewald_forces_real(grid, atoms, eta, precision) + ewald_forces_rec_pme(grid, atoms, eta, order)


def ewald_stress_pme(grid, atoms, precision=1e-8, order=10):
    eta = ewald_eta(grid, atoms, precision)
    return ewald_stress_real(grid, atoms, eta, precision) + ewald_stress_rec_pme(grid, atoms, eta, order)


# ---------------------------------------------------------------- KEDF, Hartree, LDA
def tf_stress(energy, volume):
    """``kedf/tf.py:52-62``."""
    return np.eye(3) * (-2.0 / 3.0 * energy / volume)


def vw_stress(rho, grid, y=1.0):
    """``kedf/vw.py:97-111``."""
    rhoG = grid.fft(rho)
    d = [grid.ifft(1j * grid.g[i] * rhoG) for i in range(3)]
    S = np.zeros((3, 3))
    for i in range(3):
        for j in range(i, 3):
            S[i, j] = S[j, i] = -0.25 / grid.volume * grid.dV * np.sum(d[i] * d[j] / rho) * y
    return S


def lindhard_derivative(eta, mu):
    """``kedf/kernel.py:197-211``: eta dG/deta of the inverse Lindhard function."""
    out = np.zeros_like(eta)
    atol = 1.0e-10
    c0 = np.logical_and(eta > atol, np.abs(eta - 1.0) > atol)
    c1 = eta < atol
    c2 = np.abs(eta - 1.0) < atol
    e = eta[c0]
    A = np.log(np.abs((1.0 + e) / (1.0 - e)))
    out[c0] = (0.5 / e - 0.25 * (e ** 2 + 1.0) / e ** 2 * A) / (0.5 + 0.25 * (1 - e ** 2) / e * A) ** 2 + 6.0 * e * mu
    out[c1] = -2.0 * eta[c1] * (1.0 / 3.0 - 3.0 * mu)
    out[c2] = -48
    return out * eta


def wt_stress(rho, grid, energy, y=1.0, alpha=5.0 / 6.0, beta=5.0 / 6.0):
    """``kedf/wt.py:75-105``."""
    rho0 = rho.mean()
    tkf = 2.0 * (3.0 * rho0 * np.pi ** 2) ** (1.0 / 3.0)
    rA = grid.fft(rho ** alpha) / grid.volume
    rB = np.conjugate(grid.fft(rho ** beta)) / grid.volume
    dd = lindhard_derivative(grid.q / tkf, y) * rA * rB
    m = grid.mask
    g, invgg = grid.g[:, m], grid.invgg[m]
    S = np.zeros((3, 3))
    for i in range(3):
        for j in range(i, 3):
            fac = 1.0 / 3.0 if i == j else 0.0
            S[i, j] = S[j, i] = np.sum((g[i] * g[j] * invgg - fac) * dd[m]).real
    S *= np.pi ** 2 / (alpha * beta * rho0 ** (alpha + beta - 2) * tkf / 2.0)
    S -= np.eye(3) * 2.0 / 3.0 * energy / grid.volume
    return S


def hartree_stress(rho, grid, energy):
    """``hartree.py:67-91``."""
    rhoG = grid.fft(rho)
    x = (rhoG * np.conjugate(rhoG) * grid.invgg * grid.invgg)[grid.mask]
    g = grid.g[:, grid.mask]
    S = np.zeros((3, 3))
    for i in range(3):
        for j in range(i, 3):
            S[i, j] = S[j, i] = np.sum(g[i] * g[j] * x).real * 8.0 * np.pi / grid.volume ** 2
    S -= np.eye(3) * energy / grid.volume
    return S


def lda_stress(rho, grid, energy, potential):
    """``xc/semilocal_xc.py:308-322``."""
    return np.eye(3) * (energy - np.sum(potential * rho) * grid.dV) / grid.volume


__all__ = ["ewald_forces_real", "ewald_forces_rec", "ewald_forces", "ewald_stress_real", "ewald_stress_rec",
           "ewald_stress", "vloc_deriv_line", "pp_forces", "pp_stress", "tf_stress", "vw_stress",
           "lindhard_derivative", "wt_stress", "hartree_stress", "lda_stress", "CTF"]
