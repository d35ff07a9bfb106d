"""The oracle's forces and stresses (oracle/forces.py) against the
reference's Ewald goldens and against central differences of the oracle's
own energies (CPU only)."""
import numpy as np
import pytest

from conftest import data
from oracle import forces as FO
from oracle import functionals as F
from oracle import grid as G
from oracle import io, pseudo


@pytest.fixture(scope="module")
def al_pp():
    return io.read_qepp(data("Al_fde_rho.pp"))


def test_ewald_force_stress_goldens(al_pp):
    # examples/test/test_ewald.py:21-43
    atoms, rho, lat = al_pp
    atoms.set_charges({"Al": 3})
    g = G.Grid(lat, rho.shape, full=True)
    eta = pseudo.ewald_eta(g, atoms)
    assert np.isclose(FO.ewald_forces_real(g, atoms, eta)[0, 0], 0.0, atol=1e-5)
    assert np.isclose(FO.ewald_forces_rec(g, atoms, eta)[0, 0], 0.0, atol=1e-5)
    assert np.isclose(FO.ewald_stress_real(g, atoms, eta)[0, 0], 0.0, atol=1e-5)
    assert np.isclose(FO.ewald_stress_rec(g, atoms, eta)[0, 0], 8.04967853e-03, atol=1e-5)


def _cell(seed=3, n=24):
    """fcc Al, 4 atoms, a = 7.6 bohr, displaced; a smooth positive density."""
    from dftpy_amd import systems
    rng = np.random.default_rng(seed)
    a = 7.6
    lat = np.array([[a, 0.0, 0.0], [0.3, a, 0.0], [0.0, -0.2, a]])
    frac = np.array([[0, 0, 0], [.5, .5, 0], [.5, 0, .5], [0, .5, .5]]) + rng.normal(0, 0.02, (4, 3))
    at = io.Atoms(["Al"] * 4, frac @ lat, lat)
    gp, vp, z = io.read_recpot(systems.empty_core_recpot())
    at.set_charges({"Al": z})
    g = G.Grid(lat, (n, n, n))
    x = np.indices((n, n, n)).reshape(3, -1).T / n
    rho = 12.0 / g.volume * (1.0 + 0.3 * np.cos(2 * np.pi * (x @ np.array([1, 0, 1]))) +
                             0.2 * np.sin(2 * np.pi * (x @ np.array([0, 2, 1]))))
    return g, at, rho.reshape(n, n, n), {"Al": (gp, vp)}


def test_forces_are_minus_energy_gradients():
    """PP (exact structure factors) and Ewald forces equal -dE/dR_a by central
    differences at fixed density."""
    g, at, rho, pp = _cell()
    f_pp = FO.pp_forces(rho, g, at, pp)
    f_ew = FO.ewald_forces(g, at)
    h = 1e-4
    for a in (0, 2):
        for d in range(3):
            e = []
            for s in (1.0, -1.0):
                pos = at.positions.copy()
                pos[a, d] += s * h
                at2 = io.Atoms(at.symbols, pos, at.cell, at.charges)
                e.append(((pseudo.local_pp(g, at2, pp, pme=False) * rho).sum() * g.dV,
                          pseudo.ewald_energy(g, at2, pme=False)))
            assert abs(-(e[0][0] - e[1][0]) / (2 * h) - f_pp[a, d]) < 1e-6 * max(1.0, abs(f_pp[a, d]))
            assert abs(-(e[0][1] - e[1][1]) / (2 * h) - f_ew[a, d]) < 1e-6 * max(1.0, abs(f_ew[a, d]))


def _strained(g, at, rho, eps):
    Fm = np.eye(3) + eps
    lat = g.lattice @ Fm.T
    g2 = G.Grid(lat, g.nr)
    at2 = io.Atoms(at.symbols, at.positions @ Fm.T, lat, at.charges)
    return g2, at2, rho / np.linalg.det(Fm)


def _energies(g, at, rho, pp):
    r0 = rho.mean()
    e_wt, _, _ = F.wt(rho, g, rho0=r0, kernel=F.wt_kernel(g.q, r0))
    return dict(TF=F.tf(rho, g)[0], VW=F.vw(rho, g)[0], WT=e_wt, H=F.hartree(rho, g)[0], XC=F.lda(rho, g)[0],
                PP=(pseudo.local_pp(g, at, pp, pme=False) * rho).sum() * g.dV,
                II=pseudo.ewald_energy(g, at, pme=False))


def test_stresses_are_strain_derivatives():
    """Each stress term equals (1/V) dE/d(eps_ij) by central differences
    under homogeneous strain r -> (1 + eps) r with the electron count fixed."""
    g, at, rho, pp = _cell(n=20)
    E = _energies(g, at, rho, pp)
    _, vxc = F.lda(rho, g)
    S = dict(TF=FO.tf_stress(E["TF"], g.volume), VW=FO.vw_stress(rho, g), WT=FO.wt_stress(rho, g, E["WT"]),
             H=FO.hartree_stress(rho, g, E["H"]), XC=FO.lda_stress(rho, g, E["XC"], vxc),
             PP=FO.pp_stress(rho, g, at, pp, E["PP"]), II=FO.ewald_stress(g, at))
    h = 1e-5
    for (i, j) in ((0, 0), (1, 1), (0, 1), (1, 2)):
        eps = np.zeros((3, 3))
        eps[i, j] = eps[j, i] = h
        Ep = _energies(*_strained(g, at, rho, eps), pp)
        Em = _energies(*_strained(g, at, rho, -eps), pp)
        for k in S:
            fd = (Ep[k] - Em[k]) / (2 * h) / g.volume / (1.0 if i == j else 2.0)
            assert abs(fd - S[k][i, j]) < 2e-6 * max(1e-3, np.abs(S[k]).max()), (k, i, j, fd, S[k][i, j])


@pytest.mark.parametrize("skew", [False, True])
def test_pme_forces_are_minus_pme_energy_gradients(skew):
    """PME forms (pseudo/__init__.py:475-520, ewald.py:810-864) against
    central differences of the PME energies (order-10 B-splines are smooth,
    so the PME energy is differentiable).  On an orthogonal cell the PP and
    Ewald forces are -dE/dR.  The reference contracts the stencil derivative
    d/d(frac) with the inverse lattice as matmul(dQ^T, A^-1), which equals
    A^-T A (-dE/dR) on a skewed cell (rows of A = lattice vectors): the
    restatement keeps that form and is pinned by it there.  Both stay within
    the PME-vs-exact tolerance of the reference (test_pme.py:34-36) on the
    orthogonal cell."""
    g, at, rho, pp = _cell()
    if not skew:
        A = np.diag(np.diag(g.lattice))
        g = G.Grid(A, g.nr)
        at = io.Atoms(at.symbols, at.scaled_positions() @ A, A, at.charges)
    A = g.lattice
    M = np.linalg.inv(A).T @ A
    f_pp = FO.pp_forces_pme(rho, g, at, pp)
    f_ew = FO.ewald_forces_pme(g, at)
    h = 1e-4
    for a in (0, 2):
        fd = np.zeros((2, 3))
        for d in range(3):
            e = []
            for s in (1.0, -1.0):
                pos = at.positions.copy()
                pos[a, d] += s * h
                at2 = io.Atoms(at.symbols, pos, at.cell, at.charges)
                e.append(((pseudo.local_pp(g, at2, pp, pme=True) * rho).sum() * g.dV,
                          pseudo.ewald_energy(g, at2, pme=True)))
            fd[0, d] = -(e[0][0] - e[1][0]) / (2 * h)
            fd[1, d] = -(e[0][1] - e[1][1]) / (2 * h)
        assert np.max(np.abs(M @ fd[0] - f_pp[a])) < 1e-6 * max(1.0, np.abs(f_pp[a]).max())
        assert np.max(np.abs(M @ fd[1] - f_ew[a])) < 1e-6 * max(1.0, np.abs(f_ew[a]).max())
    if not skew:
        assert np.max(np.abs(f_pp - FO.pp_forces(rho, g, at, pp))) < 1e-4
        assert np.max(np.abs(f_ew - FO.ewald_forces(g, at))) < 1e-4


def test_pme_stresses_are_strain_derivatives():
    """The PME stresses (pseudo/__init__.py:522-563, ewald.py:915-960) are
    (1/V) dE/d(eps_ij) of the PME energies under homogeneous strain."""
    g, at, rho, pp = _cell(n=20)

    def energies(g, at, rho):
        return dict(PP=(pseudo.local_pp(g, at, pp, pme=True) * rho).sum() * g.dV,
                    II=pseudo.ewald_energy(g, at, pme=True))
    E = energies(g, at, rho)
    S = dict(PP=FO.pp_stress_pme(rho, g, at, pp, E["PP"]), II=FO.ewald_stress_pme(g, at))
    h = 1e-5
    for (i, j) in ((0, 0), (1, 1), (0, 1), (1, 2)):
        eps = np.zeros((3, 3))
        eps[i, j] = eps[j, i] = h
        Ep = energies(*_strained(g, at, rho, eps))
        Em = energies(*_strained(g, at, rho, -eps))
        for k in S:
            fd = (Ep[k] - Em[k]) / (2 * h) / g.volume / (1.0 if i == j else 2.0)
            assert abs(fd - S[k][i, j]) < 2e-6 * max(1e-3, np.abs(S[k]).max()), (k, i, j, fd, S[k][i, j])
    E0 = _energies(g, at, rho, pp)
    assert np.max(np.abs(S["PP"] - FO.pp_stress(rho, g, at, pp, E0["PP"]))) < 1e-4
    assert np.max(np.abs(S["II"] - FO.ewald_stress(g, at))) < 1e-4
