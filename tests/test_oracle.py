"""The CPU oracle against the reference's own golden values and against
closed forms (CPU only)."""
import numpy as np
import pytest
from scipy.optimize._dcsrch import DCSRCH

from conftest import data
from oracle import dcsrch as D
from oracle import functionals as F
from oracle import grid as G
from oracle import density as OD
from oracle import io, optimize, pseudo, td, units


@pytest.fixture(scope="module")
def al_pp():
    atoms, rho, lat = io.read_qepp(data("Al_fde_rho.pp"))
    return atoms, rho, lat


def test_units_codata2014():
    # ase derives Bohr and Hartree from the CODATA-2014 fundamental constants;
    # the results agree with the tabulated CODATA-2014 values within their
    # standard uncertainties (0.52917721067(12) A, 27.21138602(17) eV)
    assert abs(units.Bohr - 0.52917721067) < 1.2e-10
    assert abs(units.Ha - 27.21138602) < 1.7e-7


def test_wt_golden_full_grid(al_pp):
    # examples/test/test_functional.py:23-27: the pp reader builds a full grid
    atoms, rho, lat = al_pp
    g = G.Grid(lat, rho.shape, full=True)
    e = F.tf(rho, g)[0] + F.vw(rho, g)[0] + F.wt(rho, g)[0]
    assert np.isclose(e, 2.916818700014412, rtol=1e-12, atol=0)


def test_wt_half_grid_differs_only_at_nyquist(al_pp):
    # the R2C convention changes the Nyquist planes of this non-orthogonal cell
    atoms, rho, lat = al_pp
    g = G.Grid(lat, rho.shape, full=False)
    e = F.tf(rho, g)[0] + F.vw(rho, g)[0] + F.wt(rho, g)[0]
    assert abs(e - 2.916818700014412) < 2e-3 and abs(e - 2.916818700014412) > 1e-6


@pytest.mark.parametrize("pme", [True, False])
def test_local_pp_golden(al_pp, pme):
    # examples/test/test_read_pp.py:12,31
    atoms, rho, lat = al_pp
    gp, vp, z = io.read_recpot(data("al.lda.recpot"))
    assert z == 3
    g = G.Grid(lat, rho.shape, full=True)
    v = pseudo.local_pp(g, atoms, {"Al": (gp, vp)}, pme=pme)
    assert np.isclose((v * rho).sum() * g.dV, 1.3497046, atol=1e-4)


def test_ewald_goldens(al_pp):
    # examples/test/test_ewald.py:33-35
    atoms, rho, lat = al_pp
    atoms.set_charges({"Al": 3})
    g = G.Grid(lat, rho.shape, full=True)
    eta = pseudo.ewald_eta(g, atoms)
    assert np.isclose(pseudo.ewald_corr(g, atoms, eta), -13.003839613564997, atol=1e-5)
    assert np.isclose(pseudo.ewald_real(g, atoms, eta), 0.0, atol=1e-5)
    assert np.isclose(pseudo.ewald_rec(g, atoms, eta), 7.606943756474895, atol=1e-5)
    assert np.isclose(pseudo.ewald_rec_pme(g, atoms, eta), 7.606943756474895, atol=1e-5)


def test_fcc_madelung():
    # point charges Z on an fcc lattice in a neutralising background:
    # E/atom = -alpha_WS Z^2 / (2 r_ws) ... checked through the known
    # fcc Madelung energy per ion, -0.5 * 1.79175 * Z^2 / r_s (jellium), r_s^3 = 3 V/(4 pi N)
    a = 7.6
    cell = np.eye(3) * a
    frac = np.array([[0, 0, 0], [.5, .5, 0], [.5, 0, .5], [0, .5, .5]])
    at = io.Atoms(["Al"] * 4, frac @ cell, cell)
    at.set_charges({"Al": 3.0})
    g = G.Grid(cell, (32, 32, 32))
    e = pseudo.ewald_energy(g, at, pme=False)
    rs = (3 * a ** 3 / 4 / (4 * np.pi)) ** (1 / 3)
    assert np.isclose(e / 4, -0.5 * 1.79174723 * 9 / rs, rtol=1e-5)


def test_fft_conventions_and_brute_force():
    rng = np.random.default_rng(0)
    lat = np.array([[5.0, 0.3, 0.0], [0.2, 6.0, 0.1], [0.0, 0.4, 7.0]])
    g = G.Grid(lat, (3, 5, 4))
    f = rng.random((3, 5, 4))
    F_ = g.fft(f)
    # brute-force DFT of one coefficient
    k = (1, 2, 1)
    idx = np.indices(f.shape).reshape(3, -1).T
    ph = np.exp(-2j * np.pi * (idx[:, 0] * k[0] / 3 + idx[:, 1] * k[1] / 5 + idx[:, 2] * k[2] / 4))
    assert np.isclose(F_[k], (f.reshape(-1) * ph).sum() * g.dV)
    assert np.allclose(g.ifft(F_), f)


def test_lindhard_limits():
    assert np.isclose(F.lindhard(np.array([0.0]), 1.0, 1.0)[0], 0.0)
    assert np.isclose(F.lindhard(np.array([1.0]), 1.0, 1.0)[0], 2.0 - 1.0 - 3.0)
    e = np.array([3.65 - 1e-9, 3.65 + 1e-9])
    v = F.lindhard(e, 1.0, 1.0)
    assert abs(v[0] - v[1]) < 1e-6 * abs(v[0])


def test_lda_potential_is_derivative():
    lat = np.eye(3) * 5
    g = G.Grid(lat, (4, 4, 4))
    rng = np.random.default_rng(1)
    rho = 0.01 + 0.5 * rng.random((4, 4, 4))  # spans rs < 1 and rs >= 1
    e0, v = F.lda(rho, g)
    d = np.zeros_like(rho)
    d[1, 2, 3] = 1e-6
    ep = F.lda(rho + d, g)[0]
    em = F.lda(rho - d, g)[0]
    assert np.isclose((ep - em) / (2e-6 * g.dV), v[1, 2, 3], rtol=1e-6)


def test_wt_alpha_ne_beta_is_a_derivative():
    """WT with alpha != beta (wt.py:23-25,44-56): the potential is dE/drho by
    central differences (kernel held fixed), E and V are symmetric in
    (alpha, beta) (K is even), and alpha -> beta approaches the alpha == beta
    branch.  A dropped cross term or a swapped exponent fails one of them."""
    lat = np.diag([5.0, 5.5, 6.0])
    g = G.Grid(lat, (6, 8, 10))
    rng = np.random.default_rng(3)
    rho = 0.02 + 0.02 * rng.random((6, 8, 10))
    a, b = 0.7, 1.0 - 0.7 / 3.0 + 0.4
    K = F.wt_kernel(g.q, rho.mean(), alpha=a, beta=b)
    e, v, _ = F.wt(rho, g, alpha=a, beta=b, kernel=K)
    for idx in ((1, 2, 3), (4, 7, 0), (5, 0, 9)):
        d = np.zeros_like(rho)
        d[idx] = 1e-7
        ep = F.wt(rho + d, g, alpha=a, beta=b, kernel=K)[0]
        em = F.wt(rho - d, g, alpha=a, beta=b, kernel=K)[0]
        assert np.isclose((ep - em) / (2e-7 * g.dV), v[idx], rtol=1e-6)
    e2, v2, _ = F.wt(rho, g, alpha=b, beta=a, kernel=K)
    assert np.isclose(e2, e, rtol=1e-12) and np.allclose(v2, v, rtol=1e-12)
    c = 5.0 / 6.0
    Kc = F.wt_kernel(g.q, rho.mean(), alpha=c, beta=c)
    e3, v3, _ = F.wt(rho, g, alpha=c, beta=c, kernel=Kc)
    e4, v4, _ = F.wt(rho, g, alpha=c * (1 + 1e-7), beta=c, kernel=Kc)
    assert np.isclose(e4, e3, rtol=1e-5) and np.allclose(v4, v3, rtol=1e-5)


def test_energy_density_and_v2_pins():
    """calcType 'D' integrates to each term's energy (TF, vW, WT with
    alpha == beta and alpha != beta), and 'V2' is dV/drho of the local terms
    (TF, PZ-LDA on both rs branches) by central differences."""
    lat = np.diag([5.0, 5.5, 6.0])
    g = G.Grid(lat, (6, 8, 10))
    rng = np.random.default_rng(4)
    rho = 0.01 + 0.5 * rng.random((6, 8, 10))
    for terms, e in ((("TF",), F.tf(rho, g)[0]), (("vW",), F.vw(rho, g)[0]), (("NL",), F.wt(rho, g)[0])):
        assert np.isclose(F.energy_density(rho, g, terms).sum() * g.dV, e, rtol=1e-12), terms
    a, b = (5 + np.sqrt(5)) / 6, (5 - np.sqrt(5)) / 6
    e_ab = F.wt(rho, g, alpha=a, beta=b)[0]
    assert np.isclose(F.energy_density(rho, g, ("NL",), alpha=a, beta=b).sum() * g.dV, e_ab, rtol=1e-12)
    h = 1e-6 * rho
    fd_tf = (F.tf(rho + h, g, x=0.7)[1] - F.tf(rho - h, g, x=0.7)[1]) / (2 * h)
    assert np.allclose(F.tf_v2(rho, x=0.7), fd_tf, rtol=1e-7)
    fd_xc = (F.lda(rho + h, g)[1] - F.lda(rho - h, g)[1]) / (2 * h)
    assert np.allclose(F.lda_v2(rho), fd_xc, rtol=1e-6)


def test_hartree_parseval():
    lat = np.diag([6.0, 7.0, 8.0])
    g = G.Grid(lat, (12, 14, 16))
    rng = np.random.default_rng(2)
    rho = rng.random((12, 14, 16))
    e, v = F.hartree(rho, g)
    rg = np.fft.fftn(rho) * g.dV
    gg = g.ggF
    gg[0, 0, 0] = 1
    inv = 1 / gg
    inv[0, 0, 0] = 0
    e2 = 2 * np.pi / g.volume * np.sum(np.abs(rg) ** 2 * inv)
    assert np.isclose(e, e2, rtol=1e-12)


def test_vw_sigma_on_a_plane_wave():
    """sqrt(rho) = a + b cos(G0 x): the filtered vW potential is
    (1/2) b |G0|^2 exp(-|G0|^2 sigma^2 / 4) cos(G0 x) / sqrt(rho) (vw.py:45-49)."""
    L, n, a, b, sigma = 8.0, 24, 0.2, 0.05, 0.7
    g = G.Grid(np.eye(3) * L, (n, n, n))
    x = np.arange(n) * L / n
    c = np.cos(2 * np.pi * 2 * x / L)[:, None, None] * np.ones((n, n, n))
    sq = a + b * c
    _, v = F.vw(sq * sq, g, sigma=sigma)
    G0 = 2 * np.pi * 2 / L
    assert np.allclose(v, 0.5 * b * G0 ** 2 * np.exp(-G0 ** 2 * sigma ** 2 / 4) * c / sq, rtol=1e-10, atol=1e-13)


def test_vw_uniform_is_zero():
    g = G.Grid(np.eye(3) * 5, (8, 8, 8))
    e, v = F.vw(np.full((8, 8, 8), 0.03), g)
    assert abs(e) < 1e-14 and np.abs(v).max() < 1e-12


def _ls_sequence(fun, step0, ours=True, c1=1e-4, c2=0.2, xtol=1e-12):
    trials = []

    def f(a):
        trials.append(a)
        return list(fun(a))

    if ours:
        a, task, n, _ = D.line_search(f, step0, list(fun(0.0)), c1=c1, c2=c2, amax=np.pi, amin=0.0, xtol=xtol,
                                      maxiter=50)
    else:
        ph = lambda a: f(a)[0]
        dph = lambda a: fun(a)[1]
        ls = DCSRCH(ph, dph, c1, c2, xtol, 0.0, np.pi)
        a, _, _, task = ls(step0, phi0=fun(0.0)[0], derphi0=fun(0.0)[1], maxiter=50)
    return a, trials


@pytest.mark.parametrize("case", range(4))
def test_dcsrch_matches_scipy(case):
    funs = [lambda a: (np.cos(a + 0.3) ** 2 - a * 0.1, -2 * np.cos(a + 0.3) * np.sin(a + 0.3) - 0.1),
            lambda a: ((a - 1.3) ** 2, 2 * (a - 1.3)),
            lambda a: (-np.sin(3 * a) + a ** 4, -3 * np.cos(3 * a) + 4 * a ** 3),
            lambda a: (np.exp(-a) + 0.05 * a ** 3, -np.exp(-a) + 0.15 * a ** 2)]
    for s0 in (0.01, 0.1, 1.0, 2.5):
        a1, t1 = _ls_sequence(funs[case], s0, True)
        a2, t2 = _ls_sequence(funs[case], s0, False)
        assert (a1 is None) == (a2 is None)
        assert np.allclose(t1, t2, rtol=1e-14, atol=0)


def test_tn_gaas_golden():
    # examples/test/test_density_opt.py:16-46
    atoms = io.read_poscar(data("GaAs_random.vasp"))
    nr = G.ecut2nr(atoms.cell, 0.4)
    assert list(nr) == [20, 30, 22]
    g = G.Grid(atoms.cell, nr)
    pp, z = {}, {}
    for k, f in (("Ga", "Ga_lda.oe04.recpot"), ("As", "As_lda.oe04.recpot")):
        gp, vp, zz = io.read_recpot(data(f))
        pp[k], z[k] = (gp, vp), zz
    atoms.set_charges(z)
    vreal = pseudo.local_pp(g, atoms, pp, pme=True)
    eII = pseudo.ewald_energy(g, atoms, pme=True)
    rho = np.full(tuple(nr), atoms.ncharge() / g.volume)
    cache = {}

    def ev(r, need_e):
        r0 = r.mean()
        if "k" not in cache or abs(cache["r0"] - r0) > 1e-6:
            cache["k"], cache["r0"] = F.wt_kernel(g.q, r0), r0
        e, v, _ = F.total(r, g, vreal, cache["k"], ewald=eII)
        return e, v

    res = optimize.optimize(ev, rho, g.dV, method="TN", econv=1e-6 * atoms.nat, maxiter=100, maxfun=50, maxls=30)
    assert res["converged"] == 0
    assert np.isclose(ev(res["rho"], True)[0], -43.3046, rtol=1e-5)


def _quadratic_cg_step(seed=0, shape=(2, 2, 3)):
    """One exact line search on f(x) = x.Ax/2 - b.x from x0 along -g0;
    returns (A, g0, g1, d0) with gradients and direction as fields."""
    rng = np.random.default_rng(seed)
    n = int(np.prod(shape))
    M = rng.standard_normal((n, n))
    A = M @ M.T + n * np.eye(n)
    b, x0 = rng.standard_normal(n), rng.standard_normal(n)
    g0 = A @ x0 - b
    d0 = -g0
    alpha = -(g0 @ d0) / (d0 @ A @ d0)
    g1 = A @ (x0 + alpha * d0) - b
    return A, g0.reshape(shape), g1.reshape(shape), d0.reshape(shape)


@pytest.mark.parametrize("method", ["CG-HS", "CG", "CG-PR", "CG-FR", "CG-DY"])
def test_cg_betas_conjugate_on_a_quadratic(method):
    """With an exact line search on a quadratic, the HS, PR, FR and DY betas
    all give the linear-CG direction, which is A-conjugate to the previous one."""
    A, g0, g1, d0 = _quadratic_cg_step()
    d1 = optimize.get_direction_CG([g0, g1], [d0], method)
    assert abs(d1.ravel() @ A @ d0.ravel()) < 1e-10 * np.linalg.norm(d1) * np.linalg.norm(A @ d0.ravel())
    assert np.array_equal(optimize.get_direction_CG([g0], [], method), -g0)


def test_cg_ls_and_cd_as_the_reference_writes_them():
    """math_utils.py:411-414: CG-LS divides by sum(d r0) (textbook Liu-Storey
    divides by -sum(d r0)), so on the quadratic its beta is minus the linear-CG
    beta; CG-CD's beta is the field -(r1 ** 2) / sum(d r0), whose sum is the
    textbook conjugate-descent beta, equal to Fletcher-Reeves there."""
    A, g0, g1, d0 = _quadratic_cg_step(1)
    beta_cg = np.sum(g1 ** 2) / np.sum(g0 ** 2)
    d_ls = optimize.get_direction_CG([g0, g1], [d0], "CG-LS")
    assert np.allclose(d_ls, -g1 - beta_cg * d0, rtol=1e-12, atol=1e-12)
    d_cd = optimize.get_direction_CG([g0, g1], [d0], "CG-CD")
    beta_field = (d_cd + g1) / d0
    assert np.ptp(beta_field) > 0.1 * beta_cg  # a field, not a scalar
    assert np.isclose(beta_field.sum(), beta_cg, rtol=1e-10)


@pytest.mark.parametrize("m", [0, 1, 2, 4])
@pytest.mark.parametrize("h0", [0.0, 0.7])
def test_lbfgs_two_loop_against_dense_update(m, h0):
    """The reference's two-loop recursion (math_utils.py:431-455, first loop
    over pairs m-1..1) equals the dense BFGS inverse update over pairs 1..m-1
    started from V0^T H0, V_i = I - rho_i y_i s_i^T, H0 = h0 I or gamma I."""
    rng = np.random.default_rng(10 * m + int(h0 * 10))
    shape = (3, 2, 2)
    n = int(np.prod(shape))
    lb = optimize.LBFGS(h0=h0, bound=6)
    for _ in range(m):
        s = rng.standard_normal(shape)
        y = s + 0.3 * rng.standard_normal(shape)  # s.y > 0
        lb.update(s, y)
    r = rng.standard_normal(shape)
    d = optimize.get_direction_LBFGS(r, lb)
    S = [v.ravel() for v in lb.s]
    Y = [v.ravel() for v in lb.y]
    rho = [1.0 / (a @ b) for a, b in zip(S, Y)]
    I = np.eye(n)
    if h0:
        H = h0 * I
    else:
        H = (S[-1] @ Y[-1]) / (Y[-1] @ Y[-1]) * I if m else I
    if m:
        H = (I - rho[0] * np.outer(S[0], Y[0])) @ H
    for j in range(1, m):
        V = I - rho[j] * np.outer(Y[j], S[j])
        H = V.T @ H @ V + rho[j] * np.outer(S[j], S[j])
    assert np.allclose(d.ravel(), H @ (-r.ravel()), rtol=1e-11, atol=1e-12)


def test_lbfgs_memory_keeps_bound_plus_one_pairs():
    """math_utils.py:383-388: the oldest pair is dropped only once more than
    Bound pairs are held."""
    lb = optimize.LBFGS(h0=1.0, bound=2)
    for k in range(5):
        lb.update(np.full(3, k + 1.0), np.ones(3))
    assert [v[0] for v in lb.s] == [3.0, 4.0, 5.0]
    assert np.allclose(lb.rho, [1 / 9.0, 1 / 12.0, 1 / 15.0])


def _periodic_gaussians(g, pos, sigma, z):
    """sum over atoms and neighbouring images of z * normalised Gaussians."""
    r = np.stack(np.meshgrid(*[np.arange(n) / n for n in g.nr], indexing="ij"), axis=-1) @ g.lattice
    out = np.zeros(tuple(g.nr))
    for R in pos:
        for t in np.array(np.meshgrid([-1, 0, 1], [-1, 0, 1], [-1, 0, 1], indexing="ij")).reshape(3, -1).T:
            d = np.linalg.norm(r - R - t @ g.lattice, axis=-1)
            out += z * OD.gaussian(d, sigma)
    return out


def test_gaussian_atoms_closed_forms():
    """A resolved Gaussian (sigma = 3 h, rcut < L/2) integrates to its scale and
    every placed value is the closed form at the true distance (cubic cell)."""
    g = G.Grid(np.eye(3) * 12.0, (60, 60, 60))
    pos = np.array([[3.13, 7.71, 5.02]])
    rho, nc = OD.gen_gaussian_density(g, ["Al"], pos / 12.0, {"Al": {"rcut": 5.5, "sigma": 0.6, "scale": 3.0}})
    assert nc == 3.0
    assert np.isclose(rho.sum() * g.dV, 3.0, rtol=1e-12)
    r = np.stack(np.meshgrid(*[np.arange(60) * 0.2] * 3, indexing="ij"), axis=-1)
    d = np.linalg.norm((r - pos[0] + 6.0) % 12.0 - 6.0, axis=-1)
    ref = np.where(d < 5.5, 3.0 * OD.gaussian(d, 0.6), 0.0)
    assert np.max(np.abs(rho - ref)) < 1e-12 * ref.max()


def test_recipe_of_a_gaussian_transform_is_the_periodic_gaussian():
    """get_3d_value_recipe with the table z exp(-q^2 sigma^2 / 2) (the transform
    of a normalised Gaussian) gives the periodic sum of the Gaussians; PME and
    exact structure factors agree; the real-space spline path agrees too."""
    L, sigma, z = 10.0, 0.8, 3.0
    g = G.Grid(np.eye(3) * L, (48, 48, 48))
    pos = np.array([[1.3, 2.2, 7.9], [6.1, 5.4, 0.7]])
    q = np.arange(0, 3001) * 0.01
    tab = {"Al": z * np.exp(-0.5 * (q * sigma) ** 2)}
    ref = _periodic_gaussians(g, pos, sigma, z)
    for pme in (False, True):
        rho = OD.get_3d_value_recipe(g, ["Al", "Al"], pos, pos / L, {"Al": q}, tab, 2 * z, pme=pme)
        assert np.max(np.abs(rho - ref)) < (1e-9 if not pme else 1e-6) * ref.max()
    rr = np.linspace(0.0, 6.0, 601)
    rho_a = OD.guess_rho_atom(g, ["Al", "Al"], pos / L, {"Al": rr}, {"Al": z * OD.gaussian(rr, sigma)}, 2 * z)
    assert np.max(np.abs(rho_a - ref)) < 1e-6 * ref.max()


def test_current_density_of_a_plane_wave():
    """psi = a exp(i G0.r): j = |a|^2 G0 everywhere (utils/utils.py:44-47)."""
    g = G.Grid(np.array([[7.0, 0, 0], [0.5, 6.0, 0], [0, 0.3, 8.0]]), (12, 10, 14), full=True)
    G0 = np.array([2, -1, 3]) @ g.B
    a = 0.3 - 0.2j
    psi = a * np.exp(1j * np.einsum("kxyz,k->xyz", g.r, G0))
    j = td.calc_j(psi, g)
    assert np.allclose(j, (abs(a) ** 2 * G0)[:, None, None, None], rtol=1e-11, atol=1e-13)


def test_predictor_corrector_with_a_fixed_potential_is_one_cn_step():
    """With a potential that does not depend on rho the second prediction
    repeats the first: two CN solves, converged, the state of one CN step."""
    g = G.Grid(np.eye(3) * 6.0, (12, 12, 12), full=True)
    rng = np.random.default_rng(4)
    v = 0.1 * rng.random((12, 12, 12))
    psi = np.sqrt(0.02 * (1 + 0.3 * rng.random((12, 12, 12)))) * np.exp(1j * 0.01 * g.r[0])
    out, v2, n, conv = td.predictor_corrector(psi, v, g, 0.1, lambda r: v)
    ref, _, _ = td.cn_step(psi, v, g, 0.1)
    assert n == 2 and conv and np.array_equal(out, ref) and v2 is v


@pytest.mark.parametrize("solver", ["cg", "bicg", "bicgstab", "gmres"])
def test_cn_solvers_against_a_dense_solve(solver):
    """The linear solvers of the CN step (linear_solver.py:24-110 and scipy's
    gmres) reach the dense solution of (1 + i H dt/2) psi1 = (1 - i H dt/2)
    psi0 on a 4x4x4 grid (H built column by column from the oracle's
    Hamiltonian), to their tolerance.  The reference's bicg is CG with
    conjugated sums: it needs a Hermitian positive-definite operator, so it
    is pinned on 1 + H dt/2 instead (on the CN operator it stalls, as the
    reference's does)."""
    lat = np.diag([3.0, 3.5, 4.0])
    g = G.Grid(lat, (4, 4, 4), full=True)
    rng = np.random.default_rng(8)
    v = 0.3 * rng.random((4, 4, 4))
    psi0 = (1 + 0.2 * rng.random((4, 4, 4))) * np.exp(1j * rng.random((4, 4, 4)))
    dt = 0.1
    n = 64
    Hm = np.zeros((n, n), complex)
    for j in range(n):
        e = np.zeros(n, complex)
        e[j] = 1.0
        Hm[:, j] = td.hamiltonian_apply(e.reshape(4, 4, 4), v, g).ravel()
    if solver == "bicg":
        A = np.eye(n) + 0.5 * dt * Hm
        b = psi0.ravel()
        x, info, _ = td.bicg(lambda y: y + 0.5 * dt * td.hamiltonian_apply(y, v, g), psi0, 0 * psi0, tol=1e-12,
                             maxiter=200)
        ref = np.linalg.solve(A, b).reshape(4, 4, 4)
    else:
        A = np.eye(n) + 0.5j * dt * Hm
        ref = np.linalg.solve(A, (np.eye(n) - 0.5j * dt * Hm) @ psi0.ravel()).reshape(4, 4, 4)
        x, info, _ = td.cn_step(psi0, v, g, dt, tol=1e-12, maxiter=200, solver=solver)
    assert info == 0
    assert np.max(np.abs(x - ref)) < 1e-10 * np.max(np.abs(ref))


def test_td_dipole_golden():
    # examples/test/test_propagator.py:28-81 (CN part)
    atoms, rho0, lat = io.read_xsf(data("GaAs_random.xsf"))
    rho0 = rho0 * (1.0 / units.Bohr) ** 3
    g = G.Grid(lat, rho0.shape, full=True)
    pp, z = {}, {}
    for k, f in (("Ga", "Ga_lda.oe04.recpot"), ("As", "As_lda.oe04.recpot")):
        gp, vp, zz = io.read_recpot(data(f))
        pp[k], z[k] = (gp, vp), zz
    atoms.set_charges(z)
    vreal = pseudo.local_pp(g, atoms, pp, pme=True)

    def pot(r):
        return F.tf(r, g)[1] + F.lda(r, g)[1] + F.hartree(r, g)[1] + vreal

    psi = td.initial_kick(1e-3, 0, np.sqrt(rho0), g)
    v = pot(rho0)
    E0 = td.energy(psi, v, g)
    for _ in range(10):
        psi, st, _ = td.cn_step(psi, v, g, 1e-3)
        assert st == 0
        rho = np.real(psi * np.conj(psi))
        v = pot(rho)
    assert np.isclose(td.energy(psi, v, g), E0, rtol=1e-3)
    assert np.isclose(td.dipole(rho, rho0, g)[0], 0.011423, rtol=1e-3)


# ---------------------------------------------------------------- warm-start network
def test_cnn_bf16_rounding_matches_torch():
    import torch
    from oracle import cnn
    rng = np.random.default_rng(0)
    x = np.concatenate([rng.normal(0, 3, 4000), rng.normal(0, 1e-3, 1000), [0.0, -0.0, 1.0, 3.0e38, -7.5]])
    ref = torch.tensor(x, dtype=torch.float32).to(torch.bfloat16).to(torch.float64).numpy()
    assert np.array_equal(cnn.to_bf16(x), ref)


def test_cnn_oracle_gradient_is_the_derivative():
    """Central differences of the FP64 loss along random directions, one
    parameter group at a time (dilations > 1 and uneven axes included),
    against the analytic backward.  ReLU kinks can spoil a difference
    quotient, so each group must agree at one of two small steps; a wrong,
    dropped or misplaced gradient term fails at every step."""
    from oracle import cnn
    rng = np.random.default_rng(1)
    shape = (5, 6, 7)
    dil = (1, 2, 1, 2, 1, 1)
    p = cnn.init_params(3, scale=2.0)
    for k in p:
        if k.startswith("b"):
            p[k] = rng.normal(0, 0.1, p[k].shape)
    x0 = rng.normal(0, 1, (2,) + shape)
    target = rng.normal(0, 0.1, shape)

    def L(q):
        out, _ = cnn.forward(q, x0, dil)
        return cnn.loss(out, target)[0]

    out, cache = cnn.forward(p, x0, dil)
    _, dout = cnn.loss(out, target)
    g = cnn.backward(p, cache, dout)
    for k in p:
        v = rng.normal(0, 1, p[k].shape)
        an = float((g[k] * v).sum())
        errs = []
        for eps in (1e-7, 4e-8):
            q1 = dict(p)
            q1[k] = p[k] + eps * v
            q2 = dict(p)
            q2[k] = p[k] - eps * v
            errs.append(abs((L(q1) - L(q2)) / (2 * eps) - an))
        assert min(errs) <= 2e-8 * max(1.0, abs(an)) + 5e-8, (k, an, errs)


def test_cnn_oracle_forward_is_torch_conv3d():
    """The FP64 oracle network equals torch's conv3d with circular padding
    and dilation (layer by layer: ReLU, residual sums, the 1-channel output
    layer), on an uneven grid with dilations 1, 2 and 4 that wrap around
    the cell.  A transposed weight, a mirrored tap or a wrong shift sign
    fails it."""
    import torch
    import torch.nn.functional as TF
    from oracle import cnn
    rng = np.random.default_rng(5)
    shape = (9, 10, 11)
    dil = (1, 2, 4, 4, 2, 1)
    p = cnn.init_params(7, scale=2.0)
    for k in p:
        if k.startswith("b"):
            p[k] = rng.normal(0, 0.1, p[k].shape)
    x0 = rng.normal(0, 1, (2,) + shape)
    ref, _ = cnn.forward(p, x0, dil)

    def conv(x, w, b, d):
        # x (C, nx, ny, nz); w (27, ci, co) with tap t = (i*3 + j)*3 + k
        # reading x[p + d (i-1, j-1, k-1)]
        ci, co = w.shape[1], w.shape[2]
        wt = torch.as_tensor(w).reshape(3, 3, 3, ci, co).permute(4, 3, 0, 1, 2).contiguous()
        xp = TF.pad(x[None], (d,) * 6, mode="circular")
        return TF.conv3d(xp, wt, torch.as_tensor(b), dilation=d)[0]

    x = torch.as_tensor(x0)
    h = torch.relu(conv(x, p["W1"], p["b1"], dil[0]))
    for l in range(2, 6):
        w = np.transpose(p["W%d" % l], (0, 2, 1))  # W_l[t, co, ci] -> (27, ci, co)
        h = h + torch.relu(conv(h, w, p["b%d" % l], dil[l - 1]))
    out = conv(h, p["W6"][:, :, None], p["b6"], dil[5])[0].numpy()
    assert np.abs(out - ref).max() < 1e-12 * np.abs(ref).max()


def test_lmgp_golden_and_heg(al_pp):
    """TF + vW + LMGP on Al_fde_rho.pp (examples/test/test_functional.py:29-34,
    reference tolerance atol 1e-3; the restatement reproduces it to 1e-9), and
    the homogeneous-gas limit (no nonlocal energy or potential)."""
    from oracle import lmgp as OL
    atoms, rho, lat = al_pp
    g = G.Grid(lat, rho.shape, full=True)
    tabs = OL.kernel_tables()
    e, v = OL.lmgp_nl(rho, g, tabs)
    tot = F.tf(rho, g)[0] + F.vw(rho, g)[0] + e
    assert abs(tot - 2.9145872498710492) < 1e-9
    e0, v0 = OL.lmgp_nl(np.full(rho.shape, rho.mean()), g, tabs)
    assert e0 == 0.0 and not v0.any()
