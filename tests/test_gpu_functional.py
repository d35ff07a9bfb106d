"""Fused energy/potential evaluation, local PP and Ewald on the GPU against
the oracle and the reference's goldens."""
import numpy as np
import pytest
import torch

from conftest import data
from oracle import functionals as OF
from oracle import grid as OG
from oracle import io as OIO
from oracle import pseudo as OP

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def al_pp():
    return OIO.read_qepp(data("Al_fde_rho.pp"))


def _dg(lat, nr, full, dtype=torch.float64):
    from dftpy_amd import DirectGrid
    return DirectGrid(lat, nr, full=full, dtype=dtype)


@pytest.mark.parametrize("full", [True, False])
def test_wt_golden_and_potential(al_pp, full):
    # examples/test/test_functional.py:23-27 (the reference reads the pp file onto a full grid)
    from dftpy_amd import DirectField
    from dftpy_amd.functional import Functional
    atoms, rho, lat = al_pp
    g = _dg(lat, rho.shape, full)
    out = Functional(type="KEDF", name="WT")(DirectField(g, data=rho))
    og = OG.Grid(lat, rho.shape, full=full)
    e_ref = OF.tf(rho, og)[0] + OF.vw(rho, og)[0] + OF.wt(rho, og)[0]
    v_ref = OF.tf(rho, og)[1] + OF.vw(rho, og)[1] + OF.wt(rho, og)[1]
    assert np.isclose(out.energy, e_ref, rtol=1e-12, atol=0)
    if full:
        assert np.isclose(out.energy, 2.916818700014412, rtol=1e-11, atol=0)
    v = out.potential.numpy()
    assert np.max(np.abs(v - v_ref)) < 1e-10 * max(1.0, np.max(np.abs(v_ref)))


@pytest.mark.parametrize("full,dtype", [(True, torch.float64), (False, torch.float64), (False, torch.float32)])
def test_wt_alpha_ne_beta_vs_oracle(al_pp, full, dtype):
    """WT nonlocal term with alpha != beta (wt.py:23-25,44-56; here the
    alpha + beta = 5/3 pair (5 +- sqrt 5) / 6) on Al_fde_rho.pp: energy and
    potential against the oracle (FP64 1e-11, FP32 stated), and the FP64
    stress (wt.py:75-105 with rho^alpha and rho^beta) against the oracle."""
    from dftpy_amd import DirectField
    from dftpy_amd.functional import Functional
    from oracle import forces as FO
    atoms, rho, lat = al_pp
    a, b = (5 + np.sqrt(5)) / 6, (5 - np.sqrt(5)) / 6
    g = _dg(lat, rho.shape, full, dtype)
    f = Functional(type="KEDF", name="WT-NL", alpha=a, beta=b)
    out = f(DirectField(g, data=rho))
    og = OG.Grid(lat, rho.shape, full=full)
    e_ref, v_ref, _ = OF.wt(rho, og, alpha=a, beta=b)
    v = out.potential.numpy().astype(np.float64)
    de, dv = abs(out.energy - e_ref), float(np.max(np.abs(v - v_ref)))
    print("WT alpha %.4f beta %.4f %s: E %.12f oracle %.12f |dE| %.2e max|dV| %.2e" %
          (a, b, dtype, out.energy, e_ref, de, dv))
    if dtype == torch.float64:
        assert de <= 1e-11 * abs(e_ref) and dv <= 1e-10 * max(1.0, np.max(np.abs(v_ref)))
        st = f.stress(DirectField(g, data=rho), split=True)["WT-NL"]
        st_ref = FO.wt_stress(rho, og, e_ref, alpha=a, beta=b)
        assert np.max(np.abs(st - st_ref)) <= 1e-9 * np.max(np.abs(st_ref))
    else:
        assert de <= 1e-5 * abs(e_ref) and dv <= 1e-5 * max(1.0, np.max(np.abs(v_ref)))


def _c1_density(nr, lat, seed=0):
    rng = np.random.default_rng(seed)
    og = OG.Grid(lat, nr)
    r = og.r
    rho = 0.027 * (1 + 0.4 * np.cos(2 * np.pi * r[0] / lat[0, 0]) * np.cos(2 * np.pi * r[1] / lat[1, 1]))
    rho += 0.002 * rng.random(nr)
    return rho


@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
def test_total_terms_vs_oracle(dtype):
    from dftpy_amd import DirectField, systems
    from dftpy_amd.functional import Functional, TotalFunctional
    ions, nr = systems.c1_system()
    lat = ions.cell
    rho = _c1_density(nr, lat)
    g = _dg(lat, nr, False, dtype)
    pp = Functional(type="PSEUDO", grid=g, ions=ions, PP_list={"Al": systems.empty_core_recpot()})
    tot = TotalFunctional(KE=Functional(type="KEDF", name="WT"), XC=Functional(type="XC", name="LDA"),
                          HARTREE=Functional(type="HARTREE"), PSEUDO=pp)
    out = tot(DirectField(g, data=rho))
    # oracle
    og = OG.Grid(lat, nr)
    oat = OIO.Atoms(ions.symbols, ions.positions, ions.cell)
    gp, vp, z = OIO.read_recpot(systems.empty_core_recpot())
    oat.set_charges({"Al": z})
    vreal = OP.local_pp(og, oat, {"Al": (gp, vp)}, pme=True)
    eII = OP.ewald_energy(og, oat, pme=True)
    e_ref, v_ref, parts = OF.total(rho, og, vreal, OF.wt_kernel(og.q, rho.mean()), ewald=eII)
    nat = ions.nat
    if dtype == torch.float64:
        for k, kk in (("TF", "TF"), ("vW", "vW"), ("NL", "NL"), ("XC", "XC"), ("HARTREE", "H"), ("PSEUDO", "PP"),
                      ("II", "II")):
            assert np.isclose(out.parts[k], parts[kk], rtol=1e-11, atol=1e-11), k
        assert abs(out.energy - e_ref) < 1e-9
        assert np.max(np.abs(out.potential.numpy() - v_ref)) < 1e-9
    else:
        assert abs(out.energy - e_ref) / nat < 2e-5
        assert np.max(np.abs(out.potential.numpy() - v_ref)) < 2e-4


@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
def test_energy_density_and_v2_vs_oracle(dtype):
    """calcType {'E', 'V', 'D', 'V2'} of a TotalFunctional (WT KEDF, PZ-LDA,
    Hartree, local PP): the energy density of the KEDF terms and the V2 of TF
    and LDA against the oracle; E and V equal the plain evaluation's; WT with
    alpha != beta gives its own energy density."""
    from dftpy_amd import DirectField, systems
    from dftpy_amd.functional import Functional, TotalFunctional
    ions, nr = systems.c1_system()
    lat = ions.cell
    rho = _c1_density(nr, lat)
    g = _dg(lat, nr, False, dtype)
    pp = Functional(type="PSEUDO", grid=g, ions=ions, PP_list={"Al": systems.empty_core_recpot()})
    tot = TotalFunctional(KE=Functional(type="KEDF", name="WT"), XC=Functional(type="XC", name="LDA"),
                          H=Functional(type="HARTREE"), PSEUDO=pp)
    r = DirectField(g, data=rho)
    out = tot(r, calcType={"E", "V", "D", "V2"})
    plain = tot(r, calcType={"E", "V"})
    og = OG.Grid(lat, nr)
    d_ref = OF.energy_density(rho, og, ("TF", "vW", "NL"))
    v2_ref = OF.tf_v2(rho) + OF.lda_v2(rho)
    tol = 1e-10 if dtype == torch.float64 else 2e-5
    dd = np.max(np.abs(out.energydensity.numpy() - d_ref)) / np.max(np.abs(d_ref))
    dv2 = np.max(np.abs(out.v2rho2.numpy() - v2_ref)) / np.max(np.abs(v2_ref))
    print("%s: energy density rel %.2e, v2rho2 rel %.2e" % (dtype, dd, dv2))
    assert dd < tol and dv2 < tol
    assert out.energy == plain.energy
    assert torch.equal(out.potential.data, plain.potential.data)
    a, b = (5 + np.sqrt(5)) / 6, (5 - np.sqrt(5)) / 6
    nl = Functional(type="KEDF", name="WT-NL", alpha=a, beta=b)(r, calcType={"E", "D"})
    d_ab = OF.energy_density(rho, og, ("NL",), alpha=a, beta=b)
    assert np.max(np.abs(nl.energydensity.numpy() - d_ab)) < tol * np.max(np.abs(d_ab))
    assert np.isclose(float(nl.energydensity.data.double().sum().item()) * og.dV, nl.energy,
                      rtol=1e-12 if dtype == torch.float64 else 1e-5)


@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
def test_xc_core_density_vs_oracle(dtype):
    """Non-linear core correction (xc/__init__.py:59-74): XC(core_density=c)
    evaluates PZ-LDA at rho + c (energy, potential and V2), alone and inside
    a TotalFunctional, against the oracle on the shifted density; the other
    terms still see rho."""
    from dftpy_amd import DirectField, systems
    from dftpy_amd.functional import Functional, TotalFunctional
    ions, nr = systems.c1_system()
    lat = ions.cell
    rho = _c1_density(nr, lat)
    og = OG.Grid(lat, nr)
    r = og.r
    core = 0.05 * np.exp(-((r[0] - 1.9) ** 2 + (r[1] - 3.8) ** 2 + (r[2] - 1.9) ** 2) / 0.8)
    g = _dg(lat, nr, False, dtype)
    cf = DirectField(g, data=core)
    xc = Functional(type="XC", name="LDA", core_density=cf)
    out = xc(DirectField(g, data=rho), calcType={"E", "V", "V2"})
    e_ref, v_ref = OF.lda(rho + core, og)
    tol = 1e-11 if dtype == torch.float64 else 2e-5
    print("%s: XC with core |dE| %.2e, max|dV| %.2e" % (dtype, abs(out.energy - e_ref),
                                                        np.max(np.abs(out.potential.numpy() - v_ref))))
    assert abs(out.energy - e_ref) <= tol * abs(e_ref)
    assert np.max(np.abs(out.potential.numpy() - v_ref)) <= tol * np.max(np.abs(v_ref))
    v2_ref = OF.lda_v2(rho + core)
    assert np.max(np.abs(out.v2rho2.numpy() - v2_ref)) <= tol * np.max(np.abs(v2_ref))
    tot = TotalFunctional(KE=Functional(type="KEDF", name="WT"), XC=xc, H=Functional(type="HARTREE"))
    o2 = tot(DirectField(g, data=rho))
    e2 = OF.tf(rho, og)[0] + OF.vw(rho, og)[0] + OF.wt(rho, og)[0] + e_ref + OF.hartree(rho, og)[0]
    assert abs(o2.energy - e2) <= tol * abs(e2)


@pytest.mark.parametrize("pme", [True, False])
@pytest.mark.parametrize("full", [True, False])
def test_local_pp_golden_and_vreal(al_pp, pme, full):
    # examples/test/test_read_pp.py:12,31
    from dftpy_amd import DirectField, Ions
    from dftpy_amd.functional.pseudo import LocalPseudo
    atoms, rho, lat = al_pp
    g = _dg(lat, rho.shape, full)
    ions = Ions(symbols=atoms.symbols, positions=atoms.positions, cell=lat)
    pp = LocalPseudo(grid=g, ions=ions, PP_list={"Al": data("al.lda.recpot")}, PME=pme)
    e = pp(DirectField(g, data=rho)).energy
    assert np.isclose(e, 1.3497046, atol=1e-4)
    og = OG.Grid(lat, rho.shape, full=full)
    gp, vp, _ = OIO.read_recpot(data("al.lda.recpot"))
    v_ref = OP.local_pp(og, atoms, {"Al": (gp, vp)}, pme=pme)
    assert np.max(np.abs(pp.vreal.numpy() - v_ref)) < 1e-10 * np.max(np.abs(v_ref))


@pytest.mark.parametrize("pme", [True, False])
@pytest.mark.parametrize("full", [True, False])
def test_local_pp_setup_on_fp32_plan(al_pp, pme, full):
    """The local-PP setup (ofdft_local_pp) on an FP32 plan (per-point math
    in FP64, the B-spline grid and the transforms in FP32) against the
    oracle's FP64 vreal, and the learned start's screening superposition
    (the same entry point with other radial tables) on FP32 against FP64."""
    from dftpy_amd import Ions, _lib
    from dftpy_amd.functional.pseudo import LocalPseudo
    from dftpy_amd.learned import superposition_density
    atoms, rho, lat = al_pp
    ions = Ions(symbols=atoms.symbols, positions=atoms.positions, cell=lat)
    g32 = _dg(lat, rho.shape, full, torch.float32)
    pp = LocalPseudo(grid=_dg(lat, rho.shape, full), ions=ions, PP_list={"Al": data("al.lda.recpot")}, PME=pme)
    keep, args = pp._table_args()
    v32 = torch.empty(tuple(rho.shape), dtype=torch.float32, device="cuda")
    _lib.check(_lib.lib.ofdft_local_pp(g32.plan(torch.float32).handle, *args, int(pme), 10, v32.data_ptr(),
                                       _lib.stream()))
    og = OG.Grid(lat, rho.shape, full=full)
    gp, vp, _ = OIO.read_recpot(data("al.lda.recpot"))
    v_ref = OP.local_pp(og, atoms, {"Al": (gp, vp)}, pme=pme)
    err = np.max(np.abs(v32.double().cpu().numpy() - v_ref)) / np.max(np.abs(v_ref))
    print("FP32 vreal vs oracle: %.2e" % err)
    assert err < 3e-6
    del keep
    rho0 = float(ions.get_ncharges()) / g32.volume
    s32 = superposition_density(g32, ions, pp.readpp, rho0, pme=pme).double().cpu().numpy()
    s64 = superposition_density(_dg(lat, rho.shape, full), ions, pp.readpp, rho0, pme=pme).cpu().numpy()
    err = np.max(np.abs(s32 - s64)) / np.max(np.abs(s64))
    print("superposition FP32 vs FP64: %.2e" % err)
    assert err < 3e-6


@pytest.mark.parametrize("pme", [True, False])
def test_ewald_golden(al_pp, pme):
    # examples/test/test_ewald.py:33-35
    from dftpy_amd import Ions
    from dftpy_amd.functional.pseudo import Ewald
    atoms, rho, lat = al_pp
    g = _dg(lat, rho.shape, True)
    ions = Ions(symbols=atoms.symbols, positions=atoms.positions, cell=lat)
    ions.set_charges({"Al": 3.0})
    ew = Ewald(ions, g, PME=pme)
    assert np.isclose(ew.Energy_corr(), -13.003839613564997, atol=1e-5)
    assert np.isclose(ew.Energy_real(), 0.0, atol=1e-5)
    assert np.isclose(ew.Energy_rec(), 7.606943756474895, atol=1e-5)


def test_ewald_500_atoms_vs_oracle():
    from dftpy_amd import systems
    from dftpy_amd.functional.pseudo import Ewald
    ions, nr = systems.c2_system()
    ions.set_charges({"Al": 3.0})
    lat = ions.cell
    nr = (64, 64, 64)
    g = _dg(lat, nr, False)
    og = OG.Grid(lat, nr)
    oat = OIO.Atoms(ions.symbols, ions.positions, lat)
    oat.set_charges({"Al": 3.0})
    eta = OP.ewald_eta(og, oat)
    ew = Ewald(ions, g, PME=False)
    assert np.isclose(ew.eta, eta)
    assert np.isclose(ew.Energy_real(), OP.ewald_real(og, oat, eta), rtol=1e-10, atol=1e-9)
    assert np.isclose(ew.Energy_rec(), OP.ewald_rec(og, oat, eta), rtol=1e-9)


def test_rho0_explicit_and_fp32_wt(al_pp):
    from dftpy_amd import DirectField
    from dftpy_amd.functional import Functional
    atoms, rho, lat = al_pp
    og = OG.Grid(lat, rho.shape, full=False)
    r0 = 0.03
    e_ref = OF.wt(rho, og, rho0=r0)[0]
    g = _dg(lat, rho.shape, False)
    out = Functional(type="KEDF", name="WT-NL", rho0=r0)(DirectField(g, data=rho))
    assert np.isclose(out.energy, e_ref, rtol=1e-11)
    g32 = _dg(lat, rho.shape, False, torch.float32)
    out32 = Functional(type="KEDF", name="WT")(DirectField(g32, data=rho))
    e64 = OF.tf(rho, og)[0] + OF.vw(rho, og)[0] + OF.wt(rho, og)[0]
    assert abs(out32.energy - e64) < 2e-5


@pytest.mark.parametrize("nr", [(64, 48, 32), (48, 64, 40)])
@pytest.mark.parametrize("full", [True, False])
@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
def test_fused_x_pass_matches_separate_passes(nr, full, dtype):
    """The pipeline with the G-space multiply fused into the x pass (size-
    specialised kernels) against the separate x passes + G-space kernel of
    the generic kernels, and both against the FP64 oracle."""
    from dftpy_amd import DirectField, _lib
    from dftpy_amd.functional import Functional, TotalFunctional
    lat = np.array([[9.1, 0.3, 0.0], [0.2, 8.4, 0.4], [0.0, 0.5, 7.7]])
    g = _dg(lat, nr, full, dtype)
    og = OG.Grid(lat, nr, full=full)
    r = og.r
    rho = 0.02 * (1.2 + 0.5 * np.cos(2 * np.pi * r[0] / 9.1) * np.sin(2 * np.pi * r[1] / 8.4 + 0.3)
                  + 0.2 * np.cos(4 * np.pi * r[2] / 7.7))
    tot = TotalFunctional(KE=Functional(type="KEDF", name="WT"), XC=Functional(type="XC", name="LDA"),
                          H=Functional(type="HARTREE"))
    plan = g.plan(dtype)
    outs = []
    for fixed in (1, 0):
        _lib.check(_lib.lib.ofdft_plan_set_fft_kernels(plan.handle, fixed))
        o = tot(DirectField(g, data=rho))
        outs.append((o.energy, o.potential.numpy().astype(np.float64)))
    _lib.check(_lib.lib.ofdft_plan_set_fft_kernels(plan.handle, 1))
    kern = OF.wt_kernel(og.q, rho.mean())
    e_ref = OF.tf(rho, og)[0] + OF.vw(rho, og)[0] + OF.wt(rho, og, kernel=kern)[0] + OF.lda(rho, og)[0] + \
        OF.hartree(rho, og)[0]
    v_ref = OF.tf(rho, og)[1] + OF.vw(rho, og)[1] + OF.wt(rho, og, kernel=kern)[1] + OF.lda(rho, og)[1] + \
        OF.hartree(rho, og)[1]
    # tolerances relative to the size of the individual terms (the total
    # energy is a small difference of them)
    e_scale = sum(abs(f(rho, og)[0]) for f in (OF.tf, OF.vw, OF.lda, OF.hartree)) + \
        abs(OF.wt(rho, og, kernel=kern)[0])
    etol, vtol = (1e-10, 1e-10) if dtype == torch.float64 else (2e-6, 1e-4)
    for e, v in outs:
        print("dE %.3e (scale %.3e) max dV %.3e" % (e - e_ref, e_scale, np.max(np.abs(v - v_ref))))
        assert abs(e - e_ref) <= etol * e_scale
        assert np.max(np.abs(v - v_ref)) <= vtol * np.max(np.abs(v_ref))
    assert abs(outs[0][0] - outs[1][0]) <= etol * e_scale


@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
@pytest.mark.parametrize("full", [True, False])
@pytest.mark.parametrize("nr", [(24, 24, 24), (30, 27, 25)])
def test_energy_only_matches_energy_and_potential(dtype, full, nr):
    """calcType={'E'} sums the FFT terms in G space (Parseval) instead of
    transforming back; every part equals the E+V path's (and, in FP64, the
    oracle's).  Covers even and odd sizes (fixed and generic FFTs), full and
    half grids, automatic and explicit rho0."""
    from dftpy_amd import DirectField, systems
    from dftpy_amd.functional import Functional, TotalFunctional
    ions, _ = systems.c1_system()
    lat = ions.cell
    rho = _c1_density(nr, lat, seed=3)
    g = _dg(lat, nr, full, dtype)
    pp = Functional(type="PSEUDO", grid=g, ions=ions, PP_list={"Al": systems.empty_core_recpot()})
    for ke in (Functional(type="KEDF", name="WT"), Functional(type="KEDF", name="WT", rho0=0.03),
               Functional(type="KEDF", name="TFVW")):
        tot = TotalFunctional(KE=ke, XC=Functional(type="XC", name="LDA"), HARTREE=Functional(type="HARTREE"),
                              PSEUDO=pp)
        f = DirectField(g, data=torch.as_tensor(rho, dtype=dtype, device="cuda"))
        e_only = tot(f, calcType={"E"})
        both = tot(f, calcType={"E", "V"})
        tol = 1e-11 if dtype == torch.float64 else 2e-6
        for k in ("TF", "vW", "NL", "XC", "HARTREE", "PSEUDO", "N"):
            assert abs(e_only.parts[k] - both.parts[k]) <= tol * max(1.0, abs(both.parts[k])), (k, e_only.parts[k],
                                                                                                 both.parts[k])
        assert abs(e_only.energy - both.energy) <= tol * max(1.0, abs(both.energy))
        assert not hasattr(e_only, "potential")


def test_plan_settings_follow_the_functional():
    """The plan keeps its settings between calls and skips unchanged ones:
    alternating functionals (terms, KEDF options, local potential, Ewald
    energy) on one grid gives each one's own energy every time."""
    from dftpy_amd import DirectField, systems
    from dftpy_amd.functional import Functional, TotalFunctional
    ions, nr = systems.c1_system()
    g = _dg(ions.cell, nr, False)
    rho = DirectField(g, data=torch.as_tensor(_c1_density(nr, ions.cell, seed=5), device="cuda"))
    pp = Functional(type="PSEUDO", grid=g, ions=ions, PP_list={"Al": systems.empty_core_recpot()})
    funcs = [TotalFunctional(KE=Functional(type="KEDF", name="WT"), XC=Functional(type="XC", name="LDA"),
                             HARTREE=Functional(type="HARTREE"), PSEUDO=pp),
             TotalFunctional(KE=Functional(type="KEDF", name="TFVW", y=0.2), HARTREE=Functional(type="HARTREE")),
             Functional(type="KEDF", name="WT", rho0=0.03),
             Functional(type="XC", name="LDA")]
    first = [f(rho).energy for f in funcs]
    for _ in range(2):
        for f, e in zip(reversed(funcs), reversed(first)):
            assert f(rho).energy == e


def test_vw_gaussian_sigma_vs_oracle():
    """vW with the Gaussian filter exp(-G^2 sigma^2 / 4) of vw.py:45-49 (the
    KEDF's sigma option, applied to its vW part only), FP64, WT total."""
    from dftpy_amd import DirectField, systems
    from dftpy_amd.functional import Functional
    ions, nr = systems.c1_system()
    lat = ions.cell
    rho = _c1_density(nr, lat, seed=3)
    og = OG.Grid(lat, nr)
    g = _dg(lat, nr, False)
    for name, parts in (("vW", (OF.vw,)), ("WT", (OF.tf, OF.vw, OF.wt))):
        out = Functional(type="KEDF", name=name, sigma=0.3)(DirectField(g, data=rho))
        e_ref = sum((f(rho, og, sigma=0.3) if f is OF.vw else f(rho, og))[0] for f in parts)
        v_ref = sum((f(rho, og, sigma=0.3) if f is OF.vw else f(rho, og))[1] for f in parts)
        assert np.isclose(out.energy, e_ref, rtol=1e-12, atol=0)
        assert np.max(np.abs(out.potential.numpy() - v_ref)) < 1e-10 * np.max(np.abs(v_ref))
    e0 = Functional(type="KEDF", name="vW")(DirectField(g, data=rho)).energy
    assert abs(e0 - OF.vw(rho, og)[0]) < 1e-12 * abs(e0) and abs(e0 - out.energy) > 1e-6
