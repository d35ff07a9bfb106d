"""Crank-Nicolson TD-OFDFT on the GPU against the reference's golden
dipole and the oracle."""
import numpy as np
import pytest
import torch

from conftest import data
from oracle import functionals as OF
from oracle import grid as OG
from oracle import io as OIO
from oracle import pseudo as OP
from oracle import td as OTD
from oracle import units as OU

pytestmark = pytest.mark.gpu


def test_cn_gaas_dipole():
    # examples/test/test_propagator.py:28-81, CN branch
    from dftpy_amd import DirectField, DirectGrid, formats
    from dftpy_amd.constants import Units
    from dftpy_amd.functional import Functional, TotalFunctional
    from dftpy_amd.functional.pseudo import LocalPseudo
    from dftpy_amd.td import Hamiltonian, Propagator, calc_rho, dipole, initial_kick
    ions, rho_np, lat = formats.read_xsf(data("GaAs_random.xsf"))
    rho_np = rho_np * (1.0 / Units.Bohr) ** 3
    grid = DirectGrid(lat, rho_np.shape, full=True)
    rho0 = DirectField(grid, data=rho_np)
    PP = LocalPseudo(grid=grid, ions=ions, PP_list={"Ga": data("Ga_lda.oe04.recpot"),
                                                     "As": data("As_lda.oe04.recpot")})
    ev = TotalFunctional(KE=Functional(type="KEDF", name="TF"), XC=Functional(type="XC", name="LDA"),
                         H=Functional(type="HARTREE"), PSEUDO=PP)
    ham = Hamiltonian()
    cn = Propagator(ham, 1e-3, name="crank-nicholson")
    psi = initial_kick(1e-3, 0, DirectField(grid, data=torch.sqrt(rho0.data)))
    ham.v = ev.compute(rho0, calcType=["V"]).potential
    E0 = ham.energy(psi)
    norm0 = (psi.data.abs() ** 2).sum().item()
    for _ in range(10):
        psi, info = cn(psi)
        assert info == 0
        rho = calc_rho(psi)
        ham.v = ev.compute(rho, calcType=["V"]).potential
    assert np.isclose(ham.energy(psi), E0, rtol=1e-3)
    d = dipole(rho, rho0)
    assert np.isclose(d[0], 0.011423, rtol=1e-3)
    assert np.isclose((psi.data.abs() ** 2).sum().item(), norm0, rtol=1e-9)
    # oracle, same steps
    oat, orho0, olat = OIO.read_xsf(data("GaAs_random.xsf"))
    orho0 = orho0 * (1.0 / OU.Bohr) ** 3
    og = OG.Grid(olat, orho0.shape, full=True)
    pp, z = {}, {}
    for k, f in (("Ga", "Ga_lda.oe04.recpot"), ("As", "As_lda.oe04.recpot")):
        gp, vp, zz = OIO.read_recpot(data(f))
        pp[k], z[k] = (gp, vp), zz
    oat.set_charges(z)
    vreal = OP.local_pp(og, oat, pp, pme=True)

    def pot(r):
        return OF.tf(r, og)[1] + OF.lda(r, og)[1] + OF.hartree(r, og)[1] + vreal

    ps = OTD.initial_kick(1e-3, 0, np.sqrt(orho0), og)
    v = pot(orho0)
    for _ in range(10):
        ps, st, _ = OTD.cn_step(ps, v, og, 1e-3)
        orho = np.real(ps * np.conj(ps))
        v = pot(orho)
    assert np.max(np.abs(psi.numpy() - ps)) < 1e-9
    assert np.allclose(d, OTD.dipole(orho, orho0, og), rtol=1e-7, atol=1e-12)


@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
def test_cn_batch_matches_single_runs_and_oracle(dtype):
    """Three kicked runs (x, y, z) of the 4-atom Al cell propagated together
    with ofdft_cn_step_batch, against the same runs one at a time and, for
    FP64, against the complex128 oracle."""
    from dftpy_amd import DirectField, DirectGrid, systems
    from dftpy_amd.functional import Functional, TotalFunctional
    from dftpy_amd.optimization import Optimization
    from dftpy_amd.td import CrankNicholson, CrankNicholsonBatch, Hamiltonian, calc_rho, initial_kick
    ions, nr = systems.c1_system()
    g = DirectGrid(ions.cell, nr, full=True, dtype=dtype)
    pp = Functional(type="PSEUDO", grid=g, ions=ions, PP_list={"Al": systems.empty_core_recpot()})
    ev = TotalFunctional(KE=Functional(type="KEDF", name="TF"), XC=Functional(type="XC", name="LDA"),
                         H=Functional(type="HARTREE"), PSEUDO=pp)
    rho0 = DirectField(g)
    rho0[:] = ions.get_ncharges() / g.volume
    rho0 = Optimization("TN", optimization_options={"econv": 1e-9, "maxiter": 50}, EnergyEvaluator=ev).optimize_rho(rho0)
    dt, nsteps, k = 0.1, 5, 1e-2
    tol = 1e-10 if dtype == torch.float64 else 1e-6
    psis = [initial_kick(k, d, DirectField(g, data=torch.sqrt(rho0.data))) for d in range(3)]
    batch = torch.stack([p.data for p in psis])
    vs = torch.stack([ev.compute(calc_rho(p), calcType=["V"]).potential.data for p in psis])
    cnb = CrankNicholsonBatch(dt, tol=tol, maxiter=200)
    singles = list(psis)
    ham = Hamiltonian()
    cn = CrankNicholson(ham, dt, tol=tol, maxiter=200)
    for _ in range(nsteps):
        batch, st = cnb(g, batch, vs)
        assert st == [0, 0, 0]
        for r in range(3):
            ham.v = ev.compute(calc_rho(singles[r]), calcType=["V"]).potential
            singles[r], info = cn(singles[r])
            assert info == 0
        vs = torch.stack([ev.compute(calc_rho(DirectField(g, data=batch[r], cplx=True)), calcType=["V"]).potential.data
                          for r in range(3)])
    for r in range(3):
        diff = (batch[r] - singles[r].data).abs().max().item()
        assert diff < (1e-11 if dtype == torch.float64 else 2e-5), (r, diff)
    if dtype == torch.float64:
        og = OG.Grid(ions.cell, nr, full=True)
        oat = OIO.Atoms(ions.symbols, ions.positions, ions.cell)
        gp, vp, z = OIO.read_recpot(systems.empty_core_recpot())
        oat.set_charges({"Al": z})
        vreal = OP.local_pp(og, oat, {"Al": (gp, vp)}, pme=True)

        def pot(rr):
            return OF.tf(rr, og)[1] + OF.lda(rr, og)[1] + OF.hartree(rr, og)[1] + vreal

        r0 = rho0.numpy()
        for r in range(3):
            ps = OTD.initial_kick(k, r, np.sqrt(r0), og)
            for _ in range(nsteps):
                ps, st, _ = OTD.cn_step(ps, pot(np.real(ps * np.conj(ps))), og, dt, tol=tol, maxiter=200)
            assert np.max(np.abs(batch[r].cpu().numpy() - ps)) < 1e-8


def test_predictor_corrector_and_current_vs_oracle():
    """Three runner steps (real_time_runner.py:127-144): PredictorCorrector
    with max_steps = 2 around the HIP Crank-Nicolson solve, the corrector
    potential from Hartree + PZ-LDA at (rho + rho_pred) / 2; psi, the current
    density and the potential against the complex128 oracle."""
    from dftpy_amd import DirectField, DirectGrid
    from dftpy_amd.functional import Functional, TotalFunctional
    from dftpy_amd.td import CrankNicholson, Hamiltonian, PredictorCorrector, calc_j
    from oracle import functionals as OF2
    from oracle import grid as OG2
    from oracle import td as OT
    lat = np.eye(3) * 7.6
    nr = (24, 24, 24)
    g = DirectGrid(lat, nr, full=True)
    og = OG2.Grid(lat, nr, full=True)
    r = og.r
    rho0 = 0.027 * (1 + 0.3 * np.cos(2 * np.pi * r[0] / 7.6) * np.cos(2 * np.pi * r[2] / 7.6))
    psi_np = np.sqrt(rho0) * np.exp(1j * 1e-2 * r[1])
    F = TotalFunctional(XC=Functional(type="XC", name="LDA"), H=Functional(type="HARTREE"))

    def pot_ref(rr):
        return OF2.lda(rr, og)[1] + OF2.hartree(rr, og)[1]

    v0 = F(DirectField(g, data=torch.as_tensor(rho0, device="cuda")), calcType={"V"}).potential
    H = Hamiltonian(v=v0)
    cn = CrankNicholson(H, 0.1)
    psi = DirectField(g, data=torch.as_tensor(psi_np, device="cuda"), cplx=True)
    v_np = pot_ref(rho0)
    for step in range(3):
        print("pc step", step, flush=True)
        pc = PredictorCorrector(psi, propagator=cn, functionals=F, max_steps=2)
        pc.run()
        psi = pc.psi_pred
        psi_np, v_np, n, _ = OT.predictor_corrector(psi_np, v_np, og, 0.1, pot_ref)
        assert pc.nsteps == n == 2
        assert np.max(np.abs(psi.data.cpu().numpy() - psi_np)) < 1e-9
        assert np.max(np.abs(H.v.data.cpu().numpy() - v_np)) < 1e-9 * np.max(np.abs(v_np))
    j = calc_j(psi).cpu().numpy()
    j_ref = OT.calc_j(psi_np, og)
    assert np.max(np.abs(j - j_ref)) < 1e-8 * np.max(np.abs(j_ref))


@pytest.mark.parametrize("solver,dtype", [("bicgstab", torch.float64), ("gmres", torch.float64), ("bicg", torch.float64),
                                          ("bicgstab", torch.float32), ("gmres", torch.float32)])
def test_cn_linear_solvers_vs_oracle(solver, dtype):
    """One Crank-Nicolson step of a kicked 4-atom Al orbital (32^3, dt 0.1)
    with the linear solvers of crank_nicholson.py:65-77 against the complex128
    oracle with the same solver: bicgstab follows linear_solver.py:78-110 step
    for step; GMRES(20) reaches scipy's gmres solution within the tolerance;
    bicg (CG with conjugated sums) is compared after 6 iterations, where both
    still fall short of the tolerance on this non-Hermitian operator."""
    from dftpy_amd import DirectField, DirectGrid, systems
    from dftpy_amd.td import CrankNicholson, Hamiltonian, initial_kick
    ions, nr = systems.c1_system()
    g = DirectGrid(ions.cell, nr, full=True, dtype=dtype)
    og = OG.Grid(ions.cell, nr, full=True)
    r = og.r
    rho = 0.027 * (1 + 0.3 * np.cos(2 * np.pi * r[0] / 7.6) * np.cos(2 * np.pi * r[2] / 7.6))
    v = 0.2 * np.cos(2 * np.pi * r[1] / 7.6) - 0.1
    psi = initial_kick(1e-2, 0, DirectField(g, data=torch.as_tensor(np.sqrt(rho), device="cuda")))
    ham = Hamiltonian(v=DirectField(g, data=torch.as_tensor(v, device="cuda")))
    tol = 1e-10 if dtype == torch.float64 else 1e-6
    maxiter = 6 if solver == "bicg" else 200
    cn = CrankNicholson(ham, 0.1, linearsolver=solver, tol=tol, maxiter=maxiter)
    print("cn solver", solver, dtype, flush=True)
    out, info = cn(psi)
    ps0 = OTD.initial_kick(1e-2, 0, np.sqrt(rho), og)
    ref, st, _ = OTD.cn_step(ps0, v, og, 0.1, tol=tol, maxiter=maxiter if solver != "gmres" else 10,
                             solver=solver)
    err = np.max(np.abs(out.data.cpu().numpy() - ref)) / np.max(np.abs(ref))
    print("  status %d (oracle %d), %d iterations, rel. max error %.2e" % (info, st, cn.last_info["iterations"], err))
    if solver == "bicg":
        assert info == 1 and st == 1
        assert err < 1e-10
    else:
        assert info == 0 and st == 0
        assert err < (1e-8 if dtype == torch.float64 else 1e-5)
