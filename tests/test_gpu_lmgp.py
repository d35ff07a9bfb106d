"""LMGP nonlocal KEDF on the GPU (csrc/lmgp.hip) against the oracle
(oracle/lmgp.py) and the reference golden."""
import numpy as np
import pytest
import torch

from conftest import data
from oracle import functionals as OF
from oracle import grid as OG
from oracle import io as OIO
from oracle import lmgp as OL

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def tables():
    return OL.kernel_tables()


def test_mgp_table_matches_oracle(tables):
    """The GPU t-integral (exact Lindhard, clamped like the reference's
    splev ext=3) against the oracle's spline-of-Lindhard table."""
    from dftpy_amd.functional.lmgp import lmgp_tables
    eta, k, _, _ = tables
    eta2, kc, dc = lmgp_tables()
    kg = kc.reshape(-1, 4)[:, 0]
    assert np.array_equal(eta, eta2)
    assert np.max(np.abs(kg - k[:-1])) < 1e-9 * np.max(np.abs(k))


@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
def test_lmgp_golden_and_potential(tables, dtype):
    from dftpy_amd import DirectField, DirectGrid
    from dftpy_amd.functional import Functional
    atoms, rho, lat = OIO.read_qepp(data("Al_fde_rho.pp"))
    g = DirectGrid(lat, rho.shape, full=True, dtype=dtype)
    out = Functional(type="KEDF", name="LMGP")(DirectField(g, data=rho))
    og = OG.Grid(lat, rho.shape, full=True)
    e_nl, v_nl = OL.lmgp_nl(rho, og, tables)
    e_ref = OF.tf(rho, og)[0] + OF.vw(rho, og)[0] + e_nl
    v_ref = OF.tf(rho, og)[1] + OF.vw(rho, og)[1] + v_nl
    v = out.potential.numpy().astype(np.float64)
    print("E %.12f ref %.12f golden 2.9145872498710492, max dV %.3e" % (out.energy, e_ref, np.abs(v - v_ref).max()))
    if dtype == torch.float64:
        assert abs(out.energy - 2.9145872498710492) < 1e-8
        assert abs(out.energy - e_ref) < 1e-9
        assert np.max(np.abs(v - v_ref)) < 1e-8 * np.max(np.abs(v_ref))
    else:
        assert abs(out.energy - e_ref) < 2e-5 * abs(e_ref)
        assert np.max(np.abs(v - v_ref)) < 1e-4 * np.max(np.abs(v_ref))


def test_lmgp_density_optimization(tables):
    """TN (FP64, forward-difference Hessian as the reference) with TF + vW +
    LMGP + LDA + Hartree + local PP on the 4-atom Al cell from the uniform
    density, against the oracle's run of the same algorithm: the same
    per-iteration energies, the same iteration count and the same exit
    status.  (Both stop with a line-search failure once the energy changes
    reach ~1e-6 Ha: the k_F ladder of LMGP changes between evaluations.)"""
    from dftpy_amd import DirectField, DirectGrid, systems
    from dftpy_amd.functional import Functional, TotalFunctional
    from dftpy_amd.optimization import Optimization
    from oracle import io as OI
    from oracle import optimize as OO
    from oracle import pseudo as OP
    ions, nr = systems.c1_system()
    g = DirectGrid(ions.cell, nr, full=False)
    pp = Functional(type="PSEUDO", grid=g, ions=ions, PP_list={"Al": systems.empty_core_recpot()})
    ev = TotalFunctional(ewald=type("NoEwald", (), {"energy": 0.0})(), KE=Functional(type="KEDF", name="LMGP"),
                         XC=Functional(type="XC", name="LDA"), H=Functional(type="HARTREE"), PSEUDO=pp)
    rho = DirectField(g)
    rho[:] = ions.get_ncharges() / g.volume
    opt = Optimization("TN", optimization_options={"econv": 1e-8, "maxiter": 60}, EnergyEvaluator=ev)
    opt.optimize_rho(rho)
    og = OG.Grid(ions.cell, nr)
    oat = OI.Atoms(ions.symbols, ions.positions, ions.cell)
    gp, vp, z = OI.read_recpot(systems.empty_core_recpot())
    oat.set_charges({"Al": z})
    vreal = OP.local_pp(og, oat, {"Al": (gp, vp)}, pme=True)
    state = {}

    def oev(r, calc=True):
        parts = [OF.tf(r, og), OF.vw(r, og), OL.lmgp_nl(r, og, tables, state), OF.lda(r, og), OF.hartree(r, og)]
        return sum(p[0] for p in parts) + float((vreal * r).sum() * og.dV), sum(p[1] for p in parts) + vreal

    ref = OO.optimize(oev, np.full(nr, oat.ncharge() / og.volume), og.dV, econv=1e-8, maxiter=60)
    print(opt.stats, opt.history, ref["converged"], ref["history"])
    assert opt.stats["converged"] == ref["converged"]
    assert len(opt.history) == len(ref["history"])
    # the forward-difference Hessian (eps = 1e-9) amplifies rounding
    # differences to ~1e-8 Ha mid-trajectory; both reach the same minimum
    assert np.max(np.abs(np.array(opt.history) - np.array(ref["history"]))) < 1e-7
    assert abs(opt.history[-1] - ref["history"][-1]) < 5e-9
