"""TN with the device-driven iteration (ofdft_optimize_tn_graph: control
kernels + one hipGraph launch per superstep) against the host-driven TN of
ofdft_optimize, which the C1 tests pin to the oracle, and against the oracle
itself."""
import numpy as np
import pytest
import torch

from test_gpu_optimize import _c1_oracle

pytestmark = pytest.mark.gpu


def _c1(dtype):
    from dftpy_amd import DirectField, DirectGrid, systems
    from dftpy_amd.functional import Functional, TotalFunctional
    ions, nr = systems.c1_system()
    grid = DirectGrid(ions.cell, nr, full=False, dtype=dtype)
    pp = Functional(type="PSEUDO", grid=grid, ions=ions, PP_list={"Al": systems.empty_core_recpot()})
    ev = TotalFunctional(KE=Functional(type="KEDF", name="WT"), XC=Functional(type="XC", name="LDA"),
                         H=Functional(type="HARTREE"), PSEUDO=pp)
    rho = DirectField(grid)
    rho[:] = ions.get_ncharges() / grid.volume
    return ions, ev, rho


def _run(ev, rho, **opts):
    from dftpy_amd.optimization import Optimization
    opt = Optimization(optimization_method="TN", EnergyEvaluator=ev, optimization_options=opts)
    out = opt.optimize_rho(rho)
    return opt, out


CASES = [(torch.float64, "fd", False), (torch.float64, "analytic", True), (torch.float32, "analytic", True),
         (torch.float32, "analytic", False)]


@pytest.mark.parametrize("dtype,hvp,pre", CASES)
def test_graph_tn_matches_host_tn_c1(dtype, hvp, pre):
    """C1 from the uniform density: the same status, iteration count and
    evaluation counts as the host-driven TN; energies after every step
    within 1e-9 Ha (FP64; the runs differ only in the last bits of cos/sin
    of the line-search step) or 1e-6 Ha/atom (FP32), and the oracle's
    minimum within 1e-6 Ha/atom."""
    print("case", dtype, hvp, pre, flush=True)
    ions, ev, rho = _c1(dtype)
    econv = 1e-8 if dtype == torch.float64 else 1e-6 * ions.nat
    opts = dict(econv=econv, maxiter=100, hvp=hvp, precond=pre)
    h, out_h = _run(ev, rho, **opts)
    g, out_g = _run(ev, rho, graph=True, **opts)
    print("host", h.stats, "\ngraph", g.stats, flush=True)
    assert h.converged == 0 and g.converged == 0
    assert g.stats["supersteps"] >= g.stats["niter"]
    ref = _c1_oracle(1e-10)
    assert abs(g.functional.energy - ref["energy"]) / ions.nat < (1e-6 if dtype == torch.float64 else 1e-5)
    if dtype == torch.float64:
        assert g.stats["niter"] == h.stats["niter"]
        assert g.stats["n_eval_v"] == h.stats["n_eval_v"] and g.stats["n_eval_ev"] == h.stats["n_eval_ev"]
        assert np.max(np.abs(g.history - h.history)) < 1e-9
        assert float(torch.max(torch.abs(out_g.data - out_h.data))) < 1e-9
    else:
        assert abs(g.stats["niter"] - h.stats["niter"]) <= 1
        assert abs(g.functional.energy - h.functional.energy) / ions.nat < 1e-6
        assert float(torch.max(torch.abs(out_g.data - out_h.data))) < 1e-4


def test_graph_tn_one_step_supersteps_c1():
    """One inner-CG iteration and one line-search evaluation per superstep
    (every section resumes across graph launches): the FP64 reference
    algorithm still follows the host-driven run step for step."""
    ions, ev, rho = _c1(torch.float64)
    opts = dict(econv=1e-8, maxiter=100, hvp="fd")
    h, out_h = _run(ev, rho, **opts)
    g, out_g = _run(ev, rho, graph=True, graph_cg_steps=1, graph_ls_steps=1, **opts)
    print("host", h.stats, "\ngraph", g.stats, flush=True)
    assert g.converged == h.converged == 0
    assert g.stats["niter"] == h.stats["niter"]
    assert g.stats["n_eval_v"] == h.stats["n_eval_v"] and g.stats["n_eval_ev"] == h.stats["n_eval_ev"]
    # a superstep per inner-CG iteration at least
    assert g.stats["supersteps"] >= g.stats["n_eval_v"]
    assert np.max(np.abs(g.history - h.history)) < 1e-9
    assert float(torch.max(torch.abs(out_g.data - out_h.data))) < 1e-9


def test_graph_tn_maxiter_and_history():
    """maxiter stops with status 2 after maxiter - 1 accepted steps, as the
    host loop does; the energy history matches step by step."""
    ions, ev, rho = _c1(torch.float64)
    opts = dict(econv=1e-14, maxiter=4, hvp="fd")
    h, _ = _run(ev, rho, **opts)
    g, _ = _run(ev, rho, graph=True, **opts)
    assert h.converged == g.converged == 2
    assert h.stats["niter"] == g.stats["niter"] == 3
    assert np.max(np.abs(g.history - h.history)) < 1e-9


def test_graph_tn_capture_is_reused_and_updated():
    """The superstep graph is captured once per plan and reused; a cell with
    another lattice on the same grid shape updates it in place (or captures
    anew) and still matches the host-driven TN."""
    from dftpy_amd import DirectField, DirectGrid, systems
    from dftpy_amd.functional import Functional, TotalFunctional
    from dftpy_amd.functional.pseudo import LocalPseudo, ReadPseudo
    readpp = ReadPseudo({"Al": systems.empty_core_recpot()})
    grid = None
    stats = []
    for k, strain in enumerate([0.0, 0.02]):
        ions = systems.fcc_al((2, 2, 2), sigma=0.1, strain=strain, seed=5)
        if grid is None:
            grid = DirectGrid(ions.cell, (64, 64, 64), full=False, dtype=torch.float32)
        grid.set_lattice(ions.cell)
        ev = TotalFunctional(KE=Functional(type="KEDF", name="WT"), XC=Functional(type="XC", name="LDA"),
                             H=Functional(type="HARTREE"), PSEUDO=LocalPseudo(grid=grid, ions=ions, readpp=readpp))
        rho = DirectField(grid)
        rho[:] = ions.get_ncharges() / grid.volume
        opts = dict(econv=1e-6 * ions.nat, maxiter=60, precond=True)
        h, _ = _run(ev, rho, **opts)
        for rep in range(2):  # the second run reuses the capture as it is
            g, _ = _run(ev, rho, graph=True, **opts)
            print(k, rep, "host", h.stats, "\ngraph", g.stats, flush=True)
            assert h.converged == g.converged == 0
            assert abs(g.functional.energy - h.functional.energy) / ions.nat < 1e-6
            stats.append((g.stats["graph_captures"], g.stats["graph_updates"]))
    assert stats[0][0] >= 1
    assert stats[1] == stats[0]
    # the strained cell: one more capture or in-place update, then reuse
    assert sum(stats[2]) == sum(stats[1]) + 1
    assert stats[3] == stats[2]


def test_graph_tn_needs_fixed_size_passes():
    """Grids with an axis outside the size-specialised FFT table, and LMGP,
    are refused with a message (the host-driven TN covers them)."""
    from dftpy_amd import DirectField, DirectGrid, _lib, systems
    from dftpy_amd.functional import Functional, TotalFunctional
    ions, _ = systems.c1_system()
    grid = DirectGrid(ions.cell, (30, 30, 30), full=False)
    pp = Functional(type="PSEUDO", grid=grid, ions=ions, PP_list={"Al": systems.empty_core_recpot()})
    ev = TotalFunctional(KE=Functional(type="KEDF", name="WT"), XC=Functional(type="XC", name="LDA"),
                         H=Functional(type="HARTREE"), PSEUDO=pp)
    rho = DirectField(grid)
    rho[:] = ions.get_ncharges() / grid.volume
    with pytest.raises(_lib.OFDFTError, match="size-specialised"):
        _run(ev, rho, econv=1e-8, maxiter=10, graph=True)
