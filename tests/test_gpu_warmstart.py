"""What the trained warm-start network gains over the method without it,
on held-out cells (108-atom fcc Al, 96^3, seeds outside the training range
0..1999): the start density's energy error, and the evaluations the FP32 TN
optimizer needs under the reference stop rule.  Both starts must reach the
same minimum (within 1e-6 Ha/atom): the optimizer remains the arbiter."""
import pytest

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("seed", [10001, 10002])
def test_learned_start_gains_over_superposition(seed):
    from dftpy_amd.learned import WarmStartNet
    from dftpy_amd.warmstart_train import CellFactory, converge, energy64
    net = WarmStartNet.load()
    fac = CellFactory()
    ions, ev, ne, rho0, x0 = fac.cell(seed, net)
    nat = ions.nat
    ref_rho, _, _ = converge(ev, net.start_density(fac.g32, x0, rho0, ne, use_net=True), 1e-9, nat, 100)
    estar = energy64(ev, fac.g64, ref_rho)
    res = {}
    for name, use in (("superposition", False), ("learned", True)):
        start = net.start_density(fac.g32, x0, rho0, ne, use_net=use)
        e0 = energy64(ev, fac.g64, start)
        rho, opt, secs = converge(ev, start, 1e-6, nat, 60)
        e1 = energy64(ev, fac.g64, rho)
        res[name] = dict(start_err=(e0 - estar) / nat, final_err=(e1 - estar) / nat, iters=opt.stats["niter"],
                         evals=opt.stats["n_eval_v"] + opt.stats["n_eval_ev"], seconds=secs)
    print(seed, res)
    sup, lrn = res["superposition"], res["learned"]
    assert 0 <= lrn["start_err"] < 0.05 * sup["start_err"]
    assert lrn["evals"] < sup["evals"]
    assert abs(lrn["final_err"]) <= 1e-6 and abs(sup["final_err"]) <= 1e-6


@pytest.mark.parametrize("seed", [10003])
def test_learned_start_gains_over_uniform_with_preconditioner(seed):
    """Against the uniform density, both runs with the preconditioned inner CG
    (the benchmark's form): the learned start needs fewer evaluations and
    both reach the same minimum."""
    import torch
    from dftpy_amd import DirectField
    from dftpy_amd.learned import WarmStartNet
    from dftpy_amd.optimization import Optimization
    from dftpy_amd.warmstart_train import CellFactory, converge, energy64
    net = WarmStartNet.load()
    fac = CellFactory()
    ions, ev, ne, rho0, x0 = fac.cell(seed, net)
    nat = ions.nat
    ref_rho, _, _ = converge(ev, net.start_density(fac.g32, x0, rho0, ne, use_net=True), 1e-9, nat, 100)
    estar = energy64(ev, fac.g64, ref_rho)
    uniform = DirectField(fac.g32)
    uniform[:] = ne / fac.g32.volume
    res = {}
    for name, start in (("uniform", uniform), ("learned", net.start_density(fac.g32, x0, rho0, ne, use_net=True))):
        opt = Optimization(optimization_method="TN", EnergyEvaluator=ev,
                           optimization_options={"econv": 1e-6 * nat, "maxiter": 60, "precond": True})
        torch.cuda.synchronize()
        rho = opt.optimize_rho(start)
        e1 = energy64(ev, fac.g64, rho)
        res[name] = dict(final_err=(e1 - estar) / nat, iters=opt.stats["niter"], status=opt.stats["converged"],
                         evals=opt.stats["n_eval_v"] + opt.stats["n_eval_ev"])
    print(seed, res)
    uni, lrn = res["uniform"], res["learned"]
    assert uni["status"] == 0 and lrn["status"] == 0
    assert lrn["evals"] < uni["evals"]
    assert abs(lrn["final_err"]) <= 1e-6 and abs(uni["final_err"]) <= 1e-6


@pytest.mark.parametrize("pme", [True, False])
@pytest.mark.parametrize("dtype", ["float64", "float32"])
def test_screened_superposition_matches_radial_table(pme, dtype):
    """The superposition of screened ions evaluated per G point
    (ofdft_screened_superposition: -v_s(|G|) / K(|G|)) against the same
    screening densities as interpolating radial tables through the local-PP
    setup (learned.response_table): the two interpolate one function with
    the same nodes, one as the quotient's spline, the other as the PP
    table's spline (steep, Coulomb-like at small q) over the exact K, so they
    differ by interpolation error between the nodes (measured 8.7e-6 of the
    density's fluctuation in FP64 with PME)."""
    import numpy as np
    import torch
    from dftpy_amd import DirectGrid, _lib, systems
    from dftpy_amd.functional.pseudo import ReadPseudo
    from dftpy_amd.learned import response_table, superposition_density
    dt = getattr(torch, dtype)
    ions = systems.fcc_al((2, 2, 3), sigma=0.1, strain=0.01, seed=4)
    nr = (48, 48, 72)
    g = DirectGrid(ions.cell, nr, full=False, dtype=dt)
    rp = ReadPseudo({"Al": systems.empty_core_recpot()})
    ions.set_charges(rp.zval)
    rho0 = float(ions.get_ncharges()) / g.volume
    got = superposition_density(g, ions, rp, rho0, pme=pme).double().cpu().numpy()
    t = response_table(rp.gp["Al"], rp.vp["Al"], rp.zval["Al"], rho0)
    coef = np.ascontiguousarray(t.coef.reshape(-1))
    cptr = (_lib._DP * 1)(_lib.dptr(coef))
    arrs = [np.array([t.q0]), np.array([t.dq]), np.array([t.coef.shape[0]], dtype=np.int32), np.array([t.qmax]),
            np.array(list(t.vsmall)), np.array([t.dqs])]
    pos = np.ascontiguousarray(ions.positions)
    frac = np.ascontiguousarray(ions.get_scaled_positions())
    spec = ions.species_index(["Al"])
    ref = torch.empty(nr, dtype=dt, device="cuda")
    _lib.check(_lib.lib.ofdft_local_pp(g.plan(dt).handle, 1, _lib.dptr(arrs[0]), _lib.dptr(arrs[1]),
                                       _lib.iptr(arrs[2]), cptr, _lib.dptr(arrs[3]), _lib.dptr(arrs[4]),
                                       _lib.dptr(arrs[5]), ions.nat, _lib.dptr(pos), _lib.dptr(frac),
                                       _lib.iptr(spec), int(pme), 10, ref.data_ptr(), _lib.stream()))
    ref = ref.double().cpu().numpy()
    err = np.max(np.abs(got - ref)) / np.max(np.abs(ref - ref.mean()))
    print("screened superposition vs radial table (%s, pme %s): %.2e" % (dtype, pme, err))
    assert err < 3e-5
    assert abs(got.mean() - ref.mean()) < 1e-9 * abs(ref.mean()) + (1e-6 if dtype == "float32" else 0.0)


def test_tn_preconditioner_table_on_device():
    """The TN preconditioner table built on the device (uniform-gas response
    coefficients) against the host table of optimization.tn_preconditioner."""
    import numpy as np
    import torch
    from dftpy_amd import DirectGrid, _lib, systems
    from dftpy_amd.optimization import tn_preconditioner, tn_preconditioner_coefs
    ions = systems.fcc_al((2, 2, 2), sigma=0.1, strain=-0.02, seed=1)
    g = DirectGrid(ions.cell, (40, 40, 40), full=False, dtype=torch.float32)
    plan = g.plan(torch.float32)
    kedf = dict(x=1.0, y=1.0, alpha=5.0 / 6.0, beta=5.0 / 6.0)
    terms = _lib.TERM_TF | _lib.TERM_VW | _lib.TERM_WT | _lib.TERM_XC | _lib.TERM_HARTREE | _lib.TERM_PP
    for rho0, t in ((0.0268, terms), (0.0301, terms & ~_lib.TERM_HARTREE)):
        tab, dq = tn_preconditioner(g, rho0, t, kedf)
        npt, dq2, resp = tn_preconditioner_coefs(g, rho0, t, kedf)
        assert npt == tab.size and abs(dq2 - dq) <= 1e-15 * dq
        _lib.check(_lib.lib.ofdft_set_tn_precond_uniform(plan.handle, npt, dq2, rho0, _lib.dptr(resp), _lib.stream()))
        dev = torch.empty(npt, dtype=torch.float64, device="cuda")
        _lib.check(_lib.lib.ofdft_get_tn_precond(plan.handle, dev.data_ptr(), npt, _lib.stream()))
        dev = dev.cpu().numpy()
        assert np.allclose(dev, tab, rtol=1e-12, atol=0)
