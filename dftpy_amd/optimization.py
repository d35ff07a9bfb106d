"""Optimization (optimization/optimization.py:31-542 of the reference):
direct minimisation over phi = sqrt(rho), run natively in libofdft_hip
(``ofdft_optimize``) on the fused evaluation of a TotalFunctional."""
import numpy as np
import torch

from . import _lib
from .field import DirectField
from .functional import AbstractFunctional, FunctionalOutput, configure
from .functional.total_functional import TotalFunctional
from .math_utils import uniform_response_coefs, uniform_response_kernel

METHODS = {"TN": 0, "CG-HS": 1, "CG": 1, "CG-PR": 2, "CG-FR": 3, "CG-DY": 4, "CG-CD": 5, "CG-LS": 6, "GD": 7,
           "LBFGS": 8}


def _method_code(method):
    """Method name -> ofdft_optimize code.  Any other "CG..." name takes the
    Fletcher-Reeves beta, the fall-through branch of math_utils.py:420-421."""
    if method in METHODS:
        return METHODS[method]
    if method[0:2] == "CG":
        return METHODS["CG-FR"]
    raise NotImplementedError("optimization method %s is not implemented" % method)
PRECOND_POINTS = 4096


def tn_preconditioner(grid, rho0, terms, kedf):
    """Radial table (values, dq) of M^-1(q), M(q) = 2 rho0 K(q): the Hessian
    of the TN inner solve, (V - mu) + phi K phi-linearised, at the uniform
    density rho0 (K from math_utils.uniform_response_kernel over the terms of
    the functional; M^-1(0) = 0 when Hartree is present).  Covers |G| up to
    half the sum of the reciprocal axes times the grid sizes."""
    lat = np.asarray(grid.lattice, dtype=np.float64)
    B = 2.0 * np.pi * np.linalg.inv(lat).T
    nr = [int(n) for n in grid.nrR]
    qmax = 0.5 * sum(nr[i] * np.linalg.norm(B[i]) for i in range(3)) * 1.01
    q = np.linspace(0.0, qmax, PRECOND_POINTS)
    K = uniform_response_kernel(q, rho0, tf=bool(terms & _lib.TERM_TF), vw=bool(terms & _lib.TERM_VW),
                                wt=bool(terms & (_lib.TERM_WT | _lib.TERM_LMGP)), xc=bool(terms & _lib.TERM_XC),
                                hartree=bool(terms & _lib.TERM_HARTREE), x=kedf["x"], y=kedf["y"],
                                alpha=kedf["alpha"], beta=kedf["beta"])
    M = 2.0 * rho0 * K
    with np.errstate(divide="ignore"):
        tab = np.where(np.isinf(M), 0.0, 1.0 / np.maximum(M, 1e-8))
    return np.ascontiguousarray(tab), float(q[1] - q[0])


def tn_preconditioner_coefs(grid, rho0, terms, kedf):
    """The table of tn_preconditioner as the device builds it: (points, dq,
    the response coefficients of math_utils.uniform_response_coefs)."""
    lat = np.asarray(grid.lattice, dtype=np.float64)
    B = 2.0 * np.pi * np.linalg.inv(lat).T
    nr = [int(n) for n in grid.nrR]
    qmax = 0.5 * sum(nr[i] * np.linalg.norm(B[i]) for i in range(3)) * 1.01
    resp = uniform_response_coefs(rho0, tf=bool(terms & _lib.TERM_TF), vw=bool(terms & _lib.TERM_VW),
                                  wt=bool(terms & (_lib.TERM_WT | _lib.TERM_LMGP)), xc=bool(terms & _lib.TERM_XC),
                                  hartree=bool(terms & _lib.TERM_HARTREE), x=kedf["x"], y=kedf["y"],
                                  alpha=kedf["alpha"], beta=kedf["beta"])
    return PRECOND_POINTS, qmax / (PRECOND_POINTS - 1), np.ascontiguousarray(resp)


class Optimization:
    def __init__(self, optimization_method="CG-HS", optimization_options=None, EnergyEvaluator=None, guess_rho=None):
        self.rho = guess_rho
        self.optimization_options = dict(optimization_options or {})
        defaults = {"maxcor": 6, "ftol": 1.0e-7, "xtol": 1.0e-12, "maxfun": 50, "maxiter": 200, "maxls": 30,
                    "econv": 1.0e-5, "c1": 1e-4, "c2": 0.2, "algorithm": "EMM", "vector": "Orthogonalization",
                    "ncheck": 2, "hvp": "auto", "precond": False, "graph": False, "graph_cg_steps": 0,
                    "graph_ls_steps": 0}
        for k, v in defaults.items():
            self.optimization_options.setdefault(k, v)
        if EnergyEvaluator is None:
            raise AttributeError("Must provide an energy evaluator")
        if not isinstance(EnergyEvaluator, AbstractFunctional):
            raise TypeError("EnergyEvaluator must be a dftpy_amd functional")
        _method_code(optimization_method)
        o = self.optimization_options
        if o["algorithm"] != "EMM" or o["vector"] != "Orthogonalization":
            raise NotImplementedError("only the EMM algorithm with orthogonalised rotations is implemented")
        self.EnergyEvaluator = EnergyEvaluator
        self.optimization_method = optimization_method
        self.stats = None
        self.history = None

    def optimize_rho(self, guess_rho=None, guess_phi=None, lphi=False):
        if guess_rho is None and self.rho is None:
            raise AttributeError("Must provide a guess density")
        rho_in = guess_rho if guess_rho is not None else self.rho
        if guess_phi is not None:
            rho_in = guess_phi * guess_phi
        grid = rho_in.grid
        plan = grid.plan(grid.dtype)
        ev = self.EnergyEvaluator
        funcs = list(ev.funcDict.values()) if isinstance(ev, TotalFunctional) else [ev]
        eion = ev.eion() if isinstance(ev, TotalFunctional) else 0.0
        terms_all, vloc = configure(plan, funcs, grid, eion)
        o = self.optimization_options
        # Hessian-vector product of the TN inner solve: the reference's forward
        # difference (eps = 1e-9) needs FP64; FP32 grids use the exact linear
        # response of the potential
        hvp = o["hvp"]
        lmgp = bool(ev.terms & _lib.TERM_LMGP)
        if hvp == "auto":
            hvp = "fd" if (grid.dtype == torch.float64 or lmgp) else "analytic"
        if lmgp and hvp == "fd" and grid.dtype != torch.float64:
            raise NotImplementedError("LMGP in FP32 needs the analytic Hessian product, which does not cover LMGP; "
                                      "optimise LMGP on an FP64 grid")
        if hvp not in ("fd", "analytic"):
            raise ValueError("hvp must be 'auto', 'fd' or 'analytic'")
        ks = plan.kedf_settings
        if hvp == "analytic" and getattr(plan, "has_core", False):
            raise NotImplementedError("the analytic Hessian product does not cover a core density in XC; optimise "
                                      "on an FP64 grid (forward-difference Hessian)")
        if hvp == "analytic" and ev.terms & _lib.TERM_WT and abs(ks["alpha"] - ks["beta"]) > 1e-9:
            raise NotImplementedError("the analytic Hessian product covers WT with alpha == beta only; optimise WT "
                                      "with alpha != beta on an FP64 grid (forward-difference Hessian)")
        hvp = 0 if hvp == "fd" else 1
        # preconditioned inner CG of TN (not in the reference): opt-in
        if o["precond"] and self.optimization_method == "TN":
            # rho0 = N_e / V from the ions when the functional has them (no device read)
            rho0 = None
            for f in funcs:
                if getattr(f, "type", "") == "PSEUDO":
                    rho0 = float(f.ions.get_ncharges()) / grid.volume
            if rho0 is None:
                rho0 = float(rho_in.data.mean())
            key = (float("%.8e" % rho0), int(terms_all), tuple(sorted(plan.kedf_settings.items())))
            if getattr(plan, "tn_precond_key", None) != key:
                if grid.dtype == torch.float32:
                    # FP32-bulk: the table is built on the device, in stream order
                    npt, dq, resp = tn_preconditioner_coefs(grid, key[0], key[1], plan.kedf_settings)
                    _lib.check(_lib.lib.ofdft_set_tn_precond_uniform(plan.handle, npt, dq, key[0], _lib.dptr(resp),
                                                                     _lib.stream()))
                else:
                    # FP64 keeps the host table: its forward-difference TN
                    # follows ulp-level changes of the table into other
                    # inner-CG stops (19 vs 25 evaluations at 500 atoms)
                    tab, dq = tn_preconditioner(grid, key[0], key[1], plan.kedf_settings)
                    _lib.check(_lib.lib.ofdft_set_tn_precond(plan.handle, int(tab.size), dq, _lib.dptr(tab)))
                plan.tn_precond_key = key
        elif getattr(plan, "tn_precond_key", None) is not None:
            _lib.check(_lib.lib.ofdft_set_tn_precond(plan.handle, 0, 0.0, None))
            plan.tn_precond_key = None
        h0 = 0.0
        if self.optimization_method == "LBFGS":
            # optimization.py:346 reads optimization_options["h0"], which has no default
            h0 = float(o["h0"]) if o["h0"] else 0.0
            _lib.check(_lib.lib.ofdft_plan_reserve_lbfgs(plan.handle, int(o["maxcor"])))
        opts = np.array([o["econv"], o["maxiter"], o["maxfun"], o["maxls"], o["c1"], o["c2"], o["xtol"], o["ncheck"],
                         o["maxcor"], h0], dtype=np.float64)
        stats = np.zeros(16)
        hcap = int(o["maxiter"]) + 2
        hist = np.zeros(2 * hcap)
        rho = rho_in.data.clone().contiguous()
        pot = torch.empty_like(rho)
        graph = bool(o["graph"])
        if graph and self.optimization_method != "TN":
            raise NotImplementedError("the device-driven (graph) iteration covers TN only")
        if graph:
            # device-driven TN (not in the reference): one hipGraph launch per
            # superstep of cg_steps inner-CG iterations and ls_steps line-search
            # evaluations; by default sized to the usual counts of the form
            cg_steps = int(o["graph_cg_steps"]) or (2 if (o["precond"] or hvp == 1) else 8)
            ls_steps = int(o["graph_ls_steps"]) or (1 if grid.dtype == torch.float32 else 2)
            _lib.check(_lib.lib.ofdft_optimize_tn_graph(plan.handle, rho.data_ptr(), pot.data_ptr(), hvp,
                                                        _lib.dptr(opts), cg_steps, ls_steps, _lib.dptr(stats),
                                                        _lib.dptr(hist), hcap, _lib.stream()))
        else:
            _lib.check(_lib.lib.ofdft_optimize(plan.handle, rho.data_ptr(), pot.data_ptr(),
                                               _method_code(self.optimization_method), hvp, _lib.dptr(opts),
                                               _lib.dptr(stats), _lib.dptr(hist), hcap, _lib.stream()))
        del vloc
        n = int(stats[1]) + 1
        self.history = hist[:n].copy()
        self.history_seconds = hist[hcap:hcap + n].copy()
        self.stats = dict(converged=int(stats[0]), niter=int(stats[1]), energy=stats[2], mu=stats[3],
                          n_eval_v=int(stats[4]), n_eval_ev=int(stats[5]), seconds=stats[6])
        if graph:
            self.stats.update(supersteps=int(stats[7]), graph_captures=int(stats[8]), graph_updates=int(stats[9]))
        self.converged = int(stats[0])
        self.mu = stats[3]
        self.rho = DirectField(grid, data=rho)
        # the last evaluation, as optimization.py:517 keeps it
        self.functional = FunctionalOutput(name="TOTAL", energy=stats[2], potential=DirectField(grid, data=pot))
        return self.rho

    @property
    def phi(self):
        """sqrt(rho) of the last result (computed when asked for)."""
        if self.rho is None:
            return None
        return DirectField(self.rho.grid, data=torch.sqrt(torch.clamp(self.rho.data, min=0)))

    def __call__(self, guess_rho=None, calcType={"E", "V"}, guess_phi=None, lphi=False):
        return self.optimize_rho(guess_rho=guess_rho, guess_phi=guess_phi, lphi=lphi)
