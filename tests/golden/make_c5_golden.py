"""Writes c5_oracle.json: the CPU oracle's (complex128 numpy) batched TD run
of the 256-atom fcc Al cell (4x4x4, a = 7.60 bohr, sigma = 0.1 bohr
displacements with seed 7, 128^3 full grid).

Ground state: TF + vW + WT + PZ-LDA + Hartree + the empty-core local PP
(PME), TN from the uniform density to econv 1e-10 Ha/atom.  Propagation
(td/propagator/crank_nicholson.py:147-188 with linear_solver.cg at the
reference's default tol 1e-8): six runs, delta kicks of 1e-3 and 2e-3 bohr^-1
along x, y and z (td/utils.py:10-16), dt = 0.1 a.u., 20 steps, the potential
TF + WT nonlocal + LDA + Hartree + PP re-evaluated from |psi|^2 after every
step (the kinetic operator carries vW), and the dipole int (rho - rho0) r dV
along the kick (real_time_runner.py:251-253) after every step.  The six runs
go to six worker processes.  Run time: several minutes.

    python tests/golden/make_c5_golden.py
"""
import json
import multiprocessing as mp
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from dftpy_amd import systems  # noqa: E402  (structure builder only)
from oracle import functionals as F  # noqa: E402
from oracle import grid as G  # noqa: E402
from oracle import io as OIO  # noqa: E402
from oracle import optimize as O  # noqa: E402
from oracle import pseudo as P  # noqa: E402
from oracle import td as T  # noqa: E402

KICKS = [(1e-3, 0), (1e-3, 1), (1e-3, 2), (2e-3, 0), (2e-3, 1), (2e-3, 2)]
NSTEPS, DT, TOL, MAXITER = 20, 0.1, 1e-8, 200
STATE = {}


def setup():
    ions = systems.fcc_al((4, 4, 4), sigma=0.1, seed=7)
    nr = (128, 128, 128)
    at = OIO.Atoms(ions.symbols, ions.positions, ions.cell)
    gp, vp, z = OIO.read_recpot(systems.empty_core_recpot())
    at.set_charges({"Al": z})
    gh = G.Grid(ions.cell, nr)
    gf = G.Grid(ions.cell, nr, full=True)
    return ions, at, gh, gf, {"Al": (gp, vp)}


def td_potential(rho, gf, vreal, kern):
    """TF + WT nonlocal + LDA + Hartree + local PP on the full grid."""
    return (F.tf(rho, gf)[1] + F.wt(rho, gf, kernel=kern, calc_e=False)[1] + F.lda(rho, gf)[1]
            + F.hartree(rho, gf)[1] + vreal)


def run(i):
    k, d = KICKS[i]
    gf, rho0, vreal, kern = STATE["gf"], STATE["rho0"], STATE["vreal"], STATE["kern"]
    psi = T.initial_kick(k, d, np.sqrt(rho0), gf)
    v = td_potential(rho0, gf, vreal, kern)
    dip, its = [], []
    for s in range(NSTEPS):
        psi, info, nit = T.cn_step(psi, v, gf, DT, tol=TOL, maxiter=MAXITER)
        if info != 0:
            raise RuntimeError("CG did not converge (run %d step %d)" % (i, s))
        rho = np.real(psi * np.conj(psi))
        v = td_potential(rho, gf, vreal, kern)
        dip.append(float(((rho - rho0) * gf.r[d]).sum() * gf.dV))
        its.append(int(nit))
        print("run %d step %d: dipole %.12e, %d CG iterations" % (i, s, dip[-1], nit), flush=True)
    return dip, its, float((np.abs(psi) ** 2).sum() * gf.dV)


def main(out=os.path.join(os.path.dirname(os.path.abspath(__file__)), "c5_oracle.json")):
    ions, at, gh, gf, pp = setup()
    t0 = time.time()
    vreal = P.local_pp(gh, at, pp, pme=True)
    eII = P.ewald_energy(gh, at, pme=True)
    cache = {}

    def ev(r, need_e):
        r0 = r.mean()
        if "k" not in cache or abs(cache["r0"] - r0) > 1e-6:
            cache["k"], cache["r0"] = F.wt_kernel(gh.q, r0), r0
        e, v, _ = F.total(r, gh, vreal, cache["k"], ewald=eII)
        return e, v

    rho = np.full(tuple(gh.nr), at.ncharge() / gh.volume)
    res = O.optimize(ev, rho, gh.dV, method="TN", econv=1e-10 * ions.nat, maxiter=100, maxfun=50, maxls=30,
                     log=lambda *a: print("%3d %.12f %.3e %.3e %d %d" % a, flush=True))
    print("ground state %.12f (status %d) in %.1f s" % (res["energy"], res["converged"], time.time() - t0), flush=True)
    rho0 = res["rho"]
    STATE.update(gf=gf, rho0=rho0, vreal=vreal, kern=F.wt_kernel(gf.q, rho0.mean()))
    t0 = time.time()
    with mp.get_context("fork").Pool(len(KICKS)) as pool:
        runs = pool.map(run, range(len(KICKS)))
    secs = time.time() - t0
    d = dict(system="fcc Al 4x4x4, a=7.60 bohr, sigma=0.1 seed=7, 128^3 full grid, WT+LDA+H+emptycore(Z=3,rc=1.12), "
                    "PME; TD potential TF+WT-NL+LDA+H+PP",
             nat=ions.nat, ground_energy=res["energy"], ground_status=res["converged"],
             ground_history=[float(x) for x in res["history"]], ground_rho_max=float(rho0.max()),
             ground_rho_min=float(rho0.min()), kicks=KICKS, dt=DT, nsteps=NSTEPS, cg_tol=TOL,
             dipole=[r[0] for r in runs], cg_iterations=[r[1] for r in runs], final_norm=[r[2] for r in runs],
             seconds_propagation=secs)
    with open(out, "w") as f:
        json.dump(d, f, indent=1)
    print("wrote", out, "in %.1f s" % secs)


if __name__ == "__main__":
    main()
