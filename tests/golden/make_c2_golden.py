"""Writes c2_oracle.json: the CPU oracle's (FP64 numpy) converged ground
state of the 500-atom fcc Al cell (5x5x5, a = 7.60 bohr, sigma = 0.1 bohr
displacements with seed 2024, 160^3 grid, WT + PZ-LDA + Hartree + empty-core
PP, PME structure factor), started from the uniform density with TN and
econv = 1e-9 Ha/atom.  Also records per-term energies of the uniform start
and a few potential checksums.  Run time: several minutes on one core.

    python tests/golden/make_c2_golden.py
"""
import json
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


def main(out=os.path.join(os.path.dirname(os.path.abspath(__file__)), "c2_oracle.json")):
    ions, nr = systems.c2_system()
    g = G.Grid(ions.cell, nr)
    at = OIO.Atoms(ions.symbols, ions.positions, ions.cell)
    gp, vp, z = OIO.read_recpot(systems.empty_core_recpot())
    at.set_charges({"Al": z})
    t0 = time.time()
    vreal = P.local_pp(g, at, {"Al": (gp, vp)}, pme=True)
    eII = P.ewald_energy(g, at, pme=True)
    print("setup %.1f s, E_ion %.12f" % (time.time() - t0, eII), flush=True)
    cache = {}

    def ev(r, need_e):
        r0 = r.mean()
        if "k" not in cache or abs(cache["r0"] - r0) > 1e-6:
            cache["k"], cache["r0"] = F.wt_kernel(g.q, r0), r0
        e, v, _ = F.total(r, g, vreal, cache["k"], ewald=eII)
        return e, v

    rho = np.full(tuple(nr), at.ncharge() / g.volume)
    e0, v0, parts0 = F.total(rho, g, vreal, F.wt_kernel(g.q, rho.mean()), ewald=eII)
    print("uniform start E = %.12f" % e0, flush=True)
    t0 = time.time()
    res = O.optimize(ev, rho, g.dV, method="TN", econv=1e-9 * ions.nat, maxiter=100, maxfun=50, maxls=30,
                     log=lambda *a: print("%3d %.12f %.3e %.3e %d %d" % a, flush=True))
    secs = time.time() - t0
    rho_f = res["rho"]
    e_f, v_f, parts = F.total(rho_f, g, vreal, F.wt_kernel(g.q, rho_f.mean()), ewald=eII)
    d = dict(system="fcc Al 5x5x5, a=7.60 bohr, sigma=0.1 seed=2024, 160^3, WT+LDA+H+emptycore(Z=3,rc=1.12), PME",
             nat=ions.nat, energy=e_f, energy_history=[float(x) for x in res["history"]],
             counts=[list(map(int, c)) for c in res["counts"]], converged=res["converged"], mu=res["mu"],
             parts={k: float(v) for k, v in parts.items()}, uniform_energy=e0,
             uniform_parts={k: float(v) for k, v in parts0.items()},
             uniform_potential_sum=float(v0.sum()), final_potential_sum=float(v_f.sum()),
             final_rho_max=float(rho_f.max()), final_rho_min=float(rho_f.min()), seconds=secs)
    with open(out, "w") as f:
        json.dump(d, f, indent=1)
    print("wrote", out, "E =", e_f, "in %.1f s" % secs)


if __name__ == "__main__":
    main()
