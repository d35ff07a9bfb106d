"""Many cells of uneven grid sizes optimised concurrently on one GPU.

    python -m dftpy_amd.batch [--ncell 48] [--workers 4] [--seed 0] [--sequential] [--graph]

Cells: fcc Al n x n x m supercells with n, m drawn from {2, 3, 4} (32-256 atoms),
isotropic strain U(-3%, +3%) and Gaussian displacements of sigma = 0.1 bohr,
grid spacing ~0.24 bohr rounded up to even 2/3/5-smooth sizes (64-144 per
axis); TF + vW + WT + PZ-LDA + Hartree + the empty-core local PP.  Each cell
is optimised by the FP32-bulk TN (uniform-gas preconditioned inner CG) from
the learned start with the reference stop rule (|dE| < 1e-6 Ha/atom * nat
over 2 steps); with --graph the TN iterations run device-driven, one
captured hipGraph launch per superstep (ofdft_optimize_tn_graph).

Execution: W worker threads, each with its own HIP stream, its own plans (one
per grid shape, autotuned at creation and moved to each cell's strained
lattice) and its own network buffers, take cells from a shared queue.  The
native optimizer releases the Python interpreter lock while it runs, so the
cells' kernels from the W streams run concurrently on the GPU.  (The reference
optimises one cell per process; there is no batch driver in it.)
"""
import argparse
import json
import math
import queue
import sys
import threading
import time

import numpy as np
import torch

from . import DirectField, DirectGrid, systems
from .functional import Functional, TotalFunctional
from .functional.pseudo import LocalPseudo, ReadPseudo
from .learned import WarmStartNet, superposition_density
from .math_utils import best_fft_size
from .optimization import Optimization

SPACING = 0.24


def make_cells(ncell=48, seed=0):
    """[(ions, nr)] of the batch (seeded)."""
    rng = np.random.default_rng(seed)
    cells = []
    for k in range(ncell):
        n, m = int(rng.choice([2, 3, 4])), int(rng.choice([2, 3, 4]))
        strain = float(rng.uniform(-0.03, 0.03))
        ions = systems.fcc_al((n, n, m), sigma=0.1, strain=strain, seed=seed * 1000 + k)
        lengths = np.diag(ions.cell)
        nr = tuple(int(best_fft_size(int(math.ceil(L / SPACING)), max_prime=5)) for L in lengths)
        cells.append((ions, nr))
    return cells


class _Worker:
    def __init__(self, readpp, weights=None, check64=False, graph=False):
        self.check64 = check64
        self.graph = graph
        self.stream = torch.cuda.Stream()
        self.grids = {}
        self.readpp = readpp
        self.net = WarmStartNet.load() if weights is None else WarmStartNet.load(weights)

    def grid_pair(self, ions, nr):
        if nr not in self.grids:
            self.grids[nr] = (DirectGrid(ions.cell, nr, full=False, dtype=torch.float64),
                              DirectGrid(ions.cell, nr, full=False, dtype=torch.float32))
        g64, g32 = self.grids[nr]
        g64.set_lattice(ions.cell)
        g32.set_lattice(ions.cell)
        return g64, g32

    def run(self, ions, nr):
        with torch.cuda.stream(self.stream):
            t0 = time.perf_counter()
            g64, g32 = self.grid_pair(ions, nr)
            pp = LocalPseudo(grid=g64, ions=ions, readpp=self.readpp)
            ev = TotalFunctional(KE=Functional(type="KEDF", name="WT"), XC=Functional(type="XC", name="LDA"),
                                 H=Functional(type="HARTREE"), PSEUDO=pp)
            ne = float(ions.get_ncharges())
            rho0 = ne / g64.volume
            rho_s = superposition_density(g32, ions, self.readpp, rho0)
            x0 = self.net.features(g32.plan(torch.float32), rho_s, pp.vreal_for(g32).data, rho0)
            start = self.net.start_density(g32, x0, rho0, ne)
            opt = Optimization("TN", optimization_options={"econv": 1e-6 * ions.nat, "maxiter": 60, "precond": True,
                                                           "graph": self.graph},
                               EnergyEvaluator=ev)
            rho = opt.optimize_rho(start)
            e64 = None
            if self.check64:  # the FP64 energy of the FP32 result
                e64 = float(ev(DirectField(g64, data=rho.data.double()), calcType={"E"}).energy)
            self.stream.synchronize()
            # status 0: the reference stop rule
            return dict(nat=ions.nat, nr=list(nr), energy=float(opt.stats["energy"]), niter=opt.stats["niter"],
                        status=opt.stats["converged"], evals=opt.stats["n_eval_v"] + opt.stats["n_eval_ev"],
                        seconds=time.perf_counter() - t0, energy64=e64)


def fp64_reference(ions, nr, econv_atom=1e-9):
    """The cell's ground-state energy from the FP64 TN (the reference's
    algorithm, FD Hessian) from the uniform density, converged tightly."""
    readpp = ReadPseudo({"Al": systems.empty_core_recpot()})
    g64 = DirectGrid(ions.cell, nr, full=False, dtype=torch.float64)
    ev = TotalFunctional(KE=Functional(type="KEDF", name="WT"), XC=Functional(type="XC", name="LDA"),
                         H=Functional(type="HARTREE"), PSEUDO=LocalPseudo(grid=g64, ions=ions, readpp=readpp))
    rho = DirectField(g64)
    rho[:] = ions.get_ncharges() / g64.volume
    opt = Optimization("TN", optimization_options={"econv": econv_atom * ions.nat, "maxiter": 100},
                       EnergyEvaluator=ev)
    opt.optimize_rho(rho)
    return float(opt.stats["energy"]), int(opt.stats["converged"])


def run_batch(cells, workers=4, weights=None, check64=False, graph=False):
    """Optimise all cells with `workers` concurrent streams; returns the
    per-cell records (in input order) and the wall time.  check64 adds each
    result's FP64 energy (`energy64`, inside the timed region); graph runs the
    device-driven TN."""
    readpp = ReadPseudo({"Al": systems.empty_core_recpot()})
    ws = [_Worker(readpp, weights, check64, graph) for _ in range(workers)]
    # warm every worker on every shape outside the timed region: plans
    # (autotuning, tables) and one full cell, so the device buffers of each
    # shape are in the stream's allocator cache and no allocation (which
    # synchronises the device) happens while the cells run
    shapes = {}
    for ions, nr in cells:
        shapes.setdefault(nr, ions)

    def warm(w):
        for nr, ions in sorted(shapes.items(), key=lambda kv: -int(np.prod(kv[0]))):
            with torch.cuda.stream(w.stream):
                w.grid_pair(ions, nr)[0].plan(torch.float64)
                w.grid_pair(ions, nr)[1].plan(torch.float32)
            w.run(ions, nr)

    warmers = [threading.Thread(target=warm, args=(w,)) for w in ws]
    for t in warmers:
        t.start()
    for t in warmers:
        t.join()
    torch.cuda.synchronize()
    q = queue.Queue()
    for i, c in enumerate(cells):
        q.put((i, c))
    out = [None] * len(cells)
    errors = []

    def loop(w):
        while True:
            try:
                i, (ions, nr) = q.get_nowait()
            except queue.Empty:
                return
            try:
                out[i] = w.run(ions, nr)
            except Exception as e:  # reported after the join
                errors.append((i, repr(e)))

    torch.cuda.synchronize()
    t0 = time.perf_counter()
    threads = [threading.Thread(target=loop, args=(w,)) for w in ws]
    for t in threads:
        t.start()
    for t in threads:
        t.join()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    if errors:
        raise RuntimeError("batch cells failed: %s" % errors[:3])
    return out, wall


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--ncell", type=int, default=48)
    ap.add_argument("--workers", type=int, default=4)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--sequential", action="store_true", help="also time the cells one after another")
    ap.add_argument("--graph", action="store_true", help="device-driven TN (one hipGraph launch per superstep)")
    args = ap.parse_args(argv)
    cells = make_cells(args.ncell, args.seed)
    res, wall = run_batch(cells, args.workers, graph=args.graph)
    line = dict(workload="%d fcc Al cells (n,n,m in {2,3,4}, 32-256 atoms, strain +-3%%, sigma 0.1), grids %s" %
                (len(cells), sorted({tuple(r["nr"]) for r in res})),
                workers=args.workers, graph=args.graph, wall_s=wall, cells_per_s=len(cells) / wall,
                atoms=int(sum(r["nat"] for r in res)), mean_iters=float(np.mean([r["niter"] for r in res])),
                mean_evals=float(np.mean([r["evals"] for r in res])),
                status_counts={str(k): int(sum(r["status"] == k for r in res)) for k in sorted({r["status"] for r in res})},
                all_converged=all(r["status"] == 0 for r in res))
    if args.sequential:
        res1, wall1 = run_batch(cells, 1, graph=args.graph)
        line.update(sequential_wall_s=wall1, concurrency_speedup=wall1 / wall,
                    max_energy_diff_vs_sequential=float(max(abs(a["energy"] - b["energy"]) for a, b in zip(res, res1))))
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
