"""Train the warm-start network on converged densities the method itself
generates, then measure what it gains on held-out cells.

    python -m dftpy_amd.warmstart_train --out DIR [--ntrain 2000] [--ntest 200] [--epochs 8]

Cells: fcc Al 3x3x3 (108 atoms, a = 7.60 bohr) with an isotropic strain drawn
from U(-2%, +2%) and Gaussian displacements of sigma = 0.15 bohr, on a 96^3
grid; functional TF + vW + WT + PZ-LDA + Hartree + the Ashcroft empty-core
pseudopotential.  Training seeds 0..ntrain-1, held-out seeds 10000.. .  Each
cell is converged on the GPU in FP32 from the superposition start (TN with
the exact Hessian product), the target is log(rho* / rho0) - x0[0], and the
network is trained by the project's own forward/backward kernels with Adam.
Held-out cells are then optimised from the superposition start and from the
learned start with the reference stop rule (|dE| < 1e-6 Ha/atom * nat over 2
steps), recording iterations, evaluations, wall time and the energy of each
result against the tightly converged energy.
"""
import argparse
import json
import os
import sys
import time

import numpy as np
import torch

from . import DirectField, DirectGrid, systems
from .functional import Functional, TotalFunctional
from .functional.pseudo import LocalPseudo, ReadPseudo
from .learned import WarmStartNet, metric_weights, superposition_density
from .optimization import Optimization
from . import _lib

NR = (96, 96, 96)
REPS = (3, 3, 3)
SIGMA = 0.15
STRAIN = 0.02


class _NoEwald:
    energy = 0.0  # the ion-ion energy is a constant shift for a fixed cell


def make_ions(seed):
    strain = np.random.default_rng([seed, 7]).uniform(-STRAIN, STRAIN)
    return systems.fcc_al(REPS, sigma=SIGMA, strain=strain, seed=seed)


class CellFactory:
    """Reuses one FP64 and one FP32 grid (plans) across cells of the same
    shape, moving them to each cell's lattice."""

    def __init__(self):
        self.readpp = ReadPseudo({"Al": systems.empty_core_recpot()})
        ions = make_ions(0)
        self.g64 = DirectGrid(ions.cell, NR, full=False, dtype=torch.float64)
        self.g32 = DirectGrid(ions.cell, NR, full=False, dtype=torch.float32)

    def cell(self, seed, net):
        ions = make_ions(seed)
        self.g64.set_lattice(ions.cell)
        self.g32.set_lattice(ions.cell)
        pp = LocalPseudo(grid=self.g64, ions=ions, readpp=self.readpp)
        ev = TotalFunctional(ewald=_NoEwald(), KE=Functional(type="KEDF", name="WT"),
                             XC=Functional(type="XC", name="LDA"), H=Functional(type="HARTREE"), PSEUDO=pp)
        ne = float(ions.get_ncharges())
        rho0 = ne / self.g64.volume
        rho_s = superposition_density(self.g64, ions, self.readpp, rho0)
        vloc = pp.vreal_for(self.g64).data
        x0 = net.features(self.g64.plan(torch.float64), rho_s, vloc, rho0)
        return ions, ev, ne, rho0, x0


def converge(ev, rho_start, econv_atom, nat, maxiter=80):
    opt = Optimization(optimization_method="TN", EnergyEvaluator=ev,
                       optimization_options={"econv": econv_atom * nat, "maxiter": maxiter})
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    rho = opt.optimize_rho(rho_start)
    torch.cuda.synchronize()
    return rho, opt, time.perf_counter() - t0


def energy64(ev, g64, rho):
    return ev(DirectField(g64, data=rho.data.double()), calcType={"E"}).energy


def generate(fac, net, seeds, log):
    """x0 (n, 2, N) and targets (n, N) on the GPU, plus E* per cell."""
    N = int(np.prod(NR))
    X = torch.empty((len(seeds), 2, N), dtype=torch.float32, device="cuda")
    Y = torch.empty((len(seeds), N), dtype=torch.float32, device="cuda")
    estar = np.zeros(len(seeds))
    t0 = time.perf_counter()
    for i, seed in enumerate(seeds):
        ions, ev, ne, rho0, x0 = fac.cell(seed, net)
        start = net.start_density(fac.g32, x0, rho0, ne, use_net=False)
        rho, opt, _ = converge(ev, start, 1e-9, ions.nat, maxiter=100)
        X[i].copy_(x0)
        _lib.check(_lib.lib.ofdft_cnn_target(fac.g32.plan(torch.float32).handle, rho.data.data_ptr(), x0.data_ptr(),
                                             rho0, Y[i].data_ptr(), _lib.stream()))
        estar[i] = energy64(ev, fac.g64, rho)
        if (i + 1) % 100 == 0 or i == 0:
            log("generated %d/%d cells (%.1f s), last: %d iters, status %d, E* = %.8f" %
                (i + 1, len(seeds), time.perf_counter() - t0, opt.stats["niter"], opt.stats["converged"], estar[i]))
    return X, Y, estar


def train(net, plan, X, Y, epochs, lr_max, log, Xv=None, Yv=None, w=None):
    n = X.shape[0]
    steps = epochs * n
    warm = min(200, steps // 10)
    rng = np.random.default_rng(0)
    hist = []
    t0 = time.perf_counter()
    step = 0
    for ep in range(epochs):
        order = rng.permutation(n)
        acc = []
        for j in order:
            if step < warm:
                lr = lr_max * (step + 1) / warm
            else:
                lr = 0.5 * lr_max * (1 + np.cos(np.pi * (step - warm) / max(1, steps - warm))) + 1e-6
            acc.append(net.train_step(plan, X[j], NR, Y[j], lr, w=w).clone())
            step += 1
        tl = float(torch.stack(acc).mean().item())
        vl = validate(net, plan, Xv, Yv, w) if Xv is not None else float("nan")
        hist.append(dict(epoch=ep + 1, train_loss=tl, val_loss=vl, seconds=time.perf_counter() - t0))
        log("epoch %d: train loss %.4e, held-out loss %.4e (%.1f s)" % (ep + 1, tl, vl, time.perf_counter() - t0))
    return hist


def validate(net, plan, Xv, Yv, w=None):
    tot = 0.0
    loss = torch.zeros(1, dtype=torch.float64, device="cuda")
    dout = torch.empty(Yv.shape[1], dtype=torch.float32, device="cuda")
    for j in range(Xv.shape[0]):
        out = net.forward(Xv[j], NR)
        net.loss(plan, out, Yv[j], dout, loss, w)
        tot += float(loss.item())
    return tot / Xv.shape[0]


def evaluate_gain(fac, net, seeds, estar, log):
    """Optimise each held-out cell from the superposition and the learned
    start with the reference stop rule; returns per-cell records."""
    rec = []
    for i, seed in enumerate(seeds):
        ions, ev, ne, rho0, x0 = fac.cell(seed, net)
        nat = ions.nat
        r = dict(seed=int(seed), e_star=float(estar[i]))
        for name, use in (("superposition", False), ("learned", True)):
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            start = net.start_density(fac.g32, x0, rho0, ne, use_net=use)
            torch.cuda.synchronize()
            t_start = time.perf_counter() - t0
            e_start = energy64(ev, fac.g64, start)
            rho, opt, t_opt = converge(ev, start, 1e-6, nat, maxiter=60)
            e_end = energy64(ev, fac.g64, rho)
            r[name] = dict(start_err_per_atom=(e_start - estar[i]) / nat, final_err_per_atom=(e_end - estar[i]) / nat,
                           niter=opt.stats["niter"], n_eval=opt.stats["n_eval_v"] + opt.stats["n_eval_ev"],
                           status=opt.stats["converged"], seconds=t_start + t_opt, start_seconds=t_start)
        rec.append(r)
        if (i + 1) % 20 == 0 or i == 0:
            log("held-out %d/%d: superposition %d iters %.1f ms, learned %d iters %.1f ms" %
                (i + 1, len(seeds), r["superposition"]["niter"], 1e3 * r["superposition"]["seconds"],
                 r["learned"]["niter"], 1e3 * r["learned"]["seconds"]))
    return rec


def evaluate_starts(fac, net, seeds, estar, log, starts=("uniform", "superposition", "learned")):
    """The benchmark's form on held-out cells: FP32 TN with the preconditioned
    inner CG and the reference stop rule from each start; per-cell records."""
    from . import DirectField
    rec = []
    for i, seed in enumerate(seeds):
        ions, ev, ne, rho0, x0 = fac.cell(seed, net)
        nat = ions.nat
        r = dict(seed=int(seed), e_star=float(estar[i]))
        for name in starts:
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            if name == "uniform":
                start = DirectField(fac.g32)
                start[:] = ne / fac.g32.volume
            else:
                start = net.start_density(fac.g32, x0, rho0, ne, use_net=name == "learned")
            torch.cuda.synchronize()
            t_start = time.perf_counter() - t0
            e_start = energy64(ev, fac.g64, start)
            opt = Optimization(optimization_method="TN", EnergyEvaluator=ev,
                               optimization_options={"econv": 1e-6 * nat, "maxiter": 60, "precond": True})
            t1 = time.perf_counter()
            rho = opt.optimize_rho(start)
            torch.cuda.synchronize()
            t_opt = time.perf_counter() - t1
            e_end = energy64(ev, fac.g64, rho)
            r[name] = dict(start_err_per_atom=(e_start - estar[i]) / nat, final_err_per_atom=(e_end - estar[i]) / nat,
                           niter=opt.stats["niter"], n_eval=opt.stats["n_eval_v"] + opt.stats["n_eval_ev"],
                           status=opt.stats["converged"], seconds=t_start + t_opt, start_seconds=t_start)
        rec.append(r)
        if (i + 1) % 20 == 0 or i == 0:
            log("held-out %d/%d: " % (i + 1, len(seeds)) +
                ", ".join("%s %d iters %.1f ms" % (n, r[n]["niter"], 1e3 * r[n]["seconds"]) for n in starts))
    return rec


def summarize(rec, names=("superposition", "learned")):
    out = {}
    for name in names:
        a = lambda k: np.array([r[name][k] for r in rec], dtype=float)  # noqa: E731
        out[name] = dict(niter_mean=float(a("niter").mean()), n_eval_mean=float(a("n_eval").mean()),
                         seconds_mean=float(a("seconds").mean()),
                         start_err_per_atom_median=float(np.median(a("start_err_per_atom"))),
                         final_err_per_atom_max=float(np.abs(a("final_err_per_atom")).max()),
                         within_1e6_fraction=float(np.mean(np.abs(a("final_err_per_atom")) <= 1e-6)),
                         status0_fraction=float(np.mean(a("status") == 0)) if "status" in rec[0][name] else None)
    for base in names:
        if base != "learned" and "learned" in names:
            out["time_ratio" if base == "superposition" else "time_ratio_vs_" + base] = (
                out[base]["seconds_mean"] / out["learned"]["seconds_mean"])
            out["eval_ratio" if base == "superposition" else "eval_ratio_vs_" + base] = (
                out[base]["n_eval_mean"] / out["learned"]["n_eval_mean"])
    return out


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--out", required=True)
    ap.add_argument("--ntrain", type=int, default=2000)
    ap.add_argument("--ntest", type=int, default=200)
    ap.add_argument("--epochs", type=int, default=8)
    ap.add_argument("--lr", type=float, default=5e-4)
    ap.add_argument("--dil", default="1,2,4,4,2,1")
    ap.add_argument("--c", type=float, default=0.7)
    ap.add_argument("--loss", choices=("metric", "mse"), default="metric")
    ap.add_argument("--eval-only", action="store_true",
                    help="no training: evaluate the shipped network on the held-out cells with the preconditioned "
                         "optimizer from the uniform, superposition and learned starts, and add the result to "
                         "the shipped evaluation record (written to --out)")
    args = ap.parse_args(argv)
    os.makedirs(args.out, exist_ok=True)

    def log(msg):
        print(msg, flush=True)

    if args.eval_only:
        net = WarmStartNet.load()
        fac = CellFactory()
        te_seeds = list(range(10000, 10000 + args.ntest))
        _, _, estar = generate(fac, net, te_seeds, log)
        rec = evaluate_starts(fac, net, te_seeds, estar, log)
        summ = summarize(rec, ("uniform", "superposition", "learned"))
        log(json.dumps(summ))
        src = os.path.join(os.path.dirname(os.path.abspath(__file__)), "data", "warmstart_eval.json")
        with open(src) as f:
            shipped = json.load(f)
        shipped["preconditioned"] = dict(summary=summ, cells=rec,
                                         form="FP32 TN, preconditioned inner CG, reference stop rule "
                                              "(the benchmark's form)")
        with open(os.path.join(args.out, "warmstart_eval.json"), "w") as f:
            json.dump(shipped, f, indent=1)
        return 0

    net = WarmStartNet(dil=tuple(int(d) for d in args.dil.split(",")), c=args.c, sv=1.0, seed=0)
    fac = CellFactory()
    tr_seeds = list(range(args.ntrain))
    te_seeds = list(range(10000, 10000 + args.ntest))
    X, Y, _ = generate(fac, net, tr_seeds, log)
    Xv, Yv, estar = generate(fac, net, te_seeds, log)
    plan = fac.g32.plan(torch.float32)
    ions = systems.fcc_al(REPS)
    ions.set_charges(fac.readpp.zval)
    fac.g32.set_lattice(ions.cell)
    w = metric_weights(fac.g32, float(ions.get_ncharges()) / fac.g32.volume) if args.loss == "metric" else None
    hist = train(net, plan, X, Y, args.epochs, args.lr, log, Xv, Yv, w)
    path = os.path.join(args.out, "warmstart_al_wt.npz")
    net.save(path, train_loss=hist[-1]["train_loss"], val_loss=hist[-1]["val_loss"], ntrain=args.ntrain,
             epochs=args.epochs)
    log("saved %s" % path)
    del X, Y
    rec = evaluate_gain(fac, net, te_seeds, estar, log)
    summ = summarize(rec)
    log(json.dumps(summ))
    with open(os.path.join(args.out, "warmstart_eval.json"), "w") as f:
        json.dump(dict(summary=summ, history=hist, cells=rec, config=vars(args)), f, indent=1)
    return 0


if __name__ == "__main__":
    sys.exit(main())
