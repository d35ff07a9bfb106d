#!/usr/bin/env python3
"""Headline benchmark: seconds to reach 1e-6 Ha/atom of the FP64 oracle
energy for the 500-atom fcc Al cell (5x5x5, a = 7.60 bohr, sigma = 0.1 bohr
displacements, 160^3 grid) with the WT kinetic functional + PZ-LDA +
Hartree + Ashcroft empty-core local pseudopotential, on one GPU.

One step = one complete density optimisation of the FP32-bulk form (FP32
FFTs and point-wise work, FP64 reductions / line search / normalisation,
exact linear-response Hessian-vector products) from the learned warm start
(superposition of screened-ion densities -> periodic 3D CNN, BF16 MFMA,
evaluated at 160^3 inside the timed region),
stopped by the reference's own rule (|dE| < econv * nat over 2 steps,
econv = 1e-6 Ha, optimization.py:522-535 and interface.py:168), then an
FP64 energy evaluation of the result that must lie within 1e-6 Ha/atom of
the FP64 oracle energy (FP64 TN steps are added if it does not).  Each step
also runs the plain FP64 GPU optimisation (the reference algorithm:
forward-difference Hessian, eps = 1e-9) from the uniform density with the
same stop rule; its time to the stop and its time to first reach the
threshold are both reported.

    python bench.py [--steps K] [--warmup W] [--gpus 1]

Prints one JSON line on stdout; diagnostics go to stderr.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "s to 1e-6 Ha/atom, 500-atom Al WT: FP32+learned start vs plain FP64 GPU"
THRESH = 1e-6  # Ha/atom
ECONV = 1e-6  # the reference's default stop rule, per atom
# HBM bytes per launch of the FP32 column pass on 3 spectra of 160x160x81
# (pitch 96), from the FETCH_SIZE / WRITE_SIZE counter runs of
# `bench.py --passes-only`: 28833 KiB * 2.0 + 50400 KiB * 1.0
TRAFFIC_X_C2C_FP32 = (28833 * 2.0 + 50400 * 1.0) * 1024


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def build_system(seed):
    import torch
    from dftpy_amd import DirectField, DirectGrid, systems
    from dftpy_amd.functional import Functional, TotalFunctional
    ions, nr = systems.c2_system(seed)
    g64 = DirectGrid(ions.cell, nr, full=False, dtype=torch.float64)
    g32 = DirectGrid(ions.cell, nr, full=False, dtype=torch.float32)
    pp = Functional(type="PSEUDO", grid=g64, ions=ions, PP_list={"Al": systems.empty_core_recpot()})
    ev = TotalFunctional(KE=Functional(type="KEDF", name="WT"), XC=Functional(type="XC", name="LDA"),
                         H=Functional(type="HARTREE"), PSEUDO=pp)
    rho64 = DirectField(g64)
    rho64[:] = ions.get_ncharges() / g64.volume
    rho32 = DirectField(g32)
    rho32[:] = ions.get_ncharges() / g32.volume
    return ions, nr, g64, g32, ev, rho64, rho32


def time_to_threshold(opt, wall, e_ref, tol):
    """Seconds from the call to the first iterate within tol of e_ref (the
    Python overhead outside the native loop is added in full)."""
    overhead = wall - opt.stats["seconds"]
    for e, t in zip(opt.history, opt.history_seconds):
        if abs(e - e_ref) <= tol:
            return t + overhead
    return None


def run_fp64_plain(ev, rho64, e_ref, tol, nat, precond=False):
    import torch
    from dftpy_amd.optimization import Optimization
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    opt = Optimization(optimization_method="TN", EnergyEvaluator=ev,
                       optimization_options={"econv": ECONV * nat, "maxiter": 100, "hvp": "fd", "precond": precond})
    opt.optimize_rho(rho64)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    if abs(opt.stats["energy"] - e_ref) > tol:
        log("warning: FP64 run stopped outside the threshold")
    return time_to_threshold(opt, wall, e_ref, tol), wall, opt


def run_fp32_form(ev, start_fn, g64, e_ref, tol, nat, precond=False):
    """FP32-bulk TN from the density start_fn() returns; the start density is
    built inside the timed region.  Returns (seconds, info)."""
    import torch
    from dftpy_amd import DirectField
    from dftpy_amd.optimization import Optimization
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    # the start's duration from events on the stream: no host wait between the
    # start and the optimizer
    ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    ev0.record()
    rho_start32 = start_fn()
    ev1.record()
    opt = Optimization(optimization_method="TN", EnergyEvaluator=ev,
                       optimization_options={"econv": ECONV * nat, "maxiter": 60, "hvp": "analytic",
                                             "precond": precond})
    r32 = opt.optimize_rho(rho_start32)
    r64 = DirectField(g64, data=r32.data.double())
    e64 = ev(r64, calcType={"E"}).energy
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    t_start = ev0.elapsed_time(ev1) * 1e-3
    info = dict(start_s=round(t_start, 5), iters=opt.stats["niter"], status=opt.stats["converged"],
                evals=opt.stats["n_eval_v"] + opt.stats["n_eval_ev"], fp32_energy=float(opt.stats["energy"]),
                fp64_energy_of_result=float(e64), err_per_atom=float((e64 - e_ref) / nat), polish=False)
    if abs(e64 - e_ref) <= tol:
        return t1 - t0, info
    # finish with FP64 TN steps from the FP32 density
    opt2 = Optimization(optimization_method="TN", EnergyEvaluator=ev,
                        optimization_options={"econv": 1e-8 * nat, "maxiter": 40, "hvp": "analytic"})
    t2 = time.perf_counter()
    opt2.optimize_rho(r64)
    torch.cuda.synchronize()
    wall2 = time.perf_counter() - t2
    tt = time_to_threshold(opt2, wall2, e_ref, tol)
    info.update(polish=True, polish_iters=opt2.stats["niter"])
    return (t1 - t0) + (tt if tt is not None else float("nan")), info


def learned_start_fn(ions, pp, g64, g32, net, use_net=True):
    """The learned warm start: superposition of screened-ion densities,
    network inputs, the CNN forward at the full grid and the normalised start
    density (all on the GPU)."""
    import torch
    from dftpy_amd.learned import superposition_density
    ne = float(ions.get_ncharges())
    rho0 = ne / g64.volume
    vloc = pp.vreal_for(g64).data
    plan64 = g64.plan(torch.float64)

    def fn():
        rho_s = superposition_density(g64, ions, pp.readpp, rho0)
        x0 = net.features(plan64, rho_s, vloc, rho0)
        return net.start_density(g32, x0, rho0, ne, use_net=use_net)
    return fn


def pass_timings(grid, rho, iters=50):
    """Mean duration of each pipeline pass (HIP events on torch's stream)."""
    import ctypes
    from dftpy_amd import _lib
    plan = grid.plan(grid.dtype)
    out = {}
    for which, name in ((0, "z_r2c_prologue"), (1, "y_c2c"), (2, "x_c2c"), (3, "z_c2r"), (5, "evaluation_E_V")):
        ms = ctypes.c_double()
        _lib.check(_lib.lib.ofdft_bench_pass(plan.handle, which, 3, iters, rho.data.data_ptr(), ctypes.byref(ms),
                                             _lib.stream()))
        out[name] = ms.value
    return out


def cpu_baseline(nr, seed, n_evals, budget_s=25.0):
    """The FP64 numpy oracle (pocketfft, one thread) timed per E+V evaluation
    at the same 160^3 size, projected onto the evaluation count of the FP64
    GPU run (same algorithm)."""
    from dftpy_amd import systems
    from oracle import functionals as F
    from oracle import grid as G
    from oracle import io as OIO
    from oracle import pseudo as P
    ions, nr = systems.c2_system(seed)
    g = G.Grid(ions.cell, nr)
    at = OIO.Atoms(ions.symbols, ions.positions, ions.cell)
    gp, vp, z = OIO.read_recpot(systems.empty_core_recpot())
    at.set_charges({"Al": z})
    vreal = P.local_pp(g, at, {"Al": (gp, vp)}, pme=True)
    rho = np.full(tuple(nr), at.ncharge() / g.volume)
    kern = F.wt_kernel(g.q, rho.mean())

    def per_eval(workers):
        g.workers = workers
        times = []
        t_start = time.perf_counter()
        while len(times) < 3 and time.perf_counter() - t_start < budget_s:
            t0 = time.perf_counter()
            F.total(rho, g, vreal, kern)
            times.append(time.perf_counter() - t0)
        g.workers = None
        return float(np.median(times)), len(times)

    per, nt = per_eval(None)
    ncores = len(os.sched_getaffinity(0))
    per_all, nt_all = per_eval(ncores)
    cpu_model = ""
    try:
        with open("/proc/cpuinfo") as f:
            cpu_model = next((ln.split(":", 1)[1].strip() for ln in f if ln.startswith("model name")), "")
    except OSError:
        pass
    return dict(value=per * n_evals, unit="s", cores=1, kind="port",
                sample="%d FP64 oracle E+V evaluations at 160^3 (numpy pocketfft, 1 thread), %.3f s each, x %d "
                       "evaluations of the TN run (FD Hessian, as the reference)" % (nt, per, n_evals),
                seconds_per_eval=per, value_all_cores=per_all * n_evals, cores_all=ncores,
                seconds_per_eval_all_cores=per_all,
                sample_all_cores="%d evaluations with scipy.fft on %d workers (FFTs threaded, point-wise numpy "
                                 "on one thread)" % (nt_all, ncores),
                cpu_model=cpu_model, omp_num_threads=os.environ.get("OMP_NUM_THREADS"))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--seed", type=int, default=2024)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--passes-only", action="store_true",
                    help="only launch the evaluation passes (and a reference copy kernel) for counter collection")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: this build runs the single-cell path on one device; "
              "multi-GPU runs are not part of it (requested --gpus %d)" % args.gpus, file=sys.stderr)
        sys.exit(2)
    import torch
    log("device:", torch.cuda.get_device_name(0))
    ions, nr, g64, g32, ev, rho64, rho32 = build_system(args.seed)
    if args.passes_only:
        ev(rho32, calcType={"E", "V"})
        log("fp32 passes (ms):", pass_timings(g32, rho32, iters=args.steps))
        log("fp64 passes (ms):", pass_timings(g64, rho64, iters=args.steps))
        # a plain device copy of 256 MiB, to calibrate the HBM byte counters
        a = torch.empty(64 * 1024 * 1024, dtype=torch.float32, device="cuda")
        b = torch.empty_like(a)
        for _ in range(args.steps):
            b.copy_(a)
        torch.cuda.synchronize()
        return
    nat = ions.nat
    tol = THRESH * nat
    ref_path = os.path.join(ROOT, "tests", "golden", "c2_oracle.json")
    if args.seed == 2024 and os.path.exists(ref_path):
        e_ref = json.load(open(ref_path))["energy"]
        ref_src = "CPU FP64 oracle, tests/golden/c2_oracle.json"
    else:
        log("no oracle reference for this seed: converging the FP64 GPU path tightly instead")
        from dftpy_amd.optimization import Optimization
        o = Optimization("TN", optimization_options={"econv": 1e-10 * nat, "maxiter": 100}, EnergyEvaluator=ev)
        o.optimize_rho(rho64)
        e_ref = o.stats["energy"]
        ref_src = "FP64 GPU, econv 1e-10 Ha/atom"
    t0 = time.perf_counter()
    e_setup = ev(rho64, calcType={"E"}).energy  # PP + Ewald setup, not timed
    log("setup %.3f s, E(uniform) = %.10f, E_ref = %.10f (%s)" % (time.perf_counter() - t0, e_setup, e_ref, ref_src))
    from dftpy_amd.learned import WarmStartNet
    net = WarmStartNet.load()  # trained weights shipped with the package; fails loudly if absent
    pp = ev.PSEUDO
    learned = learned_start_fn(ions, pp, g64, g32, net, use_net=True)
    # FP32 forms: (start, preconditioned inner CG)
    forms = {
        "learned": (learned, True),
        "superposition": (learned_start_fn(ions, pp, g64, g32, net, use_net=False), True),
        "uniform": (lambda: rho32, True),
        "learned_plain_cg": (learned, False),
    }
    for _ in range(args.warmup):
        run_fp64_plain(ev, rho64, e_ref, tol, nat)
        run_fp64_plain(ev, rho64, e_ref, tol, nat, precond=True)
        for fn, pre in forms.values():
            run_fp32_form(ev, fn, g64, e_ref, tol, nat, precond=pre)
    times = {k: [] for k in forms}
    infos = {k: [] for k in forms}
    t64, t64stop, opt64 = [], [], None
    t64p, t64pstop = [], []
    torch.cuda.synchronize()
    for k in range(args.steps):
        tt64, wall64, opt64 = run_fp64_plain(ev, rho64, e_ref, tol, nat)
        t64.append(tt64 if tt64 is not None else float("nan"))
        t64stop.append(wall64)
        tt64p, wall64p, opt64p = run_fp64_plain(ev, rho64, e_ref, tol, nat, precond=True)
        t64p.append(tt64p if tt64p is not None else float("nan"))
        t64pstop.append(wall64p)
        for name, (fn, pre) in forms.items():
            tt, info = run_fp32_form(ev, fn, g64, e_ref, tol, nat, precond=pre)
            times[name].append(tt)
            infos[name].append(info)
        log("step %d: fp32+learned %.4f s (%s) | fp32 superposition %.4f s | fp32 uniform %.4f s | fp32+learned "
            "plain CG %.4f s | fp64 plain %.4f s to its stop rule (%d iters, %d evals), %.4f s to threshold | fp64 "
            "preconditioned %.4f s (%d evals)" %
            (k, times["learned"][-1], infos["learned"][-1], times["superposition"][-1], times["uniform"][-1],
             times["learned_plain_cg"][-1], wall64, opt64.stats["niter"],
             opt64.stats["n_eval_v"] + opt64.stats["n_eval_ev"], t64[-1], wall64p,
             opt64p.stats["n_eval_v"] + opt64p.stats["n_eval_ev"]))
    v32 = float(np.mean(times["learned"]))
    v64 = float(np.mean(t64stop))
    v64_thr = float(np.mean(t64))
    v64p = float(np.mean(t64pstop))
    # roofline of the dominant kernel: the x-axis C2C pass on 3 spectra (FP32)
    pt32 = pass_timings(g32, rho32)
    pt64 = pass_timings(g64, rho64)
    nc = nr[0] * nr[1] * (nr[2] // 2 + 1)
    dom = max(("y_c2c", "x_c2c"), key=lambda k: pt32[k])
    bytes_dom = 6 * nc * 8  # read + write 3 complex64 half spectra
    achieved = bytes_dom / (pt32[dom] * 1e-3) / 1e9
    roof = dict(bound="hbm", achieved=round(achieved, 1), peak=8000.0, unit="GB/s", frac=round(achieved / 8000.0, 4),
                traffic=TRAFFIC_X_C2C_FP32, traffic_source="rocprofv3 FETCH_SIZE x2.0 + WRITE_SIZE x1.0 per dispatch of "
                "k_cols_f<float,160,16,false> (3 spectra), calibrated on a 256 MiB copy", kernel="k_cols<float> %s pass, 3 spectra of 160x160x81 complex64" % dom,
                algorithmic_bytes=bytes_dom, avg_ms=pt32[dom], pass_ms_fp32=pt32, pass_ms_fp64=pt64)
    n_evals_to_thr = None
    if opt64 is not None:
        n_evals_to_thr = opt64.stats["n_eval_v"] + opt64.stats["n_eval_ev"]
    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(nr, args.seed, n_evals_to_thr or 100)
        log("cpu baseline:", cpu)
    line = dict(metric=METRIC, value=round(v32, 5), unit="s", n_gpus=1, steps=args.steps, warmup=args.warmup,
                ms_per_step=round(v32 * 1e3, 3), higher_is_better=False, scaling="strong", vs_baseline=None,
                dtype="fp32 (FP64 reductions/line search)", data="synthetic",
                config=dict(workload="fcc Al 5x5x5 (500 atoms, 38.0 bohr, sigma 0.1 bohr seed %d), 160^3, WT+PZ-LDA+"
                                     "Hartree+empty-core PP (PME), TN to |E-E_ref| <= 1e-6 Ha/atom" % args.seed,
                            form="FP32-bulk TN (analytic Hessian product, inner CG preconditioned by the uniform-"
                                 "gas Hessian) from the learned warm start (CNN inference at 160^3 inside the timed "
                                 "region); plain form: the reference's FP64 TN (FD Hessian, plain inner CG) from the "
                                 "uniform density",
                            e_ref=e_ref, e_ref_source=ref_src,
                            fp64_plain_s=round(v64, 5), speedup_vs_fp64_plain=round(v64 / v32, 3),
                            fp64_plain_first_within_threshold_s=round(v64_thr, 5),
                            speedup_vs_fp64_plain_first_within_threshold=round(v64_thr / v32, 3),
                            fp64_preconditioned_s=round(v64p, 5), speedup_vs_fp64_preconditioned=round(v64p / v32, 3),
                            fp32_superposition_start_s=round(float(np.mean(times["superposition"])), 5),
                            fp32_uniform_start_s=round(float(np.mean(times["uniform"])), 5),
                            fp32_learned_plain_cg_s=round(float(np.mean(times["learned_plain_cg"])), 5),
                            stop_rule="|dE| < 1e-6 Ha/atom * nat over 2 steps (all forms)",
                            runs=infos),
                roofline=roof, cpu_baseline=cpu)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
