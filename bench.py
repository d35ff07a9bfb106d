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
# HBM bytes per launch of the CNN middle layer at 160^3, from the FETCH_SIZE /
# WRITE_SIZE counter runs of `bench.py --cnn-only --steps 5` (24 launches):
# FETCH_SIZE 267,993 KiB x 2 (a plain device copy reads half its bytes on this
# counter) + WRITE_SIZE 256,000 KiB x 1
TRAFFIC_CNN_MID = int((2 * 267993 + 256000) * 1024)
TRAFFIC_CNN_MID_SOURCE = ("rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE of bench.py --cnn-only (mean of 24 launches of "
                          "k_cnn_mid_ring), FETCH scaled by 2 and WRITE by 1 from a 256 MiB copy")


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


def run_fp32_form(ev, start_fn, g64, e_ref, tol, nat, precond=False, options=None):
    """FP32-bulk TN from the density start_fn() returns; the start density is
    built inside the timed region (options: more optimizer options, e.g. the
    device-driven iteration).  Returns (seconds, info)."""
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
                       optimization_options=dict({"econv": ECONV * nat, "maxiter": 60, "hvp": "analytic",
                                                  "precond": precond}, **(options or {})))
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
    density (all on the GPU, in FP32)."""
    import torch
    from dftpy_amd.learned import superposition_density
    ne = float(ions.get_ncharges())
    rho0 = ne / g64.volume
    vloc = pp.vreal_for(g32).data
    plan32 = g32.plan(torch.float32)

    def fn():
        rho_s = superposition_density(g32, ions, pp.readpp, rho0)
        x0 = net.features(plan32, rho_s, vloc, rho0)
        return net.start_density(g32, x0, rho0, ne, use_net=use_net)
    return fn


PASSES = ((0, "z_r2c_prologue"), (1, "y_c2c"), (6, "x_fused_gspace"), (3, "z_c2r"), (7, "post_pointwise"),
          (5, "evaluation_E_V"))


def pass_timings(grid, rho, iters=50):
    """Mean duration (ms) of each pass of the E+V pipeline with 3 spectra (vW,
    WT, Hartree), timed with HIP events on torch's stream, the stream the
    passes run on."""
    import ctypes
    from dftpy_amd import _lib
    plan = grid.plan(grid.dtype)
    out = {}
    for which, name in PASSES:
        ms = ctypes.c_double()
        _lib.check(_lib.lib.ofdft_bench_pass(plan.handle, which, 3, iters, rho.data.data_ptr(), ctypes.byref(ms),
                                             _lib.stream()))
        out[name] = ms.value
    return out


def pass_roofline(nr, ms, rs):
    """Algorithmic bytes, GB/s and fraction of the 8 TB/s HBM peak of each
    pipeline pass (3 spectra; rs = bytes per real).  Bytes per pass: each
    array of the pass's contract read once and written once."""
    n = nr[0] * nr[1] * nr[2]
    nc = nr[0] * nr[1] * (nr[2] // 2 + 1)
    cs = 2 * rs
    alg = {"z_r2c_prologue": n * rs + 3 * nc * cs,  # rho in, 3 half spectra out
           "y_c2c": 2 * 3 * nc * cs,
           "x_fused_gspace": 2 * 3 * nc * cs,  # forward, multiply and inverse in place
           "z_c2r": 3 * nc * cs + 3 * n * rs,
           "post_pointwise": 6 * n * rs,  # rho, 3 inverse outputs, vloc in; V out
           "evaluation_E_V": 29 * n * rs}  # compulsory traffic of one evaluation
    out = {}
    for k, b in alg.items():
        gbs = b / (ms[k] * 1e-3) / 1e9
        out[k] = dict(avg_ms=round(ms[k], 5), algorithmic_bytes=int(b), achieved_gbs=round(gbs, 1),
                      frac=round(gbs / 8000.0, 4))
    return out


def cnn_layer_timings(net, nr, iters=20):
    """Mean duration (ms) of each layer kernel of the warm-start network's
    inference at grid nr, on random inputs of the network's shape, timed with
    events on torch's current stream (the stream the layers run on)."""
    import torch
    from dftpy_amd import _lib
    n = int(np.prod(nr))
    nr_c = np.ascontiguousarray(nr, dtype=np.int32)
    nrp = _lib.iptr(nr_c)
    b = net._buffers(nr, False)
    h = b["h"]
    x0 = torch.randn(2, n, device="cuda", dtype=torch.float32)
    h[0].copy_(torch.randn(n, 32, device="cuda").to(torch.bfloat16))
    L = _lib.lib
    s = _lib.stream()
    calls = {"conv_in": lambda: L.ofdft_cnn_conv_in(nrp, net.dil[0], x0.data_ptr(), net.param("W1").data_ptr(),
                                                     net.param("b1").data_ptr(), h[1].data_ptr(), s)}
    for i, l in enumerate(range(2, 6)):
        calls["mid%d_d%d" % (l, net.dil[l - 1])] = (
            lambda i=i, l=l: L.ofdft_cnn_conv_mid(nrp, net.dil[l - 1], 0, h[0].data_ptr(), net.wf[i].data_ptr(),
                                                  net.param("b%d" % l).data_ptr(), None, None, h[1].data_ptr(), None,
                                                  s))
    calls["conv_out"] = lambda: L.ofdft_cnn_conv_out(nrp, net.dil[5], h[0].data_ptr(), net.param("W6").data_ptr(),
                                                      net.param("b6").data_ptr(), b["out"].data_ptr(), None, s)
    out = {}
    for name, fn in calls.items():
        _lib.check(fn())
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(iters):
            fn()
        e1.record()
        e1.synchronize()
        out[name] = e0.elapsed_time(e1) / iters
    return out


def cpu_baseline(nr, seed, gpu_evals, n_iter_sample=2):
    """The FP64 numpy oracle (pocketfft, one thread), the reference's
    algorithm (TN with the forward-difference Hessian): the first
    n_iter_sample TN iterations from the uniform density at the same 160^3
    size are timed (a bounded sample, ~10-20 s), and projected onto the
    evaluation count of the FP64 GPU run of the same algorithm to the same
    stop.  The full oracle TN to convergence is timed on the 4-atom 32^3
    cell beside it."""
    from dftpy_amd import systems
    from oracle import functionals as F
    from oracle import grid as G
    from oracle import io as OIO
    from oracle import optimize as O
    from oracle import pseudo as P

    def setup(ions, nr):
        g = G.Grid(ions.cell, nr)
        at = OIO.Atoms(ions.symbols, ions.positions, ions.cell)
        gp, vp, z = OIO.read_recpot(systems.empty_core_recpot())
        at.set_charges({"Al": z})
        vreal = P.local_pp(g, at, {"Al": (gp, vp)}, pme=True)
        eII = P.ewald_energy(g, at, pme=True)
        count = [0]
        cache = {}

        def ev(r, need_e):
            count[0] += 1
            r0 = r.mean()
            if "k" not in cache or abs(cache["r0"] - r0) > 1e-6:
                cache["k"], cache["r0"] = F.wt_kernel(g.q, r0), r0
            e, v, _ = F.total(r, g, vreal, cache["k"], ewald=eII)
            return e, v
        return g, at, ev, count

    class Stop(Exception):
        pass

    ions, nr = systems.c2_system(seed)
    g, at, ev, count = setup(ions, nr)
    rho = np.full(tuple(nr), at.ncharge() / g.volume)

    def stop_after(it, e):
        if it >= n_iter_sample:
            raise Stop()
    t0 = time.perf_counter()
    try:
        O.optimize(ev, rho, g.dV, method="TN", econv=ECONV * ions.nat, maxiter=100, callback=stop_after)
    except Stop:
        pass
    t_sample = time.perf_counter() - t0
    n_sample = count[0]
    per = t_sample / n_sample
    # the whole reference algorithm on the 4-atom cell, to its stop rule
    ions1, nr1 = systems.c1_system()
    g1, at1, ev1, count1 = setup(ions1, nr1)
    t0 = time.perf_counter()
    res1 = O.optimize(ev1, np.full(tuple(nr1), at1.ncharge() / g1.volume), g1.dV, method="TN", econv=1e-8,
                      maxiter=100)
    t_c1 = time.perf_counter() - t0
    cpu_model = ""
    try:
        with open("/proc/cpuinfo") as f:
            cpu_model = next((ln.split(":", 1)[1].strip() for ln in f if ln.startswith("model name")), "")
    except OSError:
        pass
    return dict(value=round(per * gpu_evals, 3), unit="s", cores=1, kind="port",
                sample="first %d TN iterations of the FP64 oracle (numpy pocketfft, one thread) from the uniform "
                       "density at 160^3, 500 atoms: %d evaluations in %.2f s (%.3f s each), projected onto the %d "
                       "evaluations the same algorithm (FD Hessian) takes on the GPU to the same stop" %
                       (n_iter_sample, n_sample, t_sample, per, gpu_evals),
                timed_sample_s=round(t_sample, 3), sample_evaluations=n_sample, seconds_per_eval=round(per, 4),
                c1_full_tn_s=round(t_c1, 3), c1_full_tn_evaluations=count1[0], c1_energy=res1["energy"],
                c1_workload="4-atom fcc Al, 32^3, TN from the uniform density to |dE| < 1e-8 Ha over 2 steps",
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
    ap.add_argument("--cnn-only", action="store_true",
                    help="only launch the warm-start network's layers at 160^3 (and a reference copy kernel) for "
                         "counter collection")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: this build runs the single-cell path on one device; "
              "multi-GPU runs are not part of it (requested --gpus %d)" % args.gpus, file=sys.stderr)
        sys.exit(2)
    import torch
    log("device:", torch.cuda.get_device_name(0))
    ions, nr, g64, g32, ev, rho64, rho32 = build_system(args.seed)
    if args.passes_only or args.cnn_only:
        if args.passes_only:
            ev(rho32, calcType={"E", "V"})
            log("fp32 passes (ms):", pass_timings(g32, rho32, iters=args.steps))
            log("fp64 passes (ms):", pass_timings(g64, rho64, iters=args.steps))
        else:
            from dftpy_amd.learned import WarmStartNet
            log("cnn layers (ms):", cnn_layer_timings(WarmStartNet.load(), tuple(nr), iters=args.steps))
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
    # roofline of the kernel with the most time per headline step: the CNN
    # middle layer (4 launches per step inside the learned start)
    pt32 = pass_timings(g32, rho32)
    pt64 = pass_timings(g64, rho64)
    cnn_ms = cnn_layer_timings(net, tuple(nr))
    mid_ms = float(np.mean([v for k, v in cnn_ms.items() if k.startswith("mid")]))
    npts = nr[0] * nr[1] * nr[2]
    flop_mid = 2 * 27 * 32 * 32 * npts  # 27 taps x 32 x 32 channels, multiply + add
    achieved = flop_mid / (mid_ms * 1e-3) / 1e12
    roof = dict(bound="mfma", achieved=round(achieved, 1), peak=2500.0, unit="TFLOP/s",
                frac=round(achieved / 2500.0, 4), traffic=TRAFFIC_CNN_MID,
                traffic_source=TRAFFIC_CNN_MID_SOURCE,
                kernel="k_cnn_mid_ring: warm-start CNN middle layer (32 -> 32 channels, 3^3 dilated taps, "
                       "BF16 MFMA with FP32 accumulation) at 160^3",
                algorithmic_flop=flop_mid, algorithmic_bytes=2 * 64 * npts, avg_ms=round(mid_ms, 5),
                per_step_ms=round(4 * mid_ms, 4), cnn_layer_ms={k: round(v, 5) for k, v in cnn_ms.items()},
                passes_fp32=pass_roofline(nr, pt32, 4), passes_fp64=pass_roofline(nr, pt64, 8))
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
                            speedup_learned_vs_uniform_start=round(float(np.mean(times["uniform"])) / v32, 3),
                            speedup_learned_vs_superposition_start=round(
                                float(np.mean(times["superposition"])) / v32, 3),
                            stop_rule="|dE| < 1e-6 Ha/atom * nat over 2 steps (all forms)",
                            runs=infos),
                roofline=roof, cpu_baseline=cpu)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
