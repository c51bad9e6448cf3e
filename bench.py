#!/usr/bin/env python3
"""Benchmark: NLL + gradient evaluations per second of the J/psi -> pi+ pi- pi0 fit.

Workload (one GPU): 3 decay chains (rho0 / rho- / rho+ combinatorics) each with
rho(770), rho(1450), rho(1700), plus f0(980) and f2(1270) in pi+pi-; 10 amplitudes,
10 free resonance parameters; 500k data, 2.5M phase-space and 50k background events
(weight 0.8), synthetic (GENBOD phase space, data by accept-reject on the model).

A step is one NLL+gradient evaluation at a new parameter point (couplings and
resonance masses/widths changed) through ``analysis.getNLL`` + ``torch.autograd.grad``,
events resident in HBM.  Both forms run on the same events: the mixed form (FP32
amplitude tensors, FP64 arithmetic; ``value``) and the FP64 form (``fp64``), with the
NLL difference at the same point.  Prints one JSON line.

    python bench.py [--steps K] [--warmup W] [--gpus 1] [--precision both|mixed|fp64]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import tempfile
import time

import numpy as np

METRIC = "NLL+grad evals/s @500k evts, mixed vs FP64 (|ΔNLL|≤1e-4); toy fits/hr + p99"
HBM_PEAK_GBS = 8000.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def make_samples(workdir, name, n_data, n_phsp, n_bkg):
    from ctpwa_amd import workloads as W
    t0 = time.time()
    cfg = W.write_workload(name, workdir, W.device_intensity, n_data=n_data, n_phsp=n_phsp, n_bkg=n_bkg)
    log("[bench] samples generated in %.1f s" % (time.time() - t0))
    return cfg


def cpu_model() -> str:
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(cfg_path, frac, threads, hessian=True):
    """FP64 oracle (the reference's likelihood restated in C++/OpenMP) on the host cores:
    NLL+gradient evaluations per second, the median of 5 evaluations at new parameter
    points after 2 warm-ups, on the first `frac` of every sample (1.0: the whole workload),
    scaled linearly in events when frac < 1; then one Hessian, timed alone."""
    from ctpwa_amd import workloads as W
    from ctpwa_amd.config import ConfigParser
    from ctpwa_amd.model import build_plan
    from oracle.oracle import Oracle
    d = os.path.dirname(cfg_path)
    plan = build_plan(ConfigParser(cfg_path))
    ph = np.load(os.path.join(d, "phsp.npy"))
    da = np.load(os.path.join(d, "data.npy"))
    bk = np.load(os.path.join(d, "bkg.npy")) if os.path.exists(os.path.join(d, "bkg.npy")) else None
    full = ph.shape[0] + da.shape[0] + (0 if bk is None else bk.shape[0])
    cut = lambda a: None if a is None else a[: max(1, int(a.shape[0] * frac))]
    ph, da, bk = cut(ph), cut(da), cut(bk)
    n = ph.shape[0] + da.shape[0] + (0 if bk is None else bk.shape[0])
    O = Oracle(plan.to_text(), 3, ph, da, bk, w_bkg=None if bk is None else 0.8, threads=threads)
    p0 = W.nominal_params(plan, 1)
    rng = np.random.default_rng(7)
    pts = [p0 + 1e-3 * rng.standard_normal(p0.shape) for _ in range(7)]
    ts = []
    for k, p in enumerate(pts):
        t = time.perf_counter()
        O.nll_grad(p)
        if k >= 2:
            ts.append(time.perf_counter() - t)
        log("[bench] cpu baseline evaluation %d: %.2f s" % (k, time.perf_counter() - t))
    t_sample = float(np.median(ts))
    t_full = t_sample * full / n
    try:
        nproc = len(os.sched_getaffinity(0))
    except AttributeError:
        nproc = os.cpu_count() or 1
    out = {"value": 1.0 / t_full, "unit": "evals/s", "cores": threads, "kind": "port",
           "nproc": nproc, "cpu_model": cpu_model(), "omp_threads": threads,
           "ms_per_eval": round(t_full * 1e3, 1),
           "sample": "FP64 oracle (C++/OpenMP, %d threads) NLL+gradient on %s (%d of %d events), median of 5 "
                     "evaluations at new points after 2 warm-ups = %.3f s%s"
                     % (threads, "the whole workload" if frac >= 1.0 else "the first %.0f%% of every sample" % (100 * frac),
                        n, full, t_sample, "" if frac >= 1.0 else ", scaled linearly in events to the full workload")}
    if hessian:
        log("[bench] cpu baseline Hessian")
        t = time.perf_counter()
        O.hessian(pts[-1])
        th = (time.perf_counter() - t) * full / n
        out["hessian_ms"] = round(th * 1e3, 1)
        log("[bench] cpu baseline Hessian: %.2f s" % (time.perf_counter() - t))
    return out


def run_form(ana, args, spec, torch, W, prec, sizes, label=None, hessian=True):
    """Time K NLL+gradient evaluations at perturbed points in one precision form."""
    label = label or prec
    plan = ana.plan
    p0 = W.nominal_params(plan, None if args.workload == "jpsi_rho" else spec["seed"])
    rng = np.random.default_rng(123)
    nsteps = args.warmup + args.steps
    pts = []
    for k in range(nsteps):
        p = p0.copy()
        p[: 2 * plan.n_free] += 0.01 * rng.standard_normal(2 * plan.n_free)
        p[2 * plan.n_free:] *= 1.0 + 0.002 * rng.standard_normal(plan.n_theta)
        pts.append(torch.tensor(p, dtype=torch.float64, device="cuda", requires_grad=True))

    def step(p):
        nll = ana.getNLL(p)
        (g,) = torch.autograd.grad(nll, p)
        return nll, g

    for k in range(args.warmup):
        step(pts[k])
    torch.cuda.synchronize()
    t = time.perf_counter()
    last = None
    for k in range(args.warmup, nsteps):
        last = step(pts[k])
    torch.cuda.synchronize()
    dt = time.perf_counter() - t
    ms_per_step = dt / args.steps * 1e3
    value = args.steps / dt
    nll_last = float(last[0].item())
    log("[bench] %s: %d steps in %.4f s: %.4f ms/step, NLL %.6f" % (label, args.steps, dt, ms_per_step, nll_last))
    # the rate a fit loop sees: the same steps with the loss read on the host every step
    torch.cuda.synchronize()
    t = time.perf_counter()
    for k in range(args.warmup, nsteps):
        nll, g = step(pts[k])
        float(nll.item())
    dt_sync = time.perf_counter() - t
    sync = {"value": round(args.steps / dt_sync, 3), "ms_per_step": round(dt_sync / args.steps * 1e3, 4),
            "def": "the same steps with the loss copied to the host (nll.item()) after every step"}
    log("[bench] %s synchronous: %.4f ms/step" % (label, sync["ms_per_step"]))
    # the same points through the C ABI alone (ctpwa_nll_grad into preallocated outputs):
    # the evaluation without autograd's two kernels (gradient seed, gradient x grad_out)
    h = ana._h
    outs = [torch.empty(1 + ana.getNParams(), dtype=torch.float64, device="cuda") for _ in range(4)]
    stream = torch.cuda.current_stream().cuda_stream
    raw = [p.detach() for p in pts]
    for k in range(args.warmup):
        h.nll_grad(raw[k].data_ptr(), outs[k % 4].data_ptr(), stream)
    torch.cuda.synchronize()
    t = time.perf_counter()
    for k in range(args.warmup, nsteps):
        h.nll_grad(raw[k].data_ptr(), outs[k % 4].data_ptr(), stream)
    torch.cuda.synchronize()
    dt_c = time.perf_counter() - t
    c_abi = {"value": round(args.steps / dt_c, 3), "ms_per_step": round(dt_c / args.steps * 1e3, 4),
             "def": "the same points through ctpwa_nll_grad (the C ABI) into preallocated outputs, no autograd"}
    log("[bench] %s through the C ABI: %.4f ms/step" % (label, c_abi["ms_per_step"]))

    # dominant kernel: the fused per-event kernel, timed with HIP events on the stream it runs on
    lay = ana.layout()
    kern_ms = ana.kernelTimeMs(pts[0].detach(), iters=max(10, args.steps))
    achieved = lay["bytes_per_eval"] / (kern_ms * 1e-3) / 1e9
    nev = sum(lay["n_events"])
    csl = 8 if prec == "mixed" else 16
    survey_bytes = nev * (csl * plan.npol * plan.n_amp + 8 * 3)
    # HBM traffic per launch of the fused kernel from the committed FETCH_SIZE/WRITE_SIZE
    # profile of this workload and form (rocprofv3 --pmc runs, calibrated on a copy of
    # known size)
    traffic = None
    tfile = os.path.join(os.path.dirname(os.path.abspath(__file__)), "profiles",
                         "traffic_ctpwa_eval_%s_%s.txt" % (args.workload, prec))
    if label == prec and os.path.exists(tfile) and sizes == (spec["n_data"], spec["n_phsp"], spec["n_bkg"]):
        with open(tfile) as f:
            for line in f:
                if line.startswith("bytes_per_launch"):
                    traffic = int(line.split("=")[1])
    pols = "%d stored polarisation%s%s" % (lay["n_pol_stored"], "" if lay["n_pol_stored"] == 1 else "s",
                                           " (the other is its mirror image)" if lay["n_pol_stored"] < plan.npol else "")
    scope = "" if label == prec else "; data and background events only (phase space from the stored Gram)"
    if prec == "mixed":
        bdef = ("per event: %d SL tensors x %s x 8 B (FP32 complex) + squared masses x 8 B "
                "+ other invariants x 4 B%s" % (lay["n_sl_tensors"], pols, scope))
    else:
        bdef = ("per event: %d SL tensors x %s x 16 B + %d invariants x 8 B%s"
                % (lay["n_sl_tensors"], pols, lay["n_invariants"], scope))
    roof = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
            "kernel": "ctpwa_eval", "kernel_ms": round(kern_ms, 5),
            "algorithmic_bytes": lay["bytes_per_eval"], "algorithmic_bytes_def": bdef,
            "per_amplitude_bytes_formula": survey_bytes}
    log("[bench] %s fused kernel %.4f ms, %.1f GB/s of %d B" % (label, kern_ms, achieved, lay["bytes_per_eval"]))
    hess_ms = None
    if hessian and args.hessian_reps > 0:
        # getHessian on the same workload (not part of `value`): one warm-up, then the mean
        ph = pts[0].detach()
        ana.getHessian(ph)
        torch.cuda.synchronize()
        t = time.perf_counter()
        for _ in range(args.hessian_reps):
            ana.getHessian(ph)
        torch.cuda.synchronize()
        hess_ms = (time.perf_counter() - t) / args.hessian_reps * 1e3
        log("[bench] %s getHessian %.3f ms" % (prec, hess_ms))
    return {"value": value, "ms_per_step": ms_per_step, "nll": nll_last, "roofline": roof, "hessian_ms": hess_ms,
            "sync": sync, "c_abi": c_abi}


def run_fixed_theta(args, cfg, samples, spec, torch, W, precisions, sizes):
    """The same workload and events with every resonance mass and width fixed at its
    configured value: no free resonance parameter, so the phase-space term of each
    evaluation is the quadratic form of the phase-space amplitude Gram built once when the
    analysis is created, and the fused kernel reads the data and background only."""
    from ctpwa_amd import analysis
    fixed_cfg = os.path.join(os.path.dirname(cfg), "config_fixed_theta.yml")
    with open(cfg) as f:
        txt = f.read()
    with open(fixed_cfg, "w") as f:
        f.write(W.fixed_theta_config(txt))
    res = {"config": "the headline workload's events with every resonance parameter fixed (no free theta): "
                     "phase-space term from the amplitude Gram stored at creation"}
    for prec in precisions:
        t0 = time.time()
        ana = analysis(fixed_cfg, samples=samples, precision=prec)
        torch.cuda.synchronize()
        init_s = time.time() - t0
        r = run_form(ana, args, spec, torch, W, prec, sizes, label=prec + " fixed-theta", hessian=False)
        rf = r["roofline"]
        res[prec] = {"value": round(r["value"], 3), "ms_per_step": round(r["ms_per_step"], 4), "sync": r["sync"],
                     "kernel_ms": rf["kernel_ms"], "bytes_per_eval": rf["algorithmic_bytes"],
                     "kernel_frac_of_hbm": rf["frac"], "init_s": round(init_s, 2), "nll": r["nll"]}
        del ana
        torch.cuda.empty_cache()
    return res


def hessian_flops(plan, n_data, n_phsp, n_bkg):
    """Algorithmic flops of one Hessian (the reference's formula: per data and background
    event the (2n+P)^2 outer product, per event of every class the complex n x n
    amplitude Gram over the polarisations)."""
    n, P = plan.n_free, plan.n_theta
    m = 2 * n + P
    return (n_data + n_bkg) * m * (m + 1) + 8 * plan.npol * (n_data + n_bkg + n_phsp) * n * (n + 1) // 2


def run_hessian_wide(args, W, torch, precisions):
    """getHessian on the two-family four-body cascade (149 amplitudes, 24 free resonance
    parameters, 322 x 322): 10M phase-space events and 2M "data" events, both uniform in
    phase space (the Hessian's cost does not depend on how the data are distributed, and
    accept-reject on this model's narrow phi takes minutes), seed 4; one first call
    (which compiles the generated Hessian kernel), then the mean of --hessian-wide-reps
    calls, per form."""
    from ctpwa_amd import analysis
    name = "jpsi_kkpipi_wide"
    spec = W.WORKLOADS[name]
    t0 = time.time()
    smp = {"phsp": W.phase_space(spec["n_phsp"], seed=1004, masses=spec["masses"], workers=8),
           "data": W.phase_space(spec["n_data"], seed=3004, masses=spec["masses"], workers=8)}
    t_gen = time.time() - t0
    log("[bench] %s samples generated in %.1f s" % (name, t_gen))
    d = tempfile.mkdtemp(prefix="ctpwa_wide_")
    cfg = os.path.join(d, "config.yml")
    with open(cfg, "w") as f:
        f.write(spec["config"])
    res = {"workload": "%s: J/psi -> K+ K- pi+ pi-, K1(1270), K1(1400), K1(1650) -> K*(892) pi and K rho, "
                       "K* K*bar and K* K*2(1430)bar over three K* states, phi or f2'(1525) with pi pi "
                       "resonances; %%d amplitudes, %%d free resonance "
                       "parameters; %d data (uniform in phase space) / %d phase space, seed 4" % (
                           name, spec["n_data"], spec["n_phsp"]),
           "call": "analysis.getHessian, mean of %d after one first call" % args.hessian_wide_reps,
           "sample_generation_s": round(t_gen, 1), "ms": {}, "first_call_s": {}}
    plan = None
    for prec in precisions:
        t = time.time()
        ana = analysis(cfg, samples={"phsp": smp["phsp"], "data": smp["data"]}, precision=prec)
        torch.cuda.synchronize()
        t_init = time.time() - t
        plan = ana.plan
        p = torch.tensor(W.nominal_params(plan, 4), dtype=torch.float64, device="cuda")
        log("[bench] %s %s: handle ready in %.1f s, first getHessian" % (name, prec, t_init))
        t = time.time()
        ana.getHessian(p)
        torch.cuda.synchronize()
        res["first_call_s"][prec] = round(time.time() - t, 2)
        t = time.perf_counter()
        for _ in range(args.hessian_wide_reps):
            H = ana.getHessian(p)
        torch.cuda.synchronize()
        ms = (time.perf_counter() - t) / args.hessian_wide_reps * 1e3
        res["ms"][prec] = round(ms, 2)
        sym = float((H - H.T).abs().max() / H.abs().max())
        log("[bench] %s %s: init %.1f s, first getHessian %.1f s, getHessian %.2f ms, |H - H^T|/|H| %.2e"
            % (name, prec, t_init, res["first_call_s"][prec], ms, sym))
        del ana, H
        torch.cuda.empty_cache()
    res["workload"] = res["workload"] % (plan.n_amp, plan.n_theta)
    fl = hessian_flops(plan, spec["n_data"], spec["n_phsp"], 0)
    res["algorithmic_flops"] = fl
    res["mfma_fp64_peak_tflops"] = 78.6
    res["frac_of_fp64_mfma_peak"] = {k: round(fl / (v * 1e-3) / 78.6e12, 4) for k, v in res["ms"].items()}
    import shutil
    shutil.rmtree(d, ignore_errors=True)
    return res


def _serve_once(plan, samples, x_nom, prec, method, rate, max_evals, args):
    from ctpwa_amd.serve import FitService, toy_ensemble
    jobs = toy_ensemble(plan, samples["phsp"], samples["data"], args.serving_jobs, seed=3, rate=rate,
                        x_nominal=x_nom, max_evals=max_evals)
    svc = FitService(plan.to_text(), precision=prec, max_slots=args.serving_slots, method=method)
    st = svc.run(jobs)
    svc.close()
    sm = st.summary()
    sm["arrival_rate_per_s"] = rate
    sm["arrival_span_s"] = max(j.arrival for j in jobs)
    log("[bench] serving %s %s at %.1f fits/s: %d fits (%d converged, median %d evaluations, %d Hessians) in %.2f s, "
        "%.0f fits/h, request p50 %.2f ms p99 %.2f ms, fit p50 %.2f s p99 %.2f s, mean batch %.1f"
        % (method, prec, rate, sm["fits"], sm["converged"], sm["evals_per_fit_median"], sm["hessians"], sm["wall_s"],
           sm["fits_per_hour"], sm["latency_p50_ms"], sm["latency_p99_ms"], sm["fit_latency_p50_s"],
           sm["fit_latency_p99_s"], sm["mean_batch"]))
    return {k: (round(v, 4) if isinstance(v, float) else v) for k, v in sm.items()}


def run_serving(args, samples, plan, spec, W, precisions):
    """Toy-MC ensemble through the fit service: data counts log-uniform in [5k, 500k],
    phase space 5x data, Poisson arrivals (seed 3).  Each minimiser runs twice per form: at
    a burst rate above the service's capacity (fits/hour is then the capacity) and at half
    the capacity that burst measured (a stream the service keeps up with).  L-BFGS fits
    issue NLL+gradient requests until the projected gradient is below 1e-5 or their request
    budget (200) is spent; Newton fits use the GPU Hessian and stop at EDM < 1e-5."""
    x_nom = W.nominal_params(plan, spec["seed"])
    res = {"jobs": args.serving_jobs, "slots": args.serving_slots, "seed": 3,
           "data_per_fit": "log-uniform [5000, 500000]", "phsp_per_fit": "5 x data",
           "events": "windows of the benchmark's data and phase-space samples",
           "latency_def": "per NLL+gradient request: from the fit issuing its next point to the result on the host",
           "rates": "above: burst at --serving-rate; below: half the fits/s the burst measured"}
    for method in [m for m in args.serving_methods.split(",") if m]:
        max_evals = args.serving_lbfgs_evals if method == "lbfgs" else 400
        res[method] = {"max_evals": max_evals}
        for prec in precisions:
            above = _serve_once(plan, samples, x_nom, prec, method, args.serving_rate, max_evals, args)
            cap = above["fits_per_hour"] / 3600.0
            below = _serve_once(plan, samples, x_nom, prec, method, max(0.5 * cap, 1e-3), max_evals, args)
            res[method][prec] = {"above_capacity": above, "below_capacity": below}
    return res


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--workload", default="jpsi_3pi")
    ap.add_argument("--n-data", type=int, default=None)
    ap.add_argument("--n-phsp", type=int, default=None)
    ap.add_argument("--n-bkg", type=int, default=None)
    ap.add_argument("--cpu-frac", type=float, default=1.0,
                    help="fraction of every sample the CPU baseline evaluates (1.0: the whole workload)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--workdir", default=None)
    ap.add_argument("--precision", choices=["both", "mixed", "fp64"], default="both",
                    help="both: time the mixed form (the reported value) and the FP64 form beside it")
    ap.add_argument("--serving-jobs", type=int, default=256,
                    help="toy fits run through the fit service after the timing (0: skip)")
    ap.add_argument("--serving-rate", type=float, default=1000.0,
                    help="Poisson arrival rate of the burst run (above the service's capacity, so fits/hour "
                         "measures the service); a second run streams at half the measured capacity")
    ap.add_argument("--serving-methods", default="lbfgs,newton",
                    help="minimisers of the toy fits, comma-separated (lbfgs, newton)")
    ap.add_argument("--serving-lbfgs-evals", type=int, default=200,
                    help="NLL+gradient request budget of an L-BFGS toy fit")
    ap.add_argument("--serving-slots", type=int, default=64, help="fits evaluated together at most")
    ap.add_argument("--hessian-reps", type=int, default=10,
                    help="getHessian calls timed per form on the same workload (0: skip)")
    ap.add_argument("--hessian-wide-reps", type=int, default=3,
                    help="getHessian calls timed per form on the 149-amplitude four-body cascade at "
                         "2M data / 10M phase space (0: skip)")
    ap.add_argument("--no-fixed-theta", dest="fixed_theta", action="store_false",
                    help="skip the leg that times the workload with every resonance parameter fixed")
    ap.add_argument("--copy-probe-mb", type=int, default=0,
                    help="after the run, copy a buffer of this many MiB once (a plain copy kernel of "
                         "known size that calibrates the HBM byte counters of a --pmc profile)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path (events would be "
              "sharded per GPU with one all-reduce per evaluation); run with --gpus 1", file=sys.stderr)
        sys.exit(2)

    import torch
    from ctpwa_amd import workloads as W
    from ctpwa_amd import analysis

    spec = W.WORKLOADS[args.workload]
    n_data = spec["n_data"] if args.n_data is None else args.n_data
    n_phsp = spec["n_phsp"] if args.n_phsp is None else args.n_phsp
    n_bkg = spec["n_bkg"] if args.n_bkg is None else args.n_bkg
    tmp = None
    workdir = args.workdir
    if workdir is None:
        tmp = tempfile.TemporaryDirectory(prefix="ctpwa_bench_")
        workdir = tmp.name
    cfg = make_samples(workdir, args.workload, n_data, n_phsp, n_bkg)
    precisions = ["mixed", "fp64"] if args.precision == "both" else [args.precision]
    samples = {k: np.load(os.path.join(workdir, k + ".npy")) for k in ("phsp", "data", "bkg")
               if os.path.exists(os.path.join(workdir, k + ".npy"))}
    plan = None
    results = {}
    for prec in precisions:
        t0 = time.time()
        ana = analysis(cfg, samples=samples, precision=prec)
        torch.cuda.synchronize()
        log("[bench] %s analysis initialised in %.1f s" % (prec, time.time() - t0))
        plan = ana.plan
        results[prec] = run_form(ana, args, spec, torch, W, prec, (n_data, n_phsp, n_bkg))
        del ana
        torch.cuda.empty_cache()
    primary = precisions[0]
    r = results[primary]
    value, ms_per_step, nll_last, roof = r["value"], r["ms_per_step"], r["nll"], r["roofline"]
    compare = None
    if len(precisions) == 2:
        other = results[precisions[1]]
        compare = {"precision": precisions[1], "value": round(other["value"], 3),
                   "ms_per_step": round(other["ms_per_step"], 4), "sync": other["sync"], "c_abi": other["c_abi"],
                   "roofline": other["roofline"],
                   "nll": other["nll"],
                   "speedup_mixed_vs_fp64": round(r["value"] / other["value"], 3),
                   "delta_nll_mixed_minus_fp64": r["nll"] - other["nll"],
                   "delta_nll_points": "NLL at the same last timed parameter point in both forms"}

    fixed = None
    if args.fixed_theta:
        fixed = run_fixed_theta(args, cfg, samples, spec, torch, W, precisions, (n_data, n_phsp, n_bkg))

    serving = None
    if args.serving_jobs > 0:
        serving = run_serving(args, samples, plan, spec, W, precisions)

    wide = None
    if args.hessian_wide_reps > 0:
        del samples
        wide = run_hessian_wide(args, W, torch, precisions)

    if args.copy_probe_mb > 0:
        nb = args.copy_probe_mb << 20
        src = torch.ones(nb // 8, dtype=torch.float64, device="cuda")
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        dst.copy_(src)
        torch.cuda.synchronize()
        log("[bench] copy probe: %d bytes read, %d bytes written" % (nb, nb))
        del src, dst

    cpu = None
    if not args.no_cpu_baseline:
        threads = min(os.cpu_count() or 1, int(os.environ.get("OMP_NUM_THREADS", "16") or 16))
        t0 = time.time()
        cpu = cpu_baseline(cfg, args.cpu_frac, threads, hessian=args.hessian_reps > 0)
        log("[bench] cpu baseline %.4f evals/s (%.1f s)" % (cpu["value"], time.time() - t0))

    desc = {"mixed": "mixed: FP32 amplitude (SL) tensors and invariants, FP64 squared masses, "
                     "FP64 arithmetic and sums",
            "fp64": "fp64: FP64 storage and arithmetic"}
    out = {
        "metric": METRIC, "value": round(value, 3), "unit": "evals/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64" if primary == "fp64" else "f64 arithmetic, f32 amplitude storage",
        "sync": r["sync"],
        "c_abi": r["c_abi"],
        "data": "synthetic",
        "config": {"workload": args.workload + ": J/psi -> pi+ pi- pi0, 3 chains x (rho770, rho1450, "
                   "rho1700) + f0(980) + f2(1270), %d amplitudes, %d free resonance parameters, "
                   "%d data / %d phase space / %d background (w=0.8)" % (plan.n_amp, plan.n_theta, n_data, n_phsp, n_bkg),
                   "n_data": n_data, "n_phsp": n_phsp, "n_bkg": n_bkg, "precision": desc[primary],
                   "nll": nll_last},
        "roofline": roof,
        "cpu_baseline": cpu,
    }
    if compare is not None:
        out["fp64"] = compare
    if args.hessian_reps > 0:
        fl = hessian_flops(plan, n_data, n_phsp, n_bkg)
        hms = {p: round(results[p]["hessian_ms"], 4) for p in precisions}
        out["hessian"] = {"call": "analysis.getHessian on the same workload, mean of %d after one warm-up"
                                  % args.hessian_reps,
                          "ms": hms, "algorithmic_flops": fl, "mfma_fp64_peak_tflops": 78.6,
                          "frac_of_fp64_mfma_peak": {k: round(fl / (v * 1e-3) / 78.6e12, 4) for k, v in hms.items()}}
    if fixed is not None:
        out["fixed_theta"] = fixed
    if wide is not None:
        out["hessian_wide"] = wide
    if serving is not None:
        out["serving"] = serving
    print(json.dumps(out), flush=True)
    if tmp is not None:
        tmp.cleanup()


if __name__ == "__main__":
    main()
