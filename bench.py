#!/usr/bin/env python3
"""Benchmark: RHS solves/s to 1e-5 relative residual, plus panel factor time,
FP32 fronts vs FP64 fronts, on one MI355X.

Workload: 100^3 grid, smooth random velocity in [1.5, 4.5] (32 seeded
Gaussian bumps on a background of 3.0), 10 points per wavelength at the
minimum velocity, 8 seeded Gaussian point sources as one right-hand-side
block, GMRES(20) to 1e-5.  A step is one Solve of the whole block with the
right-hand sides already in HBM.  `value` is the FP32-front / FP64-GMRES
path; the FP64-front path is timed beside it.

    python bench.py [--gpus 1] [--steps K] [--warmup W]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)


def log(*a):
    print(*a, file=sys.stderr, flush=True)


TRAFFIC_PROFILE = "k_mvp_k_mvm_front_solves_bumps100_8rhs.md"


def measured_traffic():
    """HBM bytes per front-solve product launch (FETCH_SIZE + WRITE_SIZE,
    rocprofv3 --pmc) recorded in the committed profile summary, or None."""
    import re
    path = os.path.join(ROOT, "profiles", TRAFFIC_PROFILE)
    try:
        m = re.search(r"\*\*([0-9.]+) MB per launch\*\* of HBM traffic", open(path).read())
        return float(m.group(1)) * 1e6 if m else None
    except OSError:
        return None


def cpu_model():
    """Model name of the host CPU (Linux), or None."""
    try:
        for ln in open("/proc/cpuinfo"):
            if ln.startswith("model name"):
                return ln.split(":", 1)[1].strip()
    except OSError:
        pass
    return None


def cpu_baseline(wl, out, args):
    """FP64 CPU restatement of PSP's method (the oracle, OpenMP over the host
    cores) timed on a bounded sample of the same workload, extrapolated to
    one Solve of the whole right-hand-side block:
      factor time = sum over panel classes of (#panels x one panel factored),
      solve time  = iterations x (sweep + 2 operator applications)
                    + iterations / m x one GMRES(m) Arnoldi cycle,
      sweep       = sum over the 2P-1 panel solves of the class's panel-solve
                    time, each measured once on its factored sample panel.
    The iteration count is the GPU FP64 path's (same method)."""
    from oracle.oracle import Oracle, num_threads
    d, vel, B = wl.disc, wl.velocity, wl.B
    k = B.shape[1]
    o = Oracle(d.omega, d.nx, d.ny, d.nz, d.wx, d.wy, d.wz, sigma=d.sigmax, cutoff=args.cutoff)
    cores = num_threads()
    np_ = o.num_panels
    info = [o.panel_info(p) for p in range(np_)]
    rep_panel = {}  # one sample panel per class (a full inner panel for class 0)
    for p in range(np_ - 1, -1, -1):
        rep_panel[info[p]["cls"]] = p
    t_fac = {}
    for c, p in rep_panel.items():
        t_fac[c] = o.factor_sample(vel, p, 1)
        log(f"[cpu] factored panel {p} (class {c}) in {t_fac[c]:.2f}s")
    o.fill_global(vel)
    l2n = o.local_to_natural()
    Bnd = np.asfortranarray(B[l2n])
    t_ps = {}
    for c, p in rep_panel.items():
        t0 = time.perf_counter()
        o.solve_panel(Bnd, p)
        t_ps[c] = time.perf_counter() - t0
    t0 = time.perf_counter()
    o.multiply(Bnd)
    t_mv = time.perf_counter() - t0
    t_arn = o.time_arnoldi(k, args.m)
    n_fac = sum(t_fac[info[p]["cls"]] for p in range(np_))
    t_sweep = sum(t_ps[info[p]["cls"]] * (2 if p < np_ - 1 else 1) for p in range(np_))
    its = out.get("fp64", out["fp32"])["iterations"]
    t_solve = its * (t_sweep + 2 * t_mv) + (its / args.m) * t_arn + t_sweep
    log(f"[cpu] sweep {t_sweep:.2f}s, multiply {t_mv:.3f}s, arnoldi cycle {t_arn:.2f}s -> "
        f"solve {t_solve:.1f}s, factor {n_fac:.1f}s (est.)")
    return {"value": k / t_solve, "unit": "RHS/s", "cores": cores, "kind": "port",
            "cpu_model": cpu_model(),
            "factor_s_est": n_fac, "solve_s_est": t_solve,
            "sample": (f"oracle (FP64 CPU restatement, not the PSP binary): one panel of each of "
                       f"{len(rep_panel)} panel classes factored and solved once with the "
                       f"{k}-RHS block, one operator application, one GMRES({args.m}) Arnoldi "
                       f"cycle; extrapolated to {its} iterations and {np_} panels")}


def cpu_full(wl, args):
    """The oracle's whole Solve (factor + GMRES) on the host cores, timed."""
    from oracle.oracle import Oracle, num_threads
    d, vel, B = wl.disc, wl.velocity, wl.B
    o = Oracle(d.omega, d.nx, d.ny, d.nz, d.wx, d.wy, d.wz, sigma=d.sigmax, cutoff=args.cutoff)
    t0 = time.perf_counter()
    o.factor(vel)
    t_fac = time.perf_counter() - t0
    t0 = time.perf_counter()
    x, st, rep = o.solve(B, args.m, args.tol)
    t_sol = time.perf_counter() - t0
    log(f"[cpu] factor {t_fac:.2f}s, solve {t_sol:.2f}s, {rep.iterations} its")
    return {"value": B.shape[1] / t_sol, "unit": "RHS/s", "cores": num_threads(), "kind": "port",
            "cpu_model": cpu_model(), "iterations": rep.iterations,
            "factor_s": t_fac, "solve_s": t_sol,
            "sample": "oracle (FP64 CPU restatement, not the PSP binary): the whole workload"}


def warm_library(dev):
    """Loads the library's GPU code (the first launch in a process loads the
    code object) with a small factor and solve in both front precisions, so
    that no timed Initialize pays for it."""
    import torch
    import psp_amd
    from psp_amd import workloads
    vel = workloads.bump_velocity(16, 16, 24, cfg=1)
    disc = workloads.discretization(16, 16, 24, float(vel.min()), 10.0)
    B = workloads.gaussian_sources(16, 16, 24, workloads.interior_points(16, 16, 24, 8, cfg=1))
    for prec in ("fp32", "fp64"):
        h = psp_amd.DistHelmholtz(disc, precision=prec)
        h.Initialize(vel)
        h.Solve(torch.from_numpy(B).to(dev), 20, 1e-5, raise_on_error=False)
    torch.cuda.synchronize()


def bench_batch(args, dev):
    """32 uneven problems (n in 40..96, 4..16 sources each) solved together:
    aggregate RHS/s = total right-hand sides / wall time of one SolveBatch."""
    import torch
    import psp_amd
    from psp_amd import workloads
    wls = workloads.config4()
    ktot = sum(w.B.shape[1] for w in wls)
    res = {}
    for prec in ("fp32", "fp64"):
        if prec == "fp64" and args.no_fp64:
            continue
        # FP32 fronts: all 32 problems resident and solved together.  FP64
        # fronts take twice the HBM, so they go in 4 groups of 8.
        group = len(wls) if prec == "fp32" else 8
        t_fac, t_sol, reps = 0.0, 0.0, []
        for g0 in range(0, len(wls), group):
            part = wls[g0:g0 + group]
            helms = []
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            for w in part:
                h = psp_amd.DistHelmholtz(w.disc, precision=prec)
                if args.batch_tiling:
                    h.set_solve_tiling(**{kv.split("=")[0]: int(kv.split("=")[1])
                                          for kv in args.batch_tiling.split(",")})
                h.Initialize(w.velocity)
                helms.append(h)
            torch.cuda.synchronize()
            t_fac += time.perf_counter() - t0
            if args.batch_autotune:  # tile shapes per problem, outside the timed solves
                for w, h in zip(part, helms):
                    h.Autotune(w.B.shape[1])
            srcs = [torch.from_numpy(w.B).to(dev) for w in part]
            for w_ in range(args.warmup):
                psp_amd.SolveBatch(helms, [b.clone() for b in srcs], args.m, args.tol,
                                   threads=args.batch_threads)
            times = []
            for s_ in range(args.steps):
                bufs = [b.clone() for b in srcs]
                torch.cuda.synchronize()
                t0 = time.perf_counter()
                rp = psp_amd.SolveBatch(helms, bufs, args.m, args.tol, threads=args.batch_threads)
                torch.cuda.synchronize()
                times.append(time.perf_counter() - t0)
                del bufs
            t_sol += sum(times) / len(times)
            reps += rp
            del helms, srcs
            torch.cuda.empty_cache()
        conv = all(r.status == 0 and r.final_max_rel < args.tol for r in reps)
        res[prec] = dict(factor_all_s=t_fac, solve_s=t_sol,
                         rhs_per_s=ktot / t_sol if conv else None, converged=conv,
                         problems_per_batch=group,
                         max_final_rel=max(r.final_max_rel for r in reps),
                         iterations=[r.iterations for r in reps])
        log(f"[bench] batch32 {prec} ({group} problems per SolveBatch): factor all {t_fac:.2f}s, "
            f"solve all {t_sol:.3f}s -> {ktot / t_sol:.1f} RHS/s (converged {conv}), worst final rel "
            f"{res[prec]['max_final_rel']:.2e}")
    line = {"metric": "aggregate RHS solves/s over 32 uneven problems to 1e-5 rel. residual",
            "value": res["fp32"]["rhs_per_s"], "unit": "RHS/s", "n_gpus": 1,
            "steps": args.steps, "warmup": args.warmup,
            "ms_per_step": 1e3 * res["fp32"]["solve_s"], "higher_is_better": True,
            "scaling": "weak", "vs_baseline": None, "dtype": "c64 fronts / c128 GMRES",
            "data": "synthetic",
            "config": {"workload": "32 uneven bump problems n in {40..96}, 4-16 rhs each",
                       "rhs_total": ktot, "sizes": [w.disc.nx for w in wls],
                       "rhs": [w.B.shape[1] for w in wls],
                       "batching": ("one host thread and stream per problem" if args.batch_threads
                                    else "lockstep GMRES, one launch per sweep step for all "
                                         "problems")},
            "paths": res}
    print(json.dumps(line), flush=True)


def bench_adjoint(args, dev):
    """80^3 misfit gradient with 16 sources: forward solve, adjoint solve with
    the same factors, gradient kernel; gradients/s, FP32 vs FP64 fronts."""
    import torch
    from psp_amd import workloads
    from psp_amd.adjoint import HelmholtzMisfit
    import psp_amd
    wl = workloads.config5()
    F = torch.from_numpy(wl.B).to(dev)
    rec = torch.from_numpy(wl.receivers).to(dev)
    ht = psp_amd.DistHelmholtz(wl.disc)
    ht.Initialize(1.0 / np.sqrt(wl.m_true))
    Ut = F.clone()
    ht.Solve(Ut, args.m, 1e-10)
    data = Ut[rec].clone()
    del ht, Ut
    res, grads = {}, {}
    for prec in ("fp32", "fp64"):
        opts = {"m": args.m, "tol": args.tol, "precision": prec}
        times = []
        for s_ in range(args.warmup + args.steps):
            m = torch.from_numpy(wl.m).to(dev).requires_grad_(True)
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            loss = HelmholtzMisfit.apply(m, F, rec, data, wl.disc, opts)
            loss.backward()
            torch.cuda.synchronize()
            if s_ >= args.warmup:
                times.append(time.perf_counter() - t0)
        grads[prec] = m.grad.detach()
        t = sum(times) / len(times)
        res[prec] = dict(seconds=t, gradients_per_s=1.0 / t, loss=loss.item())
        log(f"[bench] adjoint80 {prec}: {t:.3f}s per gradient (factor + 2 solves of 16 rhs)")
    err = (torch.linalg.norm(grads["fp32"] - grads["fp64"]) / torch.linalg.norm(grads["fp64"])).item()
    line = {"metric": "misfit gradients/s (factor + forward + adjoint solve, 16 sources)",
            "value": res["fp32"]["gradients_per_s"], "unit": "gradients/s", "n_gpus": 1,
            "steps": args.steps, "warmup": args.warmup,
            "ms_per_step": 1e3 * res["fp32"]["seconds"], "higher_is_better": True,
            "scaling": "weak", "vs_baseline": None, "dtype": "c64 fronts / c128 GMRES",
            "data": "synthetic",
            "config": {"workload": wl.name, "rel_tol": args.tol},
            "paths": res, "fp32_vs_fp64_grad_rel_err": err}
    print(json.dumps(line), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--n", type=int, default=100)
    ap.add_argument("--k", type=int, default=8)
    ap.add_argument("--cutoff", type=int, default=12,
                    help="nested-dissection leaf size (points of a panel's cross-section; "
                         "the reference's default is 12)")
    ap.add_argument("--m", type=int, default=20)
    ap.add_argument("--tol", type=float, default=1e-5)
    ap.add_argument("--max-its", type=int, default=-1,
                    help="GMRES iteration cap (0: none, as the reference; default: none, "
                         "100 for thrustbelt)")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-fp64", action="store_true", help="skip the FP64-front leg")
    ap.add_argument("--batch-tiling", default="",
                    help="batch32: set_solve_tiling options for every problem, e.g. "
                         "'persist=0,min_tiles=1024'")
    ap.add_argument("--batch-autotune", action="store_true",
                    help="batch32: Autotune(k) every problem before the timed solves")
    ap.add_argument("--batch-threads", action="store_true",
                    help="batch32: one host thread and stream per problem instead of the "
                         "merged sweep launches")
    ap.add_argument("--no-autotune", action="store_true",
                    help="keep the default front-solve tile shapes (no tuning pass after "
                         "Initialize)")
    ap.add_argument("--workload", default="bumps100",
                    choices=["bumps100", "unitcube40", "thrustbelt", "batch32", "adjoint80"],
                    help="bumps100 (default, the headline line): 100^3 bumps, --k sources; "
                         "unitcube40: 40^3 c=1, 1 source; thrustbelt: 200x200x100, 64 sources; "
                         "batch32: 32 uneven problems solved together; adjoint80: 80^3 "
                         "misfit gradient, 16 sources")
    ap.add_argument("--wide-k", type=int, default=64,
                    help="also time FP32 and FP64 fronts with this many sources on the "
                         "100^3 grid (0: skip)")
    ap.add_argument("--wide-steps", type=int, default=2,
                    help="timed solves of the wide block per precision")
    ap.add_argument("--fused", type=int, default=0,
                    help="1: run the tree levels below the cut depth as one fused launch per "
                         "direction (8 <= k <= 16); 0: one launch per level")
    ap.add_argument("--fused-wg", type=int, default=0,
                    help="workgroups a fused queue expects (0: the library's default)")
    ap.add_argument("--xcd-reduce", type=int, default=1,
                    help="0: split row blocks summed by a reduce launch instead of by the "
                         "product launch's last slice")
    ap.add_argument("--fused-rows", type=int, default=0,
                    help="output rows per fused tile, 16, 32 or 64 (0: the library's default)")
    ap.add_argument("--cpu-full", action="store_true",
                    help="bumps100: also run the oracle's whole Solve on the host cores (a few "
                         "minutes) and report it beside the sampled estimate")
    ap.add_argument("--calibrate", action="store_true",
                    help="copy a 1 GiB device buffer once (a known-size kernel for the "
                         "FETCH_SIZE / WRITE_SIZE counter scale)")
    args = ap.parse_args()

    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print(json.dumps({"error": "psp_amd runs on one GPU: the whole grid lives in one "
                          "MI355X's HBM; sharding over several GPUs is not implemented",
                          "n_gpus": args.gpus}))
        sys.exit(2)

    import torch
    import psp_amd
    from psp_amd import workloads

    assert torch.cuda.is_available(), "bench.py needs a GPU"
    dev = torch.device("cuda:0")
    warm_library(dev)

    if args.workload == "batch32":
        return bench_batch(args, dev)
    if args.workload == "adjoint80":
        return bench_adjoint(args, dev)
    if args.max_its < 0:
        args.max_its = 100 if args.workload == "thrustbelt" else 0
    if args.workload == "unitcube40":
        wl = workloads.config1()
    elif args.workload == "thrustbelt":
        wl = workloads.config3()
    else:
        log(f"[bench] building workload n={args.n} k={args.k}")
        wl = workloads.config2(n=args.n, k=args.k)
    disc, vel, B = wl.disc, wl.velocity, wl.B
    k = B.shape[1]
    Bsrc = torch.from_numpy(B).to(dev)

    out = {}
    if args.calibrate:
        a = torch.empty(1 << 27, dtype=torch.float64, device=dev).fill_(1.0)
        b = torch.empty_like(a)
        b.copy_(a)
        torch.cuda.synchronize()
        del a, b

    def run_path(precision, steps, warmup, src=None):
        """Initialize + Autotune, then `warmup` untimed and `steps` timed
        Solves of the block; returns the object, its factor report, the
        factor time, the timed seconds, the reports and one solution."""
        src = Bsrc if src is None else src
        kk = src.shape[1]
        h = psp_amd.DistHelmholtz(disc, precision=precision, cutoff=args.cutoff)
        if not args.xcd_reduce:
            h.set_solve_tiling(xcd_reduce=0)
        if args.fused:
            h.set_solve_tiling(fused=1, fused_wg=args.fused_wg or None,
                               fused_rows=args.fused_rows or None)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        rep = h.Initialize(vel)
        t_fac = time.perf_counter() - t0
        log(f"[bench] {precision} k={kk}: factor {t_fac:.3f}s (GPU time per phase, summed over "
            f"the panel classes: fill {rep.seconds_fill:.3f} ldl {rep.seconds_ldl:.3f} inv "
            f"{rep.seconds_inverse:.3f}), {rep.flops / 1e12:.2f} TFLOP, "
            f"{rep.bytes_factor / 1e9:.2f} GB of factors")
        # The Schur-complement GEMM rate from one more factorization with the
        # panel classes one after another (the timed one runs two of them
        # beside the third, so its GEMM events also span other classes' work).
        h.set_factor_overlap(False)
        rs = h.Initialize(vel)
        h.set_factor_overlap(True)
        rep.schur_flops, rep.seconds_schur = rs.schur_flops, rs.seconds_schur
        tuned = None
        if not args.no_autotune:
            t0 = time.perf_counter()
            before, after = h.Autotune(kk)
            tuned = dict(seconds=time.perf_counter() - t0, est_products_before_s=before,
                         est_products_after_s=after)
            log(f"[bench] {precision} k={kk}: autotuned solve tiles in {tuned['seconds']:.2f}s "
                f"(est. front-solve time per sweep {before * 1e3:.2f} -> {after * 1e3:.2f} ms)")
        bufs = [src.clone() for _ in range(max(steps, 1))]
        for w in range(warmup):
            x = src.clone()
            r = h.Solve(x, args.m, args.tol, maxIts=args.max_its, raise_on_error=False)
            log(f"[bench] {precision} k={kk} warmup {w}: {r.iterations} its, rel "
                f"{r.final_max_rel:.2e}, {r.seconds_total:.3f}s (precond {r.seconds_precond:.3f}, "
                f"multiply {r.seconds_multiply:.3f}, arnoldi {r.seconds_arnoldi:.3f})")
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        reps = []
        for s in range(steps):
            reps.append(h.Solve(bufs[s], args.m, args.tol, maxIts=args.max_its, raise_on_error=False))
        torch.cuda.synchronize()
        t = time.perf_counter() - t0
        for s, r in enumerate(reps):
            log(f"[bench] {precision} k={kk} step {s}: {r.iterations} its, rel "
                f"{r.final_max_rel:.2e}, {r.seconds_total:.3f}s"
                + ("" if r.status == 0 else f" (status {r.status})"))
        x0 = bufs[0]
        del bufs
        rep.tuned = tuned
        return h, rep, t_fac, t, reps, x0

    def summary(reps, t, t_fac, kk, f):
        conv = all(r.status == 0 and r.final_max_rel < args.tol for r in reps)
        out_ = dict(factor_s=t_fac, solve_s=t / len(reps), iterations=reps[0].iterations,
                    converged=conv, rhs_per_s=len(reps) * kk / t if conv else None,
                    final_rel=max(r.final_max_rel for r in reps), autotune=f.tuned,
                    schur_gemm_tflops=f.schur_flops / max(f.seconds_schur, 1e-12) / 1e12)
        if not conv:
            out_["rhs_per_s_unconverged"] = len(reps) * kk / t
        return out_

    h32, f32, tf32, t32, reps32, x32 = run_path("fp32", args.steps, args.warmup)
    out["fp32"] = summary(reps32, t32, tf32, k, f32)
    value = out["fp32"]["rhs_per_s"]

    # ---- roofline of the dominant kernels: the front-solve products -----
    # One more sweep of the same block, launched kernel by kernel with HIP
    # events around every front-solve product launch (the timed solves replay
    # the sweep from a hipGraph, whose kernels cannot be bracketed singly).
    launches, ksec, kbytes, kflops = h32.ProfileSweep(Bsrc.clone())
    achieved = kbytes / ksec / 1e9
    kname = ("k_mvp/k_mvm" if k <= 16 else "k_mvm") if 8 <= k < 32 else ("k_mmw" if k >= 32 else "k_mv")
    traffic = measured_traffic() if (k == 8 and args.workload == "bumps100") else None
    roofline = {"bound": "hbm", "achieved": achieved, "peak": 8000.0, "unit": "GB/s",
                "frac": achieved / 8000.0,
                "traffic": traffic,
                "traffic_source": (None if traffic is None else
                                   "profiles/" + TRAFFIC_PROFILE + " (rocprofv3 --pmc FETCH_SIZE "
                                   "+ WRITE_SIZE per launch, calibrated; not measured in this run)"),
                "kernel": f"{kname} (front-solve products, c64 fronts)",
                "launches_per_sweep": launches, "avg_launch_us": 1e6 * ksec / launches,
                "algorithmic_bytes_per_launch": kbytes / launches,
                "achieved_tflops": kflops / ksec / 1e12,
                "schur_gemm_tflops": out["fp32"]["schur_gemm_tflops"]}
    log(f"[bench] roofline: {launches} front-solve launches/sweep, avg "
        f"{1e6 * ksec / launches:.1f} us, {kbytes / launches / 1e6:.2f} MB/launch -> "
        f"{achieved:.0f} GB/s, {kflops / ksec / 1e12:.2f} TFLOP/s; Schur GEMMs "
        f"{roofline['schur_gemm_tflops']:.1f} TFLOP/s")
    del h32
    torch.cuda.empty_cache()

    if not args.no_fp64:
        # the FP64-front path with the same warm-ups and timed steps
        h64, f64, tf64, t64, reps64, x64 = run_path("fp64", args.steps, args.warmup)
        err = (torch.linalg.norm(x32 - x64) / torch.linalg.norm(x64)).item()
        out["fp64"] = summary(reps64, t64, tf64, k, f64)
        out["fp32_vs_fp64_rel_err"] = err
        if value is not None and out["fp64"]["rhs_per_s"]:
            out["fp32_over_fp64_rhs_per_s"] = value / out["fp64"]["rhs_per_s"]
        log(f"[bench] fp32 vs fp64 solution rel err {err:.3e}")
        del h64
        torch.cuda.empty_cache()

    if args.wide_k > 0:
        # the many-sources pair: the same grid with a wide block of sources
        wl_w = workloads.config2(n=args.n, k=args.wide_k) if args.workload == "bumps100" else None
        if wl_w is not None:
            Bw = torch.from_numpy(wl_w.B).to(dev)
            wide = {}
            xs = {}
            for prec in ("fp32", "fp64"):
                if prec == "fp64" and args.no_fp64:
                    continue
                hw, fw, tfw, tw, repsw, xw = run_path(prec, args.wide_steps, 1, Bw)
                wide[prec] = summary(repsw, tw, tfw, args.wide_k, fw)
                xs[prec] = xw
                del hw
                torch.cuda.empty_cache()
            if "fp64" in xs:
                wide["fp32_vs_fp64_rel_err"] = (torch.linalg.norm(xs["fp32"] - xs["fp64"]) /
                                                torch.linalg.norm(xs["fp64"])).item()
                if wide["fp32"]["rhs_per_s"] and wide["fp64"]["rhs_per_s"]:
                    wide["fp32_over_fp64_rhs_per_s"] = (wide["fp32"]["rhs_per_s"] /
                                                        wide["fp64"]["rhs_per_s"])
            out[f"k{args.wide_k}"] = wide
            del xs, Bw

    if args.no_cpu:
        cpu = None
    elif args.workload == "unitcube40":
        cpu = cpu_full(wl, args)
    elif args.workload == "bumps100":
        cpu = cpu_baseline(wl, out, args)
        if args.cpu_full:
            full = cpu_full(wl, args)
            cpu["full_run"] = {key: full[key] for key in ("value", "iterations", "factor_s",
                                                          "solve_s")}
            cpu["full_run"]["estimate_over_full"] = cpu["solve_s_est"] / full["solve_s"]
            log(f"[cpu] estimate {cpu['solve_s_est']:.1f}s against the whole solve "
                f"{full['solve_s']:.1f}s ({full['iterations']} its)")
    else:
        cpu = None  # the CPU restatement needs hours here; see the 100^3 line

    line = {
        "metric": "RHS solves/s to 1e-5 rel. residual (FP32 fronts + FP64 GMRES)",
        "value": value,
        "unit": "RHS/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * t32 / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "c64 fronts / c128 GMRES",
        "data": "synthetic",
        "config": {"workload": wl.name, "grid": [disc.nx, disc.ny, disc.nz], "rhs": k,
                   "gmres_m": args.m, "rel_tol": args.tol, "omega": disc.omega,
                   "sigma": disc.sigmax, "max_its": args.max_its,
                   "sweep_launches": "fused below the cut" if args.fused else "one per level"},
        "status": "ok" if value is not None else
                  "not converged: value is null (a timed solve ended above rel_tol)",
        "roofline": roofline,
        "cpu_baseline": cpu,
        "paths": out,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
