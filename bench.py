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


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--n", type=int, default=100)
    ap.add_argument("--k", type=int, default=8)
    ap.add_argument("--m", type=int, default=20)
    ap.add_argument("--tol", type=float, default=1e-5)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-fp64", action="store_true", help="skip the FP64-front leg")
    args = ap.parse_args()

    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print(json.dumps({"error": "this build runs on one GPU: the whole grid lives in one "
                          "MI355X's HBM; multi-GPU sharding is a design only (see DESIGN.md)",
                          "n_gpus": args.gpus}))
        sys.exit(2)

    import torch
    import psp_amd
    from psp_amd import workloads

    assert torch.cuda.is_available(), "bench.py needs a GPU"
    dev = torch.device("cuda:0")

    log(f"[bench] building workload n={args.n} k={args.k}")
    wl = workloads.config2(n=args.n, k=args.k)
    disc, vel, B = wl.disc, wl.velocity, wl.B
    k = B.shape[1]
    Bsrc = torch.from_numpy(B).to(dev)

    out = {}

    def run_path(precision, steps, warmup):
        h = psp_amd.DistHelmholtz(disc, precision=precision)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        rep = h.Initialize(vel)
        t_fac = time.perf_counter() - t0
        log(f"[bench] {precision}: factor {t_fac:.3f}s (fill {rep.seconds_fill:.3f} "
            f"ldl {rep.seconds_ldl:.3f} inv {rep.seconds_inverse:.3f}), "
            f"{rep.flops / 1e12:.2f} TFLOP, {rep.bytes_factor / 1e9:.2f} GB of factors")
        bufs = [Bsrc.clone() for _ in range(max(steps, 1))]
        for w in range(warmup):
            x = Bsrc.clone()
            r = h.Solve(x, args.m, args.tol)
            log(f"[bench] {precision} warmup {w}: {r.iterations} its, rel {r.final_max_rel:.2e}, "
                f"{r.seconds_total:.3f}s (precond {r.seconds_precond:.3f}, "
                f"multiply {r.seconds_multiply:.3f}, arnoldi {r.seconds_arnoldi:.3f})")
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        reps = []
        for s in range(steps):
            reps.append(h.Solve(bufs[s], args.m, args.tol))
        torch.cuda.synchronize()
        t = time.perf_counter() - t0
        for s, r in enumerate(reps):
            log(f"[bench] {precision} step {s}: {r.iterations} its, rel {r.final_max_rel:.2e}, "
                f"{r.seconds_total:.3f}s")
        return h, rep, t_fac, t, reps, bufs[0]

    h32, f32, tf32, t32, reps32, x32 = run_path("fp32", args.steps, args.warmup)
    value = args.steps * k / t32
    out["fp32"] = dict(factor_s=tf32, solve_s=t32 / args.steps, iterations=reps32[0].iterations,
                       rhs_per_s=value, final_rel=reps32[0].final_max_rel)

    if not args.no_fp64:
        h64, f64, tf64, t64, reps64, x64 = run_path("fp64", 1, 0)
        err = (torch.linalg.norm(x32 - x64) / torch.linalg.norm(x64)).item()
        out["fp64"] = dict(factor_s=tf64, solve_s=t64, iterations=reps64[0].iterations,
                           rhs_per_s=k / t64, final_rel=reps64[0].final_max_rel)
        out["fp32_vs_fp64_rel_err"] = err
        log(f"[bench] fp32 vs fp64 solution rel err {err:.3e}")

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
                   "gmres_m": args.m, "rel_tol": args.tol, "omega": disc.omega},
        "paths": out,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
