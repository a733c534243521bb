#!/usr/bin/env python3
"""Benchmark of the numeric phase on one MI355X.

Workload: 3D 7-point variable-coefficient Laplacian on a 64^3 grid (n = 262,144),
edge weights log-uniform in [1e-2, 1e2] (seed 1), diagonal = sum of incident
weights + 1e-6, geometric nested-dissection ordering; synthetic data.

A step is one FP64 numeric factorization with the values already resident in HBM.
`value` = factor GFLOP/s with the reference flop count F = sum_j ColCount[j]^2
(ColumnCount.h:489-499) over the analysed structure.  The same line also carries
time-to-1e-12 backward error for the FP64 and mixed paths, the roofline of the
dominant kernel (HIP events around each launch of it, on its own stream) and the
CPU oracle timed on a bounded sample of the same operator family.

    python bench.py [--steps K] [--warmup W] [--side 64] [--cpu-side 40]
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

PEAK_FP64_TFLOPS = 78.6
PEAK_FP32_TFLOPS = 157.3
PEAK_HBM_GBS = 8000.0


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def cpu_baseline(side, nthreads):
    """The CPU oracle (a restatement of the reference loop) on a side^3 sample."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import swcholesky_amd as S
    from swcholesky_amd import problems as P
    import oracle_py as O
    prob = P.laplacian3d_var(side)
    b = S.analyze_p2(prob.n, prob.colptr, prob.rowidx, prob.perm)
    a2x = b.a2_values(prob.values)
    t0 = time.perf_counter()
    _, info = O.factor(b, a2x, nthreads=nthreads)
    dt = time.perf_counter() - t0
    assert info == 0
    return {"value": b.flops / dt / 1e9, "unit": "GFLOP/s", "cores": nthreads, "kind": "port",
            "sample": f"FP64 factor of the same operator family on a {side}^3 grid "
                      f"(n={prob.n}, F={b.flops:.3e}) in {dt:.1f} s"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--side", type=int, default=64)
    ap.add_argument("--cpu-side", type=int, default=40)
    ap.add_argument("--no-cpu", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path "
              "(sharding is described in DESIGN.md)", file=sys.stderr)
        sys.exit(2)

    import torch
    import swcholesky_amd as S
    from swcholesky_amd import problems as P

    log(f"building operator {args.side}^3")
    prob = P.laplacian3d_var(args.side)
    t0 = time.perf_counter()
    b = S.analyze_p2(prob.n, prob.colptr, prob.rowidx, prob.perm)
    t_an = time.perf_counter() - t0
    log(f"analysis {t_an:.2f}s nsuper={b.nsuper} xsize={b.xsize} F={b.flops:.4e}")
    f = S.Factor(b)
    stats = f.plan_stats()
    log(f"plan {stats}")
    dvals = torch.from_numpy(prob.values).cuda()
    torch.cuda.synchronize()

    for i in range(args.warmup):
        info = f.factor(dvals)
        assert info == 0, info
        log(f"warmup {i}: {f.timings()['factor_ms']:.2f} ms")
    torch.cuda.synchronize()
    ev_ms = []
    t0 = time.perf_counter()
    for i in range(args.steps):
        f.factor(dvals)
        ev_ms.append(f.timings()["factor_ms"])
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    ms_per_step = wall * 1e3 / args.steps
    log(f"steps: wall {ms_per_step:.2f} ms/step, device {np.median(ev_ms):.2f} ms")
    gflops = b.flops / (ms_per_step * 1e-3) / 1e9

    # FP64 time to 1e-12: factor + solve (+ refinement if needed), values on device
    A = prob.scipy_full()
    xs = np.random.default_rng(0).uniform(-1, 1, prob.n)
    rhs = A @ xs
    t0 = time.perf_counter()
    f.factor(dvals)
    x = f.solve(rhs)
    t_solve_path = time.perf_counter() - t0
    r = rhs - A @ x
    eta = float(np.abs(r).max() / (abs(A).sum(axis=1).max() * np.abs(x).max() + np.abs(rhs).max()))
    log(f"fp64 solve eta={eta:.2e} factor+solve {t_solve_path * 1e3:.1f} ms (host rhs)")

    # roofline of the dominant kernel from per-launch events
    f.set_profiling(True)
    f.factor(dvals)
    tm = f.timings()
    f.set_profiling(False)
    avg_ms = tm["tile_ms"] / max(tm["tile_launches"], 1)
    flops_per_launch = tm["tile_flops"] / max(tm["tile_launches"], 1)
    achieved = flops_per_launch / (avg_ms * 1e-3) / 1e12
    log(f"tile kernel: {tm['tile_launches']:.0f} launches, avg {avg_ms:.3f} ms, "
        f"{achieved:.2f} TFLOP/s; trsm {tm['trsm_ms']:.2f} ms")

    cpu = None
    if not args.no_cpu:
        nthreads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
        log(f"cpu baseline on {args.cpu_side}^3 with {nthreads} threads")
        cpu = cpu_baseline(args.cpu_side, nthreads)

    out = {
        "metric": "factor GFLOP/s (FP64, 64^3 variable-coefficient Laplacian, 1 MI355X)",
        "value": gflops,
        "unit": "GFLOP/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"lap3d_var_{args.side}: 7-point Laplacian {args.side}^3, "
                               "weights log-uniform [1e-2,1e2] seed 1, shift 1e-6, ND ordering",
                   "n": prob.n, "nnz_tril": prob.nnz, "nsuper": b.nsuper, "flops_F": b.flops,
                   "parallelism": "single GPU"},
        "fp64": {"factor_ms_device": float(np.median(ev_ms)), "eta": eta,
                 "factor_solve_ms_wall": t_solve_path * 1e3},
        "roofline": {"bound": "mfma", "kernel": "k_tile_update_factor", "achieved": achieved,
                     "peak": PEAK_FP64_TFLOPS, "unit": "TFLOP/s",
                     "frac": achieved / PEAK_FP64_TFLOPS, "traffic": None},
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
