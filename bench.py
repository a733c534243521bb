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
reference's CPU path (its BLAS-3 loop, restated) timed on the host cores.

    python bench.py [--steps K] [--warmup W] [--side 64] [--cpu-side 64]
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


def _cpu_model():
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(side, nthreads, runs=5):
    """The reference's CPU path, restated: the left-looking supernodal loop with the
    BLAS-3 calls the reference makes (dsyrk/dgemm per descendant, dpotrf/dtrsm per
    target, sw_parallel_PB_Cholesky_05.c:300-466; MKL as in its x86 build), OpenMP
    over the targets of a level with sequential BLAS, threaded BLAS on levels with
    fewer targets than threads (as parallel_PB_Cholesky_05.h:255-264 runs its last
    level).  One warm-up, then the median of `runs` factorizations of the full
    operator."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import swcholesky_amd as S
    from swcholesky_amd import problems as P
    import oracle_py as O
    prob = P.laplacian3d_var(side)
    b = S.analyze_p2(prob.n, prob.colptr, prob.rowidx, prob.perm)
    a2x = b.a2_values(prob.values)
    avail = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else os.cpu_count()
    if not O.blas_available():
        t0 = time.perf_counter()
        _, info = O.factor(b, a2x, nthreads=nthreads)
        dt = time.perf_counter() - t0
        assert info == 0
        return {"value": b.flops / dt / 1e9, "unit": "GFLOP/s", "cores": nthreads, "kind": "port",
                "sample": f"scalar restatement (no BLAS library found), one FP64 factor of "
                          f"{side}^3 (n={prob.n}) in {dt:.1f} s", "cpu": _cpu_model(),
                "nproc": avail}
    times = []
    for k in range(runs + 1):
        t0 = time.perf_counter()
        _, info = O.factor_blas(b, a2x, nthreads=nthreads)
        dt = time.perf_counter() - t0
        assert info == 0, info
        if k > 0:
            times.append(dt)
    med = float(np.median(times))
    return {"value": b.flops / med / 1e9, "unit": "GFLOP/s", "cores": nthreads, "kind": "port",
            "sample": f"BLAS-3 restatement of the reference loop (MKL {O.BLAS_PATH}), FP64 factor "
                      f"of the full {side}^3 operator (n={prob.n}, F={b.flops:.3e}): 1 warm-up, "
                      f"median of {runs} = {med:.2f} s (runs {', '.join(f'{t:.2f}' for t in times)})",
            "cpu": _cpu_model(), "nproc": avail}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--side", type=int, default=64)
    ap.add_argument("--cpu-side", type=int, default=64)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--ordering", default="grid", choices=["grid", "nd"],
                    help="grid: geometric nested dissection of the grid (the workload's "
                         "ordering); nd: the library's graph nested dissection (swc_nd_order)")
    ap.add_argument("--quick", action="store_true",
                    help="factor steps only (for counter runs): no refinement paths, no "
                         "per-launch event pass, no CPU baseline")
    ap.add_argument("--per-launch", action="store_true",
                    help="with --quick: factor in the per-launch timing mode (each level's "
                         "updates in one launch of their own, as the roofline pass times "
                         "them), so that counters per tile-update launch match it")
    ap.add_argument("--workload", default="lap3d",
                    choices=["lap3d", "vec3d", "batch", "gmrf", "schedules"],
                    help="lap3d: the headline 64^3 operator (default); vec3d: 48^3 x 3-dof "
                         "operator, mixed precision with 1/8/32 RHS; batch: 4096 uneven 2D "
                         "instances in one launch sequence; gmrf: 256^2 GMRF gradient; "
                         "schedules: wavefront vs the reference's H-level schedule")
    ap.add_argument("--copy-bytes", type=float, default=0,
                    help="also run one device-to-device copy of this many bytes (a plain "
                         "copy kernel of known size, to scale HBM counters)")
    ap.add_argument("--params", default=None,
                    help="plan shape parameters as comma-separated integers {rlb, spw, "
                         "narrow_min_tiles, narrow_max_k, fuse_narrow, dag}, dag = 0 (default: "
                         "the tuning table's entry for the structure)")
    args = ap.parse_args()
    args.plan = tuple(int(v) for v in args.params.split(",")) if args.params else "auto"
    if args.gpus != 1:
        print("bench.py measures one GPU: the library factors each matrix on one GPU "
              "and has no multi-GPU path", file=sys.stderr)
        sys.exit(2)

    import torch
    import swcholesky_amd as S
    from swcholesky_amd import problems as P

    if args.workload != "lap3d":
        return OTHER[args.workload](args)

    log(f"building operator {args.side}^3")
    prob = P.laplacian3d_var(args.side)
    if args.ordering == "nd":
        prob.perm = P.nd_order(prob.n, prob.colptr, prob.rowidx)
    t0 = time.perf_counter()
    b = S.analyze_p2(prob.n, prob.colptr, prob.rowidx, prob.perm)
    t_an = time.perf_counter() - t0
    log(f"analysis {t_an:.2f}s nsuper={b.nsuper} xsize={b.xsize} F={b.flops:.4e}")
    f = S.Factor(b, params=args.plan)
    stats = f.plan_stats()
    log(f"plan {stats}")
    dvals = torch.from_numpy(prob.values).cuda()
    torch.cuda.synchronize()
    if args.quick and args.per_launch:
        f.set_profiling(True)

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

    if args.copy_bytes:
        nbytes = int(args.copy_bytes) // 8 * 8
        src_t = torch.ones(nbytes // 8, dtype=torch.float64, device="cuda")
        dst_t = torch.empty_like(src_t)
        torch.cuda.synchronize()
        dst_t.copy_(src_t)
        torch.cuda.synchronize()
        log(f"copy kernel: {nbytes} bytes read, {nbytes} bytes written")
    if args.quick:
        print(json.dumps({"metric": "factor GFLOP/s (quick counter run)", "value": gflops,
                          "unit": "GFLOP/s", "ms_per_step": ms_per_step}), flush=True)
        return

    # time to 1e-12 backward error: factor + solve + FP64 refinement, with the
    # values and the right-hand side resident on the device
    A = prob.scipy_full()
    rhs = A @ np.random.default_rng(0).uniform(-1, 1, prob.n)
    drhs = torch.from_numpy(rhs).cuda()
    paths = {}
    for name, mode in (("fp64", 0), ("mixed", 1), ("mixed_all", 2)):
        f.set_precision(mode)
        best = None
        for rep in range(2):
            x = drhs.clone()
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            info = f.factor(dvals)
            _, eta, iters = f.refine(x, tol=1e-12, maxiter=30)
            torch.cuda.synchronize()
            dt = (time.perf_counter() - t0) * 1e3
            tm = f.timings()
            rec = {"ms": dt, "factor_ms_device": tm["factor_ms"], "eta": float(eta[0]),
                   "iters": iters, "info": info}
            if best is None or dt < best["ms"]:
                best = rec
        # the parts: one solve with its residual evaluation (no correction), and
        # the cost of each correction (a residual, a solve and the reductions)
        x0 = drhs.clone()
        f.refine(x0, tol=0.0, maxiter=0)  # captures the solve sequence for this buffer
        x0.copy_(drhs)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        f.refine(x0, tol=0.0, maxiter=0)
        torch.cuda.synchronize()
        best["solve_and_residual_ms"] = (time.perf_counter() - t0) * 1e3
        refine_ms = best["ms"] - best["factor_ms_device"]
        best["per_correction_ms"] = ((refine_ms - best["solve_and_residual_ms"]) / best["iters"]
                                     if best["iters"] else 0.0)
        xh = x.cpu().numpy()
        r = rhs - A @ xh
        best["eta_host_check"] = float(np.abs(r).max() / (abs(A).sum(axis=1).max() * np.abs(xh).max()
                                                          + np.abs(rhs).max()))
        best["factor_gflops"] = b.flops / (best["factor_ms_device"] * 1e-3) / 1e9
        paths[name] = best
        log(f"{name}: time-to-1e-12 {best['ms']:.1f} ms (factor {best['factor_ms_device']:.1f} ms, "
            f"solve + residual {best['solve_and_residual_ms']:.1f} ms, {best['iters']} corrections "
            f"of {best['per_correction_ms']:.1f} ms, eta {best['eta']:.2e} / host "
            f"{best['eta_host_check']:.2e})")
    f.set_precision(0)

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
    # the largest fronts: the levels with the most update work (the external
    # updates of the top separators), each one tile-kernel launch
    lv = f.level_timings()
    top = np.argsort(-lv[:, 1])[:3]
    large = [{"level": int(l), "ms": float(lv[l, 0]), "gflop": float(lv[l, 1] / 1e9),
              "tflops": float(lv[l, 1] / (lv[l, 0] * 1e-3) / 1e12),
              "frac_fp64_peak": float(lv[l, 1] / (lv[l, 0] * 1e-3) / 1e12 / PEAK_FP64_TFLOPS)}
             for l in top if lv[l, 0] > 0]
    for r in large:
        log(f"large front: level {r['level']} {r['gflop']:.1f} GFLOP in {r['ms']:.2f} ms = "
            f"{r['tflops']:.1f} TFLOP/s ({100 * r['frac_fp64_peak']:.0f}% of FP64 MFMA peak)")
    traffic, traffic_src = None, None
    pmc = os.path.join(ROOT, "profiles", "k_tile_update_lap3d64_pmc.json")
    if args.side == 64 and os.path.exists(pmc):
        with open(pmc) as fh:
            rec = json.load(fh)
        traffic = rec["hbm_bytes_per_launch"]
        traffic_src = os.path.relpath(pmc, ROOT)

    cpu = None
    if not args.no_cpu:
        nthreads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
        log(f"cpu baseline on {args.cpu_side}^3 with {nthreads} threads")
        cpu = cpu_baseline(args.cpu_side, nthreads)

    out = {
        "metric": "factor GFLOP/s & time-to-1e-12 backward error, mixed vs FP64, 1 MI355X",
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
                               "weights log-uniform [1e-2,1e2] seed 1, shift 1e-6, "
                               + ("geometric grid ND ordering" if args.ordering == "grid"
                                  else "graph ND ordering (swc_nd_order)"),
                   "n": prob.n, "nnz_tril": prob.nnz, "nsuper": b.nsuper, "flops_F": b.flops,
                   "plan_params": dict(zip(("rlb", "spw", "narrow_min_tiles", "narrow_max_k",
                                            "fuse_narrow", "dag"), f.params or ())),
                   "parallelism": "single GPU"},
        "time_to_1e-12_ms": {k: v["ms"] for k, v in paths.items()},
        "mixed_speedup_vs_fp64": paths["fp64"]["ms"] / min(paths["mixed"]["ms"], paths["mixed_all"]["ms"]),
        "paths": paths,
        "roofline": {"bound": "mfma", "kernel": "k_tile_update", "achieved": achieved,
                     "peak": PEAK_FP64_TFLOPS, "unit": "TFLOP/s",
                     "frac": achieved / PEAK_FP64_TFLOPS, "traffic": traffic,
                     "traffic_source": traffic_src,
                     "hbm_gbs": (traffic / (avg_ms * 1e-3) / 1e9) if traffic else None},
        "large_fronts": large,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


def _median_factor(f, dvals, runs=5):
    """Median wall time (ms) of `runs` synchronised factorizations after one warm-up (which
    captures the launch sequence); returns (median, every run, info of the last)."""
    import torch
    f.factor(dvals)
    torch.cuda.synchronize()
    ts = []
    info = 0
    for _ in range(runs):
        t0 = time.perf_counter()
        info = f.factor(dvals)
        torch.cuda.synchronize()
        ts.append((time.perf_counter() - t0) * 1e3)
    return float(np.median(ts)), [round(t, 3) for t in ts], info


def _time_refine(f, drhs, tol=1e-12):
    import torch
    x = drhs.clone()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    _, eta, iters = f.refine(x, tol=tol, maxiter=40)
    torch.cuda.synchronize()
    return (time.perf_counter() - t0) * 1e3, eta, iters, x


def bench_vec3d(args):
    """48^3 x 3-dof operator (n = 331,776): FP64 and mixed factor, then refinement
    to 1e-12 for 1, 8 and 32 right-hand sides ~ U(-1, 1)."""
    import torch
    import swcholesky_amd as S
    from swcholesky_amd import problems as P
    side = 48 if args.side == 64 else args.side
    prob = P.vector3d_var(side)
    b = S.analyze_p2(prob.n, prob.colptr, prob.rowidx, prob.perm)
    log(f"vec3d {side}^3x3: n={prob.n} nsuper={b.nsuper} xsize={b.xsize} F={b.flops:.4e}")
    f = S.Factor(b, params=args.plan)
    dvals = torch.from_numpy(prob.values).cuda()
    rng = np.random.default_rng(7)
    res = {}
    for name, mode in (("fp64", 0), ("mixed", 1), ("mixed_all", 2)):
        f.set_precision(mode)
        tf, runs, info = _median_factor(f, dvals)
        rec = {"factor_ms": tf, "factor_runs_ms": runs, "factor_gflops": b.flops / tf / 1e6,
               "info": info}
        for nrhs in (1, 8, 32):
            B = torch.from_numpy(np.ascontiguousarray(rng.uniform(-1, 1, (nrhs, prob.n)))).cuda()
            _time_refine(f, B)  # warm-up of this shape
            ms, eta, iters, _ = _time_refine(f, B)
            rec[f"refine_{nrhs}rhs"] = {"ms": ms, "eta_max": float(np.max(eta)), "iters": iters}
        res[name] = rec
        log(f"{name}: {rec}")
    out = {"metric": "factor GFLOP/s & time-to-1e-12 backward error, mixed vs FP64, 1 MI355X",
           "workload": f"vec3d_var_{side}: 3-dof couplings w(I+uu^T/2), n={prob.n}",
           "value": res["fp64"]["factor_gflops"], "unit": "GFLOP/s", "dtype": "f64",
           "flops_F": b.flops, "paths": res}
    print(json.dumps(out), flush=True)


def bench_batch(args):
    """4096 independent 2D 5-point variable-coefficient Laplacians, side ~ U{16..128},
    weights log-uniform [1e-1, 1e1] (seed 3), shift 1e-4: factored together
    (one launch per elimination-tree level), then one RHS each refined to 1e-12."""
    import torch
    from swcholesky_amd import problems as P
    from swcholesky_amd.batch import Batch
    count = 4096 if args.side == 64 else args.side
    rng = np.random.default_rng(3)
    sides = rng.integers(16, 129, size=count)
    t0 = time.perf_counter()
    probs = [P.laplacian2d_var(int(sd), seed=None, rng=rng, wlo=1e-1, whi=1e1, shift=1e-4)
             for sd in sides]
    bt = Batch(probs)
    log(f"batch of {count}: n={bt.n} nsuper={bt.bcsc.nsuper} F={bt.bcsc.flops:.4e} "
        f"host prep {time.perf_counter() - t0:.1f} s")
    t0 = time.perf_counter()
    f = bt.factor
    log(f"plan and upload {time.perf_counter() - t0:.1f} s: {f.plan_stats()}")
    dvals = torch.from_numpy(bt.values(probs)).cuda()
    rhs = torch.from_numpy(np.random.default_rng(0).uniform(-1, 1, bt.n)).cuda()
    res = {}
    for name, mode in (("fp64", 0), ("mixed", 1)):
        f.set_precision(mode)
        tf, runs, info = _median_factor(f, dvals)
        _time_refine(f, rhs)
        ms, eta, iters, _ = _time_refine(f, rhs)
        res[name] = {"factor_ms": tf, "factor_runs_ms": runs,
                     "factor_gflops": bt.bcsc.flops / tf / 1e6, "info": info,
                     "refine_ms": ms, "eta": float(eta[0]), "iters": iters,
                     "time_to_1e-12_ms": tf + ms}
        log(f"{name}: {res[name]}")
    print(json.dumps({"metric": "factor GFLOP/s & time-to-1e-12 backward error, mixed vs FP64",
                      "workload": f"{count} uneven 2D instances", "n": bt.n,
                      "value": res["fp64"]["factor_gflops"], "unit": "GFLOP/s",
                      "flops_F": bt.bcsc.flops, "paths": res}), flush=True)


def bench_gmrf(args):
    """GMRF Q = tau(kappa^2 I + L_grid) on 256^2: factor, logdet + b^T Q^-1 b and its
    gradient w.r.t. Q's stored values (supernodal selected inverse)."""
    import torch
    import swcholesky_amd as S
    from swcholesky_amd import problems as P
    from swcholesky_amd.autograd import logdet_quad
    side = 256 if args.side == 64 else args.side
    prob = P.gmrf2d(side)
    b = S.analyze_p2(prob.n, prob.colptr, prob.rowidx, prob.perm)
    f = S.Factor(b, params=args.plan)
    vals = torch.tensor(prob.values, device="cuda", requires_grad=True)
    rhs = torch.tensor(np.random.default_rng(4).standard_normal(prob.n), device="cuda")
    fw, bw = [], []
    for rep in range(7):  # two warm-ups (graph captures, selected-inverse set-up), median of 5
        vals.grad = None
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        out = logdet_quad(vals, rhs, f)
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        out.backward()
        torch.cuda.synchronize()
        t2 = time.perf_counter()
        if rep >= 2:
            fw.append((t1 - t0) * 1e3)
            bw.append((t2 - t1) * 1e3)
    res = {"forward_ms": float(np.median(fw)), "backward_ms": float(np.median(bw)),
           "backward_runs_ms": [round(v, 3) for v in bw], "loss": float(out.detach()),
           "n": prob.n, "flops_F": b.flops}
    log(f"gmrf: {res}")
    print(json.dumps({"metric": "gradient of logdet(Q) + b^T Q^-1 b (256^2 GMRF)",
                      "workload": f"gmrf2d_{side}", "value": res["backward_ms"], "unit": "ms",
                      "higher_is_better": False, "details": res}), flush=True)


def bench_schedules(args):
    """The wavefront schedule (factor levels from the descendant sets, as the
    level sets of the supernodal elimination tree) against the reference's
    H-level schedule (its own clusters and partitions from analyze_p2, stored in
    tests/golden/full_*.npz): FP64 factor and one-RHS solve on 64^3 and on the
    48^3 x 3 vector operator."""
    import torch
    import swcholesky_amd as S
    from swcholesky_amd import problems as P
    res = {}
    for name, make in (("lap3d_var_64", lambda: P.laplacian3d_var(64)),
                       ("vec3d_var_48", lambda: P.vector3d_var(48))):
        g = np.load(os.path.join(ROOT, "tests", "golden", f"full_{name}.npz"))
        prob = make()
        b = S.analyze_p2(prob.n, prob.colptr, prob.rowidx, prob.perm)
        dvals = torch.from_numpy(prob.values).cuda()
        drhs = torch.from_numpy(np.random.default_rng(0).uniform(-1, 1, prob.n)).cuda()
        for sched_name, sched in (("wavefront", None),
                                  ("hlevel", (g["levelptr"], g["parptr"], g["partition"]))):
            f = S.Factor(b, schedule=sched)
            assert f.factor(dvals) == 0
            x = drhs.clone()
            f.solve_device(x)
            tf, ts = [], []
            for _ in range(5):
                torch.cuda.synchronize()
                t0 = time.perf_counter()
                f.factor(dvals)
                torch.cuda.synchronize()
                t1 = time.perf_counter()
                f.solve_device(x)
                torch.cuda.synchronize()
                tf.append((t1 - t0) * 1e3)
                ts.append((time.perf_counter() - t1) * 1e3)
            rec = {"levels": f.plan_stats()["levels"], "factor_ms": float(np.median(tf)),
                   "solve_ms": float(np.median(ts)),
                   "factor_gflops": b.flops / (float(np.median(tf)) * 1e-3) / 1e9}
            res[f"{name}/{sched_name}"] = rec
            log(f"{name} {sched_name}: {rec}")
            del f
    print(json.dumps({"metric": "FP64 factor time, wavefront vs H-level schedule", "unit": "ms",
                      "results": res}), flush=True)


OTHER = {"vec3d": bench_vec3d, "batch": bench_batch, "gmrf": bench_gmrf,
         "schedules": bench_schedules}

if __name__ == "__main__":
    main()
