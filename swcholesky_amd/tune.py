"""Autotuner of the plan shape parameters (run on an MI355X).

    python -m swcholesky_amd.tune [--workloads lap3d64,vec48,gmrf256,batch]
                                  [--reps 3] [--out swcholesky_amd/tuning.json]

For each workload it searches {rlb, spw, narrow_min_tiles, narrow_max_k, fuse_narrow, dag}
(swc_set_plan_params) by coordinate descent from the library defaults: one
parameter at a time over its candidate values, keeping the best, until a full
pass changes nothing.  The objective is the measured FP64 time to a solution,
the median over `reps` runs of one factorization plus one single right-hand-side
solve, with values and right-hand side resident on the device (the factor and
solve launch sequences are captured once per structure, before timing).  The
table records, per workload, the structure's class (order, widest supernode),
the chosen parameters, their times, the defaults' times and every trial; the
library looks entries up by class (swcholesky_amd/tuning.py).
"""
from __future__ import annotations

import argparse
import gc
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))

CANDIDATES = {
    "rlb": [2, 4, 8, 12, 16, 24],
    "spw": [64, 128],
    "narrow_min_tiles": [256, 1024, 4096, 1 << 30],
    "narrow_max_k": [128, 256, 512],
    "fuse_narrow": [0, 1],
    "dag": [0],  # the dependency-driven launch is disabled
}


def log(msg):
    print(f"[tune] {msg}", flush=True)


def _problem(name):
    from swcholesky_amd import problems as P
    if name == "lap3d64":
        return [P.laplacian3d_var(64)]
    if name == "vec48":
        return [P.vector3d_var(48)]
    if name == "gmrf256":
        return [P.gmrf2d(256)]
    if name.startswith("batch"):
        count = int(name[5:] or 4096)
        rng = np.random.default_rng(3)
        sides = rng.integers(16, 129, size=count)
        return [P.laplacian2d_var(int(sd), seed=None, rng=rng, wlo=1e-1, whi=1e1, shift=1e-4)
                for sd in sides]
    raise ValueError(name)


def _measure(bcsc, dvals, drhs, params, reps):
    import torch
    import swcholesky_amd as S
    f = S.Factor(bcsc, params=params)
    assert f.factor(dvals) == 0  # captures the factor sequence
    x = drhs.clone()
    f.solve_device(x)  # captures the solve sequence
    tf, ts = [], []
    for _ in range(reps):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        f.factor(dvals)
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        x.copy_(drhs)
        torch.cuda.synchronize()
        t2 = time.perf_counter()
        f.solve_device(x)
        torch.cuda.synchronize()
        t3 = time.perf_counter()
        tf.append((t1 - t0) * 1e3)
        ts.append((t3 - t2) * 1e3)
    del f
    gc.collect()
    return float(np.median(tf)), float(np.median(ts))


def tune_workload(name, reps):
    import torch
    import swcholesky_amd as S
    from swcholesky_amd.batch import Batch
    from swcholesky_amd.tuning import DEFAULTS, NAMES, structure_class
    probs = _problem(name)
    if len(probs) == 1:
        p = probs[0]
        bcsc = S.analyze_p2(p.n, p.colptr, p.rowidx, p.perm)
        vals = p.values
    else:
        bt = Batch(probs)
        bcsc, vals = bt.bcsc, bt.values(probs)
    dvals = torch.from_numpy(np.ascontiguousarray(vals)).cuda()
    drhs = torch.from_numpy(np.random.default_rng(0).uniform(-1, 1, bcsc.n)).cuda()
    cls = structure_class(bcsc)
    log(f"{name}: n={bcsc.n} nsuper={bcsc.nsuper} widest supernode {bcsc.maxSupWid}")
    trials = {}

    def run(params):
        if params not in trials:
            tf, ts = _measure(bcsc, dvals, drhs, params, reps)
            trials[params] = (tf, ts)
            log(f"{name} {dict(zip(NAMES, params))}: factor {tf:.2f} ms, solve {ts:.2f} ms")
        tf, ts = trials[params]
        return tf + ts

    # batches of small instances: no supernode is much wider than a panel, so the
    # right-looking block size does not matter there (and each trial re-plans
    # millions of columns)
    keys = NAMES[1:] if len(probs) > 1 else NAMES
    best = tuple(DEFAULTS)
    best_t = run(best)
    changed = True
    while changed:
        changed = False
        for i, key in enumerate(NAMES):
            if key not in keys:
                continue
            for v in CANDIDATES[key]:
                cand = best[:i] + (v,) + best[i + 1:]
                t = run(cand)
                if t < best_t * 0.995:  # ignore differences within timing noise
                    best, best_t, changed = cand, t, True
    tf, ts = trials[best]
    dtf, dts = trials[tuple(DEFAULTS)]
    log(f"{name}: chose {dict(zip(NAMES, best))}: {tf + ts:.2f} ms (defaults {dtf + dts:.2f} ms)")
    return {"workload": name, "class": cls, "params": dict(zip(NAMES, best)),
            "factor_ms": tf, "solve_ms": ts,
            "default": {"params": dict(zip(NAMES, DEFAULTS)), "factor_ms": dtf, "solve_ms": dts},
            "trials": [{"params": dict(zip(NAMES, p)), "factor_ms": a, "solve_ms": b}
                       for p, (a, b) in trials.items()]}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--workloads", default="lap3d64,vec48,gmrf256,batch")
    ap.add_argument("--reps", type=int, default=3)
    ap.add_argument("--out", default=os.path.join(HERE, "tuning.json"))
    args = ap.parse_args()
    import torch
    entries = [tune_workload(w, args.reps) for w in args.workloads.split(",")]
    table = {"generated_by": "python -m swcholesky_amd.tune " + " ".join(sys.argv[1:]),
             "device": torch.cuda.get_device_name(0),
             "objective": "median FP64 factor + one-RHS solve time (ms), values resident on the device",
             "entries": entries}
    with open(args.out, "w") as fh:
        json.dump(table, fh, indent=1)
    log(f"wrote {args.out}")


if __name__ == "__main__":
    main()
