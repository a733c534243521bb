"""Writes the golden fixtures of the parity tests (run here, where the reference is).

    python tests/golden/make_golden.py

For each case, runs the reference's own host analysis (oracle/_ref/ref_analyze,
built by `make -C oracle`) on a generated operator and stores its outputs: the
final permutation, column counts, supernode partition, supernodal row lists and
pointers, pruned descendant sets and A2 = tril(P A P^T).  For config 1 (2D 100^2,
diag 4+1e-3, off-diagonal -1) it also stores per-supernode Frobenius norms of the
oracle's FP64 factor and the oracle solution x of A x = A x*, x* ~ U(-1,1) seed 0.
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.dirname(HERE))

from swcholesky_amd import problems as P  # noqa: E402
import refio  # noqa: E402

CASES = {
    "c1_lap2d_const_100": lambda: P.laplacian2d_const(100),
    "lap2d_var_37": lambda: P.laplacian2d_var(37, 3),
    "lap3d_var_12": lambda: P.laplacian3d_var(12),
    "vec3d_var_6": lambda: P.vector3d_var(6),
    "gmrf2d_32": lambda: P.gmrf2d(32),
}
KEYS = ["header", "perm", "colcount", "super", "sparent", "col2sup", "pi", "s", "i_ptr", "p",
        "pruneptr", "pruneset", "levelptr", "parptr", "partition", "a2_p", "a2_i", "a2_x"]


# Full-size operators: the reference's H-level schedule (the clusters and
# partitions of getCoarseLevelSet_6) and SHA-256 digests of its analysis arrays,
# which pin the restated analysis at full size without storing the arrays.
FULL = {
    "lap3d_var_64": lambda: P.laplacian3d_var(64),
    "vec3d_var_48": lambda: P.vector3d_var(48),
}
DIGEST_KEYS = ["perm", "colcount", "super", "s", "pruneptr", "pruneset", "p", "i_ptr"]


def digest(a):
    import hashlib
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def main():
    if not refio.have_ref():
        sys.exit("build the reference harness first: make -C oracle ref")
    for name, make in FULL.items():
        prob = make()
        r = refio.run_ref(prob)
        out = {k: r[k] for k in ["header", "levelptr", "parptr", "partition"]}
        out["digest_keys"] = np.array(DIGEST_KEYS)
        out["digests"] = np.array([digest(r[k]) for k in DIGEST_KEYS])
        out["dtypes"] = np.array([str(r[k].dtype) for k in DIGEST_KEYS])
        np.savez_compressed(os.path.join(HERE, "full_" + name + ".npz"), **out)
        print("full_" + name, prob.n, {k: v.shape for k, v in out.items()})
    for name, make in CASES.items():
        prob = make()
        r = refio.run_ref(prob)
        out = {k: r[k] for k in KEYS}
        if name.startswith("c1_"):
            import swcholesky_amd as S
            import oracle_py as O
            b = S.analyze_p2(prob.n, prob.colptr, prob.rowidx, prob.perm)
            L, info = O.factor(b, b.a2_values(prob.values))
            assert info == 0
            norms = np.zeros(b.nsuper)
            for s in range(b.nsuper):
                c0, c1 = b.super[s], b.super[s + 1]
                R = b.pi[s + 1] - b.pi[s]
                norms[s] = np.linalg.norm(L[b.p[c0]:b.p[c0] + R * (c1 - c0)])
            xs = np.random.default_rng(0).uniform(-1, 1, prob.n)
            rhs = prob.scipy_full() @ xs
            out["oracle_snorm"] = norms
            out["xstar"] = xs
            out["b"] = rhs
            out["oracle_x"] = O.solve(b, L, rhs)
        np.savez_compressed(os.path.join(HERE, name + ".npz"), **out)
        print(name, prob.n, {k: v.shape for k, v in out.items()})


if __name__ == "__main__":
    main()
