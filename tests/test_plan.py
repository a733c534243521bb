"""The device plan and the kernels' index arithmetic, emulated on the CPU."""
import numpy as np
import pytest

from swcholesky_amd import api, problems as P
import emulate as E
import oracle_py as O
from helpers import relF, setup


def _amap(b):
    amap = np.zeros(len(b.a2i), np.int64)
    for j in range(b.n):
        s = b.col2Sup[j]
        rows = b.s[b.pi[s]:b.pi[s + 1]]
        q0, q1 = b.a2p[j], b.a2p[j + 1]
        amap[q0:q1] = b.p[b.super[s]] + (j - b.super[s]) * len(rows) + np.searchsorted(
            rows, b.a2i[q0:q1])
    return amap


@pytest.mark.parametrize("make", [lambda: P.laplacian2d_const(10), lambda: P.laplacian2d_var(37, 3),
                                  lambda: P.laplacian3d_var(12), lambda: P.vector3d_var(6),
                                  lambda: P.laplacian2d_const(100), lambda: P.vector3d_var(9)])
def test_emulated_kernels_match_oracle(make):
    prob = make()
    b, a2x, Lo, _ = setup(prob)
    if b.n >= 2000:  # these cases have supernodes wider than one panel
        assert b.maxSupWid > 64
    plan = api.build_plan(b)
    Le, info = E.factor(plan, b.s, b.xsize, _amap(b), a2x)
    assert info == 0
    assert relF(Le, Lo) < 1e-12
    y = np.random.default_rng(1).uniform(-1, 1, b.n)
    xe = E.solve(plan, b.s, Le, y)
    xo = O.ltsolve(b, Lo, O.lsolve(b, Lo, y))[:, 0]
    assert relF(xe, xo) < 1e-6


def test_plan_covers_every_update_once():
    """Each (target entry, source) update appears in exactly one task."""
    prob = P.laplacian3d_var(16)
    b, *_ = setup(prob)
    plan = api.build_plan(b)
    panels, tiles, tasks = plan["panels"], plan["tiles"], plan["tasks"]
    seen = {}
    for tl in tiles:
        pn = panels[tl["panel"]]
        for t in tasks[tl["task_begin"]:tl["task_end"]]:
            rows = b.s[t["rows_off"] + t["ra"]:t["rows_off"] + t["rb"]]
            cols = b.s[t["rows_off"] + t["ca"]:t["rows_off"] + t["cb"]]
            key = (int(t["src_base"]), int(t["k"]))
            for r in rows:
                for c in cols:
                    if r >= c:
                        seen.setdefault((int(r), int(c), key), 0)
                        seen[(int(r), int(c), key)] += 1
    assert all(v == 1 for v in seen.values())
    # the forward-solve panel order is sorted by level
    lp = plan["lvl_panels"]
    assert lp[0] == 0 and lp[-1] == len(panels) and np.all(np.diff(lp) >= 0)


@pytest.mark.parametrize("make", [lambda: P.laplacian3d_var(20), lambda: P.vector3d_var(10)])
def test_emulated_solve_wide_supernodes(make):
    """Solve panels of several 256-column chunks (supernodes wider than 256)."""
    prob = make()
    b, a2x, Lo, _ = setup(prob)
    assert b.maxSupWid > 256
    plan = api.build_plan(b)
    assert plan["nslevel"] < plan["nlevel"]
    y = np.random.default_rng(2).uniform(-1, 1, b.n)
    xe = E.solve(plan, b.s, Lo, y)
    xo = O.ltsolve(b, Lo, O.lsolve(b, Lo, y))[:, 0]
    assert relF(xe, xo) < 1e-12
