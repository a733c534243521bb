"""The device plan and the kernels' index arithmetic, emulated on the CPU."""
import numpy as np
import pytest

import swcholesky_amd as S
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
                                  lambda: P.laplacian2d_const(100), lambda: P.vector3d_var(9),
                                  lambda: P.dense_spd(1100)])
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


def _expected_sources(b):
    """For every pair (r, c), r >= c, of rows of L: the number of columns m < c
    whose structure holds both r and c, and the sums of m and m^2 over them,
    leaving out the columns of c's own panel (the panel factorization applies
    those)."""
    n = b.n
    cnt, s1, s2 = (np.zeros((n, n)) for _ in range(3))
    for d in range(b.nsuper):
        R = b.s[b.pi[d]:b.pi[d + 1]]
        W = b.super[d + 1] - b.super[d]
        i = np.arange(len(R))
        # columns of d before each row's column, not counting that column's panel
        m = np.where(i < W, 64 * (i // 64), W)
        m0 = b.super[d]
        ir, ic = np.tril_indices(len(R))
        r, c, k = R[ir], R[ic], m[ic]
        cnt[r, c] += k
        s1[r, c] += k * m0 + k * (k - 1) / 2
        s2[r, c] += sum((m0 + j) ** 2 * (k > j) for j in range(W))
    return cnt, s1, s2


@pytest.mark.parametrize("make", [lambda: P.vector3d_var(9), lambda: P.laplacian3d_var(12),
                                  lambda: P.dense_spd(1100)])
def test_plan_covers_every_update_once(make):
    """Every source column m of every update L(r,c) -= L(r,m) L(c,m) is applied
    by exactly one task: per target pair, the count of source columns and the
    sums of m and m^2 over the tasks' inner dimensions equal the structural
    ones.  The dense case is one supernode of 18 panels, whose right-looking
    steps apply blocks of panels as one source."""
    prob = make()
    b, *_ = setup(prob)
    plan = api.build_plan(b)
    panels, tiles, tasks = plan["panels"], plan["tiles"], plan["tasks"]
    n = b.n
    cnt, s1, s2 = (np.zeros((n, n)) for _ in range(3))
    wide = 0
    for tl in tiles:
        pn = panels[tl["panel"]]
        for t in tasks[tl["task_begin"]:tl["task_end"]]:
            rows = b.s[t["rows_off"] + t["ra"]:t["rows_off"] + t["rb"]]
            cols = b.s[t["rows_off"] + t["ca"]:t["rows_off"] + t["cb"]]
            assert np.all(cols >= pn["c0"]) and np.all(cols < pn["c0"] + pn["ncol"])
            # first global column of the source: the supernode whose rows the task indexes
            d = int(np.searchsorted(b.pi, t["rows_off"], side="right") - 1)
            assert b.pi[d] == t["rows_off"]
            ld = b.pi[d + 1] - b.pi[d]
            m0 = b.super[d] + (int(t["src_base"]) - b.p[b.super[d]]) // ld
            k = int(t["k"])
            assert m0 + k <= b.super[d + 1]
            wide += k > 64 and d == pn["sn"]  # a block of panels of the same supernode
            rr, cc = np.meshgrid(rows, cols, indexing="ij")
            keep = rr >= cc
            r, c = rr[keep], cc[keep]
            np.add.at(cnt, (r, c), k)
            np.add.at(s1, (r, c), k * m0 + k * (k - 1) / 2)
            np.add.at(s2, (r, c), sum((m0 + j) ** 2 for j in range(k)))
    ec, e1, e2 = _expected_sources(b)
    assert np.array_equal(cnt, ec) and np.array_equal(s1, e1) and np.array_equal(s2, e2)
    if b.maxSupWid > 16 * 64:
        assert wide > 0
    # the forward-solve panel order is sorted by level
    lp = plan["lvl_panels"]
    assert lp[0] == 0 and lp[-1] == len(panels) and np.all(np.diff(lp) >= 0)


@pytest.mark.parametrize("make", [lambda: P.laplacian3d_var(20), lambda: P.vector3d_var(10)])
def test_emulated_solve_wide_supernodes(make):
    """Supernodes wider than a solve panel (SPW columns) solved in several panels."""
    prob = make()
    b, a2x, Lo, _ = setup(prob)
    assert b.maxSupWid > 256
    plan = api.build_plan(b)
    assert plan["nslevel"] < plan["nlevel"]
    y = np.random.default_rng(2).uniform(-1, 1, b.n)
    xe = E.solve(plan, b.s, Lo, y)
    xo = O.ltsolve(b, Lo, O.lsolve(b, Lo, y))[:, 0]
    assert relF(xe, xo) < 1e-12


@pytest.mark.parametrize("name", ["c1_lap2d_const_100", "gmrf2d_32", "lap3d_var_12", "vec3d_var_6"])
def test_derived_descendant_sets_give_the_reference_solve_schedule(name):
    """blockedLsolve gets no descendant sets and derives them from the row lists.
    With the reference analysis's own structure (golden fixtures written by its
    analyze_p2), the solve schedule built from the derived sets equals the one built
    from the reference's pruneSet.  The derived plan is the solve-only plan
    blockedLsolve uploads (no factor levels), and a solve emulated on it matches
    the oracle's."""
    import os
    g = np.load(os.path.join(os.path.dirname(__file__), "golden", name + ".npz"))
    prob = {"c1_lap2d_const_100": lambda: P.laplacian2d_const(100),
            "gmrf2d_32": lambda: P.gmrf2d(32), "lap3d_var_12": lambda: P.laplacian3d_var(12),
            "vec3d_var_6": lambda: P.vector3d_var(6)}[name]()
    b, *_ = setup(prob)
    assert np.array_equal(g["pruneset"], b.pruneSet) and np.array_equal(g["s"], b.s)
    ref = api.build_plan(b)
    der = api.build_plan(b, derive_sets=True)
    for key in ["spanels", "sp_order", "lvl_sp", "fblocks", "lvl_fblk", "gptr", "gidx", "gcols",
                "lvl_gcol"]:
        assert np.array_equal(ref[key], der[key]), key
    assert ref["nslevel"] == der["nslevel"]
    assert der["nlevel"] == 0 and len(der["tiles"]) == 0 and len(der["tasks"]) == 0
    _, a2x, Lo, _ = setup(prob)
    y = np.random.default_rng(2).uniform(-1, 1, b.n)
    xe = E.solve(der, b.s, Lo, y)
    xo = O.ltsolve(b, Lo, O.lsolve(b, Lo, y))[:, 0]
    assert relF(xe, xo) < 1e-12


@pytest.mark.parametrize("params", [(2, 128, 1024, 256), (16, 64, 1024, 256), (8, 64, 1, 512),
                                    (1, 128, 1, 0), (64, 64, 4096, 128), (8, 128, 1, 256, 1),
                                    (8, 128, 1, 512, 0)])
@pytest.mark.parametrize("make", [lambda: P.laplacian3d_var(12), lambda: P.dense_spd(300),
                                  lambda: P.laplacian2d_var(37, 3)])
def test_plan_shape_parameters_keep_parity(params, make):
    """Every plan shape the autotuner may pick (right-looking block size, solve
    panel width, one-wave tile thresholds) factors and solves to the oracle's
    result on the emulated kernels."""
    prob = make()
    b, a2x, Lo, _ = setup(prob)
    plan = api.build_plan(b, params=params)
    Le, info = E.factor(plan, b.s, b.xsize, _amap(b), a2x)
    assert info == 0
    assert relF(Le, Lo) < 1e-12
    spw = params[1]
    assert plan["spanels"]["ncol"].max() <= spw
    y = np.random.default_rng(3).uniform(-1, 1, b.n)
    xe = E.solve(plan, b.s, Le, y)
    xo = O.ltsolve(b, Lo, O.lsolve(b, Lo, y))[:, 0]
    assert relF(xe, xo) < 1e-8  # forward error of a cond ~1e8 operator


def test_plan_shape_parameters_are_checked_and_looked_up():
    from swcholesky_amd import tuning
    prob = P.laplacian2d_var(37, 3)
    b, *_ = setup(prob)
    for bad in [(0,), (8, 96), (8, 128, 0), (8, 128, 16, -1)]:
        with pytest.raises(S.SwcError):
            api.build_plan(b, params=bad)
    entries = [{"class": {"n": 262144, "max_width": 6593}, "params": dict(zip(tuning.NAMES, (4, 64, 256, 128)))},
               {"class": {"n": 1369, "max_width": 40}, "params": dict(zip(tuning.NAMES, (16, 128, 4096, 512)))}]
    assert tuning.params_for(b, entries) == (16, 128, 4096, 512, 1, 0)  # missing keys: defaults
    assert tuning.params_for(b, []) == tuning.DEFAULTS


@pytest.mark.parametrize("name", ["c1_lap2d_const_100", "gmrf2d_32", "lap3d_var_12", "vec3d_var_6"])
def test_hlevel_schedule_plan(name):
    """The reference's H-level schedule (its own clusters and partitions, golden
    fixtures from its analyze_p2) as factor levels: clusters one after another,
    each partition's supernodes in order.  The emulated kernels on that plan give
    the oracle's factor, every supernode starts after its descendants and after
    its partition predecessor, and the wavefront plan never has more levels."""
    import os
    g = np.load(os.path.join(os.path.dirname(__file__), "golden", name + ".npz"))
    prob = {"c1_lap2d_const_100": lambda: P.laplacian2d_const(100),
            "gmrf2d_32": lambda: P.gmrf2d(32), "lap3d_var_12": lambda: P.laplacian3d_var(12),
            "vec3d_var_6": lambda: P.vector3d_var(6)}[name]()
    b, a2x, Lo, _ = setup(prob)
    sched = (g["levelptr"], g["parptr"], g["partition"])
    hl = api.build_plan(b, schedule=sched)
    wf = api.build_plan(b)
    assert wf["nlevel"] <= hl["nlevel"]
    Le, info = E.factor(hl, b.s, b.xsize, _amap(b), a2x)
    assert info == 0 and relF(Le, Lo) < 1e-12
    # level of each supernode's first panel, and of its last
    first = {}
    last = {}
    for lvl in range(hl["nlevel"]):
        for q in hl["panel_order"][hl["lvl_panels"][lvl]:hl["lvl_panels"][lvl + 1]]:
            sn = int(hl["panels"][q]["sn"])
            first[sn] = min(first.get(sn, lvl), lvl)
            last[sn] = max(last.get(sn, lvl), lvl)
    lp, pr, part = sched
    for q in range(len(pr) - 1):
        seq = part[pr[q]:pr[q + 1]]
        for a, c in zip(seq[:-1], seq[1:]):
            assert first[int(c)] > last[int(a)]
    for k in range(1, len(lp) - 1):
        prev = [int(s) for q in range(lp[k - 1], lp[k]) for s in part[pr[q]:pr[q + 1]]]
        cur = [int(s) for q in range(lp[k], lp[k + 1]) for s in part[pr[q]:pr[q + 1]]]
        assert min(first[s] for s in cur) > max(last[s] for s in prev)


def _dag_structure_checks(plan):
    """The work items of the dependency-driven launch cover each tile entry of its
    levels once; prev is the tile's entry at its previous level; a dependency on
    a panel waits for all of its factoring tiles; the diagonal scratch slots of
    the launch's panels do not overlap (they are all in flight together)."""
    items, deps, tiles = plan["dag_items"], plan["dag_deps"], plan["tiles"]
    l0 = plan["dag_level0"]
    t0 = plan["lvl_tiles"][l0]
    assert sorted(items["tile"].tolist()) == list(range(t0, len(tiles)))
    level = np.zeros(len(tiles), np.int64)
    for lvl in range(plan["nlevel"]):
        level[plan["lvl_tiles"][lvl]:plan["lvl_tiles"][lvl + 1]] = lvl
    factoring = set(int(t) for t in plan["trsm_tiles"])
    nfac = np.zeros(len(plan["panels"]), np.int64)
    for t in range(t0, len(tiles)):
        if t in factoring:
            nfac[tiles[t]["panel"]] += 1
    last = {}
    for k, d in enumerate(items):
        tl = tiles[d["tile"]]
        key = (int(tl["panel"]), int(tl["r0"]))
        if d["prev"] >= 0:
            p = items[d["prev"]]
            assert d["prev"] < k and (int(tiles[p["tile"]]["panel"]), int(tiles[p["tile"]]["r0"])) == key
            assert level[p["tile"]] < level[d["tile"]]
        assert last.get(key, -1) == d["prev"]
        last[key] = k
        assert (d["fclass"] >= 0) == (int(d["tile"]) in factoring)
        for e in range(d["dep0"], d["dep1"]):
            assert deps[e]["count"] == nfac[deps[e]["panel"]] > 0
    spans = sorted((int(tiles[t]["slot"]), int(tiles[t]["ncol"]) ** 2)
                   for t in range(t0, len(tiles)) if t in factoring and tiles[t]["kind"] == 0)
    for (a, na), (b2, _) in zip(spans, spans[1:]):
        assert a + na <= b2
    assert spans[-1][0] + spans[-1][1] <= plan["diag_scratch"]


@pytest.mark.parametrize("params", [(8, 128, 1024, 256, 1, 1), (8, 128, 1024, 256, 1, 2),
                                    (2, 128, 1 << 30, 256, 0, 1), (3, 64, 1, 256, 1, 1)])
@pytest.mark.parametrize("make", [lambda: P.laplacian3d_var(12), lambda: P.dense_spd(300),
                                  lambda: P.laplacian2d_var(37, 3), lambda: P.vector3d_var(6)])
def test_dependency_driven_launch_emulated(params, make):
    """The dependency-driven factor launch, emulated: workgroups claim items in
    the plan's order and advance in random interleavings, each item waiting as the
    kernel waits.  The result is bitwise the per-level launches' (the waits leave
    no read of a value before it is final), no wait names a later item, and no
    interleaving stalls."""
    prob = make()
    b, a2x, Lo, _ = setup(prob)
    plan = api.build_plan(b, params=params)
    assert len(plan["dag_items"]) > 0
    _dag_structure_checks(plan)
    amap = _amap(b)
    Ll, info_l = E.factor(plan, b.s, b.xsize, amap, a2x)
    assert info_l == 0 and relF(Ll, Lo) < 1e-12
    for seed, workers in ((0, 3), (1, 16)):
        Ld, info_d = E.factor_dag(plan, b.s, b.xsize, amap, a2x, seed=seed, workers=workers)
        assert info_d == info_l
        assert np.array_equal(Ld, Ll)


def test_dependency_emulation_catches_a_missing_wait():
    """The emulation has the power to see a missing wait: with the sources'
    dependencies dropped from the plan, some interleaving reads a panel before it
    is final and the factor differs."""
    prob = P.laplacian3d_var(10)
    b, a2x, Lo, _ = setup(prob)
    plan = api.build_plan(b, params=(8, 128, 1 << 30, 256, 0, 2))
    amap = _amap(b)
    Ll, _ = E.factor(plan, b.s, b.xsize, amap, a2x)
    broken = dict(plan)
    broken["dag_items"] = plan["dag_items"].copy()
    broken["dag_items"]["dep1"] = broken["dag_items"]["dep0"]
    differs = False
    with np.errstate(all="ignore"):  # values read too early may overflow
        for seed in range(4):
            Ld, _ = E.factor_dag(broken, b.s, b.xsize, amap, a2x, seed=seed, workers=16)
            differs = differs or not np.array_equal(Ld, Ll)
    assert differs


@pytest.mark.parametrize("make", [lambda: P.laplacian3d_var(64), lambda: P.dense_spd(2000)])
def test_deferred_tiles_are_last_and_not_factored_in_their_level(make):
    """Tiles a level runs beside its panel factorization (deferred) come last in
    the level's list, and none of them belongs to a panel that level factors (so
    nothing of the level reads what they write)."""
    prob = make()
    b = S.analyze_p2(prob.n, prob.colptr, prob.rowidx, prob.perm)
    plan = api.build_plan(b)
    tiles, lt, ld = plan["tiles"], plan["lvl_tiles"], plan["lvl_tiles_deferred"]
    factoring = set(int(t) for t in plan["trsm_tiles"])
    panel_level = np.zeros(len(plan["panels"]), np.int64)
    for lvl in range(plan["nlevel"]):
        panel_level[plan["panel_order"][plan["lvl_panels"][lvl]:plan["lvl_panels"][lvl + 1]]] = lvl
    assert ld.sum() > 0  # these structures have right-looking steps that defer
    for lvl in range(plan["nlevel"]):
        d = int(ld[lvl])
        fac_panels = set(int(tiles[t]["panel"]) for t in range(lt[lvl], lt[lvl + 1]) if t in factoring)
        for t in range(lt[lvl + 1] - d, lt[lvl + 1]):
            assert t not in factoring and tiles[t]["kind"] == 1
            assert int(tiles[t]["panel"]) not in fac_panels
            assert panel_level[tiles[t]["panel"]] > lvl
