"""Numpy emulation of the GPU kernels on the real device plan (test infrastructure).

Runs the same index arithmetic as the tile update, panel factor and solve
kernels of swcholesky_amd/csrc/numeric.hip, asserting that every
index stays inside its buffer, so plan and kernel indexing are checked on the CPU.
"""
import numpy as np


def _chk(lo, hi, size, what):
    assert 0 <= lo and hi <= size, f"{what}: [{lo},{hi}) outside [0,{size})"


def _bits(m):
    return np.array([i for i in range(64) if (m >> i) & 1], dtype=np.int64)


def _update_tile(plan, Ls, L, ti, xsize, scratch):
    """k_tile_update / tile_update_body: every task of tile ti, then the copy of an
    updated diagonal block into its scratch slot (scratch: slot -> (panel, block))."""
    panels, tiles, tasks = plan["panels"], plan["tiles"], plan["tasks"]
    tl = tiles[ti]
    pn = panels[tl["panel"]]
    nr, nc, ld = int(tl["nr"]), int(pn["ncol"]), int(pn["ld"])
    assert 1 <= nr <= 64 and 1 <= nc <= 64
    base = int(pn["vbase"]) + int(tl["r0"])
    _chk(base, base + (nc - 1) * ld + nr, xsize, "tile")
    ro = int(pn["rows_off"]) + int(tl["r0"])
    _chk(ro, ro + nr, len(Ls), "tile rows")
    trow = Ls[ro:ro + nr]
    idx = base + np.arange(nc)[None, :] * ld + np.arange(nr)[:, None]
    acc = L[idx].copy()
    for t in range(tl["task_begin"], tl["task_end"]):
        tk = tasks[t]
        ra, rb, ca, cb = int(tk["ra"]), int(tk["rb"]), int(tk["ca"]), int(tk["cb"])
        k, sld, sb = int(tk["k"]), int(tk["ld"]), int(tk["src_base"])
        roff = int(tk["rows_off"])
        assert ra < rb and ca < cb and rb - ra <= 64 and cb - ca <= 64
        # landing positions: the e-th set bit receives source row ra+e / ca+e
        pr = _bits(int(tk["rmask"]))
        pc = _bits(int(tk["cmask"]))
        assert len(pr) == rb - ra and len(pc) == cb - ca, "mask/row count mismatch"
        assert pr.max() < nr and pc.max() < nc, "landing outside the tile"
        _chk(roff + min(ra, ca), roff + max(rb, cb), len(Ls), "task rows")
        assert np.array_equal(trow[pr], Ls[roff + ra:roff + rb]), "row mask"
        assert np.array_equal(int(pn["c0"]) + pc, Ls[roff + ca:roff + cb]), "column mask"
        # every buffer offset of a loaded element stays inside the
        # source's k columns (row < ld), and those inside L
        assert max(rb, cb) <= sld
        _chk(sb, sb + (k - 1) * sld + max(rb, cb), xsize, "task src")
        A = L[sb + np.arange(k)[None, :] * sld + np.arange(ra, rb)[:, None]]
        B = L[sb + np.arange(k)[None, :] * sld + np.arange(ca, cb)[:, None]]
        acc[np.ix_(pr, pc)] -= A @ B.T
    keep = np.arange(nr)[:, None] >= np.arange(nc)[None, :] if tl["r0"] == 0 else \
        np.ones((nr, nc), bool)
    L[idx[keep]] = acc[keep]
    if tl["slot"] >= 0 and tl["kind"] == 0:
        _chk(int(tl["slot"]), int(tl["slot"]) + nc * nc, plan["diag_scratch"], "diag slot")
        scratch[int(tl["slot"])] = (int(tl["panel"]), np.tril(acc[:nc, :nc]).copy())


def _factor_tile(plan, L, ti, xsize, scratch):
    """k_panel_factor_regs / k_panel_factor_blk on tile ti; returns info (0 or the
    failing global column + 1) for a diagonal tile."""
    panels, tiles = plan["panels"], plan["tiles"]
    tl = tiles[ti]
    pn = panels[tl["panel"]]
    nr, nc, ld = int(tl["nr"]), int(pn["ncol"]), int(pn["ld"])
    owner, blk = scratch[int(tl["slot"])]  # the launch-invariant copy
    assert owner == int(tl["panel"]), "diagonal scratch slot holds another panel's block"
    a11 = blk.copy()
    first = nc if tl["kind"] == 0 else 0
    base = int(pn["vbase"]) + int(tl["r0"])
    if nc > 16:
        bad = panel_factor_blk(L, base, ld, a11, nc, nr, first, xsize)
    else:
        bad = _panel_factor_scalar(L, base, ld, a11, nc, nr, first, xsize)
    return int(pn["c0"]) + bad if bad and tl["kind"] == 0 else 0


def _factor_level(plan, Ls, L, lvl, xsize, scratch):
    panels, tiles = plan["panels"], plan["tiles"]
    info = 0
    t0, t1 = plan["lvl_tiles"][lvl], plan["lvl_tiles"][lvl + 1]
    s0, s1 = plan["lvl_trsm"][lvl], plan["lvl_trsm"][lvl + 1]
    for ti in range(t0, t1):
        _update_tile(plan, Ls, L, ti, xsize, scratch)
    # fused narrow panels: one wave updates (above) and factors each panel,
    # its tiles contiguous with the diagonal tile first, none in the factor lists
    if "fused" in plan:
        in_lists = set(int(t) for t in plan["trsm_tiles"][s0:s1])
        for it in plan["fused"][plan["lvl_fused"][lvl]:plan["lvl_fused"][lvl + 1]]:
            t0f, ntf = int(it["tile0"]), int(it["ntile"])
            assert t0 <= t0f and t0f + ntf <= t0 + plan["lvl_tiles_fused"][lvl]
            d = tiles[t0f]
            pn = panels[d["panel"]]
            nc, ld = int(pn["ncol"]), int(pn["ld"])
            assert d["kind"] == 0 and d["r0"] == 0 and nc <= 16 and int(pn["sw"]) == nc
            didx = int(pn["vbase"]) + np.arange(nc)[None, :] * ld + np.arange(nc)[:, None]
            a11 = np.tril(L[didx])
            for k in range(ntf):
                tl = tiles[t0f + k]
                assert tl["panel"] == d["panel"] and t0f + k not in in_lists
                assert k == 0 or tl["r0"] > tiles[t0f + k - 1]["r0"]
                first = nc if k == 0 else 0
                base = int(pn["vbase"]) + int(tl["r0"])
                assert (np.arange(64)[:, None] * 17 + np.arange(16)[None, :]).max() < 64 * 17
                bad = _panel_factor_scalar(L, base, ld, a11.copy(), nc, int(tl["nr"]), first, xsize)
                if bad and k == 0:
                    info = info or int(pn["c0"]) + bad
    # panel factor kernel: every listed tile recomputes L11 of its panel
    for ti in plan["trsm_tiles"][s0:s1]:
        info = info or _factor_tile(plan, L, int(ti), xsize, scratch)
    return info


def factor(plan, Ls, xsize, amap, vals):
    L = np.zeros(xsize)
    _chk(int(amap.min()), int(amap.max()) + 1, xsize, "amap")
    L[amap] = vals
    info = 0
    for lvl in range(plan["nlevel"]):
        info = info or _factor_level(plan, Ls, L, lvl, xsize, {})
    return L, info


def factor_dag(plan, Ls, xsize, amap, vals, seed=0, workers=8):
    """The per-level launches up to dag_level0, then k_dag_factor: `workers`
    workgroups claim the work items in array order and advance, one step at a
    time in a random interleaving, through wait -> update -> publish -> (wait for
    the diagonal tile) -> factor -> count the panel's factored tile.  A wait may
    only name items claimed before (checked), and the run must never stall; the
    diagonal scratch slots of the launch's levels are shared by all of it, so a
    slot overwritten while another tile still needs it is caught."""
    rng = np.random.default_rng(seed)
    L = np.zeros(xsize)
    L[amap] = vals
    info = 0
    l0 = int(plan["dag_level0"])
    items, deps, tiles = plan["dag_items"], plan["dag_deps"], plan["tiles"]
    if len(items) == 0:
        l0 = plan["nlevel"]
    for lvl in range(l0):
        info = info or _factor_level(plan, Ls, L, lvl, xsize, {})
    if len(items) == 0:
        return L, info
    assert int(items[0]["tile"]) >= plan["lvl_tiles"][l0]
    scratch = {}
    n = len(items)
    state = np.zeros(n, np.int64)
    pdone = np.zeros(len(plan["panels"]), np.int64)
    ticket = 0
    work = [None] * workers  # [item, phase]
    done = 0
    while done < n:
        movable = []
        for w in range(workers):
            if work[w] is None:
                if ticket < n:
                    movable.append(w)
                continue
            k, ph = work[w]
            d = items[k]
            if ph == 0:  # wait for the tile's previous writer and the source panels
                ok = d["prev"] < 0 or state[d["prev"]] >= 1
                for e in range(d["dep0"], d["dep1"]):
                    ok = ok and pdone[deps[e]["panel"]] >= deps[e]["count"]
                if ok:
                    movable.append(w)
            elif ph == 1:  # wait for the diagonal tile's updates
                if d["diag"] < 0 or state[d["diag"]] >= 1:
                    movable.append(w)
        assert movable, "the dependency-driven launch stalls"
        w = int(rng.choice(movable))
        if work[w] is None:
            work[w] = [ticket, 0]
            d = items[ticket]
            assert d["prev"] < ticket and d["diag"] < ticket
            ticket += 1
            continue
        k, ph = work[w]
        d = items[k]
        ti = int(d["tile"])
        if ph == 0:
            _update_tile(plan, Ls, L, ti, xsize, scratch)
            if d["fclass"] < 0 or tiles[ti]["kind"] == 0:
                state[k] = 1
            if d["fclass"] < 0:
                work[w] = None
                done += 1
            else:
                assert d["fclass"] == (0 if tiles[ti]["ncol"] <= 16 else 1)
                work[w][1] = 1
        else:
            info = info or _factor_tile(plan, L, ti, xsize, scratch)
            pdone[tiles[ti]["panel"]] += 1
            work[w] = None
            done += 1
    return L, info


def _panel_factor_scalar(L, base, ld, a11, nc, nr, first, xsize):
    """k_panel_factor_regs: right-looking over single columns."""
    l11 = np.zeros_like(a11)
    bad = 0
    for j in range(nc):
        d = a11[j, j]
        if not d > 0:
            bad = bad or j + 1
            d = 1.0
        l11[j, j] = np.sqrt(d) if a11[j, j] > 0 else 1.0
        l11[j + 1:, j] = a11[j + 1:, j] / l11[j, j]
        for c in range(j + 1, nc):
            a11[c:, c] -= l11[c:, j] * l11[c, j]
    if first == nc:
        keep = np.tril(np.ones((nc, nc), bool))
        didx = base + np.arange(nc)[None, :] * ld + np.arange(nc)[:, None]
        L[didx[keep]] = l11[keep]
    if nr > first:
        idx = base + np.arange(nc)[None, :] * ld + np.arange(first, nr)[:, None]
        _chk(int(idx.min()), int(idx.max()) + 1, xsize, "factor tile")
        L[idx] = np.linalg.solve(l11, L[idx].T).T
    return bad


KWLD = 2 * 64 + 1  # column stride of the LDS working matrix of k_panel_factor_blk


def panel_factor_blk(L, base, ld, a11, nc, nr, first, xsize):
    """k_panel_factor_blk with its LDS arrays W (64 columns of KWLD doubles) and
    lc (3 x 16 x 16), every LDS and global index bounds-checked: the per-wave
    column loads, per block column the sweep of waves 0-2 (lane = row; wave 0
    the diagonal block's rows, waves 1-2 the 16 pivot rows in lanes 0..15 and
    the tile rows after them), the rank-16 updates of the later block columns in
    16x16 MFMA output blocks, and the per-wave column stores."""
    W = np.full(64 * KWLD, np.nan)
    lc = np.full((3, 16, 16), np.nan)
    nb = (nc + 15) >> 4
    assert 2 <= nb <= 4
    nx = max(nr - first, 0)
    nxb = (nx + 15) >> 4
    assert nx <= 64
    lanes = np.arange(64)

    def w(c, r):
        c, r = np.asarray(c), np.asarray(r)
        if c.size and r.size:
            assert c.min() >= 0 and c.max() < 64 and r.min() >= 0 and r.max() < 128
        return c * KWLD + r

    for wave in range(4):
        c0 = 16 * wave
        for c in range(16):
            col = c0 + c
            iny = (lanes < nc) & (col < nc) & (col <= lanes)
            inx = (lanes < nx) & (col < nc)
            gy = np.where(iny, col * nc + lanes, 0)
            assert gy.min() >= 0 and gy.max() < nc * nc
            gx = base + np.where(inx, col * ld + first + lanes, 0)
            _chk(int(gx.min()), int(gx.max()) + 1, xsize, "factor load")
            y = np.where(iny, a11[np.minimum(lanes, nc - 1), min(col, nc - 1)],
                         np.where(((lanes >= nc) | (col >= nc)) & (col == lanes), 1.0, 0.0))
            x = np.where(inx, L[gx], 0.0)
            if c0 < 16 * nb:
                W[w(col, lanes)] = y
                W[w(col, 64 + lanes)] = x
    bad = 0
    for J in range(nb):
        jb = 16 * J
        sweeps = []
        for wave in range(3):
            if wave == 0:
                row = lanes.copy()
                pbase = jb
                store = (lanes >= jb) & (lanes < 16 * nb)
            else:
                dup = (lanes < 16) | ((wave == 2) & (lanes >= 32))
                row = np.where(dup, jb + (lanes & 15), 64 + 48 * (wave - 1) + (lanes - 16))
                pbase = 0
                store = (lanes >= 16) & (row - 64 < nx) & ((wave == 1) | (lanes < 32))
            # every wave reads its rows before any wave writes (a barrier)
            sweeps.append((row, pbase, store, np.stack([W[w(jb + c, row)] for c in range(16)], axis=1)))
        w(jb + 15, jb + (lanes & 15))  # wave 3 reads one row too
        for wave, (row, pbase, store, X) in enumerate(sweeps):
            assert 0 <= pbase and pbase + 15 < 64
            d = X[pbase, 0]
            for j in range(16):
                if d > 0:
                    X[:, j] *= 1.0 / np.sqrt(d)
                else:
                    bad = bad or jb + j + 1
                    X[pbase + j, j] = 1.0
                if j < 15:
                    d = X[pbase + j + 1, j + 1] - X[pbase + j + 1, j] * X[pbase + j + 1, j]
                    sel = (lanes >= pbase) & (lanes < pbase + 16)
                    idx = lanes[sel] - pbase
                    assert idx.min() >= 0 and idx.max() < 16
                    lc[wave, j, idx] = X[sel, j]
                    for c in range(j + 1, 16):
                        X[:, c] -= X[:, j] * lc[wave, j, c]
            for c in range(16):
                W[w(jb + c, row[store])] = X[store, c]
        for C in range(J + 1, nb):
            for R in list(range(C, nb)) + list(range(4, 4 + nxb)):
                for r in range(4):
                    i = 16 * R + (lanes >> 4) + 4 * r
                    jcol = 16 * C + (lanes & 15)
                    acc = W[w(jcol, i)].copy()
                    for u in range(4):
                        k = jb + (lanes >> 4) + 4 * u
                        # lane l supplies A[row 16R + l%16][k] and B[row 16C + l%16][k]
                        assert k.max() < jb + 16
                    A = np.stack([W[w(jb + kk, i)] for kk in range(16)], axis=1)
                    B = np.stack([W[w(jb + kk, jcol)] for kk in range(16)], axis=1)
                    W[w(jcol, i)] = acc - (A * B).sum(axis=1)
    for wave in range(4):
        c0 = 16 * wave
        for c in range(16):
            col = c0 + c
            if col >= nc:
                continue
            if first == nc:
                sd = (lanes < nc) & (col <= lanes)
                gi = base + col * ld + lanes[sd]
                if len(gi):
                    _chk(int(gi.min()), int(gi.max()) + 1, xsize, "factor store")
                L[gi] = W[w(col, lanes[sd])]
            sx = lanes < nx
            gi = base + col * ld + first + lanes[sx]
            if len(gi):
                _chk(int(gi.min()), int(gi.max()) + 1, xsize, "factor store")
            L[gi] = W[w(col, 64 + lanes[sx])]
    return bad


def solve(plan, Ls, L, b):
    """Forward then backward solve in the factor ordering (b is an n-vector),
    as k_fwd_gather / k_fwd_panel / k_bwd_partial / k_bwd_diag run it over the
    solve panels (up to SPW columns of one supernode): the forward sweep
    accumulates right-hand sides in x and writes the panel solutions to y; the
    backward sweep reads y and the final values of x, and overwrites x."""
    x = np.array(b, dtype=np.float64)
    n = len(x)
    y = np.full(n, np.nan)
    panels, order, lp = plan["spanels"], plan["sp_order"], plan["lvl_sp"]
    gptr, gidx, gcols, lg = plan["gptr"], plan["gidx"], plan["gcols"], plan["lvl_gcol"]
    fb, lf = plan["fblocks"], plan["lvl_fblk"]
    vsize = plan["vsize"]
    V = np.zeros(max(vsize, 1))
    xsize = len(L)
    assert len(gidx) == vsize and (vsize == 0 or (gidx.min() >= 0 and gidx.max() < vsize))
    assert np.array_equal(np.sort(gcols), np.arange(n)), "every column gathered once"

    def diag(pn):
        nc, ld = int(pn["ncol"]), int(pn["ld"])
        idx = int(pn["vbase"]) + np.arange(nc)[None, :] * ld + np.arange(nc)[:, None]
        _chk(int(idx.min()), int(idx.max()) + 1, xsize, "diag")
        return np.tril(L[idx])

    for lvl in range(plan["nslevel"]):
        for c in gcols[lg[lvl]:lg[lvl + 1]]:
            for e in range(gptr[c], gptr[c + 1]):
                x[c] -= V[gidx[e]]
        pushes = []
        for bk in fb[lf[lvl]:lf[lvl + 1]]:
            pn = panels[bk["panel"]]
            nc, ld, c0 = int(pn["ncol"]), int(pn["ld"]), int(pn["c0"])
            r0, nr = int(bk["r0"]), int(bk["nr"])
            assert 0 <= nr <= 64 and 1 <= nc <= 256 and r0 >= nc and r0 + nr <= int(pn["nrow"])
            xs = np.linalg.solve(diag(pn), x[c0:c0 + nc])
            if bk["writer"]:
                y[c0:c0 + nc] = xs
            if nr:
                blk = L[int(pn["vbase"]) + r0 + np.arange(nc)[None, :] * ld + np.arange(nr)[:, None]]
                pushes.append((pn, r0, nr, blk @ xs))
        for pn, r0, nr, yy in pushes:  # every read of x in the level precedes these writes
            p = (int(pn["c0"]) - int(pn["sc0"])) + r0 + np.arange(nr)
            intra = p < int(pn["sw"])
            tgt = int(pn["sc0"]) + p[intra]
            _chk(int(tgt.min(initial=0)), int(tgt.max(initial=0)) + 1, n, "intra")
            x[tgt] -= yy[intra]
            vi = int(pn["voff"]) + p[~intra] - int(pn["sw"])
            if vi.size:
                _chk(int(vi.min()), int(vi.max()) + 1, vsize, "V push")
            V[vi] += yy[~intra]
    assert not np.isnan(y).any(), "a panel solution was never written"
    for lvl in range(plan["nslevel"] - 1, -1, -1):
        for qi in range(lp[lvl], lp[lvl + 1]):
            pn = panels[order[qi]]
            nc, ld, c0, nrow = int(pn["ncol"]), int(pn["ld"]), int(pn["c0"]), int(pn["nrow"])
            rows = Ls[int(pn["rows_off"]) + nc:int(pn["rows_off"]) + nrow]
            blk = L[int(pn["vbase"]) + nc + np.arange(nc)[None, :] * ld + np.arange(nrow - nc)[:, None]]
            t = y[c0:c0 + nc] - blk.T @ x[rows]
            x[c0:c0 + nc] = np.linalg.solve(diag(pn).T, t)
    return x
