// Builds the panel / tile / task plan (see plan.h).
//
// The per-descendant overlap scan of the reference (lb/ub in
// sw_parallel_PB_Cholesky_05.c:247-264) becomes binary searches on the sorted
// row lists, done once here instead of once per factorization.
#include "plan.h"

#include <algorithm>
#include <cstring>
#include <queue>
#include <stdexcept>
#include <string>

namespace swc {

namespace {

int lower_idx(const int* a, int len, int v) { return int(std::lower_bound(a, a + len, v) - a); }
int upper_idx(const int* a, int len, int v) { return int(std::upper_bound(a, a + len, v) - a); }

// Work items of the dependency-driven factor launch (DagItem) for the tiles of
// levels dag_level0 on, and the order in which the launch's workgroups claim
// them.  Any topological order of the waits keeps the launch free of deadlock:
// an item waits only for items claimed before it, by workgroups that are running
// and never wait for anything later.
//  mode 2: level order; in each level the tiles of the panels it factors first,
//          each panel's diagonal tile before the others.
//  mode 1: the order in which a list schedule on the device's resident
//          workgroups would start the items, simulated here with a cost model:
//          whenever a workgroup is free it takes the ready item with the longest
//          remaining path (its "bottom level"), so the chain of a wide
//          supernode's panel factorizations is not held up behind updates that
//          can wait.  Costs are in microseconds: a fixed part plus the useful
//          update flops at about 0.1 TFLOP/s per workgroup, plus the panel
//          factorization.
void build_dag(Plan& P, const std::vector<int>& task_src, const std::vector<int>& tile_level,
               const std::vector<double>& tile_flops, const std::vector<char>& is_factor,
               int mode) {
  const int t0 = P.lvl_tiles[P.dag_level0], nt = (int)P.tiles.size();
  std::vector<int> base;
  for (int ti = t0; ti < nt; ++ti) base.push_back(ti);
  auto group = [&](int ti) { return is_factor[ti] ? (P.tiles[ti].kind == 0 ? 0 : 1) : 2; };
  std::stable_sort(base.begin(), base.end(), [&](int a, int b) {
    if (tile_level[a] != tile_level[b]) return tile_level[a] < tile_level[b];
    return group(a) < group(b);
  });
  const int N = (int)base.size();
  std::vector<int> nfac(P.npanel, 0);  // factoring tiles of each panel
  for (int ti = t0; ti < nt; ++ti)
    if (is_factor[ti]) nfac[P.tiles[ti].panel]++;
  std::vector<DagItem> it(N);
  std::vector<std::vector<int>> deps(N);
  std::vector<int64_t> rb0(P.npanel + 1, 0);  // row blocks of each panel
  for (int q = 0; q < P.npanel; ++q) rb0[q + 1] = rb0[q] + (P.panels[q].nrow + TM - 1) / TM;
  std::vector<int> last((size_t)rb0[P.npanel], -1), diag_item(P.npanel, -1);
  for (int k = 0; k < N; ++k) {
    const int ti = base[k];
    const Tile& tl = P.tiles[ti];
    DagItem& d = it[k];
    d.tile = ti;
    d.fclass = is_factor[ti] ? (tl.ncol <= 16 ? 0 : 1) : -1;
    int& lt = last[(size_t)(rb0[tl.panel] + tl.r0 / TM)];
    d.prev = lt;
    lt = k;
    d.diag = -1;
    if (is_factor[ti]) {
      if (tl.kind == 0) {
        diag_item[tl.panel] = k;
      } else {
        d.diag = diag_item[tl.panel];
        if (d.diag < 0) throw std::logic_error("factoring tile before its diagonal tile");
      }
    }
    for (int t = tl.task_begin; t < tl.task_end; ++t) {
      const int q = task_src[t];
      if (P.panel_level[q] >= tile_level[ti]) throw std::logic_error("update reads a later panel");
      if (P.panel_level[q] >= P.dag_level0) deps[k].push_back(q);
    }
    std::sort(deps[k].begin(), deps[k].end());
    deps[k].erase(std::unique(deps[k].begin(), deps[k].end()), deps[k].end());
  }
  std::vector<int> order(N);
  for (int k = 0; k < N; ++k) order[k] = k;
  const double rate = 1e5;  // useful flops per microsecond and workgroup
  auto c_upd = [&](int k) { return 3.0 + 1.4 * tile_flops[it[k].tile] / rate; };
  auto c_fac = [&](int k) { return it[k].fclass < 0 ? 0.0 : (it[k].fclass == 0 ? 5.0 : 14.0); };
  {
    // the same levels as launches: per level, the updates (longest first on 512
    // workgroups), then the factorizations, each launch with ~4 us of overhead
    P.lvl_sim_us = 0;
    for (int k0 = 0; k0 < N;) {
      int k1 = k0;
      while (k1 < N && tile_level[it[k1].tile] == tile_level[it[k0].tile]) ++k1;
      std::vector<double> c;
      double fac = 0;
      for (int k = k0; k < k1; ++k) {
        c.push_back(c_upd(k));
        fac = std::max(fac, c_fac(k));
      }
      std::sort(c.rbegin(), c.rend());
      std::priority_queue<double, std::vector<double>, std::greater<double>> w;
      for (int i = 0; i < 512; ++i) w.push(0.0);
      double span = 0;
      for (double x : c) {
        const double t = w.top() + x;
        w.pop();
        w.push(t);
        span = std::max(span, t);
      }
      P.lvl_sim_us += span + 4.0 + (fac > 0 ? fac + 4.0 : 0.0);
      k0 = k1;
    }
  }
  if (mode == 1) {
    std::vector<int> next_on_tile(N, -1);
    for (int k = 0; k < N; ++k)
      if (it[k].prev >= 0) next_on_tile[it[k].prev] = k;
    // bottom levels, from the last item back (every successor comes later in base order)
    std::vector<double> bottom(N, 0.0), pbot(P.npanel, 0.0), dbot(N, 0.0);
    for (int k = N - 1; k >= 0; --k) {
      double after = dbot[k];
      if (next_on_tile[k] >= 0) after = std::max(after, bottom[next_on_tile[k]]);
      if (it[k].fclass >= 0) after = std::max(after, c_fac(k) + pbot[P.tiles[it[k].tile].panel]);
      bottom[k] = c_upd(k) + after;
      for (int q : deps[k]) pbot[q] = std::max(pbot[q], bottom[k]);
      if (it[k].diag >= 0) dbot[it[k].diag] = std::max(dbot[it[k].diag], bottom[k]);
    }
    std::vector<int> pending(N, 0), remaining(nfac);
    std::vector<std::vector<int>> waiters(P.npanel), dwait(N);
    for (int k = 0; k < N; ++k) {
      pending[k] = (it[k].prev >= 0) + (it[k].diag >= 0) + (int)deps[k].size();
      for (int q : deps[k]) waiters[q].push_back(k);
      if (it[k].diag >= 0) dwait[it[k].diag].push_back(k);
    }
    typedef std::pair<double, int> Ready;  // (bottom level, -index): the largest first
    std::priority_queue<Ready> ready;
    struct Ev {
      double t;
      int type, k;  // type 0: updates written (diagonal tiles: releases the panel's tiles)
      bool operator<(const Ev& o) const { return t > o.t || (t == o.t && k > o.k); }
    };
    std::priority_queue<Ev> ev;
    for (int k = 0; k < N; ++k)
      if (pending[k] == 0) ready.push({bottom[k], -k});
    auto release = [&](int k) {
      if (--pending[k] == 0) ready.push({bottom[k], -k});
    };
    const int workers = 512;  // resident workgroups of the launch: 256 CUs x 2
    int free_w = workers, started = 0;
    double now = 0;
    while (started < N || !ev.empty()) {
      while (free_w > 0 && !ready.empty()) {
        const int k = -ready.top().second;
        ready.pop();
        order[started++] = k;
        --free_w;
        const double tu = now + c_upd(k);
        if (it[k].fclass >= 0) {
          ev.push({tu, 0, k});
          ev.push({tu + c_fac(k), 1, k});
        } else {
          ev.push({tu, 1, k});
        }
      }
      if (ev.empty()) throw std::logic_error("dependency simulation stalled");
      const Ev e = ev.top();
      ev.pop();
      now = e.t;
      if (e.type == 0) {
        for (int d : dwait[e.k]) release(d);
        continue;
      }
      ++free_w;
      P.dag_sim_us = std::max(P.dag_sim_us, now);
      if (next_on_tile[e.k] >= 0) release(next_on_tile[e.k]);
      if (it[e.k].fclass >= 0) {
        const int q = P.tiles[it[e.k].tile].panel;
        if (--remaining[q] == 0)
          for (int w : waiters[q]) release(w);
      }
    }
  }
  std::vector<int> pos(N), pmax(P.npanel, -1);
  for (int k = 0; k < N; ++k) pos[order[k]] = k;
  for (int k = 0; k < N; ++k)
    if (it[k].fclass >= 0) {
      int& m = pmax[P.tiles[it[k].tile].panel];
      m = std::max(m, pos[k]);
    }
  P.dag_items.resize(N);
  P.dag_deps.clear();
  P.dag_flops = 0;
  for (int k = 0; k < N; ++k) {
    DagItem d = it[order[k]];
    if (d.prev >= 0) d.prev = pos[d.prev];
    if (d.diag >= 0) d.diag = pos[d.diag];
    if (d.prev >= k || d.diag >= k) throw std::logic_error("item order is not topological");
    d.dep0 = (int)P.dag_deps.size();
    for (int q : deps[order[k]]) {
      if (pmax[q] >= k) throw std::logic_error("item order is not topological");
      P.dag_deps.push_back({q, nfac[q]});
    }
    d.dep1 = (int)P.dag_deps.size();
    P.dag_items[k] = d;
    P.dag_flops += tile_flops[d.tile];
  }
  for (int q = 0; q < P.npanel; ++q)
    if (P.panel_level[q] >= P.dag_level0) {
      const double w = P.panels[q].ncol, m = P.panels[q].nrow - P.panels[q].ncol;
      P.dag_flops += w * w * w / 3.0 + m * w * w;
    }
}

}  // namespace

void descendant_sets(int n, int nsuper, const int* super, const int64_t* pi, const int* Ls,
                     std::vector<int>& prunePtr, std::vector<int>& pruneSet) {
  std::vector<int> col2sup(n);
  for (int s = 0; s < nsuper; ++s)
    for (int c = super[s]; c < super[s + 1]; ++c) col2sup[c] = s;
  // (target, descendant) pairs, one per distinct target of each descendant's rows
  std::vector<int> cnt(nsuper + 1, 0), last(nsuper, -1);
  for (int d = 0; d < nsuper; ++d) {
    const int W = super[d + 1] - super[d];
    for (int64_t i = pi[d] + W; i < pi[d + 1]; ++i) {
      const int r = Ls[i];
      if (r < 0 || r >= n) throw std::invalid_argument("row index out of range");
      const int t = col2sup[r];
      if (t <= d) throw std::invalid_argument("row of a supernode below its own columns");
      if (last[t] != d) {
        last[t] = d;
        cnt[t + 1]++;
      }
    }
  }
  for (int s = 0; s < nsuper; ++s) cnt[s + 1] += cnt[s];
  prunePtr.assign(cnt.begin(), cnt.end());
  pruneSet.assign(cnt[nsuper], 0);
  std::fill(last.begin(), last.end(), -1);
  std::vector<int> fill(cnt.begin(), cnt.end() - 1);
  for (int d = 0; d < nsuper; ++d) {
    const int W = super[d + 1] - super[d];
    for (int64_t i = pi[d] + W; i < pi[d + 1]; ++i) {
      const int t = col2sup[Ls[i]];
      if (last[t] != d) {
        last[t] = d;
        pruneSet[fill[t]++] = d;
      }
    }
  }
}

bool plan_params_from(const int* v, int count, PlanParams& out) {
  PlanParams p;
  if (count > 0) p.rlb = v[0];
  if (count > 1) p.spw = v[1];
  if (count > 2) p.narrow_min_tiles = v[2];
  if (count > 3) p.narrow_max_k = v[3];
  if (count > 4) p.fuse_narrow = v[4];
  if (count > 5) p.dag = v[5];
  if (count < 0 || count > kPlanParams || p.rlb < 1 || p.rlb > 64 ||
      (p.spw != 64 && p.spw != SPW) || p.narrow_min_tiles < 1 || p.narrow_max_k < 0 ||
      (p.fuse_narrow != 0 && p.fuse_narrow != 1) || p.dag < 0 || p.dag > 2)
    return false;
  out = p;
  return true;
}

void build_plan(int n, int nsuper, const int* super, const int64_t* pi, const int64_t* px,
                const int* Ls, const int* prunePtr, const int* pruneSet, Plan& P, bool factor,
                const PlanParams& params, const HLevelSchedule* hl) {
  const int RLB = params.rlb, SPWP = params.spw;
  P = Plan();
  P.n = n;
  P.nsuper = nsuper;
  std::vector<int> first_panel(nsuper + 1, 0);
  std::vector<int64_t> voff(nsuper + 1, 0);
  for (int s = 0; s < nsuper; ++s) {
    const int W = super[s + 1] - super[s];
    const int64_t R = pi[s + 1] - pi[s];
    voff[s + 1] = voff[s] + (R - W);
    if (W <= 0 || R < W) throw std::invalid_argument("inconsistent supernode structure");
    if (R > 0x7fffffff) throw std::invalid_argument("supernode too tall");
    first_panel[s] = (int)P.panels.size();
    for (int q = 0; q * PW < W; ++q) {
      Panel p;
      p.vbase = px[s] + (int64_t)q * PW * R + (int64_t)q * PW;
      p.ld = (int)R;
      p.ncol = std::min(PW, W - q * PW);
      p.nrow = (int)R - q * PW;
      p.c0 = super[s] + q * PW;
      p.rows_off = (int)(pi[s] + q * PW);
      p.sn = s;
      p.voff = voff[s];
      p.sc0 = super[s];
      p.sw = W;
      P.panels.push_back(p);
    }
    P.xsize = std::max<int64_t>(P.xsize, px[s] + (int64_t)W * R);
  }
  first_panel[nsuper] = (int)P.panels.size();
  P.npanel = (int)P.panels.size();
  P.vsize = voff[nsuper];

  // ---- descendant sources of each panel ----
  struct Src { int d; int ca, cb; };  // d >= 0: supernode, d < 0: panel -(d+1) of same sn
  std::vector<std::vector<Src>> srcs(P.npanel);  // descendants touching each panel
  for (int s = 0; s < nsuper && factor; ++s) {
    const int fp = first_panel[s], lp = first_panel[s + 1];
    for (int t = prunePtr[s]; t < prunePtr[s + 1]; ++t) {
      const int d = pruneSet[t];
      if (d < 0 || d >= s) throw std::invalid_argument("pruneSet entry is not a descendant");
      const int* rows = Ls + pi[d];
      const int len = (int)(pi[d + 1] - pi[d]);
      for (int q = fp; q < lp; ++q) {
        const Panel& pn = P.panels[q];
        const int ca = lower_idx(rows, len, pn.c0);
        const int cb = lower_idx(rows, len, pn.c0 + pn.ncol);
        if (ca < cb) srcs[q].push_back({d, ca, cb});
      }
    }
  }

  auto describe = [&](int q, const Src& sr, Task& t, int nsrc = 1) {
    const Panel& pn = P.panels[q];
    if (sr.d >= 0) {
      const int d = sr.d;
      t.src_base = px[d];
      t.ld = (int)(pi[d + 1] - pi[d]);
      t.k = super[d + 1] - super[d];
      t.rows_off = (int)pi[d];
    } else {
      // panels [-(sr.d + 1), -(sr.d + 1) + nsrc) of the same supernode: their
      // columns are consecutive in the supernode's column-major block
      const Panel& sp = P.panels[-(sr.d + 1)];
      const int s = pn.sn;
      const int qoff = sp.c0 - super[s];
      t.src_base = px[s] + (int64_t)qoff * pn.ld;
      t.ld = pn.ld;
      t.k = 0;
      for (int i = 0; i < nsrc; ++i) t.k += P.panels[-(sr.d + 1) + i].ncol;
      t.rows_off = (int)pi[s];
    }
    t.ca = sr.ca;
    t.cb = sr.cb;
    t.kind = sr.d >= 0 ? 0 : 1;
  };

  // ---- forward-solve gathers: column c takes row c of every descendant's
  // contribution vector, in increasing supernode order ----
  P.gptr.assign((size_t)n + 1, 0);
  for (int d = 0; d < nsuper; ++d) {
    const int W = super[d + 1] - super[d];
    for (int64_t i = pi[d] + W; i < pi[d + 1]; ++i) P.gptr[(size_t)Ls[i] + 1]++;
  }
  for (int c = 0; c < n; ++c) P.gptr[c + 1] += P.gptr[c];
  if (P.gptr[n] != P.vsize) throw std::logic_error("gather lists do not cover V");
  P.gidx.resize((size_t)P.vsize);
  {
    std::vector<int64_t> fill(P.gptr.begin(), P.gptr.end() - 1);
    for (int d = 0; d < nsuper; ++d) {
      const int W = super[d + 1] - super[d];
      for (int64_t i = pi[d] + W; i < pi[d + 1]; ++i)
        P.gidx[(size_t)fill[Ls[i]]++] = voff[d] + (i - pi[d] - W);
    }
  }

  // ---- levels: tiles of each level form one update launch, followed by one
  // TRSM launch over the tiles whose panel diagonal that level factored ----
  struct PTile { Tile t; int level; bool trsm; };
  std::vector<PTile> ptiles;
  std::vector<std::vector<Task>> ptasks;
  std::vector<std::vector<int>> psrc;  // per task: the last source panel it reads
  P.panel_level.assign(P.npanel, 0);
  std::vector<int> sn_level(nsuper, -1);  // level at which a supernode is complete
  auto add_tile = [&](int q, int r0, int kind, int level, bool trsm) {
    const Panel& pn = P.panels[q];
    PTile pt;
    pt.t.vbase = pn.vbase;
    pt.t.panel = q;
    pt.t.r0 = r0;
    pt.t.nr = std::min(TM, pn.nrow - r0);
    pt.t.kind = kind;
    pt.t.slot = -1;
    pt.t.ld = pn.ld;
    pt.t.ncol = pn.ncol;
    pt.t.c0 = pn.c0;
    pt.t.task_begin = pt.t.task_end = 0;
    pt.level = level;
    pt.trsm = trsm;
    ptiles.push_back(pt);
    ptasks.emplace_back();
    psrc.emplace_back();
    return (int)ptiles.size() - 1;
  };
  // descendant tasks of tile (panel q, rows r0..r0+nr) in pruneSet order
  auto add_desc_tasks = [&](int ti) {
    const Tile& tl = ptiles[ti].t;
    const int q = tl.panel;
    const Panel& pn = P.panels[q];
    const int vfirst = Ls[pn.rows_off + tl.r0];
    const int vlast = Ls[pn.rows_off + tl.r0 + tl.nr - 1];
    for (const Src& sr : srcs[q]) {
      Task t;
      describe(q, sr, t);
      const int* rows = Ls + t.rows_off;
      const int len = (int)(pi[sr.d + 1] - pi[sr.d]);
      const int ra = std::max(lower_idx(rows, len, vfirst), sr.ca);
      const int rb = upper_idx(rows, len, vlast);
      if (ra >= rb) continue;
      t.ra = ra;
      t.rb = rb;
      ptasks[ti].push_back(t);
      psrc[ti].push_back(first_panel[sr.d + 1] - 1);  // a supernode is final with its last panel
    }
  };
  // task of tile (panel q, rows r0..) from the earlier panels q2 .. q2+nsrc-1
  // of the same supernode (one inner dimension over all their columns)
  auto add_panel_task = [&](int ti, int q2, int nsrc) {
    const Tile& tl = ptiles[ti].t;
    const int q = tl.panel;
    const Panel& pn = P.panels[q];
    const int off = pn.c0 - super[pn.sn];
    Task t;
    describe(q, {-(q2 + 1), off, off + pn.ncol}, t, nsrc);
    t.ra = off + tl.r0;
    t.rb = t.ra + tl.nr;
    ptasks[ti].push_back(t);
    psrc[ti].push_back(q2 + nsrc - 1);
  };

  // Supernode s starts at the first level after its pruned descendants complete
  // and not before `floor`; returns the level at which it completes.
  auto place = [&](int s, int floor) {
    const int fp = first_panel[s], lp = first_panel[s + 1], np = lp - fp;
    int le = floor;
    for (int t = prunePtr[s]; t < prunePtr[s + 1]; ++t) {
      if (sn_level[pruneSet[t]] < 0)
        throw std::invalid_argument("schedule places a supernode before its descendants");
      le = std::max(le, sn_level[pruneSet[t]] + 1);
    }
    // external phase (and, for one panel, the whole factorization) at level le
    for (int J = 0; J < np; ++J) {
      const Panel& pn = P.panels[fp + J];
      for (int r0 = 0; r0 < pn.nrow; r0 += TM) {
        const bool diag = r0 == 0;
        const int kind = (diag && J == 0) ? 0 : 1;
        const int ti = add_tile(fp + J, r0, kind, le, !diag && J == 0);
        add_desc_tasks(ti);
      }
    }
    P.panel_level[fp] = le;
    // right-looking steps over blocks of RLB panels: step p at level le+p+1
    // updates the later panels of p's block from panel p; the block's last step
    // updates every panel after the block from all of the block's panels at
    // once (inner dimension up to RLB * PW).  Panel p+1 is updated at step p
    // either way, and is factored at that level.
    for (int p = 0; p + 1 < np; ++p) {
      const int lvl = le + p + 1;
      const int b0 = p - p % RLB, b1 = std::min(b0 + RLB, np);
      const bool last = p + 1 == b1;
      for (int J = p + 1; J < (last ? np : b1); ++J) {
        const Panel& pn = P.panels[fp + J];
        for (int r0 = 0; r0 < pn.nrow; r0 += TM) {
          const bool diag = r0 == 0;
          const int kind = (diag && J == p + 1) ? 0 : 1;
          const int ti = add_tile(fp + J, r0, kind, lvl, !diag && J == p + 1);
          if (last)
            add_panel_task(ti, fp + b0, b1 - b0);
          else
            add_panel_task(ti, fp + p, 1);
        }
      }
      P.panel_level[fp + p + 1] = lvl;
    }
    sn_level[s] = le + np - 1;
    return sn_level[s];
  };
  if (factor && !hl) {
    // wavefront: every supernode as early as its descendants allow (the level
    // sets of the supernodal elimination tree, TreeUtils.h:119-170)
    for (int s = 0; s < nsuper; ++s) place(s, 0);
  } else if (factor) {
    // the reference's H-level schedule: clusters one after another (a barrier
    // between them), the partitions of a cluster side by side, each partition's
    // supernodes one after another in its order (sw_parallel_PB_Cholesky_05.c:
    // the level-cluster loop, one thread per partition)
    const int ncl = (int)hl->level_ptr.size() - 1;
    if (ncl < 0 || hl->par_ptr.empty() || hl->level_ptr[0] != 0 ||
        hl->level_ptr[ncl] != (int)hl->par_ptr.size() - 1 || hl->par_ptr[0] != 0 ||
        hl->par_ptr.back() != nsuper || (int)hl->partition.size() != nsuper)
      throw std::invalid_argument("H-level schedule does not cover the supernodes");
    std::vector<char> seen(nsuper, 0);
    int done = -1;  // last level of the clusters placed so far
    for (int k = 0; k < ncl; ++k) {
      int cl_end = done;
      for (int q = hl->level_ptr[k]; q < hl->level_ptr[k + 1]; ++q) {
        int prev = done;
        for (int i = hl->par_ptr[q]; i < hl->par_ptr[q + 1]; ++i) {
          const int sn = hl->partition[i];
          if (sn < 0 || sn >= nsuper || seen[sn])
            throw std::invalid_argument("H-level schedule lists a supernode twice or out of range");
          seen[sn] = 1;
          prev = place(sn, prev + 1);
          cl_end = std::max(cl_end, prev);
        }
      }
      done = cl_end;
    }
  }
  P.nlevel = 0;
  for (const PTile& pt : ptiles) P.nlevel = std::max(P.nlevel, pt.level + 1);

  // ---- panels grouped by factor level (a solve-only plan has no factor levels) ----
  P.lvl_panels.assign(P.nlevel + 1, 0);
  if (factor) {
    for (int q = 0; q < P.npanel; ++q) P.lvl_panels[P.panel_level[q] + 1]++;
    for (int l = 0; l < P.nlevel; ++l) P.lvl_panels[l + 1] += P.lvl_panels[l];
    P.panel_order.resize(P.npanel);
    std::vector<int> fill(P.lvl_panels.begin(), P.lvl_panels.end() - 1);
    for (int q = 0; q < P.npanel; ++q) P.panel_order[fill[P.panel_level[q]]++] = q;
  }

  // ---- solve panels: up to SPW columns of one supernode; their own levels ----
  // A supernode's first solve panel follows the last solve panel of every
  // supernode in its pruned set (its descendants); its later ones follow it.
  {
    std::vector<int> last_sp_level(nsuper, 0);
    std::vector<int> sp_level;
    for (int s = 0; s < nsuper; ++s) {
      int first = 0;
      for (int t = prunePtr[s]; t < prunePtr[s + 1]; ++t)
        first = std::max(first, last_sp_level[pruneSet[t]] + 1);
      const int W = super[s + 1] - super[s];
      const int64_t R = pi[s + 1] - pi[s];
      int lvl = first;
      for (int c = 0; c < W; c += SPWP, ++lvl) {
        Panel sp;
        sp.vbase = px[s] + (int64_t)c * R + c;
        sp.ld = (int)R;
        sp.ncol = std::min(SPWP, W - c);
        sp.nrow = (int)R - c;
        sp.c0 = super[s] + c;
        sp.rows_off = (int)(pi[s] + c);
        sp.sn = s;
        sp.voff = voff[s];
        sp.sc0 = super[s];
        sp.sw = W;
        P.spanels.push_back(sp);
        sp_level.push_back(lvl);
      }
      last_sp_level[s] = lvl - 1;
      P.nslevel = std::max(P.nslevel, lvl);
    }
    const int nsp = (int)P.spanels.size();
    P.lvl_sp.assign(P.nslevel + 1, 0);
    for (int q = 0; q < nsp; ++q) P.lvl_sp[sp_level[q] + 1]++;
    for (int l = 0; l < P.nslevel; ++l) P.lvl_sp[l + 1] += P.lvl_sp[l];
    // within a level, narrow solve panels (<= SPW_NARROW columns, one wave each)
    // first, then the others
    P.sp_order.resize(nsp);
    P.lvl_sp_narrow.assign(P.nslevel, 0);
    for (int q = 0; q < nsp; ++q)
      if (P.spanels[q].ncol <= SPW_NARROW) P.lvl_sp_narrow[sp_level[q]]++;
    std::vector<int> fill(P.lvl_sp.begin(), P.lvl_sp.end() - 1);
    std::vector<int> fill_wide(P.nslevel);
    for (int l = 0; l < P.nslevel; ++l) fill_wide[l] = P.lvl_sp[l] + P.lvl_sp_narrow[l];
    for (int q = 0; q < nsp; ++q) {
      const int l = sp_level[q];
      P.sp_order[P.spanels[q].ncol <= SPW_NARROW ? fill[l]++ : fill_wide[l]++] = q;
    }

    // row blocks below each solve panel's diagonal block, by level
    P.lvl_fblk.assign(P.nslevel + 1, 0);
    P.lvl_fblk_narrow.assign(P.nslevel, 0);
    P.fblk_ptr.assign(nsp + 1, 0);
    for (int l = 0; l < P.nslevel; ++l) {
      for (int qi = P.lvl_sp[l]; qi < P.lvl_sp[l + 1]; ++qi) {
        if (qi == P.lvl_sp[l] + P.lvl_sp_narrow[l])
          P.lvl_fblk_narrow[l] = (int)P.fblocks.size() - P.lvl_fblk[l];
        const int q = P.sp_order[qi];
        const Panel& sp = P.spanels[q];
        P.fblk_ptr[qi] = (int)P.fblocks.size();
        if (sp.nrow == sp.ncol) {
          P.fblocks.push_back({q, sp.ncol, 0, 1});
        } else {
          for (int r0 = sp.ncol; r0 < sp.nrow; r0 += TM)
            P.fblocks.push_back({q, r0, std::min(TM, sp.nrow - r0), r0 == sp.ncol ? 1 : 0});
        }
      }
      P.lvl_fblk[l + 1] = (int)P.fblocks.size();
      if (P.lvl_sp_narrow[l] == P.lvl_sp[l + 1] - P.lvl_sp[l])
        P.lvl_fblk_narrow[l] = P.lvl_fblk[l + 1] - P.lvl_fblk[l];
      P.max_lvl_fblk = std::max(P.max_lvl_fblk, P.lvl_fblk[l + 1] - P.lvl_fblk[l]);
    }
    P.fblk_ptr[nsp] = (int)P.fblocks.size();

    // columns gathered when their supernode's first solve panel starts
    P.lvl_gcol.assign(P.nslevel + 1, 0);
    std::vector<int> first_level(nsuper);
    for (int q = nsp - 1; q >= 0; --q) first_level[P.spanels[q].sn] = sp_level[q];
    for (int sn = 0; sn < nsuper; ++sn) P.lvl_gcol[first_level[sn] + 1] += super[sn + 1] - super[sn];
    for (int l = 0; l < P.nslevel; ++l) P.lvl_gcol[l + 1] += P.lvl_gcol[l];
    P.gcols.resize(n);
    std::vector<int> gf(P.lvl_gcol.begin(), P.lvl_gcol.end() - 1);
    for (int sn = 0; sn < nsuper; ++sn) {
      int& f = gf[first_level[sn]];
      for (int c = super[sn]; c < super[sn + 1]; ++c) P.gcols[f++] = c;
    }
  }

  // ---- flatten: tiles by level (stable), tasks, TRSM lists ----
  // Within a level the most expensive tiles go first (longest-processing-time
  // order), so that the level does not end on a long tile started last.  The
  // estimate is the operand volume plus a fixed latency per task.
  std::vector<double> cost(ptiles.size(), 0.0);
  for (size_t i = 0; i < ptiles.size(); ++i)
    for (const Task& t : ptasks[i])
      cost[i] += (double)t.k * ((t.rb - t.ra) + (t.cb - t.ca)) + 2048.0;
  std::vector<int> order(ptiles.size());
  for (size_t i = 0; i < order.size(); ++i) order[i] = (int)i;
  // Tiles of panels at most 16 wide whose tasks have a short total inner
  // dimension (at most narrow_max_k) come first in their level, where the level
  // has at least narrow_min_tiles of them: they run one wave each in their own
  // launch.  The others keep the four-wave split of the inner dimension.
  std::vector<char> is_narrow(ptiles.size(), 0);
  for (size_t i = 0; i < ptiles.size(); ++i) {
    int64_t sumk = 0;
    for (const Task& t : ptasks[i]) sumk += t.k;
    is_narrow[i] = P.panels[ptiles[i].t.panel].ncol <= 16 && sumk <= params.narrow_max_k;
  }
  {
    // a separate launch only pays where a level has many of them
    std::vector<int> cnt(P.nlevel, 0);
    for (size_t i = 0; i < ptiles.size(); ++i) cnt[ptiles[i].level] += is_narrow[i];
    for (size_t i = 0; i < ptiles.size(); ++i)
      if (cnt[ptiles[i].level] < params.narrow_min_tiles) is_narrow[i] = 0;
  }
  // Panels (<= 16 columns, so single-panel supernodes, factored at the level of
  // their updates) whose tiles are all narrow are fused: one wave updates and
  // factors the whole panel, and they leave the narrow and factor launches.
  std::vector<char> fused_panel(P.npanel, 0);
  if (params.fuse_narrow) {
    std::vector<char> all_narrow(P.npanel, 1), has_diag(P.npanel, 0);
    for (size_t i = 0; i < ptiles.size(); ++i) {
      const int q = ptiles[i].t.panel;
      if (!is_narrow[i]) all_narrow[q] = 0;
      if (ptiles[i].t.kind == 0) has_diag[q] = 1;
    }
    for (int q = 0; q < P.npanel; ++q)
      fused_panel[q] = all_narrow[q] && has_diag[q] && P.panels[q].ncol <= 16 &&
                       P.panels[q].sw == P.panels[q].ncol;
  }
  auto fused = [&](int i) { return fused_panel[ptiles[i].t.panel] != 0; };
  auto narrow = [&](int i) { return is_narrow[i] != 0; };
  // the dependency-driven launch takes the levels after the last one with fused
  // or one-wave tiles (those keep their own launches)
  P.dag_level0 = P.nlevel;
  if (params.dag && factor) {
    int last = -1;
    for (size_t i = 0; i < ptiles.size(); ++i)
      if (fused((int)i) || narrow((int)i)) last = std::max(last, ptiles[i].level);
    P.dag_level0 = last + 1;
  }
  // Then, in levels where it pays, the tiles of the panels the level factors
  // before the others (later panels of a wide supernode): only the former hold
  // up the level's panel factorization, so the latter can run in the same
  // launch as it.  It pays where the former are light (a right-looking step's
  // next panel) and not where they weigh as much as the rest (a wide
  // supernode's external updates): per level, with launch durations estimated
  // as max(longest tile, total / 512 workgroups + half the longest tile) from
  // each tile's update flops (3 us + 1.4 flops / 1e5 per us and workgroup) and
  // ~20 us for the panel factorization (measured on MI355X), the level defers
  // when critical + max(factor, deferred) beats all tiles + factor by 5%.
  auto critical = [&](int i) { return ptiles[i].trsm || ptiles[i].t.kind == 0; };
  std::vector<char> defer_level(P.nlevel, 0);
  {
    std::vector<double> tus(ptiles.size(), 3.0);
    for (size_t i = 0; i < ptiles.size(); ++i)
      for (const Task& t : ptasks[i])
        tus[i] += 1.4 * 2.0 * (double)(t.rb - t.ra) * (t.cb - t.ca) * t.k / 1e5;
    std::vector<double> cmax(P.nlevel, 0), csum(P.nlevel, 0), dmax(P.nlevel, 0), dsum(P.nlevel, 0);
    std::vector<char> has_blk(P.nlevel, 0);
    for (size_t i = 0; i < ptiles.size(); ++i) {
      const int l = ptiles[i].level;
      if (fused((int)i) || narrow((int)i)) continue;
      if (critical((int)i)) {
        cmax[l] = std::max(cmax[l], tus[i]);
        csum[l] += tus[i];
        if (P.panels[ptiles[i].t.panel].ncol > 16) has_blk[l] = 1;
      } else {
        dmax[l] = std::max(dmax[l], tus[i]);
        dsum[l] += tus[i];
      }
    }
    for (int l = 0; l < P.nlevel; ++l) {
      if (dsum[l] == 0 || !has_blk[l]) continue;
      auto span = [](double mx, double sum) { return std::max(mx, sum / 512 + 0.5 * mx); };
      const double tc = span(cmax[l], csum[l]), td = span(dmax[l], dsum[l]);
      const double tcd = span(std::max(cmax[l], dmax[l]), csum[l] + dsum[l]);
      const double tf = 20.0;
      defer_level[l] = tc + std::max(tf, td) < 0.95 * (tcd + tf);
    }
  }
  std::stable_sort(order.begin(), order.end(), [&](int a, int b) {
    if (ptiles[a].level != ptiles[b].level) return ptiles[a].level < ptiles[b].level;
    if (fused(a) != fused(b)) return fused(a);
    if (fused(a)) {  // grouped by panel, the diagonal tile first
      if (ptiles[a].t.panel != ptiles[b].t.panel) return ptiles[a].t.panel < ptiles[b].t.panel;
      return ptiles[a].t.r0 < ptiles[b].t.r0;
    }
    if (narrow(a) != narrow(b)) return narrow(a);
    if (!narrow(a) && defer_level[ptiles[a].level] && critical(a) != critical(b))
      return critical(a);
    return cost[a] > cost[b];
  });
  P.lvl_tiles_fused.assign(P.nlevel, 0);
  P.lvl_tiles_narrow.assign(P.nlevel, 0);
  P.lvl_tiles_deferred.assign(P.nlevel, 0);
  for (size_t i = 0; i < ptiles.size(); ++i) {
    if (fused((int)i)) P.lvl_tiles_fused[ptiles[i].level]++;
    else if (is_narrow[i]) P.lvl_tiles_narrow[ptiles[i].level]++;
    else if (defer_level[ptiles[i].level] && !critical((int)i))
      P.lvl_tiles_deferred[ptiles[i].level]++;
  }
  P.lvl_fused.assign(P.nlevel + 1, 0);
  P.lvl_tiles.assign(P.nlevel + 1, 0);
  P.lvl_trsm.assign(P.nlevel + 1, 0);
  P.lvl_tile_flops.assign(P.nlevel, 0.0);
  std::vector<std::vector<int>> trsm_by_level(P.nlevel);
  std::vector<int> task_src;  // per task of P.tasks: the last source panel it reads
  std::vector<int> tile_level;
  std::vector<double> tile_flops;  // per tile: useful flops of its updates
  for (int oi : order) {
    PTile& pt = ptiles[oi];
    pt.t.task_begin = (int)P.tasks.size();
    tile_level.push_back(pt.level);
    double tile_fl = 0;
    task_src.insert(task_src.end(), psrc[oi].begin(), psrc[oi].end());
    const Panel& pn = P.panels[pt.t.panel];
    const int* trow = Ls + pn.rows_off + pt.t.r0;
    for (Task t : ptasks[oi]) {
      // landing masks: source rows merged against the tile's sorted rows
      const int* srows = Ls + t.rows_off;
      t.rmask = 0;
      for (int e = t.ra, i = 0; e < t.rb; ++e) {
        while (i < pt.t.nr && trow[i] < srows[e]) ++i;
        if (i == pt.t.nr || trow[i] != srows[e]) throw std::logic_error("source row not in tile");
        t.rmask |= 1ull << i;
      }
      t.cmask = 0;
      for (int e = t.ca; e < t.cb; ++e) {
        const int j = srows[e] - pn.c0;
        if (j < 0 || j >= pn.ncol) throw std::logic_error("source row not in panel");
        t.cmask |= 1ull << j;
      }
      P.tasks.push_back(t);
      P.update_flops += 2.0 * (t.rb - t.ra) * (double)(t.cb - t.ca) * t.k;
      // useful products: pairs (row, col) with row >= col (both sorted)
      const int* rows = Ls + t.rows_off;
      double pairs = 0;
      for (int c = t.ca, r = t.ra; c < t.cb; ++c) {
        while (r < t.rb && rows[r] < rows[c]) ++r;
        pairs += t.rb - r;
      }
      P.tile_kernel_flops += 2.0 * pairs * t.k;
      P.lvl_tile_flops[pt.level] += 2.0 * pairs * t.k;
      tile_fl += 2.0 * pairs * t.k;
    }
    tile_flops.push_back(tile_fl);
    pt.t.task_end = (int)P.tasks.size();
    P.lvl_tiles[pt.level + 1]++;
    if (fused(oi)) {
      if (pt.t.kind == 0) {
        P.fused.push_back({(int)P.tiles.size(), 0});
        P.lvl_fused[pt.level + 1]++;
      }
      P.fused.back().ntile++;
    } else if (pt.trsm || pt.t.kind == 0) {
      trsm_by_level[pt.level].push_back((int)P.tiles.size());
    }
    P.tiles.push_back(pt.t);
  }
  for (int l = 0; l < P.nlevel; ++l) P.lvl_fused[l + 1] += P.lvl_fused[l];
  for (int l = 0; l < P.nlevel; ++l) P.lvl_tiles[l + 1] += P.lvl_tiles[l];
  // scratch slots for the diagonal blocks each level factors: the update kernel
  // copies the diagonal tile's block there, every factor-kernel tile reads it.
  // The levels of the dependency-driven launch are in flight together, so their
  // slots do not overlap.
  int64_t off = 0;
  for (int l = 0; l < P.nlevel; ++l) {
    if (l <= P.dag_level0) off = 0;
    std::vector<std::pair<int, int>> diag_slots;  // panel, slot
    for (int ti : trsm_by_level[l]) {
      Tile& tl = P.tiles[ti];
      if (tl.kind == 0) {
        const int nc = P.panels[tl.panel].ncol;
        tl.slot = (int)off;
        diag_slots.push_back({tl.panel, (int)off});
        off += (int64_t)nc * nc;
      }
    }
    std::sort(diag_slots.begin(), diag_slots.end());
    for (int ti : trsm_by_level[l]) {
      Tile& tl = P.tiles[ti];
      if (tl.kind == 0) continue;
      auto it = std::lower_bound(diag_slots.begin(), diag_slots.end(), std::make_pair(tl.panel, -1));
      if (it == diag_slots.end() || it->first != tl.panel)
        throw std::logic_error("panel factored without its diagonal tile");
      tl.slot = it->second;
    }
    if (off > 0x7fffffff) throw std::invalid_argument("diagonal scratch too large");
    P.diag_scratch = std::max(P.diag_scratch, off);
  }
  std::vector<char> is_factor(P.tiles.size(), 0);
  for (int l = 0; l < P.nlevel; ++l)
    for (int ti : trsm_by_level[l]) is_factor[ti] = 1;
  // each level's factor list grouped by panel width class (one launch per class)
  P.lvl_trsm_cls.assign((size_t)4 * P.nlevel, 0);
  auto width_class = [&](int ti) {
    const int nc = P.panels[P.tiles[ti].panel].ncol;
    return nc <= 16 ? 0 : (nc <= 32 ? 1 : 2);
  };
  for (int l = 0; l < P.nlevel; ++l) {
    std::vector<int>& v = trsm_by_level[l];
    std::stable_sort(v.begin(), v.end(),
                     [&](int a, int b) { return width_class(a) < width_class(b); });
    int* cls = &P.lvl_trsm_cls[(size_t)4 * l];
    const int base = (int)P.trsm_tiles.size();
    int counts[3] = {0, 0, 0};
    for (int ti : v) counts[width_class(ti)]++;
    cls[0] = base;
    cls[1] = cls[0] + counts[0];
    cls[2] = cls[1] + counts[1];
    cls[3] = cls[2] + counts[2];
    P.trsm_tiles.insert(P.trsm_tiles.end(), v.begin(), v.end());
    P.lvl_trsm[l + 1] = (int)P.trsm_tiles.size();
  }
  P.ftiles.reserve(P.trsm_tiles.size());
  for (int ti : P.trsm_tiles) P.ftiles.push_back(P.tiles[ti]);
  if (P.dag_level0 < P.nlevel)
    build_dag(P, task_src, tile_level, tile_flops, is_factor, params.dag);
  for (const Panel& pn : P.panels) {
    const double w = pn.ncol, m = pn.nrow - pn.ncol;
    P.factor_flops += w * w * w / 3.0 + m * w * w;
    P.trsm_kernel_flops += w * w * w / 3.0 + m * w * w;  // POTRF + TRSM of the panel
  }
}

}  // namespace swc

// ---- host-only planning introspection (C ABI) ----
#include "capi_common.h"

struct swc_plan {
  swc::Plan P;
};

extern "C" {

int swc_plan_build(int n, const int* blockSet, int supNo, const size_t* lC, const int* Ls,
                   const size_t* Li_ptr, const int* prunePtr, const int* pruneSet,
                   swc_plan** out) {
  return swc_plan_build_params(n, blockSet, supNo, lC, Ls, Li_ptr, prunePtr, pruneSet, nullptr, 0,
                               0, nullptr, nullptr, nullptr, out);
}

int swc_plan_build_params(int n, const int* blockSet, int supNo, const size_t* lC, const int* Ls,
                          const size_t* Li_ptr, const int* prunePtr, const int* pruneSet,
                          const int* params, int nparams, int nLevels, const int* levelPtr,
                          const int* parPtr, const int* partition, swc_plan** out) {
  swc::PlanParams prm;
  if (params && !swc::plan_params_from(params, nparams, prm))
    return swc_set_error(SWC_EINVAL, "swc_plan_build: plan parameter out of range");
  swc::HLevelSchedule hl;
  if (nLevels > 0) {
    if (!levelPtr || !parPtr || !partition || levelPtr[0] != 0 || levelPtr[nLevels] < 0 ||
        parPtr[0] != 0 || parPtr[levelPtr[nLevels]] < 0)
      return swc_set_error(SWC_EINVAL, "swc_plan_build: bad H-level schedule");
    hl.level_ptr.assign(levelPtr, levelPtr + nLevels + 1);
    hl.par_ptr.assign(parPtr, parPtr + levelPtr[nLevels] + 1);
    hl.partition.assign(partition, partition + parPtr[levelPtr[nLevels]]);
  }
  // pruneSet may be null when it is empty (no supernode has descendants); with
  // prunePtr null both are derived from the row lists and only the solve part of
  // the plan is built (what blockedLsolve does)
  if (!out || !blockSet || !lC || !Ls || !Li_ptr || supNo <= 0 ||
      (prunePtr && !pruneSet && prunePtr[supNo] != 0))
    return swc_set_error(SWC_EINVAL, "swc_plan_build: bad argument");
  *out = nullptr;
  try {
    std::vector<int64_t> pi(supNo + 1), px(supNo);
    for (int s = 0; s <= supNo; ++s) pi[s] = (int64_t)Li_ptr[blockSet[s]];
    for (int s = 0; s < supNo; ++s) px[s] = (int64_t)lC[blockSet[s]];
    std::vector<int> pp, ps;
    const bool factor = prunePtr != nullptr;
    if (!factor) {
      swc::descendant_sets(n, supNo, blockSet, pi.data(), Ls, pp, ps);
      prunePtr = pp.data();
      pruneSet = ps.data();
    }
    auto* h = new swc_plan;
    swc::build_plan(n, supNo, blockSet, pi.data(), px.data(), Ls, prunePtr, pruneSet, h->P, factor,
                    prm, nLevels > 0 ? &hl : nullptr);
    *out = h;
    return SWC_OK;
  } catch (const std::exception& e) {
    return swc_set_error(SWC_EINVAL, std::string("swc_plan_build: ") + e.what());
  }
}

void swc_plan_free(swc_plan* p) { delete p; }

int swc_plan_struct_sizes(int64_t* out) {
  if (!out) return swc_set_error(SWC_EINVAL, "swc_plan_struct_sizes: null argument");
  out[0] = sizeof(swc::Panel);
  out[1] = sizeof(swc::Tile);
  out[2] = sizeof(swc::Task);
  out[3] = sizeof(swc::SolveBlock);
  return SWC_OK;
}

int swc_plan_sizes(const swc_plan* p, int64_t* out, double* flops) {
  if (!p || !out) return swc_set_error(SWC_EINVAL, "swc_plan_sizes: null argument");
  const swc::Plan& P = p->P;
  out[0] = P.npanel;
  out[1] = (int64_t)P.tiles.size();
  out[2] = (int64_t)P.tasks.size();
  out[3] = P.nlevel;
  out[4] = P.n;
  out[5] = P.xsize;
  out[6] = (int64_t)P.trsm_tiles.size();
  out[7] = (int64_t)P.fblocks.size();
  out[8] = P.vsize;
  out[9] = (int64_t)P.spanels.size();
  out[10] = P.diag_scratch;
  out[11] = P.nslevel;
  out[12] = (int64_t)P.fused.size();
  out[13] = (int64_t)P.dag_items.size();
  out[14] = P.dag_level0;
  out[15] = (int64_t)P.dag_deps.size();
  if (flops) {
    flops[0] = P.update_flops;
    flops[1] = P.factor_flops;
    flops[2] = P.dag_flops;
    flops[3] = P.dag_sim_us;
    flops[4] = P.lvl_sim_us;
  }
  return SWC_OK;
}

int swc_plan_export(const swc_plan* p, int which, void* dst) {
  if (!p || !dst) return swc_set_error(SWC_EINVAL, "swc_plan_export: null argument");
  const swc::Plan& P = p->P;
  auto cp = [&](const auto& v) {
    if (!v.empty()) std::memcpy(dst, v.data(), v.size() * sizeof(v[0]));
    return SWC_OK;
  };
  switch (which) {
    case 0: return cp(P.panels);
    case 1: return cp(P.tiles);
    case 2: return cp(P.tasks);
    case 3: return cp(P.lvl_tiles);
    case 4: return cp(P.lvl_trsm);
    case 5: return cp(P.panel_order);
    case 6: return cp(P.lvl_panels);
    case 7: return cp(P.gptr);
    case 8: return cp(P.gidx);
    case 9: return cp(P.trsm_tiles);
    case 17: return cp(P.fused);
    case 18: return cp(P.lvl_fused);
    case 19: return cp(P.lvl_tiles_fused);
    case 20: return cp(P.lvl_tiles_narrow);
    case 21: return cp(P.dag_items);
    case 22: return cp(P.dag_deps);
    case 23: return cp(P.lvl_tiles_deferred);
    case 10: return cp(P.fblocks);
    case 11: return cp(P.lvl_fblk);
    case 12: return cp(P.gcols);
    case 13: return cp(P.lvl_gcol);
    case 14: return cp(P.spanels);
    case 15: return cp(P.sp_order);
    case 16: return cp(P.lvl_sp);
    default: return swc_set_error(SWC_EINVAL, "swc_plan_export: unknown array id");
  }
}

}  // extern "C"
