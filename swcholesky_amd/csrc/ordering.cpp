// Nested-dissection fill-reducing ordering of a general sparse symmetric
// pattern: the ordering the reference takes from METIS_NodeND
// (cholesky/LSparsity.h:597, x86 path) or reads from a perm file computed
// elsewhere (software/README.md:45), restated as a self-contained heuristic so
// that any operator can be ordered where METIS is absent.
//
// A part of the graph is split by a vertex separator taken from a breadth-first
// level structure rooted at a pseudo-peripheral vertex (George and Liu's
// level-structure bisection): among the levels that leave 35-65% of the part's
// vertices before them, the smallest.  Separator vertices that touch only one
// side are then moved to that side.  Both sides are ordered recursively (each
// connected component on its own), then the separator; parts of at most
// kLeaf vertices are ordered by reverse Cuthill-McKee.
#include <algorithm>
#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

#include "capi_common.h"

namespace swc {
namespace {

constexpr int kLeaf = 64;

struct Graph {
  std::vector<int64_t> ptr;
  std::vector<int> adj;
  int n = 0;
};

// Full adjacency (no self loops) from the lower triangle in CSC.
Graph from_lower(int n, const int* Ap, const int* Ai) {
  Graph g;
  g.n = n;
  std::vector<int64_t> deg(n + 1, 0);
  for (int j = 0; j < n; ++j)
    for (int q = Ap[j]; q < Ap[j + 1]; ++q) {
      const int i = Ai[q];
      if (i < 0 || i >= n) throw std::invalid_argument("row index out of range");
      if (i == j) continue;
      deg[i + 1]++;
      deg[j + 1]++;
    }
  for (int i = 0; i < n; ++i) deg[i + 1] += deg[i];
  g.ptr = deg;
  g.adj.resize(g.ptr[n]);
  std::vector<int64_t> fill(g.ptr.begin(), g.ptr.end() - 1);
  for (int j = 0; j < n; ++j)
    for (int q = Ap[j]; q < Ap[j + 1]; ++q) {
      const int i = Ai[q];
      if (i == j) continue;
      g.adj[fill[i]++] = j;
      g.adj[fill[j]++] = i;
    }
  return g;
}

class Dissector {
 public:
  Dissector(const Graph& g) : g_(g), part_(g.n, -1), level_(g.n, -1) {}

  // Orders the vertices `verts` (all with part_ == id) into out, separators last.
  void order(std::vector<int>& verts, int id, std::vector<int>& out) {
    if (verts.empty()) return;
    if ((int)verts.size() <= kLeaf) {
      rcm(verts, id, out);
      return;
    }
    // connected components first: each is dissected on its own
    std::vector<std::vector<int>> comps = components(verts, id);
    if (comps.size() > 1) {
      for (auto& c : comps) {
        const int cid = next_id_++;
        for (int v : c) part_[v] = cid;
        order(c, cid, out);
      }
      return;
    }
    const int root = pseudo_peripheral(verts, id);
    std::vector<int> lvl_start;
    std::vector<int> lvl_verts = bfs(root, id, lvl_start);
    const int nl = (int)lvl_start.size() - 1;
    const int total = (int)verts.size();
    int best = -1;
    for (int l = 1; l + 1 < nl; ++l) {
      const int before = lvl_start[l];
      if (before < 0.35 * total || before > 0.65 * total) continue;
      if (best < 0 || lvl_start[l + 1] - lvl_start[l] < lvl_start[best + 1] - lvl_start[best])
        best = l;
    }
    if (best < 0) {
      // no level in the window (a path-like or tiny-diameter part): the level
      // closest to the middle
      int target = total / 2;
      best = 1;
      for (int l = 1; l + 1 < nl; ++l)
        if (std::abs(lvl_start[l] - target) < std::abs(lvl_start[best] - target)) best = l;
      if (nl < 3) {  // diameter too small to separate: order the part directly
        rcm(verts, id, out);
        return;
      }
    }
    const int a_id = next_id_++, b_id = next_id_++, s_id = next_id_++;
    for (int l = 0; l < nl; ++l)
      for (int k = lvl_start[l]; k < lvl_start[l + 1]; ++k)
        part_[lvl_verts[k]] = l < best ? a_id : (l == best ? s_id : b_id);
    // thin the separator: a vertex with no neighbour on one side joins the other
    std::vector<int> sep;
    for (int k = lvl_start[best]; k < lvl_start[best + 1]; ++k) {
      const int v = lvl_verts[k];
      bool na = false, nb = false;
      for (int64_t e = g_.ptr[v]; e < g_.ptr[v + 1]; ++e) {
        const int u = g_.adj[e];
        na = na || part_[u] == a_id;
        nb = nb || part_[u] == b_id;
      }
      if (!nb) part_[v] = a_id;
      else if (!na) part_[v] = b_id;
      else sep.push_back(v);
    }
    std::vector<int> va, vb;
    for (int v : lvl_verts) {
      if (part_[v] == a_id) va.push_back(v);
      else if (part_[v] == b_id) vb.push_back(v);
    }
    order(va, a_id, out);
    order(vb, b_id, out);
    for (int v : sep) out.push_back(v);
  }

  int fresh_id() { return next_id_++; }
  std::vector<int>& part() { return part_; }

 private:
  // Breadth-first level structure of the part `id` from root; returns the
  // vertices by level, lvl_start[l] the offset of level l.
  std::vector<int> bfs(int root, int id, std::vector<int>& lvl_start) {
    std::vector<int> q;
    q.push_back(root);
    level_[root] = 0;
    lvl_start.assign(1, 0);
    size_t head = 0;
    while (head < q.size()) {
      const size_t end = q.size();
      lvl_start.push_back((int)end);
      for (; head < end; ++head) {
        const int v = q[head];
        for (int64_t e = g_.ptr[v]; e < g_.ptr[v + 1]; ++e) {
          const int u = g_.adj[e];
          if (part_[u] == id && level_[u] < 0) {
            level_[u] = level_[v] + 1;
            q.push_back(u);
          }
        }
      }
    }
    for (int v : q) level_[v] = -1;
    // lvl_start ends with q.size() pushed once more by the last (empty) round
    lvl_start.back() = (int)q.size();
    return q;
  }

  int pseudo_peripheral(const std::vector<int>& verts, int id) {
    int root = verts[0];
    for (int v : verts)  // start from a vertex of least degree
      if (g_.ptr[v + 1] - g_.ptr[v] < g_.ptr[root + 1] - g_.ptr[root]) root = v;
    int depth = -1;
    for (int it = 0; it < 8; ++it) {
      std::vector<int> ls;
      std::vector<int> lv = bfs(root, id, ls);
      const int nl = (int)ls.size() - 1;
      if (nl <= depth) break;
      depth = nl;
      // next root: a vertex of least degree in the last level
      int cand = lv[ls[nl - 1]];
      for (int k = ls[nl - 1]; k < ls[nl]; ++k) {
        const int v = lv[k];
        if (g_.ptr[v + 1] - g_.ptr[v] < g_.ptr[cand + 1] - g_.ptr[cand]) cand = v;
      }
      if (cand == root) break;
      root = cand;
    }
    return root;
  }

  std::vector<std::vector<int>> components(const std::vector<int>& verts, int id) {
    std::vector<std::vector<int>> comps;
    const int mark = next_id_++;
    for (int s : verts) {
      if (part_[s] != id) continue;
      std::vector<int> c{s};
      part_[s] = mark;
      for (size_t h = 0; h < c.size(); ++h) {
        const int v = c[h];
        for (int64_t e = g_.ptr[v]; e < g_.ptr[v + 1]; ++e) {
          const int u = g_.adj[e];
          if (part_[u] == id) {
            part_[u] = mark;
            c.push_back(u);
          }
        }
      }
      comps.push_back(std::move(c));
    }
    for (auto& c : comps)
      for (int v : c) part_[v] = id;
    return comps;
  }

  // Reverse Cuthill-McKee of the part `id`, component by component.
  void rcm(const std::vector<int>& verts, int id, std::vector<int>& out) {
    const int done = next_id_++;
    std::vector<int> seq;
    std::vector<int> sorted(verts);
    std::sort(sorted.begin(), sorted.end(), [&](int a, int b) {
      const int64_t da = g_.ptr[a + 1] - g_.ptr[a], db = g_.ptr[b + 1] - g_.ptr[b];
      return da != db ? da < db : a < b;
    });
    for (int s : sorted) {
      if (part_[s] != id) continue;
      const size_t base = seq.size();
      seq.push_back(s);
      part_[s] = done;
      for (size_t h = base; h < seq.size(); ++h) {
        const int v = seq[h];
        std::vector<int> nb;
        for (int64_t e = g_.ptr[v]; e < g_.ptr[v + 1]; ++e)
          if (part_[g_.adj[e]] == id) nb.push_back(g_.adj[e]);
        std::sort(nb.begin(), nb.end(), [&](int a, int b) {
          const int64_t da = g_.ptr[a + 1] - g_.ptr[a], db = g_.ptr[b + 1] - g_.ptr[b];
          return da != db ? da < db : a < b;
        });
        for (int u : nb)
          if (part_[u] == id) {
            part_[u] = done;
            seq.push_back(u);
          }
      }
    }
    out.insert(out.end(), seq.rbegin(), seq.rend());
  }

  const Graph& g_;
  std::vector<int> part_, level_;
  int next_id_ = 1;
};

}  // namespace

void nd_order(int n, const int* Ap, const int* Ai, std::vector<int64_t>& perm) {
  Graph g = from_lower(n, Ap, Ai);
  Dissector d(g);
  std::vector<int> all(n);
  const int id = d.fresh_id();
  for (int i = 0; i < n; ++i) {
    all[i] = i;
    d.part()[i] = id;
  }
  std::vector<int> out;
  out.reserve(n);
  d.order(all, id, out);
  if ((int)out.size() != n) throw std::logic_error("ordering lost vertices");
  perm.assign(out.begin(), out.end());
}

}  // namespace swc

extern "C" {

int swc_nd_order(int n, const int* Ap, const int* Ai, int64_t* perm) {
  if (n <= 0 || !Ap || !Ai || !perm) return swc_set_error(SWC_EINVAL, "swc_nd_order: bad argument");
  try {
    std::vector<int64_t> p;
    swc::nd_order(n, Ap, Ai, p);
    std::copy(p.begin(), p.end(), perm);
    return SWC_OK;
  } catch (const std::bad_alloc&) {
    return swc_set_error(SWC_ENOMEM, "swc_nd_order: out of host memory");
  } catch (const std::exception& e) {
    return swc_set_error(SWC_EINVAL, std::string("swc_nd_order: ") + e.what());
  }
}

}  // extern "C"
