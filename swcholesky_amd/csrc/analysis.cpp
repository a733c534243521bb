// Host symbolic analysis: a clean-room restatement of the reference's analyze_p2
// (software/swCholesky/cholesky/sw_LSparsity.h:248-702) for a GIVEN ordering.
//
// The numeric phase consumes exactly what that routine produces: the final
// permutation, the relaxed supernode partition, the supernodal row lists, the
// per-column pointers and the pruned descendant sets.  Each stage below is
// written from the published algorithm it implements; where the reference fixes
// an order or a tie-break that the result depends on, the comment names it.
//
//   stage                         reference
//   tril/triu of P A P^T          Transpose.h:554-719 (ptranspose), sorted rows
//   elimination tree              Etree.h:56 (Liu, path compression)
//   column counts                 ColumnCount.h:141-500 (Gilbert-Ng-Peyton)
//   weighted postorder            PostOrder.h:11-160, sw_LSparsity.h:526-579
//   relaxed supernodes            Inspection_BlockC.h:116-504 (nrelax/zrelax rule)
//   supernodal row lists          Inspection_BlockC.h:631-752 (row subtrees)
//   BCSC pointers                 sw_LSparsity.h:608-638
//   pruned descendant sets        Inspection_Prune.h:24-56, Reach.h:121-152
#include "analysis.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <climits>
#include <new>
#include <stdexcept>

namespace swc {

namespace {

// tril(P A P^T) with rows sorted inside each column, plus for every output entry
// the index of the input entry it came from.  perm[new] = old.
void permuted_lower(int n, const int* ap, const int* ai, const int* pinv,
                    std::vector<int>& cp, std::vector<int>& ci, std::vector<int64_t>& src) {
  const int64_t nnz = ap[n];
  std::vector<int> rr(nnz), cc(nnz);
  for (int j = 0; j < n; ++j) {
    for (int64_t q = ap[j]; q < ap[j + 1]; ++q) {
      int a = pinv[ai[q]], b = pinv[j];
      rr[q] = std::max(a, b);
      cc[q] = std::min(a, b);
    }
  }
  // counting sort by row, then a stable counting sort by column
  std::vector<int64_t> cnt(n + 1, 0);
  for (int64_t q = 0; q < nnz; ++q) cnt[rr[q] + 1]++;
  for (int i = 0; i < n; ++i) cnt[i + 1] += cnt[i];
  std::vector<int64_t> byrow(nnz);
  for (int64_t q = 0; q < nnz; ++q) byrow[cnt[rr[q]]++] = q;
  cp.assign(n + 1, 0);
  for (int64_t q = 0; q < nnz; ++q) cp[cc[q] + 1]++;
  for (int j = 0; j < n; ++j) cp[j + 1] += cp[j];
  std::vector<int> next(cp.begin(), cp.end() - 1);
  ci.resize(nnz);
  src.resize(nnz);
  for (int64_t t = 0; t < nnz; ++t) {
    int64_t q = byrow[t];
    int64_t dst = next[cc[q]]++;
    ci[dst] = rr[q];
    src[dst] = q;
  }
  for (int j = 0; j < n; ++j) {
    for (int64_t q = cp[j] + 1; q < cp[j + 1]; ++q) {
      if (ci[q] == ci[q - 1]) throw std::invalid_argument("duplicate entry in A");
    }
  }
}

// Transpose of a sorted lower CSC: the upper triangle by columns, rows sorted.
void transpose_pattern(int n, const std::vector<int>& cp, const std::vector<int>& ci,
                       std::vector<int>& up, std::vector<int>& ui) {
  up.assign(n + 1, 0);
  for (int v : ci) up[v + 1]++;
  for (int i = 0; i < n; ++i) up[i + 1] += up[i];
  std::vector<int> next(up.begin(), up.end() - 1);
  ui.resize(ci.size());
  for (int j = 0; j < n; ++j)
    for (int q = cp[j]; q < cp[j + 1]; ++q) ui[next[ci[q]]++] = j;
}

// Upper triangle of P A P^T by columns, with the entries of each column in the
// order a one-pass permuted symmetric transpose leaves them (Transpose.h:83-109):
// new columns j are visited in order; each lower entry (iold >= jold) of input
// column jold = perm[j] lands in column max(i,j) as row min(i,j).  Column c thus
// holds first the rows j < c with perm[j] < perm[c] (ascending j), then c itself,
// then the rows j < c with perm[j] > perm[c] in ascending perm[j].  The pruned
// descendant sets are listed in the order this column order induces.
void permuted_upper_unsorted(int n, const int* ap, const int* ai, const std::vector<int>& perm,
                             const std::vector<int>& pinv, std::vector<int>& up,
                             std::vector<int>& ui) {
  up.assign(n + 1, 0);
  for (int j = 0; j < n; ++j) {
    const int jold = perm[j];
    for (int q = ap[jold]; q < ap[jold + 1]; ++q) up[std::max(pinv[ai[q]], j) + 1]++;
  }
  for (int i = 0; i < n; ++i) up[i + 1] += up[i];
  std::vector<int> next(up.begin(), up.end() - 1);
  ui.resize(up[n]);
  for (int j = 0; j < n; ++j) {
    const int jold = perm[j];
    for (int q = ap[jold]; q < ap[jold + 1]; ++q) {
      const int i = pinv[ai[q]];
      if (i > j) ui[next[i]++] = j;
      else ui[next[j]++] = i;
    }
  }
}

// Elimination tree of a symmetric matrix from its upper triangle (by columns):
// parent[i] = min{ k > i : L(k,i) != 0 }, via Liu's ancestor path compression.
void etree_upper(int n, const std::vector<int>& up, const std::vector<int>& ui,
                 std::vector<int>& parent) {
  parent.assign(n, -1);
  std::vector<int> anc(n, -1);
  for (int k = 0; k < n; ++k) {
    for (int q = up[k]; q < up[k + 1]; ++q) {
      int i = ui[q];
      while (i != -1 && i < k) {
        int nxt = anc[i];
        anc[i] = k;
        if (nxt == -1) { parent[i] = k; break; }
        i = nxt;
      }
    }
  }
}

// Depth-first postorder of a forest.  Children are visited in the order of the
// `kids` lists; roots in ascending index.  Iterative, with an explicit stack.
void postorder_lists(int n, const std::vector<int>& parent, const std::vector<int>& head,
                     const std::vector<int>& next, std::vector<int>& post) {
  post.resize(n);
  std::vector<int> hd(head), stack(n);
  int k = 0;
  for (int r = 0; r < n; ++r) {
    if (parent[r] != -1) continue;
    int top = 0;
    stack[0] = r;
    while (top >= 0) {
      int v = stack[top];
      int c = hd[v];
      if (c == -1) {
        post[k++] = v;
        --top;
      } else {
        hd[v] = next[c];
        stack[++top] = c;
      }
    }
  }
  if (k != n) throw std::runtime_error("elimination tree is not a forest");
}

// Unweighted postorder: children in ascending index.
void postorder_plain(int n, const std::vector<int>& parent, std::vector<int>& post) {
  std::vector<int> head(n, -1), next(n, -1);
  for (int j = n - 1; j >= 0; --j) {
    int p = parent[j];
    if (p >= 0) { next[j] = head[p]; head[p] = j; }
  }
  postorder_lists(n, parent, head, next, post);
}

// Postorder with children sorted by ascending weight, ties by ascending index
// (the lightest child is ordered first, the heaviest last), weights clamped to
// [0, n-1]; this is the tie-break of PostOrder.h:96-137.
void postorder_weighted(int n, const std::vector<int>& parent, const std::vector<int>& weight,
                        std::vector<int>& post) {
  std::vector<int> bucket_head(n, -1), bucket_next(n, -1);
  // bucket lists hold nodes in descending index (insert at head, ascending scan)
  for (int j = 0; j < n; ++j) {
    if (parent[j] < 0) continue;
    int w = std::min(std::max(weight[j], 0), n - 1);
    bucket_next[j] = bucket_head[w];
    bucket_head[w] = j;
  }
  std::vector<int> head(n, -1), next(n, -1);
  // heaviest buckets first, each bucket in descending index, pushed at the head of
  // the parent's list: the list ends up ascending by (weight, index).
  for (int w = n - 1; w >= 0; --w) {
    for (int j = bucket_head[w]; j != -1; j = bucket_next[j]) {
      int p = parent[j];
      next[j] = head[p];
      head[p] = j;
    }
  }
  postorder_lists(n, parent, head, next, post);
}

// Column counts of L (diagonal included) by the skeleton-matrix / row-subtree
// leaf method of Gilbert, Ng and Peyton (1994), from the lower triangle.
void column_counts(int n, const std::vector<int>& lp, const std::vector<int>& li,
                   const std::vector<int>& parent, const std::vector<int>& post,
                   std::vector<int>& count) {
  std::vector<int> first(n, -1), maxfirst(n, -1), prevleaf(n, -1), anc(n);
  count.assign(n, 0);
  std::vector<int>& delta = count;
  for (int k = 0; k < n; ++k) {
    int j = post[k];
    delta[j] = (first[j] == -1) ? 1 : 0;  // a leaf starts with its diagonal
    for (; j != -1 && first[j] == -1; j = parent[j]) first[j] = k;
  }
  for (int i = 0; i < n; ++i) anc[i] = i;
  for (int k = 0; k < n; ++k) {
    int j = post[k];
    if (parent[j] != -1) delta[parent[j]]--;
    for (int q = lp[j]; q < lp[j + 1]; ++q) {
      int i = li[q];
      if (i <= j || first[j] <= maxfirst[i]) continue;  // j is not a leaf of T_i
      maxfirst[i] = first[j];
      int jprev = prevleaf[i];
      prevleaf[i] = j;
      delta[j]++;  // A(i,j) is in the skeleton
      if (jprev != -1) {
        // least common ancestor of jprev and j, with path compression
        int q2 = jprev;
        while (q2 != anc[q2]) q2 = anc[q2];
        for (int s = jprev; s != q2;) {
          int sp = anc[s];
          anc[s] = q2;
          s = sp;
        }
        delta[q2]--;
      }
    }
    if (parent[j] != -1) anc[j] = parent[j];
  }
  for (int j = 0; j < n; ++j)
    if (parent[j] != -1) count[parent[j]] += count[j];
}

}  // namespace

void analyze(int n, const int* ap, const int* ai, const int64_t* perm_in, const int nrelax[3],
             const double zrelax[3], Symbolic& S) {
  if (n <= 0) throw std::invalid_argument("n must be positive");
  S.n = n;
  std::vector<int> perm(n), pinv(n, -1);
  for (int k = 0; k < n; ++k) {
    int64_t o = perm_in[k];
    if (o < 0 || o >= n || pinv[o] != -1) throw std::invalid_argument("perm is not a bijection");
    perm[k] = (int)o;
    pinv[o] = k;
  }
  for (int j = 0; j < n; ++j) {
    for (int q = ap[j]; q < ap[j + 1]; ++q) {
      if (ai[q] < j || ai[q] >= n) throw std::invalid_argument("A must be lower triangular");
    }
  }

  // ---- etree, postorder and column counts of P A P^T (analyze_ordering) ----
  std::vector<int> lp, li, up, ui, parent, post, colcount;
  std::vector<int64_t> src;
  permuted_lower(n, ap, ai, pinv.data(), lp, li, src);
  transpose_pattern(n, lp, li, up, ui);
  etree_upper(n, up, ui, parent);
  postorder_plain(n, parent, post);
  column_counts(n, lp, li, parent, post, colcount);

  // ---- combine the ordering with the colcount-weighted postorder ----
  std::vector<int> wpost;
  postorder_weighted(n, parent, colcount, wpost);
  std::vector<int> invpost(n);
  for (int k = 0; k < n; ++k) invpost[wpost[k]] = k;
  S.perm.resize(n);
  S.colcount.resize(n);
  S.parent.resize(n);
  for (int k = 0; k < n; ++k) {
    S.perm[k] = perm[wpost[k]];
    S.colcount[k] = colcount[wpost[k]];
    int op = parent[wpost[k]];
    S.parent[k] = op == -1 ? -1 : invpost[op];
  }
  for (int k = 0; k < n; ++k) pinv[S.perm[k]] = k;
  permuted_lower(n, ap, ai, pinv.data(), S.a2p, S.a2i, S.a2src);
  transpose_pattern(n, S.a2p, S.a2i, up, ui);
  S.flops = 0;
  for (int j = 0; j < n; ++j) S.flops += (double)S.colcount[j] * (double)S.colcount[j];

  // ---- fundamental supernodes ----
  const std::vector<int>& Parent = S.parent;
  const std::vector<int>& CC = S.colcount;
  std::vector<int> nchild(n, 0);
  for (int j = 0; j < n; ++j)
    if (Parent[j] != -1) nchild[Parent[j]]++;
  std::vector<int> fsuper;
  fsuper.push_back(0);
  for (int j = 1; j < n; ++j) {
    // j starts a new supernode unless it is the only child-chain continuation of
    // j-1 with a nested pattern
    if (Parent[j - 1] != j || CC[j - 1] != CC[j] + 1 || nchild[j] > 1) fsuper.push_back(j);
  }
  const int nf = (int)fsuper.size();
  fsuper.push_back(n);
  std::vector<int> smap(n);
  for (int s = 0; s < nf; ++s)
    for (int k = fsuper[s]; k < fsuper[s + 1]; ++k) smap[k] = s;
  std::vector<int> fparent(nf);
  for (int s = 0; s < nf; ++s) {
    int p = Parent[fsuper[s + 1] - 1];
    fparent[s] = p == -1 ? -1 : smap[p];
  }

  // ---- relaxed amalgamation (merge s into s+1 when s+1 is its parent) ----
  std::vector<int> merged(nf, -1), nscol(nf), zeros(nf, 0), snz(nf);
  for (int s = 0; s < nf; ++s) {
    nscol[s] = fsuper[s + 1] - fsuper[s];
    snz[s] = CC[fsuper[s]];
  }
  for (int s = nf - 2; s >= 0; --s) {
    if (fparent[s] == -1) continue;
    // follow merges to the current representative of the parent, compressing
    int rep = fparent[s];
    while (merged[rep] != -1) rep = merged[rep];
    for (int t = fparent[s]; merged[t] != -1;) {
      int nx = merged[t];
      merged[t] = rep;
      t = nx;
    }
    if (rep != s + 1) continue;
    const int c0 = nscol[s], c1 = nscol[s + 1], ns = c0 + c1;
    int totzeros = zeros[s + 1];
    const double lnz1 = (double)snz[s + 1];
    bool merge;
    if (ns <= nrelax[0]) {
      merge = true;  // small enough: merge unconditionally, zero count unchanged
    } else {
      const double lnz0 = (double)snz[s];
      const double newz = (double)c0 * (lnz1 + c0 - lnz0);
      if (newz == 0) {
        merge = true;
      } else {
        const double tz = (double)totzeros + newz;
        const double xns = (double)ns;
        const double tsize = xns * (xns + 1) / 2 + xns * (lnz1 - c1);
        const double z = tz / tsize;
        totzeros += c0 * (snz[s + 1] + c0 - snz[s]);
        merge = ((ns <= nrelax[1] && z < zrelax[0]) || (ns <= nrelax[2] && z < zrelax[1]) ||
                 (z < zrelax[2])) &&
                (tsize < (double)INT_MAX / sizeof(double));
      }
    }
    if (merge) {
      zeros[s] = totzeros;
      merged[s + 1] = s;
      snz[s] = c0 + snz[s + 1];
      nscol[s] += nscol[s + 1];
    }
  }
  S.super.clear();
  std::vector<int> rsnz;
  for (int s = 0; s < nf; ++s) {
    if (merged[s] == -1) {
      S.super.push_back(fsuper[s]);
      rsnz.push_back(snz[s]);
    }
  }
  const int ns = (int)S.super.size();
  S.nsuper = ns;
  S.super.push_back(n);
  S.col2sup.resize(n);
  for (int s = 0; s < ns; ++s)
    for (int k = S.super[s]; k < S.super[s + 1]; ++k) S.col2sup[k] = s;
  S.sparent.resize(ns);
  for (int s = 0; s < ns; ++s) {
    int p = Parent[S.super[s + 1] - 1];
    S.sparent[s] = p == -1 ? -1 : S.col2sup[p];
  }

  // ---- supernodal row lists from the row subtrees of triu(P A P^T) ----
  S.pi.assign(ns + 1, 0);
  for (int s = 0; s < ns; ++s) S.pi[s + 1] = S.pi[s] + rsnz[s];
  S.ssize = S.pi[ns];
  S.s.assign(S.ssize, 0);
  std::vector<int64_t> fill(S.pi.begin(), S.pi.end() - 1);
  std::vector<int> flag(ns, -1);
  int stamp = 0;
  for (int s = 0; s < ns; ++s) {
    const int k1 = S.super[s], k2 = S.super[s + 1];
    for (int k = k1; k < k2; ++k) S.s[fill[s]++] = k;
    for (int k = k1; k < k2; ++k) {
      ++stamp;
      flag[s] = stamp;
      for (int q = up[k]; q < up[k + 1]; ++q) {
        int i = ui[q];
        if (i >= k1) break;  // rows are sorted; only A(0:k1-1, k)
        for (int t = S.col2sup[i]; t != -1 && flag[t] != stamp; t = S.sparent[t]) {
          if (fill[t] >= S.pi[t + 1]) throw std::runtime_error("supernode row count mismatch");
          S.s[fill[t]++] = k;
          flag[t] = stamp;
        }
      }
    }
  }
  for (int s = 0; s < ns; ++s)
    if (fill[s] != S.pi[s + 1]) throw std::runtime_error("supernode row count mismatch");

  // ---- per-column pointers (Li_ptr and lC of the numeric interface) ----
  S.i_ptr.assign(n + 1, 0);
  S.p.assign(n + 1, 0);
  S.max_sup_wid = 0;
  S.max_col = 0;
  S.xsize = 0;
  int64_t col_len = 0;
  for (int s = 0; s < ns; ++s) {
    const int c0 = S.super[s], c1 = S.super[s + 1];
    col_len = S.pi[s + 1] - S.pi[s];
    S.max_sup_wid = std::max(S.max_sup_wid, c1 - c0);
    S.max_col = std::max<int64_t>(S.max_col, col_len);
    for (int c = c0; c < c1; ++c) {
      S.i_ptr[c] = S.pi[s];
      S.p[c + 1] = S.p[c0] + (int64_t)(c + 1 - c0) * col_len;
    }
    S.xsize += (int64_t)(c1 - c0) * col_len;
  }
  // the reference leaves the final entries as i_ptr[n] = pi[nsuper] and
  // p[n] = p[n-1] + (row count of the last supernode) (sw_LSparsity.h:637-638)
  S.i_ptr[n] = S.pi[ns];
  S.p[n] = S.p[n - 1] + col_len;

  // ---- pruned descendant sets: supernodal ereach, in its stack order ----
  permuted_upper_unsorted(n, ap, ai, S.perm, pinv, up, ui);
  S.pruneptr.assign(ns + 1, 0);
  S.pruneset.clear();
  std::vector<int> stack(ns), path(ns), seen(ns, -1);
  for (int s = 0; s < ns; ++s) {
    int top = ns;
    seen[s] = s;
    for (int k = S.super[s]; k < S.super[s + 1]; ++k) {
      for (int q = up[k]; q < up[k + 1]; ++q) {
        int len = 0;
        for (int t = S.col2sup[ui[q]]; t != -1 && seen[t] != s; t = S.sparent[t]) {
          path[len++] = t;
          seen[t] = s;
        }
        while (len > 0) stack[--top] = path[--len];
      }
    }
    S.pruneset.insert(S.pruneset.end(), stack.begin() + top, stack.end());
    S.pruneptr[s + 1] = (int)S.pruneset.size();
  }
}

}  // namespace swc
