// C ABI of the host symbolic analysis (see include/swcholesky_amd.h).
#include <algorithm>
#include <atomic>
#include <cstring>
#include <exception>
#include <new>
#include <string>
#include <thread>
#include <vector>

#include "analysis.h"
#include "capi_common.h"

struct swc_symbolic {
  swc::Symbolic S;
};

extern "C" {

int swc_analyze(int n, const int* Ap, const int* Ai, const int64_t* perm, const int* nrelax,
                const double* zrelax, swc_symbolic** out) {
  if (!out || !Ap || !Ai || !perm) return swc_set_error(SWC_EINVAL, "swc_analyze: null argument");
  *out = nullptr;
  static const int kNrelax[3] = {4, 16, 48};
  static const double kZrelax[3] = {0.8, 0.1, 0.05};
  try {
    auto* h = new swc_symbolic;
    swc::analyze(n, Ap, Ai, perm, nrelax ? nrelax : kNrelax, zrelax ? zrelax : kZrelax, h->S);
    *out = h;
    return SWC_OK;
  } catch (const std::bad_alloc&) {
    return swc_set_error(SWC_ENOMEM, "swc_analyze: out of host memory");
  } catch (const std::exception& e) {
    return swc_set_error(SWC_EINVAL, std::string("swc_analyze: ") + e.what());
  }
}

// Independent instances analysed one by one on nthreads host threads, then
// concatenated as the analysis of the block-diagonal operator they form:
// instance k's columns, supernodes, row lists, values, descendant sets and
// input entries follow those of instances 0..k-1, with every index offset by
// theirs.  The result equals swc_analyze on the block-diagonal matrix (its
// elimination forest has one tree per instance, and the postorder keeps each
// tree's columns in place).
int swc_analyze_batch(int count, const int* n, const int* const* Ap, const int* const* Ai,
                      const int64_t* const* perm, const int* nrelax, const double* zrelax,
                      int nthreads, swc_symbolic** out) {
  if (!out || count <= 0 || !n || !Ap || !Ai || !perm)
    return swc_set_error(SWC_EINVAL, "swc_analyze_batch: bad argument");
  *out = nullptr;
  for (int k = 0; k < count; ++k)
    if (n[k] <= 0 || !Ap[k] || !Ai[k] || !perm[k])
      return swc_set_error(SWC_EINVAL, "swc_analyze_batch: bad instance " + std::to_string(k));
  static const int kNrelax[3] = {4, 16, 48};
  static const double kZrelax[3] = {0.8, 0.1, 0.05};
  const int* nr = nrelax ? nrelax : kNrelax;
  const double* zr = zrelax ? zrelax : kZrelax;
  try {
    std::vector<swc::Symbolic> parts(count);
    std::vector<std::exception_ptr> err(count);
    std::atomic<int> next(0);
    auto work = [&]() {
      for (int k = next++; k < count; k = next++) {
        try {
          swc::analyze(n[k], Ap[k], Ai[k], perm[k], nr, zr, parts[k]);
        } catch (...) {
          err[k] = std::current_exception();
        }
      }
    };
    int nt = nthreads > 0 ? nthreads : (int)std::max(1u, std::thread::hardware_concurrency());
    nt = std::min(nt, count);
    std::vector<std::thread> pool;
    for (int t = 1; t < nt; ++t) pool.emplace_back(work);
    work();
    for (auto& t : pool) t.join();
    for (int k = 0; k < count; ++k)
      if (err[k]) std::rethrow_exception(err[k]);

    auto* h = new swc_symbolic;
    swc::Symbolic& S = h->S;
    int64_t N = 0, NS = 0, SS = 0, X = 0, NNZ = 0, NP = 0;
    for (const auto& P : parts) {
      N += P.n; NS += P.nsuper; SS += P.ssize; X += P.xsize;
      NNZ += (int64_t)P.a2i.size(); NP += (int64_t)P.pruneset.size();
    }
    if (N > 0x7fffffff || NS > 0x7fffffff || NNZ > 0x7fffffff || NP > 0x7fffffff) {
      delete h;
      return swc_set_error(SWC_EINVAL, "swc_analyze_batch: batch too large for 32-bit indices");
    }
    S.n = (int)N;
    S.nsuper = (int)NS;
    S.ssize = SS;
    S.xsize = X;
    S.perm.reserve(N); S.colcount.reserve(N); S.parent.reserve(N); S.col2sup.reserve(N);
    S.super.reserve(NS + 1); S.sparent.reserve(NS); S.pi.reserve(NS + 1); S.s.reserve(SS);
    S.i_ptr.reserve(N + 1); S.p.reserve(N + 1); S.pruneptr.reserve(NS + 1); S.pruneset.reserve(NP);
    S.a2p.reserve(N + 1); S.a2i.reserve(NNZ); S.a2src.reserve(NNZ);
    int c0 = 0, s0 = 0;
    int64_t r0 = 0, x0 = 0, e0 = 0, q0 = 0, in0 = 0;  // rows, values, a2 entries, prune, input
    for (int k = 0; k < count; ++k) {
      const swc::Symbolic& P = parts[k];
      auto sh = [](int v, int o) { return v < 0 ? v : v + o; };
      for (int j = 0; j < P.n; ++j) {
        S.perm.push_back(P.perm[j] + c0);
        S.colcount.push_back(P.colcount[j]);
        S.parent.push_back(sh(P.parent[j], c0));
        S.col2sup.push_back(P.col2sup[j] + s0);
        S.i_ptr.push_back(P.i_ptr[j] + r0);
        S.p.push_back(P.p[j] + x0);
        S.a2p.push_back((int)(P.a2p[j] + e0));
      }
      for (int s = 0; s < P.nsuper; ++s) {
        S.super.push_back(P.super[s] + c0);
        S.sparent.push_back(sh(P.sparent[s], s0));
        S.pi.push_back(P.pi[s] + r0);
        S.pruneptr.push_back((int)(P.pruneptr[s] + q0));
      }
      for (int v : P.s) S.s.push_back(v + c0);
      for (int v : P.pruneset) S.pruneset.push_back(v + s0);
      for (int v : P.a2i) S.a2i.push_back(v + c0);
      for (int64_t v : P.a2src) S.a2src.push_back(v + in0);
      S.max_sup_wid = std::max(S.max_sup_wid, P.max_sup_wid);
      S.max_col = std::max(S.max_col, P.max_col);
      S.flops += P.flops;
      c0 += P.n; s0 += P.nsuper; r0 += P.ssize; x0 += P.xsize;
      e0 += (int64_t)P.a2i.size(); q0 += (int64_t)P.pruneset.size();
      in0 += (int64_t)Ap[k][P.n];
    }
    // closing entries; p[n] follows the reference's convention for the last
    // supernode (p[n-1] + its row count), as the last instance has it
    const swc::Symbolic& last = parts[count - 1];
    S.super.push_back(c0);
    S.pi.push_back(r0);
    S.pruneptr.push_back((int)q0);
    S.i_ptr.push_back(r0);
    S.p.push_back(last.p[last.n] + (x0 - last.xsize));
    S.a2p.push_back((int)e0);
    *out = h;
    return SWC_OK;
  } catch (const std::bad_alloc&) {
    return swc_set_error(SWC_ENOMEM, "swc_analyze_batch: out of host memory");
  } catch (const std::exception& e) {
    return swc_set_error(SWC_EINVAL, std::string("swc_analyze_batch: ") + e.what());
  }
}

void swc_symbolic_free(swc_symbolic* s) { delete s; }

int swc_symbolic_sizes(const swc_symbolic* sym, int64_t* sizes, double* flops) {
  if (!sym || !sizes) return swc_set_error(SWC_EINVAL, "swc_symbolic_sizes: null argument");
  const swc::Symbolic& S = sym->S;
  sizes[0] = S.n;
  sizes[1] = S.nsuper;
  sizes[2] = S.ssize;
  sizes[3] = S.xsize;
  sizes[4] = S.max_sup_wid;
  sizes[5] = S.max_col;
  sizes[6] = (int64_t)S.a2i.size();
  sizes[7] = (int64_t)S.pruneset.size();
  if (flops) *flops = S.flops;
  return SWC_OK;
}

int swc_symbolic_copy(const swc_symbolic* sym, int which, void* dst) {
  if (!sym || !dst) return swc_set_error(SWC_EINVAL, "swc_symbolic_copy: null argument");
  const swc::Symbolic& S = sym->S;
  auto cp = [&](const auto& v) {
    if (!v.empty()) std::memcpy(dst, v.data(), v.size() * sizeof(v[0]));
    return SWC_OK;
  };
  switch (which) {
    case SWC_SYM_PERM: return cp(S.perm);
    case SWC_SYM_COLCOUNT: return cp(S.colcount);
    case SWC_SYM_PARENT: return cp(S.parent);
    case SWC_SYM_SUPER: return cp(S.super);
    case SWC_SYM_SPARENT: return cp(S.sparent);
    case SWC_SYM_COL2SUP: return cp(S.col2sup);
    case SWC_SYM_PI: return cp(S.pi);
    case SWC_SYM_S: return cp(S.s);
    case SWC_SYM_I_PTR: return cp(S.i_ptr);
    case SWC_SYM_P: return cp(S.p);
    case SWC_SYM_PRUNEPTR: return cp(S.pruneptr);
    case SWC_SYM_PRUNESET: return cp(S.pruneset);
    case SWC_SYM_A2P: return cp(S.a2p);
    case SWC_SYM_A2I: return cp(S.a2i);
    case SWC_SYM_A2SRC: return cp(S.a2src);
    default: return swc_set_error(SWC_EINVAL, "swc_symbolic_copy: unknown array id");
  }
}

// Reference export of the factor to CSC (common/Util.h:311-338, MKL_INT as a
// 32-bit int, the reference's LP64 build): column j of supernode [c0, c1) takes
// the rows of the supernode's row list from its own diagonal down, with their
// stored values (explicit zeros of relaxed amalgamation included).
int bcsc2csc(size_t n, size_t nBlocks, size_t* Ap, int* Ai, size_t* AiP, int* sup2col, double* Ax,
             int* Cp, int* Ci, double* Cx) {
  if (!Ap || !Ai || !AiP || !sup2col || !Ax || !Cp || !Ci || !Cx) {
    swc_set_error(SWC_EINVAL, "bcsc2csc: null argument");
    return 0;
  }
  int64_t nnz = 0;
  Cp[0] = 0;
  for (size_t b = 0; b < nBlocks; ++b) {
    const int c0 = sup2col[b], c1 = sup2col[b + 1];
    if (c1 <= c0 || (size_t)c1 > n) {
      swc_set_error(SWC_EINVAL, "bcsc2csc: bad supernode partition");
      return 0;
    }
    for (int j = c0; j < c1; ++j) {
      size_t kk = AiP[c0] + (size_t)(j - c0);
      for (size_t k = Ap[j] + (size_t)(j - c0); k < Ap[j + 1]; ++k, ++kk) {
        if (Ai[kk] < 0 || (size_t)Ai[kk] >= n || nnz >= 0x7fffffffLL) {
          swc_set_error(SWC_EINVAL, "bcsc2csc: row index out of range or factor too large");
          return 0;
        }
        Cx[nnz] = Ax[k];
        Ci[nnz] = Ai[kk];
        ++nnz;
      }
      Cp[j + 1] = (int)nnz;
    }
  }
  return 1;
}

}  // extern "C"
