// C ABI of the host symbolic analysis (see include/swcholesky_amd.h).
#include <cstring>
#include <exception>
#include <new>
#include <string>

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
    std::memcpy(dst, v.data(), v.size() * sizeof(v[0]));
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

}  // extern "C"
