/*
 * CPU oracle -- TEST INFRASTRUCTURE ONLY (see oracle.h).  Checker for the
 * parity tests and the CPU baseline of bench.py; never part of the product.
 */
#include "oracle.h"

#include <dlfcn.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* Lower Cholesky of the w x w block at a (leading dimension lda), column by column
 * (left-looking, as MyBLAS.h:11-27 Cholesky_col computes it).  Returns 0 or j+1
 * for the first non-positive (or NaN) pivot j. */
static int potrf_lower(int w, double* a, int64_t lda) {
  for (int j = 0; j < w; ++j) {
    double* cj = a + (int64_t)j * lda;
    for (int k = 0; k < j; ++k) {
      const double* ck = a + (int64_t)k * lda;
      const double ljk = ck[j];
      for (int i = j; i < w; ++i) cj[i] -= ck[i] * ljk;
    }
    const double d = cj[j];
    if (!(d > 0.0)) return j + 1;
    const double r = sqrt(d);
    cj[j] = r;
    for (int i = j + 1; i < w; ++i) cj[i] /= r;
  }
  return 0;
}

/* B <- B L^{-T} for the m x w block B (ldb), L the w x w lower block at a (lda):
 * dtrsm('R','L','T','N') as sw_parallel_PB_Cholesky_05.c:461-467 calls it. */
static void trsm_rltn(int m, int w, const double* a, int64_t lda, double* b, int64_t ldb) {
  for (int j = 0; j < w; ++j) {
    double* bj = b + (int64_t)j * ldb;
    for (int k = 0; k < j; ++k) {
      const double ljk = a[(int64_t)k * lda + j];
      const double* bk = b + (int64_t)k * ldb;
      for (int i = 0; i < m; ++i) bj[i] -= bk[i] * ljk;
    }
    const double inv = 1.0 / a[(int64_t)j * lda + j];
    for (int i = 0; i < m; ++i) bj[i] *= inv;
  }
}

/* contribs[i + j*ldc] = sum_k S[i,k] S[j,k] for 0 <= j < n1, j <= i < nr, where S is
 * the nr x k block at src (lds): the dsyrk ('L','N') + dgemm ('N','C') pair of
 * sw_parallel_PB_Cholesky_05.c:223-224 and 265-266 written as one lower trapezoid. */
static void update_block(int nr, int n1, int k, const double* src, int64_t lds, double* cb,
                         int ldc, int fp32) {
  for (int j = 0; j < n1; ++j) memset(cb + (int64_t)j * ldc, 0, sizeof(double) * nr);
  if (!fp32) {
    for (int kk = 0; kk < k; ++kk) {
      const double* sk = src + (int64_t)kk * lds;
      for (int j = 0; j < n1; ++j) {
        const double sj = sk[j];
        double* cj = cb + (int64_t)j * ldc;
        for (int i = j; i < nr; ++i) cj[i] += sk[i] * sj;
      }
    }
  } else {
    /* operands rounded to float, products accumulated in float */
    float* acc = (float*)calloc((size_t)nr * n1, sizeof(float));
    for (int kk = 0; kk < k; ++kk) {
      const double* sk = src + (int64_t)kk * lds;
      for (int j = 0; j < n1; ++j) {
        const float sj = (float)sk[j];
        float* aj = acc + (int64_t)j * nr;
        for (int i = j; i < nr; ++i) aj[i] = fmaf((float)sk[i], sj, aj[i]);
      }
    }
    for (int j = 0; j < n1; ++j)
      for (int i = j; i < nr; ++i) cb[(int64_t)j * ldc + i] = (double)acc[(int64_t)j * nr + i];
    free(acc);
  }
}

/* One target supernode: copy-in of A, descendant updates in pruneSet order,
 * POTRF of the diagonal block, TRSM of the panel (_pthread_inner.c:128-390). */
static int factor_target(int s, const int* c, const int* r, const double* values,
                         const int64_t* lC, const int* Ls, const int64_t* Li_ptr,
                         double* lValues, const int* blockSet, const int* prunePtr,
                         const int* pruneSet, int* map, double** cbuf, size_t* cap, int fp32) {
  const int curCol = blockSet[s], nxtCol = blockSet[s + 1];
  const int supWdt = nxtCol - curCol;
  const int nSupR = (int)(Li_ptr[nxtCol] - Li_ptr[curCol]);
  int cnt = 0;
  for (int64_t i = Li_ptr[curCol]; i < Li_ptr[nxtCol]; ++i) map[Ls[i]] = cnt++;
  for (int i = curCol; i < nxtCol; ++i)
    for (int j = c[i]; j < c[i + 1]; ++j) lValues[lC[i] + map[r[j]]] = values[j];
  double* cur = lValues + lC[curCol];
  for (int t = prunePtr[s]; t < prunePtr[s + 1]; ++t) {
    const int d = pruneSet[t];
    const int cSN = blockSet[d], cNSN = blockSet[d + 1];
    const int64_t p0 = Li_ptr[cSN], p1 = Li_ptr[cNSN];
    const int nSNRCur = (int)(p1 - p0);
    const int supWdts = cNSN - cSN;
    /* rows of d inside [curCol, nxtCol): lb first, ub last (rows are sorted) */
    int lb = -1, ub = -1;
    for (int64_t j = p0; j < p1; ++j) {
      if (Ls[j] >= nxtCol) break;
      if (Ls[j] >= curCol) {
        if (lb < 0) lb = (int)(j - p0);
        ub = (int)(j - p0);
      }
    }
    if (lb < 0) continue; /* not reachable for a pruned descendant */
    const int nSupRs = nSNRCur - lb;
    const int ndrow1 = ub - lb + 1;
    const size_t need = (size_t)nSupRs * ndrow1;
    if (need > *cap) {
      free(*cbuf);
      *cap = need * 2;
      *cbuf = (double*)malloc(sizeof(double) * (*cap));
      if (!*cbuf) return -1;
    }
    double* contribs = *cbuf;
    update_block(nSupRs, ndrow1, supWdts, lValues + lC[cSN] + lb, nSNRCur, contribs, nSupRs,
                 fp32);
    for (int i2 = 0; i2 < ndrow1; ++i2) {
      const int col = map[Ls[p0 + i2 + lb]];
      for (int j2 = i2; j2 < nSupRs; ++j2) {
        const int cRow = Ls[p0 + j2 + lb];
        cur[(int64_t)col * nSupR + map[cRow]] -= contribs[(int64_t)i2 * nSupRs + j2];
      }
    }
  }
  const int info = potrf_lower(supWdt, cur, nSupR);
  if (info) return curCol + info; /* 1-based global column */
  trsm_rltn(nSupR - supWdt, supWdt, cur, nSupR, cur + supWdt, nSupR);
  return 0;
}

int oracle_factor(int n, const int* c, const int* r, const double* values, const int64_t* lC,
                  const int* Ls, const int64_t* Li_ptr, double* lValues, const int* blockSet,
                  int supNo, const int* prunePtr, const int* pruneSet, int fp32_updates,
                  int nthreads) {
  /* level of each supernode in the descendant DAG (descendants have lower ids) */
  int* lev = (int*)calloc(supNo, sizeof(int));
  int nlev = 0;
  for (int s = 0; s < supNo; ++s) {
    int l = 0;
    for (int t = prunePtr[s]; t < prunePtr[s + 1]; ++t)
      if (lev[pruneSet[t]] + 1 > l) l = lev[pruneSet[t]] + 1;
    lev[s] = l;
    if (l + 1 > nlev) nlev = l + 1;
  }
  int* lptr = (int*)calloc(nlev + 1, sizeof(int));
  int* order = (int*)malloc(sizeof(int) * supNo);
  for (int s = 0; s < supNo; ++s) lptr[lev[s] + 1]++;
  for (int l = 0; l < nlev; ++l) lptr[l + 1] += lptr[l];
  int* fillp = (int*)malloc(sizeof(int) * (nlev + 1));
  memcpy(fillp, lptr, sizeof(int) * (nlev + 1));
  for (int s = 0; s < supNo; ++s) order[fillp[lev[s]]++] = s;
  free(fillp);

  int result = 0;
#ifdef _OPENMP
  if (nthreads <= 0) nthreads = omp_get_max_threads();
#else
  nthreads = 1;
#endif
  for (int l = 0; l < nlev && result == 0; ++l) {
    int first_bad = 0;
#pragma omp parallel num_threads(nthreads)
    {
      int* map = (int*)malloc(sizeof(int) * (size_t)n);
      double* cbuf = NULL;
      size_t cap = 0;
#pragma omp for schedule(dynamic, 1)
      for (int q = lptr[l]; q < lptr[l + 1]; ++q) {
        int info = factor_target(order[q], c, r, values, lC, Ls, Li_ptr, lValues, blockSet,
                                 prunePtr, pruneSet, map, &cbuf, &cap, fp32_updates);
        if (info) {
#pragma omp critical
          {
            if (first_bad == 0 || (info > 0 && info < first_bad) || info < 0) first_bad = info;
          }
        }
      }
      free(cbuf);
      free(map);
    }
    result = first_bad;
  }
  free(lev);
  free(lptr);
  free(order);
  return result;
}

void oracle_lsolve(int n, const int64_t* lC, const int* Ls, const int64_t* Li_ptr,
                   const double* lValues, const int* blockSet, int supNo, double* x, int nrhs,
                   int ldx) {
  (void)n;
  for (int s = 0; s < supNo; ++s) {
    const int c0 = blockSet[s], c1 = blockSet[s + 1], w = c1 - c0;
    const int nr = (int)(Li_ptr[c1] - Li_ptr[c0]);
    const double* L = lValues + lC[c0];
    const int* rows = Ls + Li_ptr[c0];
    for (int q = 0; q < nrhs; ++q) {
      double* xq = x + (int64_t)q * ldx;
      for (int j = 0; j < w; ++j) {
        const double v = xq[c0 + j] / L[(int64_t)j * nr + j];
        xq[c0 + j] = v;
        for (int i = j + 1; i < w; ++i) xq[c0 + i] -= L[(int64_t)j * nr + i] * v;
      }
      for (int i = w; i < nr; ++i) {
        double t = 0.0;
        for (int j = 0; j < w; ++j) t += L[(int64_t)j * nr + i] * xq[c0 + j];
        xq[rows[i]] -= t;
      }
    }
  }
}

void oracle_ltsolve(int n, const int64_t* lC, const int* Ls, const int64_t* Li_ptr,
                    const double* lValues, const int* blockSet, int supNo, double* x, int nrhs,
                    int ldx) {
  (void)n;
  for (int s = supNo - 1; s >= 0; --s) {
    const int c0 = blockSet[s], c1 = blockSet[s + 1], w = c1 - c0;
    const int nr = (int)(Li_ptr[c1] - Li_ptr[c0]);
    const double* L = lValues + lC[c0];
    const int* rows = Ls + Li_ptr[c0];
    for (int q = 0; q < nrhs; ++q) {
      double* xq = x + (int64_t)q * ldx;
      for (int j = 0; j < w; ++j) {
        double t = 0.0;
        for (int i = w; i < nr; ++i) t += L[(int64_t)j * nr + i] * xq[rows[i]];
        xq[c0 + j] -= t;
      }
      for (int j = w - 1; j >= 0; --j) {
        double v = xq[c0 + j];
        for (int i = j + 1; i < w; ++i) v -= L[(int64_t)j * nr + i] * xq[c0 + i];
        xq[c0 + j] = v / L[(int64_t)j * nr + j];
      }
    }
  }
}

void oracle_spmv(int n, const int* rp, const int* ci, const double* v, const double* x,
                 double* y) {
  for (int i = 0; i < n; ++i) {
    double t = 0.0;
    for (int q = rp[i]; q < rp[i + 1]; ++q) t += v[q] * x[ci[q]];
    y[i] = t;
  }
}

double oracle_backward_error(int n, const int* rp, const int* ci, const double* v,
                             const double* x, const double* b) {
  double rmax = 0.0, amax = 0.0, xmax = 0.0, bmax = 0.0;
  for (int i = 0; i < n; ++i) {
    double t = 0.0, arow = 0.0;
    for (int q = rp[i]; q < rp[i + 1]; ++q) {
      t += v[q] * x[ci[q]];
      arow += fabs(v[q]);
    }
    const double ri = fabs(b[i] - t);
    if (ri > rmax) rmax = ri;
    if (arow > amax) amax = arow;
    if (fabs(x[i]) > xmax) xmax = fabs(x[i]);
    if (fabs(b[i]) > bmax) bmax = fabs(b[i]);
  }
  return rmax / (amax * xmax + bmax);
}

/* ---- BLAS-3 restatement (the reference's XMATH / MKL path) ------------------
 * The same left-looking loop as factor_target, with the dense kernels the
 * reference calls per descendant and per target (sw_parallel_PB_Cholesky_05.c:
 * dsyrk_ 300-306, dgemm_ 330-337, contribution scatter 346-367, dpotrf_ 421-427,
 * dtrsm_ 457-466; the x86 twin parallel_PB_Cholesky_05.h:148-210 makes the same
 * calls).  BLAS/LAPACK come from a shared library loaded at run time (Fortran
 * calling convention, 32-bit integers). */
typedef void (*dsyrk_fn)(const char*, const char*, const int*, const int*, const double*,
                         const double*, const int*, const double*, double*, const int*);
typedef void (*dgemm_fn)(const char*, const char*, const int*, const int*, const int*,
                         const double*, const double*, const int*, const double*, const int*,
                         const double*, double*, const int*);
typedef void (*dpotrf_fn)(const char*, const int*, double*, const int*, int*);
typedef void (*dtrsm_fn)(const char*, const char*, const char*, const char*, const int*,
                         const int*, const double*, const double*, const int*, double*,
                         const int*);
typedef void (*set_threads_fn)(int);
static dsyrk_fn p_dsyrk;
static dgemm_fn p_dgemm;
static dpotrf_fn p_dpotrf;
static dtrsm_fn p_dtrsm;
static set_threads_fn p_set_threads;

int oracle_blas_load(const char* path) {
  if (p_dsyrk && p_dgemm && p_dpotrf && p_dtrsm) return 0;
  /* MKL's runtime dispatcher: run its threads on the GNU OpenMP runtime this
   * library already uses (one OpenMP runtime per process) */
  setenv("MKL_THREADING_LAYER", "GNU", 0);
  void* h = dlopen(path, RTLD_NOW | RTLD_LOCAL);
  if (!h) return -1;
  p_dsyrk = (dsyrk_fn)dlsym(h, "dsyrk_");
  p_dgemm = (dgemm_fn)dlsym(h, "dgemm_");
  p_dpotrf = (dpotrf_fn)dlsym(h, "dpotrf_");
  p_dtrsm = (dtrsm_fn)dlsym(h, "dtrsm_");
  p_set_threads = (set_threads_fn)dlsym(h, "MKL_Set_Num_Threads");
  if (!p_dsyrk || !p_dgemm || !p_dpotrf || !p_dtrsm) {
    p_dsyrk = NULL; p_dgemm = NULL; p_dpotrf = NULL; p_dtrsm = NULL;
    return -2;
  }
  return 0;
}

/* One target supernode through BLAS-3 (follows sw_parallel_PB_Cholesky_05.c:209-466). */
static int factor_target_blas(int s, const int* c, const int* r, const double* values,
                              const int64_t* lC, const int* Ls, const int64_t* Li_ptr,
                              double* lValues, const int* blockSet, const int* prunePtr,
                              const int* pruneSet, int* map, double** cbuf, size_t* cap) {
  const double one = 1.0, zero = 0.0;
  const int curCol = blockSet[s], nxtCol = blockSet[s + 1];
  const int supWdt = nxtCol - curCol;
  const int nSupR = (int)(Li_ptr[nxtCol] - Li_ptr[curCol]);
  int cnt = 0;
  for (int64_t i = Li_ptr[curCol]; i < Li_ptr[nxtCol]; ++i) map[Ls[i]] = cnt++;
  for (int i = curCol; i < nxtCol; ++i)
    for (int j = c[i]; j < c[i + 1]; ++j) lValues[lC[i] + map[r[j]]] = values[j];
  double* cur = lValues + lC[curCol];
  for (int t = prunePtr[s]; t < prunePtr[s + 1]; ++t) {
    const int d = pruneSet[t];
    const int cSN = blockSet[d], cNSN = blockSet[d + 1];
    const int64_t p0 = Li_ptr[cSN], p1 = Li_ptr[cNSN];
    const int nSNRCur = (int)(p1 - p0);
    const int supWdts = cNSN - cSN;
    int lb = -1, ub = -1;
    for (int64_t j = p0; j < p1; ++j) {
      if (Ls[j] >= nxtCol) break;
      if (Ls[j] >= curCol) {
        if (lb < 0) lb = (int)(j - p0);
        ub = (int)(j - p0);
      }
    }
    if (lb < 0) continue;
    const int nSupRs = nSNRCur - lb;
    const int ndrow1 = ub - lb + 1;
    const int ndrow3 = nSupRs - ndrow1;
    const size_t need = (size_t)nSupRs * ndrow1;
    if (need > *cap) {
      free(*cbuf);
      *cap = need * 2;
      *cbuf = (double*)malloc(sizeof(double) * (*cap));
      if (!*cbuf) return -1;
    }
    double* contribs = *cbuf;
    const double* src = lValues + lC[cSN] + lb;
    const double* srcL = lValues + lC[cSN] + ub + 1;
    p_dsyrk("L", "N", &ndrow1, &supWdts, &one, src, &nSNRCur, &zero, contribs, &nSupRs);
    if (ndrow3 > 0)
      p_dgemm("N", "C", &ndrow3, &ndrow1, &supWdts, &one, srcL, &nSNRCur, src, &nSNRCur, &zero,
              contribs + ndrow1, &nSupRs);
    for (int i2 = 0; i2 < ndrow1; ++i2) {
      const int col = map[Ls[p0 + i2 + lb]];
      for (int j2 = i2; j2 < nSupRs; ++j2) {
        const int cRow = Ls[p0 + j2 + lb];
        cur[(int64_t)col * nSupR + map[cRow]] -= contribs[(int64_t)i2 * nSupRs + j2];
      }
    }
  }
  int info = 0;
  p_dpotrf("L", &supWdt, cur, &nSupR, &info);
  if (info > 0) return curCol + info;
  if (info < 0) return -1;
  const int rowNo = nSupR - supWdt;
  if (rowNo > 0) p_dtrsm("R", "L", "C", "N", &rowNo, &supWdt, &one, cur, &nSupR, cur + supWdt, &nSupR);
  return 0;
}

int oracle_factor_blas(int n, const int* c, const int* r, const double* values, const int64_t* lC,
                       const int* Ls, const int64_t* Li_ptr, double* lValues, const int* blockSet,
                       int supNo, const int* prunePtr, const int* pruneSet, int nthreads) {
  if (!p_dsyrk) return -2;
  int* lev = (int*)calloc(supNo, sizeof(int));
  int nlev = 0;
  for (int s = 0; s < supNo; ++s) {
    int l = 0;
    for (int t = prunePtr[s]; t < prunePtr[s + 1]; ++t)
      if (lev[pruneSet[t]] + 1 > l) l = lev[pruneSet[t]] + 1;
    lev[s] = l;
    if (l + 1 > nlev) nlev = l + 1;
  }
  int* lptr = (int*)calloc(nlev + 1, sizeof(int));
  int* order = (int*)malloc(sizeof(int) * (supNo > 0 ? supNo : 1));
  for (int s = 0; s < supNo; ++s) lptr[lev[s] + 1]++;
  for (int l = 0; l < nlev; ++l) lptr[l + 1] += lptr[l];
  int* fillp = (int*)malloc(sizeof(int) * (nlev + 1));
  memcpy(fillp, lptr, sizeof(int) * (nlev + 1));
  for (int s = 0; s < supNo; ++s) order[fillp[lev[s]]++] = s;
  free(fillp);
#ifdef _OPENMP
  if (nthreads <= 0) nthreads = omp_get_max_threads();
#else
  nthreads = 1;
#endif
  int result = 0;
  int* map0 = (int*)malloc(sizeof(int) * (size_t)n);
  double* cbuf0 = NULL;
  size_t cap0 = 0;
  for (int l = 0; l < nlev && result == 0; ++l) {
    const int cnt = lptr[l + 1] - lptr[l];
    if (cnt >= nthreads) {
      /* many targets: one per thread, sequential BLAS inside the parallel region
       * (the reference's partitions of a level cluster) */
      int first_bad = 0;
      if (p_set_threads) p_set_threads(1);
#pragma omp parallel num_threads(nthreads)
      {
        int* map = (int*)malloc(sizeof(int) * (size_t)n);
        double* cbuf = NULL;
        size_t cap = 0;
#pragma omp for schedule(dynamic, 1)
        for (int q = lptr[l]; q < lptr[l + 1]; ++q) {
          const int info = factor_target_blas(order[q], c, r, values, lC, Ls, Li_ptr, lValues,
                                              blockSet, prunePtr, pruneSet, map, &cbuf, &cap);
          if (info) {
#pragma omp critical
            {
              if (first_bad == 0 || info < 0 || (first_bad > 0 && info < first_bad)) first_bad = info;
            }
          }
        }
        free(cbuf);
        free(map);
      }
      result = first_bad;
    } else {
      /* few (large) targets: one after another with threaded BLAS, as the
       * reference runs its last level (parallel_PB_Cholesky_05.h:255-264) */
      if (p_set_threads) p_set_threads(nthreads);
      for (int q = lptr[l]; q < lptr[l + 1] && result == 0; ++q)
        result = factor_target_blas(order[q], c, r, values, lC, Ls, Li_ptr, lValues, blockSet,
                                    prunePtr, pruneSet, map0, &cbuf0, &cap0);
    }
  }
  free(cbuf0);
  free(map0);
  free(lev);
  free(lptr);
  free(order);
  return result;
}
