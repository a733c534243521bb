/*
 * CPU oracle for the parity tests -- TEST INFRASTRUCTURE ONLY.
 *
 * A plain-C restatement of the reference's supernodal left-looking numeric
 * factorization and of the triangular solves around it.  Only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg load it, always as the
 * checker; the product (swcholesky_amd) never links or calls it.
 *
 * Array conventions are those of the reference BCSC factor
 * (software/swCholesky/common/def.h:123-210): lC = L->p (per-column value
 * offsets), Ls = L->s, Li_ptr = L->i_ptr, blockSet = L->super.
 *
 * Parity pins (what each function is checked against, see tests/):
 *   oracle_factor  : ||P A P^T - L L^T||_F / ||A||_F at rounding level; equality
 *                    with numpy's dense Cholesky on tiny grids; the reference's
 *                    BLAS-sw known-answer shapes for the dense kernels.
 *   oracle_lsolve / oracle_ltsolve : b = L*1 => x = 1 (Util.h:277-306); dense
 *                    triangular solves on tiny cases.
 *   oracle_spmv / oracle_backward_error : scipy sparse product; closed form.
 */
#ifndef SWC_ORACLE_H
#define SWC_ORACLE_H
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Left-looking supernodal LL^T (sw_parallel_PB_Cholesky_05.c:193-485, non-PTHREAD
 * XMATH path).  lValues must be zeroed by the caller.  Targets run level by level
 * of the descendant DAG (OpenMP over a level); each target's arithmetic is
 * sequential in pruneSet order, so results do not depend on the thread count.
 * fp32_updates != 0 rounds the descendant SYRK/GEMM operands to float and
 * accumulates each update in float before it is subtracted in double (the
 * mixed-precision scheme).  Returns 0, or k+1 for the first failing column k. */
int oracle_factor(int n, const int* c, const int* r, const double* values, const int64_t* lC,
                  const int* Ls, const int64_t* Li_ptr, double* lValues, const int* blockSet,
                  int supNo, const int* prunePtr, const int* pruneSet, int fp32_updates,
                  int nthreads);

/* BLAS-3 form of the same loop: the reference's dsyrk_/dgemm_ per descendant,
 * dpotrf_/dtrsm_ per target (sw_parallel_PB_Cholesky_05.c:300-466), from the
 * BLAS/LAPACK library loaded by oracle_blas_load (Fortran ABI, LP64).  Levels of
 * the descendant DAG with at least nthreads targets run one target per thread
 * with sequential BLAS; smaller levels run their targets in turn with threaded
 * BLAS.  Returns 0, k+1 for the first failing column k, -1 on allocation
 * failure, -2 when no BLAS is loaded. */
int oracle_blas_load(const char* path);
int oracle_factor_blas(int n, const int* c, const int* r, const double* values, const int64_t* lC,
                       const int* Ls, const int64_t* Li_ptr, double* lValues, const int* blockSet,
                       int supNo, const int* prunePtr, const int* pruneSet, int nthreads);

/* Forward solve L y = b in place (Triangular_BCSC.h:14-49, blockedLsolve), nrhs
 * columns of length n with leading dimension ldx. */
void oracle_lsolve(int n, const int64_t* lC, const int* Ls, const int64_t* Li_ptr,
                   const double* lValues, const int* blockSet, int supNo, double* x, int nrhs,
                   int ldx);
/* Backward solve L^T x = y in place (no reference counterpart). */
void oracle_ltsolve(int n, const int64_t* lC, const int* Ls, const int64_t* Li_ptr,
                    const double* lValues, const int* blockSet, int supNo, double* x, int nrhs,
                    int ldx);
/* y = A x with A given by its full CSR pattern (both triangles). */
void oracle_spmv(int n, const int* rp, const int* ci, const double* v, const double* x,
                 double* y);
/* eta = ||b - A x||_inf / (||A||_inf ||x||_inf + ||b||_inf). */
double oracle_backward_error(int n, const int* rp, const int* ci, const double* v,
                             const double* x, const double* b);

#ifdef __cplusplus
}
#endif
#endif
