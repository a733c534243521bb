/*
 * swcholesky_amd: MI355X (gfx950) numeric phase of swCholesky's supernodal sparse
 * Cholesky -- C ABI.
 *
 * Every entry point has C linkage and takes plain pointers and sizes.  Unless a
 * function says otherwise, pointers are host pointers, arrays use the layouts of
 * the reference's BCSC factor (software/swCholesky/common/def.h:123-210) and
 * integer status codes are returned (SWC_OK == 0; see swc_last_error()).
 */
#ifndef SWCHOLESKY_AMD_H
#define SWCHOLESKY_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ------------------------------------------------------ */
#define SWC_OK 0
#define SWC_EINVAL -1   /* bad argument or inconsistent structure */
#define SWC_ENOMEM -2   /* host or device allocation failed */
#define SWC_EDEVICE -3  /* HIP runtime error */
#define SWC_ENOTPOS -4  /* matrix not positive definite (see swc_factor's info) */
#define SWC_ESTATE -5   /* call out of order (e.g. solve before factor) */

/* Message of the calling thread's last failing call ("" if none). */
const char* swc_last_error(void);

/* ---- host symbolic analysis ---------------------------------------------
 * Restates analyze_p2 (cholesky/sw_LSparsity.h:248-702) for a GIVEN ordering:
 * elimination tree, column counts, colcount-weighted postorder, relaxed
 * supernodes (nrelax/zrelax as sw_choleskyTest01.cpp:112-113 when NULL),
 * supernodal row lists, BCSC pointers and pruned descendant sets
 * (cholesky/Inspection_Prune.h:24-56).  A is the lower triangle in CSC with
 * sorted rows; perm[new] = old.  The H-level partition (InspectionLevel_06.h)
 * is not produced: the numeric phase schedules from the descendant sets. */
typedef struct swc_symbolic swc_symbolic;

int swc_analyze(int n, const int* Ap, const int* Ai, const int64_t* perm, const int* nrelax,
                const double* zrelax, swc_symbolic** out);
/* Batches (no counterpart in the reference, which analyses one matrix per
 * process): count independent instances (instance k: n[k] columns, lower
 * triangle Ap[k]/Ai[k], ordering perm[k]) analysed on nthreads host threads (0:
 * all cores) and concatenated into the symbolic factor of their block-diagonal
 * operator, equal to swc_analyze of that operator.  A2SRC indexes the
 * instances' input entries concatenated in instance order. */
int swc_analyze_batch(int count, const int* n, const int* const* Ap, const int* const* Ai,
                      const int64_t* const* perm, const int* nrelax, const double* zrelax,
                      int nthreads, swc_symbolic** out);
void swc_symbolic_free(swc_symbolic* sym);
/* sizes[8] = {n, nsuper, ssize, xsize, maxSupWid, maxCol, nnz(tril A), |pruneSet|};
 * *flops = sum_j ColCount[j]^2 (ColumnCount.h:489-499). */
int swc_symbolic_sizes(const swc_symbolic* sym, int64_t* sizes, double* flops);

enum swc_symbolic_array {
  SWC_SYM_PERM = 0,      /* int32[n]    L->Perm                         */
  SWC_SYM_COLCOUNT = 1,  /* int32[n]    L->ColCount                     */
  SWC_SYM_PARENT = 2,    /* int32[n]    column elimination tree         */
  SWC_SYM_SUPER = 3,     /* int32[nsuper+1]  L->super (blockSet)        */
  SWC_SYM_SPARENT = 4,   /* int32[nsuper]    L->sParent                 */
  SWC_SYM_COL2SUP = 5,   /* int32[n]    L->col2Sup                      */
  SWC_SYM_PI = 6,        /* int64[nsuper+1]  L->pi                      */
  SWC_SYM_S = 7,         /* int32[ssize]     L->s (Ls)                  */
  SWC_SYM_I_PTR = 8,     /* int64[n+1]  L->i_ptr (Li_ptr)               */
  SWC_SYM_P = 9,         /* int64[n+1]  L->p (lC)                       */
  SWC_SYM_PRUNEPTR = 10, /* int32[nsuper+1]                             */
  SWC_SYM_PRUNESET = 11, /* int32[|pruneSet|]                           */
  SWC_SYM_A2P = 12,      /* int32[n+1]  tril(P A P^T) column pointers   */
  SWC_SYM_A2I = 13,      /* int32[nnz]  tril(P A P^T) row indices       */
  SWC_SYM_A2SRC = 14     /* int64[nnz]  entry of A each A2 entry comes from */
};
int swc_symbolic_copy(const swc_symbolic* sym, int which, void* dst);

/* Fill-reducing nested-dissection ordering of a symmetric pattern given as its
 * lower triangle in CSC (Ap[n+1], Ai): perm[new] = old (int64), to be passed to
 * swc_analyze as the GIVEN ordering.  Stands in for the METIS_NodeND call of the
 * reference's x86 path (cholesky/LSparsity.h:597) and the perm files it reads
 * (software/README.md:45); the ordering is this library's own heuristic, not
 * METIS's. */
int swc_nd_order(int n, const int* Ap, const int* Ai, int64_t* perm);

/* ---- reference numeric entry point ----------------------------------------
 * Drop-in for sw_cholesky_left_par_05 (cholesky/sw_parallel_PB_Cholesky_05.h:52-66,
 * PRUNE variant; called at examples/sw_choleskyTest01.cpp:223-232).  A2 = tril(P A
 * P^T) in CSC (c, r, values); lValues (xsize doubles, zeroed by the caller) receives
 * L in the BCSC layout.  Factors on the current HIP device.  The H-level schedule
 * arguments are accepted and ignored.  Returns 1 on success and 0 on failure (non
 * SPD or device error; the reference returns 1 even for non-SPD input,
 * sw_parallel_PB_Cholesky_05.c:818).  timing[0] receives the device time (s). */
int sw_cholesky_left_par_05(int n, int* c, int* r, double* values, size_t* lC, int* Ls,
                            size_t* Li_ptr, double* lValues, int* blockSet, int supNo,
                            double* timing, int* prunePtr, int* pruneSet, int nLevels,
                            int* levelPtr, int* levelSet, int nPar, int* parPtr, int* partition,
                            int chunk, int threads, int super_max, int col_max, double* nodCost);

/* ---- reference forward-solve entry point -----------------------------------
 * Drop-in for blockedLsolve (triangularSolve/Triangular_BCSC.h:14-15; called at
 * examples/triangularTest02.cpp:195 as blockedLsolve(n, L->p, L->s, valL, nnz,
 * L->i_ptr, L->col2Sup, L->super, L->nsuper, x)): solves L y = x in place on the
 * host vector x on the current HIP device.  sup2col has supNo+1 entries; NNZ and
 * col2sup are not read (nor does the reference read them).  The solve schedule is
 * derived from the row lists and kept for later calls with an equal structure.
 * Returns 1 on success and 0 on failure, as the reference does (0 when Lp, Li or
 * x is null; swc_last_error() says why otherwise). */
int blockedLsolve(int n, size_t* Lp, int* Li, double* Lx, int NNZ, size_t* Li_ptr, int* col2sup,
                  int* sup2col, int supNo, double* x);

/* ---- reference CSC export -------------------------------------------------
 * Drop-in for bcsc2csc (common/Util.h:311-338), host arrays in and out, MKL_INT
 * as a 32-bit int (the reference's LP64 build): the BCSC factor (Ap = L->p, Ai =
 * L->s, AiP = L->i_ptr, sup2col = L->super, Ax = values) to CSC Cp[n+1], Ci, Cx.
 * Returns 1, or 0 on a bad argument or a factor with more than 2^31-1 entries. */
int bcsc2csc(size_t n, size_t nBlocks, size_t* Ap, int* Ai, size_t* AiP, int* sup2col, double* Ax,
             int* Cp, int* Ci, double* Cx);

/* ---- handle API (structure uploaded once, refactor / solve many times) ------ */
typedef struct swc_handle swc_handle;

/* Binds to the current HIP device and creates the handle's stream. */
int swc_create(swc_handle** h);
int swc_destroy(swc_handle* h);

/* Uploads the symbolic structure (arrays exactly as analyze_p2 returns them,
 * sw_LSparsity.h:248-702) and the pattern of A2 = tril(P A P^T).  With a2src
 * non-NULL, A2 entry e holds input value a2src[e] and swc_factor takes nvals
 * values in that input order; otherwise it takes the A2 values in A2 order. */
int swc_upload(swc_handle* h, int n, const int* blockSet, int supNo, const size_t* lC,
               const int* Ls, const size_t* Li_ptr, const int* prunePtr, const int* pruneSet,
               const int* A2p, const int* A2i, const int64_t* a2src, int64_t nvals);
/* Same, from the host analysis (values then in the analysed A's entry order, and
 * the permutation set for solves in the original ordering). */
int swc_upload_symbolic(swc_handle* h, const swc_symbolic* sym);
/* perm[new] = old: solves take and return vectors in the original ordering.
 * NULL clears it (solves then work in the factor's ordering). */
int swc_set_perm(swc_handle* h, const int* perm);

/* Shape parameters of the plan built by the next structure upload (swc_upload,
 * swc_upload_symbolic), as {rlb, spw, narrow_min_tiles, narrow_max_k, fuse_narrow,
 * dag} (count <= 6, leading entries; NULL restores the defaults {8, 128, 1024, 256,
 * 1, 0}): rlb panels per block of a wide supernode's right-looking steps, spw = 64
 * or 128 columns per solve panel, the one-wave tile kernel for levels with at least
 * narrow_min_tiles tiles of panels <= 16 columns whose inner dimensions sum to
 * <= narrow_max_k, with fuse_narrow = 1 such panels (all tiles narrow) updated
 * and factored by one wave each in one persistent launch per level.  dag must be
 * 0: the dependency-driven launch of the later levels (dag = 1, 2) is disabled and
 * rejected with SWC_EINVAL, after one run of it gave a wrong factor whose cause
 * was not found.  No reference counterpart (the reference's Sunway BLAS fixes
 * its blocking, BLAS-sw/dpotrf_cluster_master.c:11-63); the values come from the
 * autotuner's table (swcholesky_amd/tuning.json). */
int swc_set_plan_params(swc_handle* h, const int* params, int count);

/* Schedule of the factor levels of the next structure upload.  nLevels = 0 (the
 * default): the wavefront schedule, every supernode at the first level its pruned
 * descendants allow (the level sets of the supernodal elimination tree,
 * common/TreeUtils.h:119-170).  Otherwise the reference's H-level schedule as
 * analyze_p2 returns it (InspectionLevel_06.h:18-391; levelPtr[nLevels+1],
 * parPtr[levelPtr[nLevels]+1], partition[nsuper], the arguments the reference
 * passes to sw_cholesky_left_par_05): clusters one after another, the partitions
 * of a cluster side by side, each partition's supernodes one after another.
 * Offered to compare the two; the upload fails if it does not cover the
 * structure or places a supernode before its descendants. */
int swc_set_schedule(swc_handle* h, int nLevels, const int* levelPtr, const int* parPtr,
                     const int* partition);

/* Precision of the next swc_factor calls.  0 (default): FP64 throughout.
 * 1: descendant SYRK/GEMM updates on FP32 MFMA with FP32 accumulation, each
 *    update subtracted into the FP64 target; POTRF, TRSM and the updates inside
 *    a supernode's own diagonal block stay FP64.
 * 2: as 1, and the updates inside wide supernodes also use FP32 MFMA; only the
 *    64x64 diagonal-tile POTRF and the TRSMs stay FP64.
 * L is stored in FP64 in every mode; swc_refine recovers FP64 accuracy. */
int swc_set_precision(swc_handle* h, int mode);
/* Numeric factorization.  values: host or device pointer (values_on_device).
 * *info = 0, or k+1 for the first column k whose pivot is not positive
 * (LAPACK convention); returns SWC_ENOTPOS in that case. */
int swc_factor(swc_handle* h, const double* values, int values_on_device, int* info);
/* Copies L (xsize doubles, BCSC layout) to host or device memory. */
int swc_get_factor(swc_handle* h, double* lValues, int on_device);
/* The factor in CSC, as the reference's bcsc2csc lays it out (common/Util.h:
 * 311-338): column j holds its supernode's rows from its own diagonal down
 * (explicit zeros of relaxed amalgamation included), rows ascending.
 * swc_factor_nnz gives the entry count; swc_factor_csc fills Cp[n+1] (int64),
 * Ci[nnz], Cx[nnz] in host or device memory, converting on the GPU. */
int swc_factor_nnz(swc_handle* h, int64_t* nnz);
int swc_factor_csc(swc_handle* h, int64_t* Cp, int* Ci, double* Cx, int on_device);
/* Solves A X = B in place for nrhs columns (leading dimension ldb >= n):
 * permutation, forward L y = P b (blockedLsolve, triangularSolve/Triangular_BCSC.h:14-49),
 * backward L^T z = y (no reference counterpart), inverse permutation. */
int swc_solve(swc_handle* h, double* B, int nrhs, int64_t ldb, int on_device);
/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc): swc_solve on device columns (B on the current device,
 * ldb >= n) that calls no hipMalloc/hipFree and captures no graph.  Its
 * workspace is a fixed 128 MiB device array per GPU, held until the process
 * ends; a call that needs more (about (2n + vsize + 128 max_lvl_fblk) * 8 bytes
 * per right-hand side) fails with SWC_EINVAL.  Calls on one GPU share that
 * array: calls that overlap in time on one GPU race. */
int swc_solve_noalloc(swc_handle* h, double* B, int nrhs, int64_t ldb);
/* Solves A X = B to backward error eta = ||b - A x||_inf / (||A||_inf ||x||_inf +
 * ||b||_inf) <= tol per column, by FP64 refinement on the current factor: the
 * residual is an FP64 SpMV with the values given to swc_factor, corrections use
 * the factor.  method 0: classical refinement; 1: conjugate gradients
 * preconditioned by L L^T; -1: 1 after a mixed-precision factor, else 0.
 * B and X (n x nrhs, leading dimension ld) may alias; eta[nrhs] receives the
 * final backward errors, *iters the number of corrections.  No reference
 * counterpart (the reference has no solve after factorization). */
int swc_refine(swc_handle* h, const double* B, double* X, int nrhs, int64_t ld, int on_device,
               double tol, int maxiter, int method, double* eta, int* iters);
/* eta[c] of a given X for A X = B (FP64, on the GPU). */
int swc_residual(swc_handle* h, const double* B, const double* X, int nrhs, int64_t ld,
                 int on_device, double* eta);
/* ---- adjoint (no reference counterpart: the reference has no gradients) ----
 * Selected inverse Z = A^{-1} on the pattern of L (L's BCSC layout, factor
 * ordering), by the supernodal reverse sweep of the Cholesky adjoint; Z may be
 * NULL to compute without copying out. */
int swc_selected_inverse(swc_handle* h, double* Z, int on_device);
/* logdet(A) = 2 sum_j log L_jj of the current factor. */
int swc_logdet(swc_handle* h, double* logdet);
/* grad[q] = w_q (alpha Z_ij - beta x_i x_j) for every input value q (an entry
 * a_ij, i >= j, of the lower triangle, standing for a_ij and a_ji; w_q = 1 on the
 * diagonal, 2 off it), i.e. the gradient of alpha logdet(A) + beta b^T A^{-1} b
 * with respect to the stored values when x = A^{-1} b (original ordering). */
int swc_gradient(swc_handle* h, const double* x, int x_on_device, double alpha, double beta,
                 double* grad, int grad_on_device);
/* out[7] = {panels, tiles, tasks, levels, xsize, work items of the dependency-driven
 * factor launch (0: none), its first level}; flops[5] = {update, factor, useful
 * flops of the dependency-driven launch, its simulated makespan and that of the
 * same levels as per-level launches (microseconds, the plan's cost model)}. */
int swc_plan_stats(const swc_handle* h, int64_t* out, double* flops);
/* Per factor level l < cap: out[2l] = duration (ms) of its tile-update launch in
 * the last profiled factorization (swc_set_profiling), out[2l+1] = that
 * launch's useful (lower-triangle) update flops; *nlevel = number of levels.
 * No counterpart in the reference: instrumentation of the GPU schedule. */
int swc_level_timings(const swc_handle* h, double* out, int cap, int* nlevel);
/* With graphs on (default), swc_factor and the solves replay launch sequences
 * captured once per structure (hipGraph) instead of launching level by level. */
int swc_set_graphs(swc_handle* h, int on);
/* With profiling on, every kernel launch of swc_factor is bracketed by HIP events
 * on the handle's stream (for measurement runs; adds launch gaps). */
int swc_set_profiling(swc_handle* h, int on);
/* out[8] = {last factor ms, last solve ms (device events), tile-kernel total ms,
 * tile-kernel launches, trsm-kernel total ms, trsm launches (profiling only),
 * tile-kernel algorithmic flops, trsm-kernel algorithmic flops}. */
int swc_get_timings(const swc_handle* h, double* out);

/* ---- planning introspection (host only, no reference counterpart) ----------
 * Builds the panel/tile/task plan the GPU kernels run, without touching the
 * device, for tools and tests.  sizes[16] = {panels, tiles, tasks, levels, n,
 * xsize, factor-list tiles, solve row blocks, contribution-vector size, solve
 * panels, diagonal-block scratch doubles, solve levels, fused narrow panels,
 * dependency-driven work items, first level of the dependency-driven launch,
 * its dependency entries}; flops[5] = {update flops, factor flops, useful flops
 * of the dependency-driven launch, its simulated makespan and that of the same
 * levels as per-level launches (microseconds, cost model)}; swc_plan_export copies
 * array `which` (0 panels, 1 tiles, 2 tasks, 3 level->tile offsets, 4
 * level->factor-list offsets, 5 panels by level, 6 level->panel offsets, 7
 * column->gather offsets (int64, n+1), 8 gathered contribution offsets (int64,
 * vsize), 9 factor-list tiles, 10 solve row blocks, 11 solve level->row-block
 * offsets, 12 gathered columns by solve level, 13 solve level->gathered-column
 * offsets, 14 solve panels (Panel structs), 15 solve panels by level, 16 solve
 * level->solve-panel offsets, 17 fused narrow panels (int pairs: first tile, tile
 * count), 18 level->fused-panel offsets, 19 per level: tiles of fused panels, 20
 * per level: one-wave tiles, 21 dependency-driven work items (DagItem), 22 their
 * dependencies (int pairs: panel, tile count), 23 per level: tiles of panels the
 * level does not factor, last in the level) as the packed structs and arrays of
 * swcholesky_amd/csrc/plan.h.  With prunePtr NULL the descendant sets are derived
 * from the row lists (supernode s depends on every d whose rows below its own
 * columns meet s's columns) and only the solve part of the plan is built, as
 * blockedLsolve builds it (the factor lists stay empty). */
typedef struct swc_plan swc_plan;
int swc_plan_build(int n, const int* blockSet, int supNo, const size_t* lC, const int* Ls,
                   const size_t* Li_ptr, const int* prunePtr, const int* pruneSet,
                   swc_plan** out);
/* The same with plan shape parameters (as swc_set_plan_params, but dag = 1 or 2 is
 * accepted here: the plan's dependency-driven work items can be inspected, although
 * the launch that would run them is disabled) and, with nLevels > 0, the H-level
 * schedule (as swc_set_schedule). */
int swc_plan_build_params(int n, const int* blockSet, int supNo, const size_t* lC, const int* Ls,
                          const size_t* Li_ptr, const int* prunePtr, const int* pruneSet,
                          const int* params, int nparams, int nLevels, const int* levelPtr,
                          const int* parPtr, const int* partition, swc_plan** out);
void swc_plan_free(swc_plan* p);
int swc_plan_sizes(const swc_plan* p, int64_t* sizes, double* flops);
int swc_plan_export(const swc_plan* p, int which, void* dst);
/* Byte sizes of the exported structs {Panel, Tile, Task, SolveBlock}. */
int swc_plan_struct_sizes(int64_t* out);

#ifdef __cplusplus
}
#endif
#endif /* SWCHOLESKY_AMD_H */
