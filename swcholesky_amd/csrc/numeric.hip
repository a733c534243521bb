// MI355X (gfx950) numeric phase: supernodal left-looking Cholesky, triangular
// solves, residual and refinement, behind the C ABI of include/swcholesky_amd.h.
//
// Work is organised by the plan of plan.h: panels of <= 64 columns, tiles of
// <= 64 rows, each tile's accumulator held in MFMA registers.  One level of the
// panel DAG is one launch of the tile kernel (every update of every tile of
// the level) followed by the panel factor launches (POTRF of the diagonal
// block of each panel the level completes and TRSM of its tiles).  The solves
// walk solve panels of <= SPW columns, level by level, forward then backward.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <memory>
#include <mutex>
#include <new>
#include <string>
#include <vector>

#include "analysis.h"
#include "capi_common.h"
#include "plan.h"

using swc::Panel;
using swc::Plan;
using swc::Task;
using swc::Tile;

namespace {

constexpr int kThreads = 256;
constexpr int kLdsLd = swc::TM + 1;  // padded column stride of LDS tiles (doubles)

#define SWC_HIP(call)                                                                      \
  do {                                                                                     \
    hipError_t e_ = (call);                                                                \
    if (e_ != hipSuccess)                                                                  \
      return swc_set_error(SWC_EDEVICE, std::string(#call ": ") + hipGetErrorString(e_));  \
  } while (0)

// ---------------------------------------------------------------------------
// kernels
// ---------------------------------------------------------------------------

__global__ void k_fill(double* __restrict__ x, int64_t n, double v) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (; i < n; i += stride) x[i] = v;
}

// L[amap[q]] = vals[q]: copy-in of A into the supernodal storage
// (_pthread_inner.c:139-147, with the map lookup precomputed).
__global__ void k_assemble(double* __restrict__ L, const int64_t* __restrict__ amap,
                           const double* __restrict__ vals, int64_t nnz) {
  int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (; q < nnz; q += stride) L[amap[q]] = vals[q];
}

__global__ void k_set_int(int* p, int v) { *p = v; }

// Workgroup barrier for LDS traffic only: it waits for this wave's LDS
// operations but not for its global loads, which stay in flight across it
// (__syncthreads() waits for every outstanding memory operation).  Used where no
// thread reads global data that another thread of the workgroup has written.
__device__ inline void lds_barrier() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
  __builtin_amdgcn_s_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

typedef double d4 __attribute__((ext_vector_type(4)));
typedef float f4 __attribute__((ext_vector_type(4)));

// Lane layout of the 16x16x4 MFMA builtins on gfx950 (fixed by an on-device
// probe of both): lane l feeds A[l%16][k0 + l/16] and B^T[k0 + l/16][l%16];
// v_mfma_f64_16x16x4f64 leaves D[l/16 + 4r][l%16] in component r, the f32
// variant D[4(l/16) + r][l%16].

// Raw buffer resource over [base, base + bytes) (gfx9 descriptor word 3): a
// buffer load whose byte offset falls outside returns zero.
__device__ inline __amdgpu_buffer_rsrc_t make_rsrc(const void* base, int64_t bytes) {
  const int n = (int)(bytes < 0x7fffffffLL ? bytes : 0x7fffffffLL);
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(base), (short)0, n, 0x00020000);
}

// The whole offset goes in the VGPR operand: the range check covers only the
// VGPR offset (plus the instruction offset), not the SGPR one.
template <typename T>
__device__ inline T load_elem(__amdgpu_buffer_rsrc_t r, unsigned voff);
template <>
__device__ inline double load_elem<double>(__amdgpu_buffer_rsrc_t r, unsigned voff) {
  return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(r, (int)voff, 0, 0));
}
template <>
__device__ inline float load_elem<float>(__amdgpu_buffer_rsrc_t r, unsigned voff) {
  return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, (int)voff, 0, 0));
}

// Byte offset of the source row feeding tile position p (a tile row or a panel
// column), elements of es bytes: the e-th set bit of mask receives source row
// base + e.  Positions receiving nothing get an offset past every buffer, so
// their loads read zero.
__device__ inline unsigned landing_off(uint64_t mask, int p, int base, unsigned es) {
  return ((mask >> p) & 1ull) ? (unsigned)(base + __popcll(mask & ((1ull << p) - 1ull))) * es
                              : 0x80000000u;
}

// acc += A B^T for one task, A = source rows landing on the tile's rows, B =
// source rows landing on the panel's columns, both placed at their landing
// positions, over this wave's share of the inner dimension: the 16-column
// chunks starting at 16 w + 64 g.  acc[bi][bj] is the 16x16 block (bi, bj) of
// the 64x64 tile in the f64 output layout.  A block is computed only if the
// task lands rows in block row bi and columns in block column bj (uniform over
// the wave) and, on the diagonal tile, bi >= bj (blocks above the diagonal are
// never stored).  Operands come through buffer loads over the chunk's columns,
// so rows that land nowhere (all rows of an inactive block) and columns past k
// read as zero without branches: a branch around a load would make the
// compiler wait for it at the join instead of overlapping all of them.
// T = float reads the FP32 copy of L and multiplies on v_mfma_f32_16x16x4f32,
// whose A rows are fed in the order pi(x) = x/4 + 4(x%4) so that its output
// (D[4(l/16) + r][l%16]) lands in the f64 layout; each 16-column chunk is
// accumulated in FP32 and then added to the FP64 accumulator.
template <typename T, int NBJ = 4, int KSTEP = 64>
__device__ inline void task_update(const T* __restrict__ src, int ld, int k, int ra, int ca,
                                   uint64_t rmask, uint64_t cmask, bool diag, int wave, int lane,
                                   d4 (&acc)[4][NBJ]) {
  constexpr bool kF32 = sizeof(T) == 4;
  const int lr = lane & 15, lk = lane >> 4;
  const int arow = kF32 ? (lr >> 2) + 4 * (lr & 3) : lr;
  const unsigned kstep = (unsigned)ld * sizeof(T);  // bytes between source columns
  unsigned ao[4], bo[NBJ];
  bool ab[4], bb[NBJ];
#pragma unroll
  for (int b = 0; b < 4; ++b) {
    ao[b] = landing_off(rmask, 16 * b + arow, ra, sizeof(T)) + lk * kstep;
    ab[b] = ((rmask >> (16 * b)) & 0xffffull) != 0;
  }
#pragma unroll
  for (int b = 0; b < NBJ; ++b) {
    bo[b] = landing_off(cmask, 16 * b + lr, ca, sizeof(T)) + lk * kstep;
    bb[b] = ((cmask >> (16 * b)) & 0xffffull) != 0;
  }
  for (int kb = 16 * wave; kb < k; kb += KSTEP) {
    const __amdgpu_buffer_rsrc_t rs = make_rsrc((const char*)src + (int64_t)kb * kstep,
                                                (int64_t)(k - kb) * kstep);
    T a[4][4], bv[NBJ][4];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
#pragma unroll
      for (int b = 0; b < 4; ++b) a[b][u] = load_elem<T>(rs, ao[b] + 4 * u * kstep);
#pragma unroll
      for (int b = 0; b < NBJ; ++b) bv[b][u] = load_elem<T>(rs, bo[b] + 4 * u * kstep);
    }
    if constexpr (!kF32) {
#pragma unroll
      for (int bi = 0; bi < 4; ++bi)
#pragma unroll
        for (int bj = 0; bj < NBJ; ++bj)
          if (ab[bi] && bb[bj] && !(diag && bi < bj)) {
#pragma unroll
            for (int u = 0; u < 4; ++u)
              acc[bi][bj] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[bi][u], bv[bj][u], acc[bi][bj], 0, 0, 0);
          }
    } else {
#pragma unroll
      for (int bi = 0; bi < 4; ++bi) {
        if (!ab[bi]) continue;
        f4 c[NBJ];
#pragma unroll
        for (int bj = 0; bj < NBJ; ++bj) c[bj] = f4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int bj = 0; bj < NBJ; ++bj)
          if (bb[bj] && !(diag && bi < bj)) {
#pragma unroll
            for (int u = 0; u < 4; ++u)
              c[bj] = __builtin_amdgcn_mfma_f32_16x16x4f32(a[bi][u], bv[bj][u], c[bj], 0, 0, 0);
          }
#pragma unroll
        for (int bj = 0; bj < NBJ; ++bj)
#pragma unroll
          for (int r = 0; r < 4; ++r) acc[bi][bj][r] += (double)c[bj][r];
      }
    }
  }
}

// acc (f64 MFMA layout) -> a 64x64 LDS tile (column j at buf[j * kLdsLd])
__device__ inline void acc_store(const d4 (&acc)[4][4], double* buf, int lane) {
#pragma unroll
  for (int bi = 0; bi < 4; ++bi)
#pragma unroll
    for (int bj = 0; bj < 4; ++bj)
#pragma unroll
      for (int r = 0; r < 4; ++r)
        buf[(16 * bj + (lane & 15)) * kLdsLd + 16 * bi + (lane >> 4) + 4 * r] = acc[bi][bj][r];
}

// the LDS tile += acc, element by element
__device__ inline void acc_add(const d4 (&acc)[4][4], double* buf, int lane) {
#pragma unroll
  for (int bi = 0; bi < 4; ++bi)
#pragma unroll
    for (int bj = 0; bj < 4; ++bj)
#pragma unroll
      for (int r = 0; r < 4; ++r)
        buf[(16 * bj + (lane & 15)) * kLdsLd + 16 * bi + (lane >> 4) + 4 * r] += acc[bi][bj][r];
}

// Tile kernel: applies all source updates of one tile, in task order.  Every
// wave accumulates the products of a fixed share of each task's inner
// dimension over the whole 64x64 tile in MFMA registers; the four partial sums
// are combined in a fixed order, (w0 + w2) + (w1 + w3), through LDS and
// subtracted from the tile, so results do not depend on scheduling.  The
// factorization of panels completed by this level follows in the panel factor kernels.
// MODE 0: FP64 throughout; 1: descendant (kind 0) tiles on the FP32 copy,
// intra-supernode tiles in FP64 (a tile's tasks all have one kind); 2: FP32
// for every tile.
// The body of one tile (a workgroup of kThreads threads; red: 2 x PW x kLdsLd
// doubles of LDS), shared by k_tile_update and the dependency-driven kernel.
template <int MODE>
__device__ inline void tile_update_body(double* __restrict__ L, const Tile& tl,
                                        const Task* __restrict__ tasks,
                                        double* __restrict__ dscratch,
                                        const float* __restrict__ Lf, double* red_lds,
                                        int tid) {
  double(*red)[swc::PW * kLdsLd] = reinterpret_cast<double(*)[swc::PW * kLdsLd]>(red_lds);
  // the wave index is uniform over the wave: say so, so that the per-chunk
  // buffer resources stay in scalar registers
  const int lane = tid & 63, wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int nr = tl.nr, nc = tl.ncol, ld = tl.ld;
  const bool diag = tl.r0 == 0;
  double* gtile = L + tl.vbase + tl.r0;
  d4 acc[4][4];
#pragma unroll
  for (int bi = 0; bi < 4; ++bi)
#pragma unroll
    for (int bj = 0; bj < 4; ++bj) acc[bi][bj] = d4{0.0, 0.0, 0.0, 0.0};
  const bool f32 = MODE == 2 || (MODE == 1 && tl.task_end > tl.task_begin &&
                                   tasks[tl.task_begin].kind == 0);
  if (f32) {
    for (int t = tl.task_begin; t < tl.task_end; ++t) {
      const Task tk = tasks[t];
      task_update<float>(Lf + tk.src_base, tk.ld, tk.k, tk.ra, tk.ca, tk.rmask, tk.cmask, diag,
                         wave, lane, acc);
    }
  } else {
    for (int t = tl.task_begin; t < tl.task_end; ++t) {
      const Task tk = tasks[t];
      task_update<double>(L + tk.src_base, tk.ld, tk.k, tk.ra, tk.ca, tk.rmask, tk.cmask, diag,
                          wave, lane, acc);
    }
  }
  // the tile's current values, loaded before the reduction so that their
  // latency overlaps it: thread t owns elements t, t + 256, ... of the
  // column-major tile
  double cur[16];
#pragma unroll
  for (int q = 0; q < 16; ++q) {
    const int e = tid + q * kThreads, i = e & 63, j = e >> 6;
    cur[q] = (i < nr && j < nc) ? gtile[(int64_t)j * ld + i] : 0.0;
  }
  // partial sums (w0 + w2) and (w1 + w3) in LDS, added element-wise by all
  // threads below
  if (wave < 2) acc_store(acc, red[wave], lane);
  lds_barrier();
  if (wave >= 2) acc_add(acc, red[wave - 2], lane);
  lds_barrier();
  // write back tile - sum; rows 0..nc-1 of a tile at row offset 0 are the
  // panel's diagonal block, whose strict upper part stays untouched (zero).
  // When this level factors the panel, the diagonal block also goes to the
  // launch-invariant copy that every factor-kernel tile of the panel reads.
  const bool copy = tl.kind == 0 && tl.slot >= 0;
#pragma unroll
  for (int q = 0; q < 16; ++q) {
    const int e = tid + q * kThreads, i = e & 63, j = e >> 6;
    if (i < nr && j < nc) {
      if (!(diag && i < j)) {
        const double v = cur[q] - (red[0][j * kLdsLd + i] + red[1][j * kLdsLd + i]);
        gtile[(int64_t)j * ld + i] = v;
        if (copy && i < nc) dscratch[tl.slot + j * nc + i] = v;
      } else if (copy) {
        dscratch[tl.slot + j * nc + i] = 0.0;
      }
    }
  }
}

template <int MODE>
__global__ void __launch_bounds__(kThreads) __attribute__((amdgpu_waves_per_eu(2)))
k_tile_update(double* __restrict__ L, const Tile* __restrict__ tiles,
              const Task* __restrict__ tasks, int tile0, double* __restrict__ dscratch,
              const float* __restrict__ Lf) {
  __shared__ double red[2 * swc::PW * kLdsLd];
  const Tile tl = tiles[tile0 + blockIdx.x];
  tile_update_body<MODE>(L, tl, tasks, dscratch, Lf, red, threadIdx.x);
}

// Tiles of panels of at most 16 columns with a short total inner dimension: one
// wave per tile over the whole inner dimension, a 64x16 accumulator (four 16x16
// blocks), written back from the registers.  Same tasks, modes and write-back
// rules as k_tile_update.
template <int MODE>
__global__ void __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(4)))
k_tile_narrow(double* __restrict__ L, const Tile* __restrict__ tiles,
              const Task* __restrict__ tasks, int tile0, double* __restrict__ dscratch,
              const float* __restrict__ Lf) {
  const Tile tl = tiles[tile0 + blockIdx.x];
  const int lane = threadIdx.x;
  const int nr = tl.nr, nc = tl.ncol, ld = tl.ld;
  const bool diag = tl.r0 == 0;
  double* gtile = L + tl.vbase + tl.r0;
  d4 acc[4][1];
#pragma unroll
  for (int bi = 0; bi < 4; ++bi) acc[bi][0] = d4{0.0, 0.0, 0.0, 0.0};
  const bool f32 = MODE == 2 || (MODE == 1 && tl.task_end > tl.task_begin &&
                                   tasks[tl.task_begin].kind == 0);
  if (f32) {
    for (int t = tl.task_begin; t < tl.task_end; ++t) {
      const Task tk = tasks[t];
      task_update<float, 1, 16>(Lf + tk.src_base, tk.ld, tk.k, tk.ra, tk.ca, tk.rmask, tk.cmask,
                                diag, 0, lane, acc);
    }
  } else {
    for (int t = tl.task_begin; t < tl.task_end; ++t) {
      const Task tk = tasks[t];
      task_update<double, 1, 16>(L + tk.src_base, tk.ld, tk.k, tk.ra, tk.ca, tk.rmask, tk.cmask,
                                 diag, 0, lane, acc);
    }
  }
  const bool copy = tl.kind == 0 && tl.slot >= 0;
  const int j = lane & 15;
#pragma unroll
  for (int bi = 0; bi < 4; ++bi)
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int i = 16 * bi + (lane >> 4) + 4 * r;
      if (i < nr && j < nc) {
        if (!(diag && i < j)) {
          const double v = gtile[(int64_t)j * ld + i] - acc[bi][0][r];
          gtile[(int64_t)j * ld + i] = v;
          if (copy && i < nc) dscratch[tl.slot + j * nc + i] = v;
        } else if (copy) {
          dscratch[tl.slot + j * nc + i] = 0.0;
        }
      }
    }
}

// Broadcast of lane src's value (src uniform over the wave).
__device__ inline double bcast(double v, int src) {
  const int lo = __builtin_amdgcn_readlane(__double2loint(v), src);
  const int hi = __builtin_amdgcn_readlane(__double2hiint(v), src);
  return __hiloint2double(hi, lo);
}

// Panel factorization for panels of at most 16 columns (NC), one wave per tile
// and registers only, for levels with many narrow panels (the leaves, batched
// instances), where how many workgroups fit on a CU matters more than one
// tile's latency: POTRF of the panel's nc x nc
// diagonal block (dpotrf 'L', sw_parallel_PB_Cholesky_05.c:430) in registers --
// lane i holds row i, the pivot and the multipliers of column j are read from
// their lanes -- then B <- B L11^{-T} for the tile's rows below the diagonal
// block (dtrsm 'R','L','T','N', .c:461-467), lane i substituting its own row.
// Both apply the updates of an entry in increasing column order, as the
// left-looking loops of the reference do.  Every tile of a panel computes the
// same L11 from the same launch-invariant copy (bitwise identical), so no tile
// waits for another; the diagonal tile stores it and reports info.  Entries
// above the diagonal of the register copy are never read by lanes on or below
// it, so they are left unmasked.
template <int NC>
__device__ inline void panel_factor_regs_body(double* __restrict__ L, const Tile& tl,
                                              const double* __restrict__ dscratch,
                                              int* __restrict__ info_out, float* __restrict__ Lf,
                                              int lane) {
  const int nr = tl.nr, nc = tl.ncol;
  const double* gdiag = dscratch + tl.slot;  // updated diagonal block, ld nc
  double y[NC];                              // row `lane` of the diagonal block
#pragma unroll
  for (int c = 0; c < NC; ++c) y[c] = (lane < nc && c < nc && c <= lane) ? gdiag[c * nc + lane] : 0.0;
  int badcol = 0;
#pragma unroll
  for (int j = 0; j < NC; ++j) {
    if (j < nc) {
      const double d = bcast(y[j], j);
      const bool ok = d > 0.0;
      if (!ok && badcol == 0) badcol = j + 1;
      // 1/sqrt(d) once, then a multiply per entry (dpotf2 scales the column by
      // 1/ajj); a non-positive pivot continues with 1 so no NaNs spread
      double r = rsqrt(d);
      asm volatile("" : "+v"(r));  // computed whatever d is: a select, not a branch
      const double rs = ok ? r : 1.0;
      if (lane == j) y[j] = ok ? d * rs : 1.0;
      else if (lane > j) y[j] *= rs;
#pragma unroll
      for (int c = j + 1; c < NC; ++c)
        if (c < nc) y[c] -= y[j] * bcast(y[j], c);
    }
  }
  double* gtile = L + tl.vbase + tl.r0;
  float* ftile = Lf ? Lf + tl.vbase + tl.r0 : nullptr;
  if (tl.kind == 0) {
    if (badcol && lane == 0) atomicMin(info_out, tl.c0 + badcol);
#pragma unroll
    for (int c = 0; c < NC; ++c)
      if (lane < nc && c <= lane) {
        gtile[(int64_t)c * tl.ld + lane] = y[c];
        if (ftile) ftile[(int64_t)c * tl.ld + lane] = (float)y[c];
      }
  }
  // rows of this tile below the diagonal block: tile rows [first, nr)
  const int first = tl.kind == 0 ? nc : 0;
  const bool mine = lane >= first && lane < nr;
  double x[NC];
#pragma unroll
  for (int c = 0; c < NC; ++c) x[c] = (mine && c < nc) ? gtile[(int64_t)c * tl.ld + lane] : 0.0;
#pragma unroll
  for (int j = 0; j < NC; ++j) {
    if (j < nc) {
      x[j] *= 1.0 / bcast(y[j], j);
#pragma unroll
      for (int c = j + 1; c < NC; ++c)
        if (c < nc) x[c] -= x[j] * bcast(y[j], c);
    }
  }
#pragma unroll
  for (int c = 0; c < NC; ++c)
    if (mine && c < nc) {
      gtile[(int64_t)c * tl.ld + lane] = x[c];
      if (ftile) ftile[(int64_t)c * tl.ld + lane] = (float)x[c];
    }
}

template <int NC>
__global__ void __launch_bounds__(64)
k_panel_factor_regs(double* __restrict__ L, const Tile* __restrict__ ftiles, int list0,
                    const double* __restrict__ dscratch, int* __restrict__ info_out,
                    float* __restrict__ Lf) {
  const Tile tl = ftiles[list0 + blockIdx.x];
  panel_factor_regs_body<NC>(L, tl, dscratch, info_out, Lf, threadIdx.x);
}


// Fused narrow panels of a level: one persistent launch whose waves take chunks
// of consecutive panels from a device worklist (a counter claimed with a vector
// atomic, the next claim issued while the current chunk is worked on) until the
// level's list is exhausted.  A panel (<= 16 columns, a single-panel supernode whose tiles all
// have short updates) is done by one wave from start to finish: every tile's
// updates as k_tile_narrow applies them, the updated tile handed from the MFMA
// layout to row-per-lane registers through the wave's own LDS buffer, the
// diagonal block factored and the rows solved as k_panel_factor_regs does it
// (the same operations in the same order, so the result does not depend on
// which wave takes which panel), no copy of the diagonal block through global
// memory and no second launch.  MODE as k_tile_update; Lf: the FP32 copy of L,
// read by FP32 updates and written with the factor when write_f32 is set.
template <int MODE>
__global__ void __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(4)))
k_narrow_fused(double* __restrict__ L, const Tile* __restrict__ tiles,
               const Task* __restrict__ tasks, const swc::FusedPanel* __restrict__ items,
               int nitems, int chunk, int* __restrict__ counter, int* __restrict__ info_out,
               float* Lf, int write_f32) {
  constexpr int NC = 16;
  __shared__ double T[swc::TM * (NC + 1)];  // updated tile, row i at T[i * (NC + 1)]
  __shared__ double L11[NC * (NC + 1)];     // the panel's factored diagonal block, row-major
  const int lane = threadIdx.x;
  // the worklist hands out chunks of `chunk` consecutive panels
  int claim = 0;
  if (lane == 0) claim = atomicAdd(counter, 1);
  int item = __builtin_amdgcn_readfirstlane(claim) * chunk;
  int end = min(item + chunk, nitems);
  if (lane == 0) claim = atomicAdd(counter, 1);  // the next chunk, used after this one
  while (item < nitems) {
    if (item == end) {
      item = __builtin_amdgcn_readfirstlane(claim) * chunk;
      if (item >= nitems) break;
      end = min(item + chunk, nitems);
      if (lane == 0) claim = atomicAdd(counter, 1);
    }
    const swc::FusedPanel it = items[item];
    for (int t = 0; t < it.ntile; ++t) {
      const Tile tl = tiles[it.tile0 + t];
      const int nr = tl.nr, nc = tl.ncol, ld = tl.ld;
      const bool diag = tl.r0 == 0;
      double* gtile = L + tl.vbase + tl.r0;
      // (1) the tile's updates (k_tile_narrow)
      d4 acc[4][1];
#pragma unroll
      for (int bi = 0; bi < 4; ++bi) acc[bi][0] = d4{0.0, 0.0, 0.0, 0.0};
      const bool f32 = MODE == 2 || (MODE == 1 && tl.task_end > tl.task_begin &&
                                       tasks[tl.task_begin].kind == 0);
      if (f32) {
        for (int q = tl.task_begin; q < tl.task_end; ++q) {
          const Task tk = tasks[q];
          task_update<float, 1, 16>(Lf + tk.src_base, tk.ld, tk.k, tk.ra, tk.ca, tk.rmask,
                                    tk.cmask, diag, 0, lane, acc);
        }
      } else {
        for (int q = tl.task_begin; q < tl.task_end; ++q) {
          const Task tk = tasks[q];
          task_update<double, 1, 16>(L + tk.src_base, tk.ld, tk.k, tk.ra, tk.ca, tk.rmask,
                                     tk.cmask, diag, 0, lane, acc);
        }
      }
      {
        const int j = lane & 15;
#pragma unroll
        for (int bi = 0; bi < 4; ++bi)
#pragma unroll
          for (int r = 0; r < 4; ++r) {
            const int i = 16 * bi + (lane >> 4) + 4 * r;
            const bool in = i < nr && j < nc && !(diag && i < j);
            const double g = gtile[in ? (int64_t)j * ld + i : 0];
            T[i * (NC + 1) + j] = in ? g - acc[bi][0][r] : 0.0;
          }
      }
      float* ftile = write_f32 ? Lf + tl.vbase + tl.r0 : nullptr;
      if (t == 0) {
        // (2) POTRF of the diagonal block (k_panel_factor_regs), kept in LDS
        double y[NC];
#pragma unroll
        for (int c = 0; c < NC; ++c)
          y[c] = (lane < nc && c < nc && c <= lane) ? T[lane * (NC + 1) + c] : 0.0;
        int badcol = 0;
#pragma unroll
        for (int j = 0; j < NC; ++j) {
          if (j < nc) {
            const double d = bcast(y[j], j);
            const bool ok = d > 0.0;
            if (!ok && badcol == 0) badcol = j + 1;
            double r = rsqrt(d);
            asm volatile("" : "+v"(r));  // a select, not a branch
            const double rs = ok ? r : 1.0;
            if (lane == j) y[j] = ok ? d * rs : 1.0;
            else if (lane > j) y[j] *= rs;
#pragma unroll
            for (int c = j + 1; c < NC; ++c)
              if (c < nc) y[c] -= y[j] * bcast(y[j], c);
          }
        }
        if (badcol && lane == 0) atomicMin(info_out, tl.c0 + badcol);
#pragma unroll
        for (int c = 0; c < NC; ++c) {
          if (lane < nc && c <= lane) {
            gtile[(int64_t)c * ld + lane] = y[c];
            if (ftile) ftile[(int64_t)c * ld + lane] = (float)y[c];
          }
          if (lane < NC) L11[lane * (NC + 1) + c] = y[c];
        }
      }
      // (3) the tile's rows below the diagonal block: x <- x L11^{-T}
      const int first = t == 0 ? nc : 0;
      const bool mine = lane >= first && lane < nr;
      double x[NC];
#pragma unroll
      for (int c = 0; c < NC; ++c) x[c] = (mine && c < nc) ? T[lane * (NC + 1) + c] : 0.0;
#pragma unroll
      for (int j = 0; j < NC; ++j) {
        if (j < nc) {
          x[j] *= 1.0 / L11[j * (NC + 1) + j];
#pragma unroll
          for (int c = j + 1; c < NC; ++c)
            if (c < nc) x[c] -= x[j] * L11[c * (NC + 1) + j];
        }
      }
#pragma unroll
      for (int c = 0; c < NC; ++c)
        if (mine && c < nc) {
          gtile[(int64_t)c * ld + lane] = x[c];
          if (ftile) ftile[(int64_t)c * ld + lane] = (float)x[c];
        }
    }
    ++item;
  }
}

__global__ void k_fill_int(int* __restrict__ x, int n, int v) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) x[i] = v;
}

// Column stride (doubles) of the LDS working matrix of k_panel_factor_blk: rows
// 0..63 hold the panel's diagonal block, rows 64..127 the tile rows being solved.
// An odd stride spreads the 16 columns an MFMA output block touches over all
// LDS banks.
constexpr int kWld = 2 * swc::TM + 1;

// acc (16x16 block of the f64 MFMA output layout: lane l holds rows l/16 + 4r,
// column l%16) -= A B^T over 16 columns, A = W rows [16 ra, 16 ra + 16),
// B = W rows [16 rb, 16 rb + 16), both from columns [kb, kb + 16) of W.  The
// MFMA accumulates its products one after another, as a chain of FMAs would.
__device__ inline d4 lds_block_update(const double* __restrict__ W, d4 acc, int ra, int rb, int kb,
                                      int lane) {
  const int lr = lane & 15, lk = lane >> 4;
#pragma unroll
  for (int u = 0; u < 4; ++u) {
    const double* col = W + (kb + lk + 4 * u) * kWld;
    acc = __builtin_amdgcn_mfma_f64_16x16x4f64(-col[16 * ra + lr], col[16 * rb + lr], acc, 0, 0, 0);
  }
  return acc;
}

// Panel factorization of the panels completed at this level for panels of
// 17..64 columns, one workgroup of four waves per tile: POTRF of the panel's
// nc x nc diagonal block (dpotrf 'L', sw_parallel_PB_Cholesky_05.c:430) and
// B <- B L11^{-T} for the tile's rows below it (dtrsm 'R','L','T','N',
// .c:461-467), blocked in 16 columns and right-looking, as the reference's
// dpotrf_cluster_master blocks its POTRF (BLAS-sw/dpotrf_cluster_master.c:11-63).
// The diagonal block (from the launch-invariant copy, identity past nc) and the
// tile rows sit in LDS (W) from the first load to the final store.  Per block
// column J:
//  (1) one sweep over the 16 columns factors the 16x16 diagonal block and
//      solves every row below it at once, lane = row: wave 0 takes the diagonal
//      block's rows, waves 1 and 2 the tile rows, each of them also carrying
//      the 16 pivot rows in lanes 0..15 (the same operations on the same
//      values, so all three compute the same pivots bitwise).  Each pivot's
//      reciprocal square root scales its column; the next pivot is read from
//      its lane (readlane) and the later columns' multipliers from LDS, so the
//      pivot chain does not wait for LDS;
//  (2) the later block columns of the diagonal block (blocks on or below its
//      diagonal) and of the tile rows take the rank-16 update on FP64 MFMA,
//      the 16x16 output blocks spread over the four waves;
//  (3) while waves 0-2 sweep block column J, wave 3 stores block column J-1,
//      which is final, to L, so that only the last block column is stored
//      after the loop.
// Both apply each entry's updates in increasing column order, so the result is
// bitwise the one of a column-by-column right-looking factorization.  Every
// tile of a panel computes the same L11, so no tile waits for another; the
// diagonal tile stores it and reports info.
// LDS of k_panel_factor_blk: W, then per sweeping wave column j's multipliers
constexpr int kBlkLds = swc::PW * kWld + 3 * 16 * 16;

__device__ inline void panel_factor_blk_body(double* __restrict__ L, const Tile& tl,
                                             const double* __restrict__ dscratch,
                                             int* __restrict__ info_out, float* __restrict__ Lf,
                                             double* lds, int tid) {
  double* W = lds;
  double(*lc)[16][16] = reinterpret_cast<double(*)[16][16]>(lds + swc::PW * kWld);
  const int lane = tid & 63, wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int nc = tl.ncol;
  const int nb = (nc + 15) >> 4;              // 16-column blocks (2..4)
  const int first = tl.kind == 0 ? nc : 0;    // tile rows [first, nr) are solved ...
  const int nx = tl.nr > first ? tl.nr - first : 0;  // ... into W rows 64 .. 64 + nx
  const int nxb = (nx + 15) >> 4;
  const double* gdiag = dscratch + tl.slot;  // updated diagonal block, ld nc
  double* gtile = L + tl.vbase + tl.r0;
  float* ftile = Lf ? Lf + tl.vbase + tl.r0 : nullptr;
  {
    // wave w moves columns 16w..16w+15 of diagonal-block row `lane` and of tile
    // row `lane`; every lane loads from an address inside both buffers and then
    // selects, so all its loads are in flight together
    const int c0 = 16 * wave;
    double y[16], x[16];
#pragma unroll
    for (int c = 0; c < 16; ++c) {
      const int col = c0 + c;
      const bool iny = lane < nc && col < nc && col <= lane, inx = lane < nx && col < nc;
      const double vy = gdiag[iny ? col * nc + lane : 0];
      const double vx = gtile[inx ? (int64_t)col * tl.ld + first + lane : 0];
      y[c] = iny ? vy : ((lane >= nc || col >= nc) && col == lane ? 1.0 : 0.0);
      x[c] = inx ? vx : 0.0;
    }
    if (c0 < 16 * nb) {
#pragma unroll
      for (int c = 0; c < 16; ++c) {
        W[(c0 + c) * kWld + lane] = y[c];
        W[(c0 + c) * kWld + 64 + lane] = x[c];
      }
    }
  }
  // Final columns c0 + cstep c, c < ncols (<= 16), of W to L and to the FP32
  // copy: the diagonal tile's rows of L11 and the solved tile rows.  All LDS
  // reads are issued before the first store.
  auto write_back = [&](int c0, int cstep, int ncols) {
    double vd[16], vt[16];
#pragma unroll
    for (int c = 0; c < 16; ++c) {
      if (c < ncols) {
        const int col = c0 + cstep * c;  // < 64
        vd[c] = W[col * kWld + lane];
        vt[c] = W[col * kWld + 64 + lane];
      }
    }
#pragma unroll
    for (int c = 0; c < 16; ++c) {
      const int col = c0 + cstep * c;
      if (c < ncols && col < nc) {
        if (tl.kind == 0 && lane < nc && col <= lane) {
          gtile[(int64_t)col * tl.ld + lane] = vd[c];
          if (ftile) ftile[(int64_t)col * tl.ld + lane] = (float)vd[c];
        }
        if (lane < nx) {
          gtile[(int64_t)col * tl.ld + first + lane] = vt[c];
          if (ftile) ftile[(int64_t)col * tl.ld + first + lane] = (float)vt[c];
        }
      }
    }
  };
  lds_barrier();
  int badcol = 0;
  // (1) this lane's W row in block column jb/16 (wave 3 only reads one); the
  // lane holding pivot row jb + j is pbase + j.  Every wave has read its rows
  // before any is written back (waves 1 and 2 read the pivot rows that wave 0
  // rewrites): a barrier follows every read.
  auto sweep_row = [&](int jb) {
    return wave == 0 ? lane
                     : (wave == 3 || lane < 16 || (wave == 2 && lane >= 32)
                            ? jb + (lane & 15)
                            : 64 + 48 * (wave - 1) + (lane - 16));
  };
  double x[16];
  auto read_row = [&](int jb) {
    const int row = sweep_row(jb);
#pragma unroll
    for (int c = 0; c < 16; ++c) x[c] = W[(jb + c) * kWld + row];
  };
  read_row(0);
  lds_barrier();
  for (int J = 0; J < nb; ++J) {
    const int jb = 16 * J;
    const int row = sweep_row(jb);
    if (wave < 3) {
      const int pbase = wave == 0 ? jb : 0;
      const bool store = wave == 0 ? (lane >= jb && lane < 16 * nb)
                                   : (lane >= 16 && row - 64 < nx && (wave == 1 || lane < 32));
      double(*mult)[16] = lc[wave];
      double d = bcast(x[0], pbase);
      // Software-pipelined over the columns: column j's multiplier for column
      // j + 1 (the entry (j+1, j), needed by the next pivot) comes by readlane;
      // its multipliers for the columns after that go through LDS and are
      // applied one column later, so no LDS round trip sits on the pivot chain.
      // Every entry still takes its updates in increasing column order.
      double pm[16];  // column j-1's multipliers, read from LDS during column j-1
#pragma unroll
      for (int c = 0; c < 16; ++c) pm[c] = 0.0;
#pragma unroll
      for (int j = 0; j < 16; ++j) {
        // x[j] has every update from the columns before j
        // 1/sqrt(d) once, then a multiply per entry (dpotf2 scales the column
        // by 1/ajj).  A non-positive pivot is reported and continues as 1, so
        // no NaNs spread.  Selects rather than a branch: the test on d runs
        // beside the reciprocal square root instead of before it.
        double r = rsqrt(d);
        asm volatile("" : "+v"(r));  // computed whatever d is, not under a branch
        const bool pos = d > 0.0;
        badcol = pos || badcol != 0 ? badcol : jb + j + 1;
        x[j] = pos ? x[j] * r : (lane == pbase + j ? 1.0 : x[j]);
        if (j < 15) {
          const double m1 = bcast(x[j], pbase + j + 1);  // entry (j+1, j)
          if (j >= 1) x[j + 1] -= x[j - 1] * pm[j + 1];
          x[j + 1] -= x[j] * m1;
          // the next pivot: its lane's x[j+1], now updated by every earlier column
          d = bcast(x[j + 1], pbase + j + 1);
          if (lane >= pbase && lane < pbase + 16) mult[j][lane - pbase] = x[j];
          if (j >= 1) {
#pragma unroll
            for (int c = j + 2; c < 16; ++c) x[c] -= x[j - 1] * pm[c];
          }
#pragma unroll
          for (int c = j + 2; c < 16; ++c) pm[c] = mult[j][c];
        }
      }
      if (store) {
#pragma unroll
        for (int c = 0; c < 16; ++c) W[(jb + c) * kWld + row] = x[c];
      }
    } else if (J > 0) {
      // block column J - 1 is final (the updates after its sweep touch later
      // columns only): wave 3 writes it back while the others sweep
      write_back(jb - 16, 1, 16);
    }
    lds_barrier();
    if (J + 1 < nb) {
      // (2) block columns C > J: rows of the diagonal block from C on, then the
      // tile rows (W row blocks 4 .. 4 + nxb)
      int idx = 0;
      for (int C = J + 1; C < nb; ++C) {
        const int nrb = (nb - C) + nxb;
        for (int q = 0; q < nrb; ++q, ++idx) {
          if ((idx & 3) != wave) continue;
          const int R = q < nb - C ? C + q : 4 + (q - (nb - C));
          double* out = W + (16 * C + (lane & 15)) * kWld + 16 * R + (lane >> 4);
          d4 acc = d4{out[0], out[4], out[8], out[12]};
          acc = lds_block_update(W, acc, R, C, jb, lane);
          out[0] = acc[0];
          out[4] = acc[1];
          out[8] = acc[2];
          out[12] = acc[3];
        }
      }
      lds_barrier();
      read_row(jb + 16);
      lds_barrier();
    }
  }
  if (tl.kind == 0 && badcol && badcol <= nc && tid == 0) atomicMin(info_out, tl.c0 + badcol);
  // the last block column, four columns per wave
  write_back(16 * (nb - 1) + wave, 4, 4);
}

// The tiles of a level's panels of at most 16 columns, four to a workgroup (one
// wave each, as k_panel_factor_regs<16>), in a launch of 256-thread workgroups
// whose first ceil(nregs / 4) workgroups take them: the level's panels of both
// width classes are then factored in one launch.  Returns false for the
// workgroups after those.
__device__ inline bool regs_tiles_in_block(double* __restrict__ L,
                                           const Tile* __restrict__ ftiles, int rlist0,
                                           int nregs, const double* __restrict__ dscratch,
                                           int* __restrict__ info_out, float* __restrict__ Lf) {
  if ((int)blockIdx.x >= (nregs + 3) >> 2) return false;
  const int t = 4 * (int)blockIdx.x + __builtin_amdgcn_readfirstlane((int)threadIdx.x >> 6);
  if (t < nregs) {
    const Tile tl = ftiles[rlist0 + t];
    panel_factor_regs_body<16>(L, tl, dscratch, info_out, Lf, threadIdx.x & 63);
  }
  return true;
}

// Workgroups [0, ceil(nregs / 4)): the level's panels of at most 16 columns
// (ftiles[rlist0 ..]); the others: one tile of a 17..64-column panel each
// (ftiles[list0 ..]).
__global__ void __launch_bounds__(256)
k_panel_factor_blk(double* __restrict__ L, const Tile* __restrict__ ftiles, int list0,
                   const double* __restrict__ dscratch, int* __restrict__ info_out,
                   float* __restrict__ Lf, int rlist0, int nregs) {
  __shared__ double lds[kBlkLds];
  if (regs_tiles_in_block(L, ftiles, rlist0, nregs, dscratch, info_out, Lf)) return;
  const Tile tl = ftiles[list0 + blockIdx.x - ((nregs + 3) >> 2)];
  panel_factor_blk_body(L, tl, dscratch, info_out, Lf, lds, threadIdx.x);
}

// ---- dependency-driven factorization ----------------------------------------
// One persistent launch for the plan's levels from dag_level0 on (swc::DagItem):
// as many workgroups as the device holds at once claim work items from a ticket
// counter, in the plan's item order.  An item applies one level's updates to one
// tile with tile_update_body and, when that level completes the tile's panel,
// factors the tile's rows with the panel factor bodies; the arithmetic is that of
// the per-level launches, so the factor is bitwise the same.  Before it starts, an
// item waits for the tile's previous writer and for every panel its updates read
// to be factored (all of the panel's tiles); a factoring tile below the diagonal
// also waits for its panel's diagonal tile to be updated (the copy of the
// diagonal block it factors from).  Item order is a topological order of these
// waits, so an item only waits for items claimed earlier by running workgroups
// that wait for nothing later: the launch cannot deadlock.  Every wait is also
// bounded (kDagWaitTicks of the constant 100 MHz clock); a wait that runs out
// sets the abort flag, every workgroup then stops, and the host reports the
// failure.  Publication: every wave waits for its stores to complete, a barrier,
// then one thread's agent-scope release store of the flag (which writes the
// XCD's cache back first: the device's L2 caches are per XCD); the waiting side
// polls with relaxed agent-scope loads, then the polling wave takes an
// agent-scope acquire fence (invalidating the CU's and the XCD's caches) before
// the barrier after which the workgroup loads.
// sync: [0] ticket counter, [1] abort flag, [2, 2 + nitems) item states (1: the
// tile's updates are written), then per panel the count of its factored tiles.
constexpr int kDagLds = kBlkLds > 2 * swc::PW * kLdsLd ? kBlkLds : 2 * swc::PW * kLdsLd;
constexpr uint64_t kDagWaitTicks = 500000000ull;  // 5 s

__device__ inline int ld_agent(const int* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// One wave polls until the tile's previous writer (prev >= 0) has written it
// and the panels deps[dep0, dep1) are factored; false when the launch aborts.
__device__ bool dag_poll(const int* __restrict__ state, const int* __restrict__ pdone,
                         const swc::DagDep* __restrict__ deps, int prev, int dep0, int dep1,
                         int* __restrict__ abort_flag, int lane) {
  const uint64_t t0 = wall_clock64();
  for (;;) {
    bool ok = !(lane == 0 && prev >= 0) || ld_agent(state + prev) >= 1;
    for (int e = dep0 + lane; e < dep1; e += 64) {
      const swc::DagDep dp = deps[e];
      ok = ok && ld_agent(pdone + dp.panel) >= dp.count;
    }
    if (__builtin_amdgcn_ballot_w64(!ok) == 0) return true;
    if (ld_agent(abort_flag) != 0) return false;
    if (wall_clock64() - t0 > kDagWaitTicks) {
      if (lane == 0) __hip_atomic_store(abort_flag, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      return false;
    }
    __builtin_amdgcn_s_sleep(2);
  }
}

template <int MODE>
__global__ void __launch_bounds__(kThreads) __attribute__((amdgpu_waves_per_eu(2)))
k_dag_factor(double* __restrict__ L, const Tile* __restrict__ tiles,
             const Task* __restrict__ tasks, const swc::DagItem* __restrict__ items, int nitems,
             const swc::DagDep* __restrict__ deps, int* __restrict__ sync,
             double* __restrict__ dscratch, int* __restrict__ info_out, float* Lf, int write_f32) {
  __shared__ double lds[kDagLds];
  __shared__ int s_claim, s_ok;
  int* const ticket = sync;
  int* const abort_flag = sync + 1;
  int* const state = sync + 2;
  int* const pdone = sync + 2 + nitems;
  const int tid = threadIdx.x, lane = tid & 63, wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  float* const lf = write_f32 ? Lf : nullptr;
  for (;;) {
    if (tid == 0) {
      const int c = atomicAdd(ticket, 1);
      s_claim = ld_agent(abort_flag) != 0 ? nitems : c;
    }
    __syncthreads();
    const int k = s_claim;
    if (k >= nitems) break;
    {
      const swc::DagItem d = items[k];
      if (wave == 0) {
        const bool ok = dag_poll(state, pdone, deps, d.prev, d.dep0, d.dep1, abort_flag, lane);
        // one acquire for the workgroup: it invalidates the CU's and the XCD's
        // caches, and the other waves load only after the barrier below
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
        if (lane == 0) s_ok = ok;
      }
      __syncthreads();
      if (!s_ok) break;
      // the thread index is made opaque per item, so that the compiler does not
      // hoist the bodies' per-lane address arithmetic out of the loop and keep it
      // in registers across items
      int t1 = tid;
      asm volatile("" : "+v"(t1));
      tile_update_body<MODE>(L, tiles[d.tile], tasks, dscratch, Lf, lds, t1);
    }
    // every wave's stores complete before the barrier; the one release below then
    // writes them back from the XCD's cache before the flag is set
    __builtin_amdgcn_s_waitcnt(0);
    __syncthreads();
    // the item and its tile are read again rather than kept in registers over the
    // updates (an index the compiler cannot see through)
    int k2 = k;
    asm volatile("" : "+s"(k2));
    const swc::DagItem d = items[k2];
    const Tile tl = tiles[d.tile];
    if (d.fclass < 0 || tl.kind == 0) {  // someone waits for these updates
      if (tid == 0) __hip_atomic_store(state + k2, 1, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
    }
    if (d.fclass >= 0) {
      if (wave == 0) {
        const bool ok = d.diag < 0 || dag_poll(state, pdone, deps, d.diag, 0, 0, abort_flag, lane);
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
        if (lane == 0) s_ok = ok;
      }
      __syncthreads();
      if (!s_ok) break;
      int t2 = tid;
      asm volatile("" : "+v"(t2));
      if (d.fclass == 0) {
        if (wave == 0) panel_factor_regs_body<16>(L, tl, dscratch, info_out, lf, t2 & 63);
      } else {
        panel_factor_blk_body(L, tl, dscratch, info_out, lf, lds, t2);
      }
      __builtin_amdgcn_s_waitcnt(0);
      __syncthreads();
      if (tid == 0)
        __hip_atomic_fetch_add(pdone + tl.panel, 1, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
    }
    __syncthreads();
  }
}

// A level's panel factorization (panels of at most 16 columns in the first
// ceil(nregs / 4) workgroups, four tiles each; then panels of 17..64 columns,
// nfac workgroups: k_panel_factor_blk) in the same launch as the level's updates
// of tiles whose panels it does not factor (the remaining workgroups:
// k_tile_update), so that those updates, which nothing of this level waits for,
// overlap the factorization instead of delaying it.  Same bodies, same results.
template <int MODE>
__global__ void __launch_bounds__(kThreads) __attribute__((amdgpu_waves_per_eu(2)))
k_factor_and_update(double* __restrict__ L, const Tile* __restrict__ ftiles, int list0, int nfac,
                    const Tile* __restrict__ tiles, const Task* __restrict__ tasks, int tile0,
                    double* __restrict__ dscratch, int* __restrict__ info_out, float* Lf_factor,
                    const float* __restrict__ Lf_update, int rlist0, int nregs) {
  __shared__ double lds[kDagLds];
  if (regs_tiles_in_block(L, ftiles, rlist0, nregs, dscratch, info_out, Lf_factor)) return;
  const int b = (int)blockIdx.x - ((nregs + 3) >> 2);
  if (b < nfac) {
    const Tile tl = ftiles[list0 + b];
    panel_factor_blk_body(L, tl, dscratch, info_out, Lf_factor, lds, threadIdx.x);
  } else {
    const Tile tl = tiles[tile0 + (b - nfac)];
    tile_update_body<MODE>(L, tl, tasks, dscratch, Lf_update, lds, threadIdx.x);
  }
}

// CSC export of the factor (the layout of the reference's bcsc2csc,
// common/Util.h:311-338): column j's entries are its supernode's rows from its
// own diagonal down.  col[j] = {value offset of L(j, j), offset in Ls of row j,
// entry count, offset in the output}; one wave per column.
struct CscColumn {
  int64_t v0, r0, len, c0;
};

__global__ void k_factor_csc(const double* __restrict__ L, const int* __restrict__ Ls,
                             const CscColumn* __restrict__ col, int n, int64_t* __restrict__ Cp,
                             int* __restrict__ Ci, double* __restrict__ Cx) {
  const int j = blockIdx.x * (kThreads / 64) + (threadIdx.x >> 6), lane = threadIdx.x & 63;
  if (j >= n) return;
  const CscColumn c = col[j];
  for (int64_t e = lane; e < c.len; e += 64) {
    Cx[c.c0 + e] = L[c.v0 + e];
    Ci[c.c0 + e] = Ls[c.r0 + e];
  }
  if (lane == 0) {
    Cp[j + 1] = c.c0 + c.len;
    if (j == 0) Cp[0] = 0;
  }
}

// ---- triangular solves --------------------------------------------------
// Forward L y = b (blockedLsolve, Triangular_BCSC.h:14-49, restated level by
// level).  The reference pushes L21(s) x_s into the rows below each supernode
// s; here every supernode s keeps that product as its contribution vector V_s,
// and column c takes the contributions of all its descendants in increasing
// supernode order (the order of the pushes) when its supernode starts.  Each
// panel then solves its diagonal block; its row blocks push L21 x into the
// supernode's later columns (in place) or into V_s.  The right-hand sides are
// accumulated in X and the panel solutions written to Y, so that no launch
// reads what it writes.  Backward L^T x = y: per panel, L21^T x over the rows
// below (ancestors, already final in X) and the transposed diagonal solve,
// written to X.

// x[c] -= V[gidx[e]] for e in [gptr[c], gptr[c+1]), one thread per gathered column.
__global__ void __launch_bounds__(kThreads)
k_fwd_gather(const int* __restrict__ gcols, int g0, int ng, const int64_t* __restrict__ gptr,
             const int64_t* __restrict__ gidx, const double* __restrict__ V, int64_t vld,
             double* __restrict__ X, int nrhs, int64_t ldx) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  const int r = blockIdx.y;  // right-hand side
  if (t >= ng || r >= nrhs) return;
  const int c = gcols[g0 + t];
  const int64_t e0 = gptr[c], e1 = gptr[c + 1];
  const double* v = V + (int64_t)r * vld;
  double acc = X[(int64_t)r * ldx + c];
#pragma unroll 8
  for (int64_t e = e0; e < e1; ++e) acc -= v[gidx[e]];
  X[(int64_t)r * ldx + c] = acc;
}

// Substitution of one 64-column chunk for RB right-hand sides at once, lane i
// holding row i of each: x_j = t_j / L_jj, then t_i -= L_ij x_j (forward,
// row[j] = L_ij, j increasing) or t_i -= L_ji x_j (backward, row[j] = L_ji, the
// transposed chunk, j decreasing).  row[j] must be zero wherever lane i does
// not depend on x_j (j >= i forward, j <= i backward, and rows outside the
// chunk), so every lane applies every step unmasked and its own entry is left
// as t_i after the steps it depends on: x_i = t_i / L_ii at the end.  The RB
// chains are independent and interleaved step by step, so their latencies
// overlap.
template <bool FORWARD, int RB>
__device__ inline void substitute64(const double (&row)[64], double rinv, double (&ti)[RB]) {
#pragma unroll
  for (int s = 0; s < 64; ++s) {
    const int j = FORWARD ? s : 63 - s;
    double xj[RB];
#pragma unroll
    for (int q = 0; q < RB; ++q) xj[q] = bcast(ti[q] * rinv, j);  // lane j scales its entry
#pragma unroll
    for (int q = 0; q < RB; ++q) ti[q] = fma(-row[j], xj[q], ti[q]);
  }
#pragma unroll
  for (int q = 0; q < RB; ++q) ti[q] *= rinv;
}

// Forward push of a row block for RB right-hand sides: part[w][q][i] holds wave
// w's share of (L21 x_q)(row i); the shares are added in wave order and the sum
// goes to a later column of the supernode (subtracted from x) or to a row of an
// ancestor (added to the supernode's contribution vector V).
template <int RB>
__device__ inline void push_rows(const double (&part)[kThreads / 64][RB][swc::TM],
                                 const Panel& sp, const swc::SolveBlock& bk, int rb0, int nq,
                                 double* __restrict__ X, int64_t ldx, double* __restrict__ V,
                                 int64_t vld) {
  for (int e = threadIdx.x; e < RB * swc::TM; e += kThreads) {
    const int q = e / swc::TM, i = e % swc::TM;
    if (q < nq && i < bk.nr) {
      const double y = ((part[0][q][i] + part[1][q][i]) + part[2][q][i]) + part[3][q][i];
      const int p = (sp.c0 - sp.sc0) + bk.r0 + i;  // position in the supernode's rows
      if (p < sp.sw) X[(int64_t)(rb0 + q) * ldx + sp.sc0 + p] -= y;
      else V[(int64_t)(rb0 + q) * vld + sp.voff + (p - sp.sw)] += y;
    }
  }
}

// One forward row block of a solve panel (up to SPW = 128 columns of one
// supernode, two 64-column chunks) for a group of RB right-hand sides
// (blockIdx.y; the last group may hold fewer, padded with zero chains): solve
// the panel's diagonal block (every block of the panel, from the same inputs),
// write the solution if this is the panel's writer block, then push L21 x of
// the block's rows.  L is read once per group, not once per right-hand side.
// Each wave loads what its step needs at the start, so all loads are in flight
// at once: wave 0 the first chunk's triangle (lane = row), wave 1 the second
// chunk's rows of the first chunk's columns, wave 2 the second chunk's
// triangle, and every wave its share of the block's rows.  Then wave 0
// substitutes the first chunk, wave 1 subtracts its products from the second
// chunk's entries, wave 2 substitutes the second chunk.
static_assert(swc::SPW == 128, "the solve kernels split a solve panel in two 64-column chunks");
template <int RB>
__global__ void __launch_bounds__(kThreads)
k_fwd_panel(const double* __restrict__ L, const Panel* __restrict__ spanels,
            const swc::SolveBlock* __restrict__ blocks, int b0, double* __restrict__ V,
            int64_t vld, double* __restrict__ X, double* __restrict__ Y, int nrhs, int64_t ldx,
            int64_t ldy) {
  __shared__ double xs[RB][swc::SPW];
  __shared__ double part[kThreads / 64][RB][swc::TM];
  const swc::SolveBlock bk = blocks[b0 + blockIdx.x];
  const Panel sp = spanels[bk.panel];
  const int tid = threadIdx.x, lane = tid & 63, wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int nc = sp.ncol, na0 = min(64, nc), na1 = nc - 64;  // chunk widths (na1 <= 0: one chunk)
  const double* D = L + sp.vbase;  // the panel's diagonal block, column j at D + j * ld
  const int rb0 = blockIdx.y * RB, nq = min(RB, nrhs - rb0);  // this group's right-hand sides
  // xs[q][c] = x_q of the panel's columns (zero for padded right-hand sides)
  // The right-hand sides are loaded into registers before the panel's entries
  // and written to LDS after them: loads complete in issue order, so this way
  // the write waits only for its own loads, and the panel's entries arrive while
  // the substitution runs (the barriers below wait for LDS traffic only).
  constexpr int NX = (RB * swc::SPW + kThreads - 1) / kThreads;
  double xv[NX];
#pragma unroll
  for (int u = 0; u < NX; ++u) {
    const int e = tid + u * kThreads, q = e / swc::SPW, c = e % swc::SPW;
    const bool in = e < RB * swc::SPW && c < nc && q < nq;
    const double v = X[in ? (int64_t)(rb0 + q) * ldx + sp.c0 + c : 0];
    xv[u] = in ? v : 0.0;
  }
  auto stage = [&]() {
#pragma unroll
    for (int u = 0; u < NX; ++u) {
      const int e = tid + u * kThreads, q = e / swc::SPW, c = e % swc::SPW;
      if (e < RB * swc::SPW && c < nc) xs[q][c] = xv[u];
    }
  };
  auto store = [&]() {
    for (int e = tid; e < RB * swc::SPW; e += kThreads) {
      const int q = e / swc::SPW, c = e % swc::SPW;
      if (q < nq && c < nc) Y[(int64_t)(rb0 + q) * ldy + sp.c0 + c] = xs[q][c];
    }
  };
  if (nc <= 64) {
    // one chunk: wave 0 substitutes it, then the block's rows are pushed
    stage();
    lds_barrier();
    if (wave == 0) {
      // row `lane` (clamped into the diagonal block), columns uniform over the wave
      const int rl = min(lane, nc - 1);
      const double dval = D[(int64_t)rl * sp.ld + rl];
      double row[64];
#pragma unroll
      for (int c = 0; c < 64; ++c) {
        const bool in = lane < nc && c < lane;
        const double val = (D + (int64_t)min(c, nc - 1) * sp.ld)[rl];
        row[c] = in ? val : 0.0;
      }
      const double rinv = lane < nc ? 1.0 / dval : 1.0;
      double ti[RB];
#pragma unroll
      for (int q = 0; q < RB; ++q) ti[q] = lane < nc ? xs[q][lane] : 0.0;
      substitute64<true, RB>(row, rinv, ti);
      if (lane < nc) {
#pragma unroll
        for (int q = 0; q < RB; ++q) xs[q][lane] = ti[q];
      }
    }
    lds_barrier();
    if (bk.writer) store();
    if (bk.nr > 0) {
      double acc[RB] = {};
      const int rb = bk.r0 + min(lane, bk.nr - 1);  // row clamped into the block
#pragma unroll
      for (int k = 0; k < 16; ++k) {
        const int j = wave + 4 * k;
        const bool in = lane < bk.nr && j < nc;
        const double val = (D + (int64_t)min(j, nc - 1) * sp.ld)[rb];
        const double lv = in ? val : 0.0;
#pragma unroll
        for (int q = 0; q < RB; ++q) acc[q] = fma(lv, j < nc ? xs[q][j] : 0.0, acc[q]);
      }
#pragma unroll
      for (int q = 0; q < RB; ++q) part[wave][q][lane] = acc[q];
      lds_barrier();
      push_rows<RB>(part, sp, bk, rb0, nq, X, ldx, V, vld);
    }
    return;
  }
  // Loads are issued in the order they are needed (they complete in issue
  // order): this wave's diagonal entry, its step's 64 columns, then its share of
  // the block's rows, which are used last.
  // this wave's step: element (row0 + lane, col0 + c), c < 64, where valid
  // Every load takes a column base that is uniform over the wave (scalar
  // arithmetic) and a row clamped into the panel once per lane; entries outside
  // the wanted range are read from a valid address and replaced by zero.
  const int row0 = wave == 0 ? 0 : 64, col0 = wave == 2 ? 64 : 0;
  const int nrow = wave == 0 ? na0 : na1, ncol = wave == 2 ? na1 : na0;
  const int rlast = sp.nrow - 1;  // last row of the panel's columns
  const bool dg = (wave == 0 || wave == 2) && lane < nrow;
  const int rs = min(row0 + lane, rlast);  // this lane's row of the step
  const double dval = D[(int64_t)min(row0 + lane, nc - 1) * sp.ld + rs];
  double m[64];
#pragma unroll
  for (int c = 0; c < 64; ++c) {
    const bool in = wave < 3 && lane < nrow && c < ncol && (wave == 1 || c < lane);
    const double val = (D + (int64_t)min(col0 + c, nc - 1) * sp.ld)[rs];
    m[c] = in ? val : 0.0;
  }
  // the block's rows: wave w holds columns j = w, w+4, ... of row `lane`
  constexpr int NQ = swc::SPW / 4;
  double bv[NQ];
  const int rb = min(bk.r0 + lane, rlast);
#pragma unroll
  for (int k = 0; k < NQ; ++k) {
    const int j = wave + 4 * k;
    const bool in = lane < bk.nr && j < nc;
    const double val = (D + (int64_t)min(j, nc - 1) * sp.ld)[rb];
    bv[k] = in ? val : 0.0;
  }
  // reciprocal of the diagonal entry; 1 beyond the chunk (padded unknowns stay 0)
  const double rinv = dg ? 1.0 / dval : 1.0;
  stage();
  lds_barrier();
  if (wave == 0) {
    double ti[RB];
#pragma unroll
    for (int q = 0; q < RB; ++q) ti[q] = lane < na0 ? xs[q][lane] : 0.0;
    substitute64<true, RB>(m, rinv, ti);
    if (lane < na0) {
#pragma unroll
      for (int q = 0; q < RB; ++q) xs[q][lane] = ti[q];
    }
  }
  lds_barrier();
  if (na1 > 0) {
    if (wave == 1 && lane < na1) {
#pragma unroll
      for (int q = 0; q < RB; ++q) {  // one right-hand side at a time, as in the push
        double acc = 0.0;
#pragma unroll
        for (int c = 0; c < 64; ++c) acc = fma(m[c], xs[q][c], acc);  // na0 = 64 here
        xs[q][64 + lane] -= acc;
        __builtin_amdgcn_sched_barrier(0);
      }
    }
    lds_barrier();
    if (wave == 2) {
      double ti[RB];
#pragma unroll
      for (int q = 0; q < RB; ++q) ti[q] = lane < na1 ? xs[q][64 + lane] : 0.0;
      substitute64<true, RB>(m, rinv, ti);
      if (lane < na1) {
#pragma unroll
        for (int q = 0; q < RB; ++q) xs[q][64 + lane] = ti[q];
      }
    }
    lds_barrier();
  }
  if (bk.writer) store();
  if (bk.nr > 0) {
    // wave w sums columns j = w, w+4, ... for row `lane`, in increasing j
    // (one right-hand side at a time: the barrier keeps the scheduler from
    // hoisting every right-hand side's LDS reads at once, which would not fit
    // the registers)
#pragma unroll
    for (int q = 0; q < RB; ++q) {
      double acc = 0.0;
#pragma unroll
      for (int k = 0; k < NQ; ++k) {
        const int j = wave + 4 * k;
        acc = fma(bv[k], j < nc ? xs[q][j] : 0.0, acc);
      }
      part[wave][q][lane] = acc;
      __builtin_amdgcn_sched_barrier(0);
    }
    lds_barrier();
    push_rows<RB>(part, sp, bk, rb0, nq, X, ldx, V, vld);
  }
}

// Narrow solve panels (at most SPW_NARROW = 16 columns): one wave per row block
// and per panel, everything in registers.  Forward: substitute the panel's
// triangle (lane i holds row i), store the solution from the writer block, and
// push L21 x of the block's rows (lane = row).
__global__ void __launch_bounds__(64)
k_fwd_narrow(const double* __restrict__ L, const Panel* __restrict__ spanels,
             const swc::SolveBlock* __restrict__ blocks, int b0, double* __restrict__ V,
             int64_t vld, double* __restrict__ X, double* __restrict__ Y, int64_t ldx,
             int64_t ldy) {
  constexpr int NW = swc::SPW_NARROW;
  const swc::SolveBlock bk = blocks[b0 + blockIdx.x];
  const Panel sp = spanels[bk.panel];
  const int lane = threadIdx.x, r = blockIdx.y;
  const int nc = sp.ncol;
  double* x = X + (int64_t)r * ldx;
  double* v = V + (int64_t)r * vld;
  const double* D = L + sp.vbase;
  const double* blk = D + bk.r0;
  double row[NW], lv[NW];
#pragma unroll
  for (int c = 0; c < NW; ++c) {
    row[c] = (lane < nc && c < lane) ? D[(int64_t)c * sp.ld + lane] : 0.0;
    lv[c] = (lane < bk.nr && c < nc) ? blk[(int64_t)c * sp.ld + lane] : 0.0;
  }
  const double rinv = lane < nc ? 1.0 / D[(int64_t)lane * sp.ld + lane] : 1.0;
  double ti = lane < nc ? x[sp.c0 + lane] : 0.0;
#pragma unroll
  for (int j = 0; j < NW; ++j)  // row[j] = 0 for j >= lane: unmasked steps
    ti = fma(-row[j], bcast(ti * rinv, j), ti);  // lane j scales its own entry
  ti *= rinv;
  if (bk.writer && lane < nc) Y[(int64_t)r * ldy + sp.c0 + lane] = ti;
  if (bk.nr > 0) {
    double acc = 0.0;
#pragma unroll
    for (int c = 0; c < NW; ++c) acc = fma(lv[c], bcast(ti, c), acc);
    if (lane < bk.nr) {
      const int p = (sp.c0 - sp.sc0) + bk.r0 + lane;  // position in the supernode's rows
      if (p < sp.sw) x[sp.sc0 + p] -= acc;
      else v[sp.voff + (p - sp.sw)] += acc;
    }
  }
}

// Backward row block of a narrow panel: partial[j] = sum over the block's rows
// of L(row, j) x(row).  Lane i forms row i's products and stores them in LDS;
// lane l then sums column l mod 16 over rows 16g..16g+15 (g = l / 16) in order,
// and two butterfly steps add the four row groups in a fixed order.
__global__ void __launch_bounds__(64)
k_bwd_partial_narrow(const double* __restrict__ L, const int* __restrict__ Ls,
                     const Panel* __restrict__ spanels,
                     const swc::SolveBlock* __restrict__ blocks, int b0, int bl,
                     double* __restrict__ Pb, int64_t pld, const double* __restrict__ X,
                     int64_t ldx) {
  constexpr int NW = swc::SPW_NARROW;
  const swc::SolveBlock bk = blocks[b0 + blockIdx.x];  // nr = 0 writes zeros
  const Panel sp = spanels[bk.panel];
  const int lane = threadIdx.x, r = blockIdx.y;
  const int nc = sp.ncol;
  const double* blk = L + sp.vbase + bk.r0;
  const double xr = lane < bk.nr ? X[(int64_t)r * ldx + Ls[sp.rows_off + bk.r0 + lane]] : 0.0;
  static_assert(NW == 16, "lane l sums column l mod 16 of a quarter of the rows");
  __shared__ double pr[64][NW + 1];  // pr[row][column] (padded against bank conflicts)
#pragma unroll
  for (int c = 0; c < NW; ++c)
    pr[lane][c] = (lane < bk.nr && c < nc) ? blk[(int64_t)c * sp.ld + lane] * xr : 0.0;
  lds_barrier();
  const int cl = lane % NW, g = lane / NW;
  double sum = 0.0;
#pragma unroll
  for (int rr = 0; rr < 16; ++rr) sum += pr[16 * g + rr][cl];
  sum += __shfl_xor(sum, 16);
  sum += __shfl_xor(sum, 32);
  double* out = Pb + (int64_t)r * pld + (int64_t)(b0 + blockIdx.x - bl) * swc::SPW;
  if (lane < nc) out[lane] = sum;  // lanes 0..15 hold columns 0..15
}

// Backward narrow panel: t = y_p - (sum of its row blocks' partials, in block
// order), then the transposed substitution (lane i holds column i of L11).
__global__ void __launch_bounds__(64)
k_bwd_diag_narrow(const double* __restrict__ L, const Panel* __restrict__ spanels,
                  const int* __restrict__ order, int p0, const int* __restrict__ fblk_ptr,
                  int bl, const double* __restrict__ Pb, int64_t pld,
                  const double* __restrict__ Y, double* __restrict__ X, int64_t ldx,
                  int64_t ldy) {
  constexpr int NW = swc::SPW_NARROW;
  const int qi = p0 + blockIdx.x;
  const Panel sp = spanels[order[qi]];
  const int lane = threadIdx.x, r = blockIdx.y;
  const int nc = sp.ncol;
  const int f0 = fblk_ptr[qi], f1 = fblk_ptr[qi + 1];
  const double* D = L + sp.vbase;
  double col[NW];
#pragma unroll
  for (int j = 0; j < NW; ++j) col[j] = (lane < nc && j > lane && j < nc) ? D[(int64_t)lane * sp.ld + j] : 0.0;
  const double rinv = lane < nc ? 1.0 / D[(int64_t)lane * sp.ld + lane] : 1.0;
  double ti = 0.0;
  if (lane < nc) {
    const double* pb = Pb + (int64_t)r * pld;
    double sum = 0.0;
    for (int b = f0; b < f1; ++b) sum += pb[(int64_t)(b - bl) * swc::SPW + lane];
    ti = Y[(int64_t)r * ldy + sp.c0 + lane] - sum;
  }
#pragma unroll
  for (int j = NW - 1; j >= 0; --j)  // col[j] = 0 for j <= lane: unmasked steps
    ti = fma(-col[j], bcast(ti * rinv, j), ti);  // lane j scales its own entry
  ti *= rinv;
  if (lane < nc) X[(int64_t)r * ldx + sp.c0 + lane] = ti;
}

// Backward row block: partial[j] = sum over the block's rows of L(row, j) x(row)
// for the solve panel's columns j, into Pb at the block's slot in its level
// (its index minus the level's first block bl), for a group of RB right-hand
// sides (blockIdx.y), so the block of L is read once per group.
// Thread t takes column t mod SPW and half t / SPW of the block's rows, reads
// them (contiguous in memory) all at once, and sums them in row order against
// x(rows) staged in LDS; the two halves are then added, first half first.
static_assert(kThreads == 2 * swc::SPW && swc::TM % 2 == 0, "two threads per column");
template <int RB>
__global__ void __launch_bounds__(kThreads)
k_bwd_partial(const double* __restrict__ L, const int* __restrict__ Ls,
              const Panel* __restrict__ spanels, const swc::SolveBlock* __restrict__ blocks,
              int b0, int bl, double* __restrict__ Pb, int64_t pld, const double* __restrict__ X,
              int nrhs, int64_t ldx) {
  constexpr int H = swc::TM / 2;
  __shared__ double xr[RB][swc::TM];
  __shared__ double part[2][RB][swc::SPW];
  const swc::SolveBlock bk = blocks[b0 + blockIdx.x];  // nr = 0 writes zeros
  const Panel sp = spanels[bk.panel];
  const int tid = threadIdx.x, c = tid % swc::SPW, h = tid / swc::SPW;
  const int nc = sp.ncol, nr = bk.nr;
  const int rb0 = blockIdx.y * RB, nq = min(RB, nrhs - rb0);  // this group's right-hand sides
  // column c, rows [H h, H h + H) of the block (entries outside it read the
  // panel's first diagonal entry, which always exists, and are replaced by zero)
  // (the column clamped into the panel; a full block reads its 32 rows from one
  // base address, a partial one clamps each row into the block)
  const double* colp = L + sp.vbase + (int64_t)min(c, nc - 1) * sp.ld + bk.r0;
  double lv[H];
  if (nr == swc::TM) {
#pragma unroll
    for (int i = 0; i < H; ++i) {
      const double val = colp[H * h + i];
      lv[i] = c < nc ? val : 0.0;
    }
  } else if (nr > 0) {
#pragma unroll
    for (int i = 0; i < H; ++i) {
      const bool in = c < nc && H * h + i < nr;
      const double val = colp[min(H * h + i, nr - 1)];
      lv[i] = in ? val : 0.0;
    }
  } else {
#pragma unroll
    for (int i = 0; i < H; ++i) lv[i] = 0.0;
  }
  for (int e = tid; e < RB * swc::TM; e += kThreads) {
    const int q = e / swc::TM, i = e % swc::TM;
    if (i < nr)
      xr[q][i] = q < nq ? X[(int64_t)(rb0 + q) * ldx + Ls[sp.rows_off + bk.r0 + i]] : 0.0;
  }
  lds_barrier();
  // one right-hand side at a time (the barrier bounds the LDS reads in flight
  // to what the registers hold)
#pragma unroll
  for (int q = 0; q < RB; ++q) {
    double acc = 0.0;
#pragma unroll
    for (int i = 0; i < H; ++i) acc = fma(lv[i], H * h + i < nr ? xr[q][H * h + i] : 0.0, acc);
    part[h][q][c] = acc;
    __builtin_amdgcn_sched_barrier(0);
  }
  lds_barrier();
  const int64_t slot = (int64_t)(b0 + blockIdx.x - bl) * swc::SPW;
  for (int e = tid; e < RB * swc::SPW; e += kThreads) {
    const int q = e / swc::SPW, cc = e % swc::SPW;
    if (q < nq && cc < nc) Pb[(int64_t)(rb0 + q) * pld + slot + cc] = part[0][q][cc] + part[1][q][cc];
  }
}

// t = y - (sum of the panel's row-block partials) for RB right-hand sides, into
// ts: thread t sums half t / SPW of the blocks for column t mod SPW, in block
// order, kGather loads in flight per step; the halves are added first half first.
// Every column of ts is written (zero past nc and for padded right-hand sides).
constexpr int kGather = 32;
template <int RB>
__device__ inline void gather_partials(double (&ts)[RB][swc::SPW], const Panel& sp, int f0,
                                       int f1, int b0, const double* __restrict__ Pb, int64_t pld,
                                       const double* __restrict__ Y, int64_t ldy, int rb0,
                                       int nq) {
  const int tid = threadIdx.x, c = tid % swc::SPW, h = tid / swc::SPW, nc = sp.ncol;
  const int fm = f0 + (f1 - f0 + 1) / 2, fa = h == 0 ? f0 : fm, fb = h == 0 ? fm : f1;
  double sum[RB] = {};
  if (c < nc) {
    for (int b = fa; b < fb; b += kGather) {
#pragma unroll
      for (int q = 0; q < RB; ++q) {
        if (q < nq) {  // uniform
          const double* pb = Pb + (int64_t)(rb0 + q) * pld + c;
          double pv[kGather];
#pragma unroll
          for (int u = 0; u < kGather; ++u) {
            // clamped to the last block, so every load is inside the partials
            const double val = pb[(int64_t)(min(b + u, fb - 1) - b0) * swc::SPW];
            pv[u] = b + u < fb ? val : 0.0;
          }
#pragma unroll
          for (int u = 0; u < kGather; ++u) sum[q] += pv[u];
        }
        __builtin_amdgcn_sched_barrier(0);  // kGather loads in flight, not kGather RB
      }
    }
  }
  if (h == 1) {
#pragma unroll
    for (int q = 0; q < RB; ++q) ts[q][c] = sum[q];
  }
  lds_barrier();
  if (h == 0 && c < nc) {
#pragma unroll
    for (int q = 0; q < RB; ++q)
      ts[q][c] = q < nq ? Y[(int64_t)(rb0 + q) * ldy + sp.c0 + c] - (sum[q] + ts[q][c]) : 0.0;
  }
}

// Backward solve panel: t = y_p - (sum of its row blocks' partials, in block
// order), then L11^T x_p = t chunk by chunk from the last, for a group of RB
// right-hand sides (blockIdx.y).  As in the forward
// kernel each wave loads its step's operands first: wave 0 the second
// chunk's transposed triangle (lane = column), wave 1 the second chunk's rows
// of the first chunk's columns (lane = column), wave 2 the first chunk's
// transposed triangle.  Wave 0 substitutes the second chunk, wave 1 subtracts
// its products from the first chunk's entries, wave 2 substitutes the first.
template <int RB>
__global__ void __launch_bounds__(kThreads)
k_bwd_diag(const double* __restrict__ L, const Panel* __restrict__ spanels,
           const int* __restrict__ order, int p0, const int* __restrict__ fblk_ptr, int b0,
           const double* __restrict__ Pb, int64_t pld, const double* __restrict__ Y,
           double* __restrict__ X, int nrhs, int64_t ldx, int64_t ldy) {
  __shared__ double ts[RB][swc::SPW];
  const int qi = p0 + blockIdx.x;
  const Panel sp = spanels[order[qi]];
  const int tid = threadIdx.x, lane = tid & 63, wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int nc = sp.ncol, na0 = min(64, nc), na1 = nc - 64;  // chunk widths (na1 <= 0: one chunk)
  const int f0 = fblk_ptr[qi], f1 = fblk_ptr[qi + 1];
  const double* D = L + sp.vbase;
  const int rb0 = blockIdx.y * RB, nq = min(RB, nrhs - rb0);  // this group's right-hand sides
  auto store = [&]() {
    for (int e = tid; e < RB * swc::SPW; e += kThreads) {
      const int q = e / swc::SPW, c = e % swc::SPW;
      if (q < nq && c < nc) X[(int64_t)(rb0 + q) * ldx + sp.c0 + c] = ts[q][c];
    }
  };
  if (nc <= 64) {
    // one chunk: wave 0 loads its transposed triangle while the partials are
    // summed, then substitutes it
    double col[64];
    double rinv = 1.0;
    if (wave == 0) {
      // column `lane` (clamped into the panel), rows uniform over the wave
      const double* colp = D + (int64_t)min(lane, nc - 1) * sp.ld;
      const double dval = colp[min(lane, nc - 1)];
#pragma unroll
      for (int j = 0; j < 64; ++j) {
        const bool in = lane < nc && j > lane && j < nc;
        const double val = colp[min(j, nc - 1)];
        col[j] = in ? val : 0.0;
      }
      rinv = lane < nc ? 1.0 / dval : 1.0;
    }
    gather_partials<RB>(ts, sp, f0, f1, b0, Pb, pld, Y, ldy, rb0, nq);
    lds_barrier();
    if (wave == 0) {
      double ti[RB];
#pragma unroll
      for (int q = 0; q < RB; ++q) ti[q] = lane < nc ? ts[q][lane] : 0.0;
      substitute64<false, RB>(col, rinv, ti);
      if (lane < nc) {
#pragma unroll
        for (int q = 0; q < RB; ++q) ts[q][lane] = ti[q];
      }
    }
    lds_barrier();
    store();
    return;
  }
  // this wave's step: m[j] = L(rowc + j, colc + lane) where valid
  const int colc = wave == 0 ? 64 : 0, rowc = wave == 2 ? 0 : 64;
  const int ncl = wave == 0 ? na1 : na0, nrw = wave == 2 ? na0 : na1;
  // column colc + lane (clamped into the panel), rows uniform over the wave;
  // the diagonal entry first, as loads complete in issue order
  const double* colp = D + (int64_t)min(colc + lane, nc - 1) * sp.ld;
  const bool dg = (wave == 0 || wave == 2) && lane < ncl;
  const double dval = colp[min(colc + lane, nc - 1)];
  double m[64];
#pragma unroll
  for (int j = 0; j < 64; ++j) {
    const bool in = wave < 3 && lane < ncl && j < nrw && (wave == 1 || j > lane);
    const double val = colp[min(rowc + j, nc - 1)];
    m[j] = in ? val : 0.0;
  }
  const double rinv = dg ? 1.0 / dval : 1.0;
  gather_partials<RB>(ts, sp, f0, f1, b0, Pb, pld, Y, ldy, rb0, nq);
  lds_barrier();
  if (na1 > 0) {
    if (wave == 0) {
      double ti[RB];
#pragma unroll
      for (int q = 0; q < RB; ++q) ti[q] = lane < na1 ? ts[q][64 + lane] : 0.0;
      substitute64<false, RB>(m, rinv, ti);
      if (lane < na1) {
#pragma unroll
        for (int q = 0; q < RB; ++q) ts[q][64 + lane] = ti[q];
      }
    }
    lds_barrier();
    // first-chunk entries: t_i -= sum_j L(64 + j, i) x_(64 + j); ts is zero
    // past nc (gather_partials writes every column), so no read is masked
    if (wave == 1) {
#pragma unroll
      for (int q = 0; q < RB; ++q) {  // one right-hand side at a time (register bound)
        double acc = 0.0;
#pragma unroll
        for (int j = 0; j < 64; ++j) acc = fma(m[j], ts[q][64 + j], acc);
        ts[q][lane] -= acc;  // na0 = 64 here
        __builtin_amdgcn_sched_barrier(0);
      }
    }
    lds_barrier();
  }
  if (wave == 2) {
    double ti[RB];
#pragma unroll
    for (int q = 0; q < RB; ++q) ti[q] = lane < na0 ? ts[q][lane] : 0.0;
    substitute64<false, RB>(m, rinv, ti);
    if (lane < na0) {
#pragma unroll
      for (int q = 0; q < RB; ++q) ts[q][lane] = ti[q];
    }
  }
  lds_barrier();
  store();
}

// y[k] = b[perm[k]] (to the factor's ordering) or x[perm[k]] = y[k] (back).
__global__ void k_permute(const double* __restrict__ in, double* __restrict__ out,
                          const int* __restrict__ perm, int n, int nrhs, int64_t ldin,
                          int64_t ldout, int inverse) {
  const int64_t total = (int64_t)n * nrhs;
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < total;
       e += (int64_t)gridDim.x * blockDim.x) {
    const int k = (int)(e % n);
    const int r = (int)(e / n);
    if (!inverse) out[r * ldout + k] = in[r * ldin + perm[k]];
    else out[r * ldout + perm[k]] = in[r * ldin + k];
  }
}

}  // namespace

// ---- residual, norms and vector updates (FP64) ---------------------------

constexpr int kRedBlocks = 512;

// Workspace of swc_solve_noalloc, the entry point that allocates nothing: one
// copy of this array per GPU, held from the library's load until the process
// ends.  Host code never uses its name as an address: swc_create looks up each
// device's copy once (scratch_ptr) and everything else uses that pointer.
constexpr int64_t kScratchDoubles = (int64_t)16 << 20;  // 128 MiB
static __device__ __align__(256) double scratch[kScratchDoubles];

// Per-device address of `scratch`, looked up once per device with that device
// current (hipGetSymbolAddress); null until then.
std::mutex g_scratch_mu;
std::vector<double*> g_scratch_ptr;

int bind_scratch(int device) {
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if ((int)g_scratch_ptr.size() <= device) g_scratch_ptr.resize(device + 1, nullptr);
  if (g_scratch_ptr[device]) return SWC_OK;
  void* ptr = nullptr;
  const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
  if (e != hipSuccess || !ptr)
    return swc_set_error(SWC_EDEVICE, std::string("no-alloc workspace: hipGetSymbolAddress: ") +
                                          hipGetErrorString(e));
  g_scratch_ptr[device] = (double*)ptr;
  return SWC_OK;
}

double* scratch_ptr(int device) {
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  return device < (int)g_scratch_ptr.size() ? g_scratch_ptr[device] : nullptr;
}

// Af[e] = vals[vmap[e]]: the full P A P^T values in CSR order, from the caller's values.
__global__ void k_gather(double* __restrict__ out, const double* __restrict__ vals,
                         const int64_t* __restrict__ vmap, int64_t nnz) {
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < nnz;
       e += (int64_t)gridDim.x * blockDim.x)
    out[e] = vals[vmap[e]];
}

// Y = A X, or R = B - A X when B is given, for all nrhs columns (CSR): a group
// of 8 lanes per row, lane k taking the row's entries k, k+8, ... (consecutive
// lanes read consecutive entries), right-hand sides 4 at a time against the
// same entries, each sum completed by a fixed butterfly over the group.  Whole
// waves step through the rows together (the butterfly needs every lane).
constexpr int kMvGroup = 8;
__global__ void __launch_bounds__(kThreads)
k_csr_mv(const int* __restrict__ rp, const int* __restrict__ ci, const double* __restrict__ av,
         const double* __restrict__ X, const double* __restrict__ B, double* __restrict__ Y, int n,
         int64_t ld, int nrhs) {
  constexpr int G = kMvGroup, RPW = 64 / G;  // rows per wave
  const int lane = threadIdx.x & 63, k = lane % G;
  const int64_t wid = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) / 64;
  const int64_t nw = (int64_t)gridDim.x * blockDim.x / 64;
  for (int64_t base = wid * RPW; base < n; base += nw * RPW) {
    const int64_t i = base + lane / G;
    const bool valid = i < n;
    const int e0 = valid ? rp[i] : 0, e1 = valid ? rp[i + 1] : 0;
    for (int c0 = 0; c0 < nrhs; c0 += 4) {
      double acc[4] = {0.0, 0.0, 0.0, 0.0};
      for (int e = e0 + k; e < e1; e += G) {
        const double a = av[e];
        const int64_t col = ci[e];
#pragma unroll
        for (int q = 0; q < 4; ++q)
          if (c0 + q < nrhs) acc[q] = fma(a, X[(int64_t)(c0 + q) * ld + col], acc[q]);
      }
#pragma unroll
      for (int q = 0; q < 4; ++q)
#pragma unroll
        for (int off = G / 2; off >= 1; off /= 2) acc[q] += __shfl_xor(acc[q], off, G);
      if (valid && k < 4 && c0 + k < nrhs) {
        // lane k of the group writes right-hand side c0 + k
        const double t = k == 0 ? acc[0] : k == 1 ? acc[1] : k == 2 ? acc[2] : acc[3];
        const int64_t o = (int64_t)(c0 + k) * ld + i;
        Y[o] = B ? B[o] - t : t;
      }
    }
  }
}

// Partial maxima of the row sums of |A| (||A||_inf), one per block of a fixed grid.
__global__ void k_norm_inf(const int* __restrict__ rp, const double* __restrict__ av, int n,
                           double* __restrict__ part) {
  __shared__ double red[kThreads];
  double m = 0.0;
  for (int i = blockIdx.x * kThreads + threadIdx.x; i < n; i += kRedBlocks * kThreads) {
    double t = 0.0;
    for (int e = rp[i]; e < rp[i + 1]; ++e) t += fabs(av[e]);
    m = fmax(m, t);
  }
  red[threadIdx.x] = m;
  __syncthreads();
  for (int w = kThreads / 2; w > 0; w >>= 1) {
    if (threadIdx.x < w) red[threadIdx.x] = fmax(red[threadIdx.x], red[threadIdx.x + w]);
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}

// Per-column partial reductions over a fixed grid (deterministic): op 0 max|x|,
// op 1 dot(x, y).  part[c * kRedBlocks + block].
__global__ void k_reduce_partial(const double* __restrict__ X, const double* __restrict__ Y,
                                 int n, int64_t ld, int op, double* __restrict__ part) {
  __shared__ double red[kThreads];
  const int64_t c = blockIdx.y;
  const double* x = X + c * ld;
  const double* y = Y ? Y + c * ld : nullptr;
  double acc = 0.0;
  for (int i = blockIdx.x * kThreads + threadIdx.x; i < n; i += kRedBlocks * kThreads)
    acc = op == 0 ? fmax(acc, fabs(x[i])) : fma(x[i], y[i], acc);
  red[threadIdx.x] = acc;
  __syncthreads();
  for (int w = kThreads / 2; w > 0; w >>= 1) {
    if (threadIdx.x < w)
      red[threadIdx.x] = op == 0 ? fmax(red[threadIdx.x], red[threadIdx.x + w])
                                 : red[threadIdx.x] + red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[c * kRedBlocks + blockIdx.x] = red[0];
}

__global__ void k_reduce_final(const double* __restrict__ part, int op, double* __restrict__ out) {
  __shared__ double red[kThreads];
  const int c = blockIdx.x;
  double acc = 0.0;
  for (int i = threadIdx.x; i < kRedBlocks; i += kThreads) {
    const double v = part[(int64_t)c * kRedBlocks + i];
    acc = op == 0 ? fmax(acc, v) : acc + v;
  }
  red[threadIdx.x] = acc;
  __syncthreads();
  for (int w = kThreads / 2; w > 0; w >>= 1) {
    if (threadIdx.x < w)
      red[threadIdx.x] = op == 0 ? fmax(red[threadIdx.x], red[threadIdx.x + w])
                                 : red[threadIdx.x] + red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[c] = red[0];
}

// Y[:, c] += alpha[c] * X[:, c]   (sign folded into alpha by the caller)
__global__ void k_axpy_cols(double* __restrict__ Y, const double* __restrict__ X,
                            const double* __restrict__ alpha, int n, int64_t ld) {
  const int64_t c = blockIdx.y;
  const double a = alpha[c];
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
    Y[c * ld + i] = fma(a, X[c * ld + i], Y[c * ld + i]);
}

// P[:, c] = Z[:, c] + beta[c] * P[:, c]
__global__ void k_xpby_cols(double* __restrict__ P, const double* __restrict__ Z,
                            const double* __restrict__ beta, int n, int64_t ld) {
  const int64_t c = blockIdx.y;
  const double b = beta[c];
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
    P[c * ld + i] = fma(b, P[c * ld + i], Z[c * ld + i]);
}

// ---- adjoint: selected inverse on the pattern of L ------------------------
// Z = A^{-1} restricted to the pattern of L (in L's BCSC layout), by the
// supernodal Takahashi recurrences, one level of supernodes at a time (four
// launches), root level first.  For supernode J with L = [L11; L21] (rows R
// below):
//   T   = L11^{-1}
//   U   = L21 T                       (recomputed from L, not stored by the factor)
//   Z21 = -Z(R,R) U                   (Z(R,R) gathered from the ancestors' Z)
//   Z11 = T^T T - U^T Z21
// This is the reverse sweep of the blocked Cholesky adjoint for
// d logdet(A) / dA = A^{-1} (no reference counterpart).

struct SelinvNode {
  int64_t vbase;    // value offset of the supernode (row 0, column 0)
  int64_t scratch;  // offset of its Tt | Ut | Zrr scratch
  int w, nrow;      // width, rows
  int c0;           // first column
  int rows_off;     // offset into Ls
};

__device__ inline int lower_pos(const int* a, int len, int v) {
  int lo = 0, hi = len;
  while (lo < hi) {
    const int mid = (lo + hi) >> 1;
    if (a[mid] < v) lo = mid + 1;
    else hi = mid;
  }
  return lo;
}

// The sweep of one level runs in four launches, each spread over blockIdx.y
// workgroups per supernode (a level near the root holds one or a few large
// supernodes).  Scratch per supernode: Tt (W x W, Tt[k W + x] = T(k, x), row
// major), Ut (m x W, Ut[j W + x] = U(j, x), row major) and Zrr (m x m), so that
// the threads of a wave (consecutive x or i) read consecutive addresses, and
// the operand shared by the wave is one broadcast load.  Every entry is
// accumulated in the same order as a one-thread-per-entry loop would.

// T = L11^{-1}, column j per lane, one wave per workgroup (64 columns from
// j0): T(i, j) = 0 above the diagonal, T(j, j) = 1 / L(j, j), and below it
// T(i, j) = -(sum_{k=j}^{i-1} L(i, k) T(k, j)) / L(i, i), the sum taken in
// increasing k (the wave's columns start it at j0; the terms k < j are zero).
// Rows go in blocks of 64: a lane keeps its column's sums for the block's rows
// in registers (x[r]); the finished rows above the block (k < i0) are applied
// 64 columns of L at a time from LDS (each load of T(k, j) serves 64 rows),
// then the block's own triangle, row by row, turns x[r] into T(i0 + r, j).
__global__ void __launch_bounds__(64)
k_selinv_T(const double* __restrict__ L, const SelinvNode* __restrict__ nodes, int node0,
           double* __restrict__ arena) {
  __shared__ double lb[64][64];  // lb[q][r] = L(i0 + r, k0 + q)
  __shared__ double tl[64][64];  // tl[q][lane] = T(k0 + q, j)
  const SelinvNode nd = nodes[node0 + blockIdx.x];
  const int W = nd.w, R = nd.nrow;
  const int lane = (int)threadIdx.x;
  const int j0 = blockIdx.y * 64, j = j0 + lane;
  if (j0 >= W) return;  // the whole workgroup
  const double* L11 = L + nd.vbase;  // ld R
  double* Tt = arena + nd.scratch;   // Tt[k W + j] = T(k, j)
  const bool mine = j < W;
  for (int i0 = j0; i0 < W; i0 += 64) {
    const int nb = min(64, W - i0);
    double x[64];
#pragma unroll
    for (int r = 0; r < 64; ++r) x[r] = 0.0;
    for (int k0 = j0; k0 < i0; k0 += 64) {  // k0 + 63 < i0: whole chunks
      __syncthreads();  // the previous chunk has been read
#pragma unroll 16
      for (int q = 0; q < 64; ++q) {
        lb[q][lane] = lane < nb ? L11[(int64_t)(k0 + q) * R + i0 + lane] : 0.0;
        tl[q][lane] = mine ? Tt[(int64_t)(k0 + q) * W + j] : 0.0;
      }
      __syncthreads();
#pragma unroll 1
      for (int q = 0; q < 64; ++q) {
        const double t = tl[q][lane];
#pragma unroll
        for (int r = 0; r < 64; ++r) {
          x[r] = fma(lb[q][r], t, x[r]);
          if (r % 16 == 15) __builtin_amdgcn_sched_barrier(0);  // 16 LDS reads in flight
        }
      }
    }
    // the block's triangle: lb[c][r] = L(i0 + r, i0 + c)
    __syncthreads();
#pragma unroll 8
    for (int q = 0; q < 64; ++q)
      lb[q][lane] = lane < nb && i0 + q < W ? L11[(int64_t)(i0 + q) * R + i0 + lane] : 0.0;
    __syncthreads();
#pragma unroll
    for (int r = 0; r < 64; ++r) {
      const int i = i0 + r;
      double sum = x[r];
#pragma unroll
      for (int c = 0; c < r; ++c) {
        sum = fma(lb[c][r], x[c], sum);
        if (c % 16 == 15) __builtin_amdgcn_sched_barrier(0);  // 16 LDS reads in flight
      }
      __builtin_amdgcn_sched_barrier(0);
      const double d = lb[r][r];  // L(i, i) (0 past the block's rows)
      x[r] = i < j ? 0.0 : (i == j ? 1.0 / d : -sum / d);
    }
    if (mine) {
#pragma unroll
      for (int r = 0; r < 64; ++r)
        if (r < nb) Tt[(int64_t)(i0 + r) * W + j] = x[r];
    }
  }
}

// U = L21 T (entry (i, c): sum over k >= c of L21(i, k) T(k, c)), and Zrr = Z(R, R),
// both triangles, gathered from the ancestors' selected inverse.
__global__ void __launch_bounds__(kThreads)
k_selinv_UZ(const double* __restrict__ L, const double* __restrict__ Z,
            const int* __restrict__ Ls, const int* __restrict__ col2sup,
            const int64_t* __restrict__ spx, const int* __restrict__ ssuper,
            const int64_t* __restrict__ spi, const SelinvNode* __restrict__ nodes, int node0,
            double* __restrict__ arena) {
  const SelinvNode nd = nodes[node0 + blockIdx.x];
  const int W = nd.w, R = nd.nrow, m = R - W;
  const double* L21 = L + nd.vbase + W;  // ld R
  const double* Tt = arena + nd.scratch;
  double* Ut = arena + nd.scratch + (int64_t)W * W;
  double* Zrr = Ut + (int64_t)m * W;
  const int* rows = Ls + nd.rows_off;
  const int64_t nU = (int64_t)m * W, nZ = (int64_t)m * m;
  for (int64_t e = (int64_t)blockIdx.y * kThreads + threadIdx.x; e < nU + nZ;
       e += (int64_t)gridDim.y * kThreads) {
    if (e < nU) {
      const int i = (int)(e % m), c = (int)(e / m);
      double sum = 0.0;
      for (int k = c; k < W; ++k) sum = fma(L21[(int64_t)k * R + i], Tt[(int64_t)k * W + c], sum);
      Ut[(int64_t)i * W + c] = sum;
    } else {
      const int64_t f = e - nU;
      const int i = (int)(f % m), j = (int)(f / m);
      if (i < j) continue;
      const int ri = rows[W + i], rj = rows[W + j];
      const int K = col2sup[rj];
      const int rk0 = (int)spi[K];
      const int lenK = (int)(spi[K + 1] - spi[K]);
      const int pos = lower_pos(Ls + rk0, lenK, ri);
      const double z = Z[spx[K] + (int64_t)(rj - ssuper[K]) * lenK + pos];
      Zrr[(int64_t)j * m + i] = z;
      Zrr[(int64_t)i * m + j] = z;
    }
  }
}

// Z21 = -Zrr U, written into Z (rows W.. of the supernode).
__global__ void __launch_bounds__(kThreads)
k_selinv_Z21(double* __restrict__ Z, const SelinvNode* __restrict__ nodes, int node0,
             const double* __restrict__ arena) {
  const SelinvNode nd = nodes[node0 + blockIdx.x];
  const int W = nd.w, R = nd.nrow, m = R - W;
  const double* Ut = arena + nd.scratch + (int64_t)W * W;
  const double* Zrr = Ut + (int64_t)m * W;
  double* Zs = Z + nd.vbase;
  for (int64_t e = (int64_t)blockIdx.y * kThreads + threadIdx.x; e < (int64_t)m * W;
       e += (int64_t)gridDim.y * kThreads) {
    const int i = (int)(e % m), c = (int)(e / m);
    double sum = 0.0;
    for (int j = 0; j < m; ++j) sum = fma(Zrr[(int64_t)j * m + i], Ut[(int64_t)j * W + c], sum);
    Zs[(int64_t)c * R + W + i] = -sum;
  }
}

// Z11 = T^T T - U^T Z21 (lower triangle; the strict upper part stays zero).
__global__ void __launch_bounds__(kThreads)
k_selinv_Z11(double* __restrict__ Z, const SelinvNode* __restrict__ nodes, int node0,
             const double* __restrict__ arena) {
  const SelinvNode nd = nodes[node0 + blockIdx.x];
  const int W = nd.w, R = nd.nrow, m = R - W;
  const double* Tt = arena + nd.scratch;
  const double* Ut = Tt + (int64_t)W * W;
  double* Zs = Z + nd.vbase;
  for (int64_t e = (int64_t)blockIdx.y * kThreads + threadIdx.x; e < (int64_t)W * W;
       e += (int64_t)gridDim.y * kThreads) {
    const int i = (int)(e % W), c = (int)(e / W);
    if (i < c) {
      Zs[(int64_t)c * R + i] = 0.0;
      continue;
    }
    double sum = 0.0;
    for (int k = i; k < W; ++k) sum = fma(Tt[(int64_t)k * W + i], Tt[(int64_t)k * W + c], sum);
    double sub = 0.0;
    for (int j = 0; j < m; ++j) sub = fma(Ut[(int64_t)j * W + i], Zs[(int64_t)c * R + W + j], sub);
    Zs[(int64_t)c * R + i] = sum - sub;
  }
}

// logdet partials: sum of log L_jj over the diagonal entries (fixed grid).
__global__ void k_logdet_partial(const double* __restrict__ L, const int64_t* __restrict__ diag,
                                 int n, double* __restrict__ part) {
  __shared__ double red[kThreads];
  double acc = 0.0;
  for (int j = blockIdx.x * kThreads + threadIdx.x; j < n; j += kRedBlocks * kThreads)
    acc += log(L[diag[j]]);
  red[threadIdx.x] = acc;
  __syncthreads();
  for (int w = kThreads / 2; w > 0; w >>= 1) {
    if (threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}

// Gradient w.r.t. the input values of A (each stored lower entry a_q stands for
// a_ij and a_ji):  g_q = w_q (alpha Z_ij - beta x_i x_j),  w_q = 1 on the
// diagonal, 2 off it; Z at zpos[q] (offset in L's layout), x in factor order.
__global__ void k_grad_values(const double* __restrict__ Zv, const int64_t* __restrict__ zpos,
                              const int* __restrict__ qi, const int* __restrict__ qj,
                              const double* __restrict__ x, double alpha, double beta,
                              int64_t nvals, double* __restrict__ g) {
  for (int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; q < nvals;
       q += (int64_t)gridDim.x * blockDim.x) {
    const int i = qi[q], j = qj[q];
    const double w = i == j ? 1.0 : 2.0;
    double v = 0.0;
    if (alpha != 0.0) v += alpha * Zv[zpos[q]];
    if (beta != 0.0) v -= beta * x[i] * x[j];
    g[q] = w * v;
  }
}

// ---------------------------------------------------------------------------
// handle
// ---------------------------------------------------------------------------

struct swc_handle {
  int device = 0;
  hipStream_t stream = nullptr;
  Plan plan;
  swc::PlanParams params;  // shape parameters of the next structure upload
  swc::HLevelSchedule hlevel;  // non-empty: factor levels follow this schedule
  int n = 0, nsuper = 0;
  int64_t nvals = 0;  // number of values swc_factor expects
  std::vector<int> super, Ls;
  std::vector<int64_t> pi, px;
  std::vector<int> perm;  // perm[new] = old, empty if not set
  std::vector<int> pruneptr, pruneset;
  // device buffers
  double* dL = nullptr;
  float* dLf = nullptr;  // FP32 shadow of L, written with it when FP32 updates are on
  int64_t lf_cap = 0;    // elements allocated at dLf
  int* dLs = nullptr;
  Panel* dPanels = nullptr;
  Tile* dTiles = nullptr;
  Task* dTasks = nullptr;
  int64_t* dGptr = nullptr;  // forward gathers: per column, offsets into dGidx
  int64_t* dGidx = nullptr;
  int* dGcols = nullptr;     // gathered columns by level
  swc::SolveBlock* dFblk = nullptr;
  int* dFblkPtr = nullptr;  // per panel (level order): its forward/backward row blocks
  double* dV = nullptr;   // contribution vectors, vsize * work_cols
  double* dPb = nullptr;  // backward partial sums, max_lvl_fblk * SPW * work_cols
  Panel* dSpanels = nullptr;  // solve panels
  int* dSpOrder = nullptr;    // solve panels by level
  double* dY = nullptr;   // forward solutions, n * work_cols
  int* dOrder = nullptr;
  Tile* dFtiles = nullptr;  // factor-kernel tiles (copies), by level and width class
  swc::FusedPanel* dFused = nullptr;  // fused narrow panels, by level
  int* dCounters = nullptr;  // per level: the fused kernel's worklist counter
  int fused_cap[3] = {0, 0, 0};  // per precision mode: fused-kernel workgroups resident at once
  swc::DagItem* dDagItems = nullptr;  // dependency-driven launch: work items, dependencies,
  swc::DagDep* dDagDeps = nullptr;    // and its ticket / abort / per-item / per-panel flags
  int* dDagSync = nullptr;
  int dag_cap[3] = {0, 0, 0};  // per precision mode: its workgroups resident at once
  double* dDiag = nullptr;  // per-level copies of the diagonal blocks being factored
  int64_t* dAmap = nullptr;
  int* dPerm = nullptr;
  double* dVals = nullptr;  // staging for host values
  double* dWork = nullptr;  // solve workspace, n * work_cols doubles
  int work_cols = 0;
  int* dInfo = nullptr;
  // full P A P^T (factor ordering) for residuals: CSR pattern, value map, values
  int* dRp = nullptr;
  int* dCi = nullptr;
  int64_t* dVmap = nullptr;
  double* dAf = nullptr;
  int64_t nnz_full = 0;
  // small device scalars written by captured launch sequences ([0] ||A||_inf,
  // [1] logdet): allocated with the structure and never moved while a graph
  // that writes them exists
  double* dNorm = nullptr;
  double* dScal = nullptr;   // refinement scalars, sized by ensure_work (eager launches only)
  double* dNormPart = nullptr;  // kRedBlocks partial maxima of ||A||_inf
  double* dPart = nullptr;   // reduction partials, kRedBlocks * work_cols
  double* dRef = nullptr;    // refinement vectors, 5 * n * work_cols
  double normA = 0;
  // adjoint (selected inverse) state, built on first use
  double* dZ = nullptr;          // selected inverse, xsize
  double* dArena = nullptr;      // per-level T | U | Zrr scratch
  SelinvNode* dNodes = nullptr;
  int* dCol2sup = nullptr;
  int64_t* dSpx = nullptr;
  int* dSsuper = nullptr;
  int64_t* dSpi = nullptr;
  int64_t* dDiagPos = nullptr;   // offset of L(j,j), j in factor order
  int* dQi = nullptr;            // factor-order row / column of every input value
  int* dQj = nullptr;
  std::vector<int> sel_lvl;      // level -> node offsets (root level first)
  std::vector<int> sel_wmax;     // per level: widest supernode
  std::vector<int64_t> sel_emax; // per level: most scratch entries of one supernode
  bool zvalid = false;
  int precision = 0;  // 0 FP64; 1 FP32 descendant updates; 2 FP32 for every update
  int fp32_mask = 0;  // bit k: tasks of kind k use FP32 MFMA
  int factored_precision = 0;
  bool factored = false;
  bool has_values = false;  // the A2 pattern (value map, residual CSR) is uploaded
  int last_info = 0;
  // captured launch sequences (hipGraph): factor per precision mode, solves per
  // (buffer, nrhs, ld); rebuilt when the structure changes
  bool use_graphs = true;
  hipGraphExec_t fgraph[3] = {nullptr, nullptr, nullptr};
  struct SolveGraph {
    const double* x;
    int nrhs;
    int64_t ld;
    hipGraphExec_t exec;
  };
  std::vector<SolveGraph> sgraphs;
  // timing (device events on `stream`)
  hipEvent_t ev[4] = {nullptr, nullptr, nullptr, nullptr};
  bool profile = false;
  std::vector<hipEvent_t> pev;  // per-launch pairs when profiling
  double t_factor_ms = 0, t_solve_ms = 0, t_tile_ms = 0, t_trsm_ms = 0;
  std::vector<double> lvl_tile_ms;  // per level, when profiling
  int n_tile = 0, n_trsm = 0;

  ~swc_handle() { release(); }
  void drop_graphs() {
    for (auto& g : fgraph)
      if (g) (void)hipGraphExecDestroy(g);
    for (auto& g : fgraph) g = nullptr;
    for (auto& g : sgraphs) (void)hipGraphExecDestroy(g.exec);
    sgraphs.clear();
  }
  void release() {
    drop_graphs();
    (void)hipFree(dL); (void)hipFree(dLf); (void)hipFree(dLs); (void)hipFree(dPanels); (void)hipFree(dTiles); (void)hipFree(dTasks);
    (void)hipFree(dGptr); (void)hipFree(dGidx); (void)hipFree(dGcols); (void)hipFree(dFblk); (void)hipFree(dFblkPtr); (void)hipFree(dV); (void)hipFree(dPb); (void)hipFree(dY); (void)hipFree(dSpanels); (void)hipFree(dSpOrder); (void)hipFree(dOrder); (void)hipFree(dFtiles); (void)hipFree(dFused); (void)hipFree(dCounters); (void)hipFree(dDagItems); (void)hipFree(dDagDeps); (void)hipFree(dDagSync); (void)hipFree(dDiag); (void)hipFree(dAmap); (void)hipFree(dPerm);
    (void)hipFree(dVals); (void)hipFree(dWork); (void)hipFree(dInfo);
    (void)hipFree(dRp); (void)hipFree(dCi); (void)hipFree(dVmap); (void)hipFree(dAf); (void)hipFree(dNorm); (void)hipFree(dScal); (void)hipFree(dPart); (void)hipFree(dNormPart);
    (void)hipFree(dRef);
    (void)hipFree(dZ); (void)hipFree(dArena); (void)hipFree(dNodes); (void)hipFree(dCol2sup); (void)hipFree(dSpx);
    (void)hipFree(dSsuper); (void)hipFree(dSpi); (void)hipFree(dDiagPos); (void)hipFree(dQi); (void)hipFree(dQj);
    dL = nullptr; dLf = nullptr; dLs = nullptr; dPanels = nullptr; dTiles = nullptr; dTasks = nullptr;
    dGptr = nullptr; dGidx = nullptr; dGcols = nullptr; dFblk = nullptr; dFblkPtr = nullptr; dV = nullptr; dPb = nullptr; dY = nullptr; dSpanels = nullptr; dSpOrder = nullptr; dOrder = nullptr; dFtiles = nullptr; dFused = nullptr; dCounters = nullptr; dDagItems = nullptr; dDagDeps = nullptr; dDagSync = nullptr; dDiag = nullptr; dAmap = nullptr; dPerm = nullptr;
    dVals = nullptr; dWork = nullptr; dInfo = nullptr;
    dRp = nullptr; dCi = nullptr; dVmap = nullptr; dAf = nullptr; dNorm = nullptr; dScal = nullptr; dPart = nullptr; dNormPart = nullptr;
    lf_cap = 0;
    dRef = nullptr;
    dZ = nullptr; dArena = nullptr; dNodes = nullptr; dCol2sup = nullptr; dSpx = nullptr;
    dSsuper = nullptr; dSpi = nullptr; dDiagPos = nullptr; dQi = nullptr; dQj = nullptr;
    for (auto& e : ev)
      if (e) (void)hipEventDestroy(e);
    for (auto& e : pev) (void)hipEventDestroy(e);
    pev.clear();
    if (stream) (void)hipStreamDestroy(stream);
    stream = nullptr;
  }
};

namespace {

template <class T>
int upload(T** dst, const T* src, size_t count, hipStream_t s) {
  (void)hipFree(*dst);
  *dst = nullptr;
  const size_t alloc = count == 0 ? 1 : count;
  SWC_HIP(hipMalloc((void**)dst, alloc * sizeof(T)));
  if (src && count) SWC_HIP(hipMemcpyAsync(*dst, src, count * sizeof(T), hipMemcpyHostToDevice, s));
  return SWC_OK;
}

int grid_for(int64_t n) { return (int)std::max<int64_t>(1, std::min<int64_t>((n + kThreads - 1) / kThreads, 65536)); }

// Value offset inside L of every entry of tril(P A P^T) given as CSC (a2p, a2i).
// With src != null, entry e of A2 came from input entry src[e] and the map is
// indexed by input entries (so swc_factor takes values in the caller's order).
int build_amap(const swc_handle* h, const int* a2p, const int* a2i, const int64_t* src,
               int64_t nvals, std::vector<int64_t>& amap) {
  const int n = h->n;
  const int64_t nnz = a2p[n];
  amap.assign(src ? nvals : nnz, -1);
  std::vector<int> col2sup(n);
  for (int s = 0; s < h->nsuper; ++s)
    for (int c = h->super[s]; c < h->super[s + 1]; ++c) col2sup[c] = s;
  for (int j = 0; j < n; ++j) {
    const int s = col2sup[j];
    const int* rows = h->Ls.data() + h->pi[s];
    const int len = (int)(h->pi[s + 1] - h->pi[s]);
    for (int64_t q = a2p[j]; q < a2p[j + 1]; ++q) {
      const int r = a2i[q];
      const int pos = int(std::lower_bound(rows, rows + len, r) - rows);
      if (r < j || pos >= len || rows[pos] != r)
        return swc_set_error(SWC_EINVAL, "A2 entry outside the structure of L");
      const int64_t off = h->px[s] + (int64_t)(j - h->super[s]) * len + pos;
      const int64_t dst = src ? src[q] : q;
      if (dst < 0 || dst >= (int64_t)amap.size())
        return swc_set_error(SWC_EINVAL, "value map index out of range");
      amap[dst] = off;
    }
  }
  for (int64_t v : amap)
    if (v < 0) return swc_set_error(SWC_EINVAL, "value map does not cover every value");
  return SWC_OK;
}

int ensure_work(swc_handle* h, int cols) {
  if (h->work_cols >= cols) return SWC_OK;
  for (auto& g : h->sgraphs) (void)hipGraphExecDestroy(g.exec);  // they captured the old buffers
  h->sgraphs.clear();
  (void)hipFree(h->dWork);
  (void)hipFree(h->dV);
  (void)hipFree(h->dY);
  (void)hipFree(h->dPb);
  (void)hipFree(h->dPart);
  (void)hipFree(h->dRef);
  h->dWork = h->dV = h->dY = h->dPb = h->dPart = h->dRef = nullptr;
  h->work_cols = 0;
  const Plan& P = h->plan;
  SWC_HIP(hipMalloc((void**)&h->dPart, sizeof(double) * (size_t)kRedBlocks * cols));
  (void)hipFree(h->dScal);
  h->dScal = nullptr;
  SWC_HIP(hipMalloc((void**)&h->dScal, sizeof(double) * (size_t)(64 + 4 * cols)));
  SWC_HIP(hipMalloc((void**)&h->dRef, sizeof(double) * (size_t)5 * h->n * cols));
  SWC_HIP(hipMalloc((void**)&h->dWork, sizeof(double) * (size_t)h->n * cols));
  SWC_HIP(hipMalloc((void**)&h->dV, sizeof(double) * (size_t)std::max<int64_t>(P.vsize, 1) * cols));
  SWC_HIP(hipMalloc((void**)&h->dY, sizeof(double) * (size_t)h->n * cols));
  SWC_HIP(hipMalloc((void**)&h->dPb, sizeof(double) * (size_t)std::max(P.max_lvl_fblk, 1) * swc::SPW * cols));
  h->work_cols = cols;
  return SWC_OK;
}

// Enqueues the whole numeric factorization of the staged values h->dVals.
void enqueue_factor(swc_handle* h, bool profile) {
  hipStream_t st = h->stream;
  const Plan& P = h->plan;
  const double* dvals = h->dVals;
  hipLaunchKernelGGL(k_fill, dim3(grid_for(P.xsize)), dim3(kThreads), 0, st, h->dL, P.xsize, 0.0);
  hipLaunchKernelGGL(k_assemble, dim3(grid_for(h->nvals)), dim3(kThreads), 0, st, h->dL,
                     (const int64_t*)h->dAmap, dvals, h->nvals);
  hipLaunchKernelGGL(k_gather, dim3(grid_for(h->nnz_full)), dim3(kThreads), 0, st, h->dAf, dvals,
                     (const int64_t*)h->dVmap, h->nnz_full);
  hipLaunchKernelGGL(k_norm_inf, dim3(kRedBlocks), dim3(kThreads), 0, st, (const int*)h->dRp,
                     (const double*)h->dAf, h->n, h->dNormPart);
  hipLaunchKernelGGL(k_reduce_final, dim3(1), dim3(kThreads), 0, st, (const double*)h->dNormPart, 0,
                     h->dNorm);
  hipLaunchKernelGGL(k_set_int, dim3(1), dim3(1), 0, st, h->dInfo, 0x7fffffff);
  hipLaunchKernelGGL(k_fill_int, dim3((std::max(P.nlevel, 1) + kThreads - 1) / kThreads),
                     dim3(kThreads), 0, st, h->dCounters, std::max(P.nlevel, 1), 0);
  // the levels from dag_level0 on run as one dependency-driven launch (per-level
  // launches when profiling, so that each level can be timed)
  const bool dag = !profile && !P.dag_items.empty();
  const int nsync = 2 + (int)P.dag_items.size() + P.npanel;
  if (dag)
    hipLaunchKernelGGL(k_fill_int, dim3((nsync + kThreads - 1) / kThreads), dim3(kThreads), 0, st,
                       h->dDagSync, nsync, 0);
  for (int l = 0; l < (dag ? P.dag_level0 : P.nlevel); ++l) {
    const int t0 = P.lvl_tiles[l], nt = P.lvl_tiles[l + 1] - t0;
    const int s0 = P.lvl_trsm[l], ns = P.lvl_trsm[l + 1] - s0;
    hipEvent_t* pe = profile ? &h->pev[(size_t)4 * l] : nullptr;
    if (pe) (void)hipEventRecord(pe[0], st);
    // the level's tiles: fused narrow panels first, then one-wave tiles, then the rest
    const int ntf = P.lvl_tiles_fused[l], ntn = P.lvl_tiles_narrow[l];
    const int f0 = P.lvl_fused[l], nf = P.lvl_fused[l + 1] - f0;
    if (nf > 0) {
      auto kern = h->precision == 0 ? k_narrow_fused<0> : (h->precision == 1 ? k_narrow_fused<1>
                                                                              : k_narrow_fused<2>);
      // about four chunks per workgroup: few claims on the one counter, and the
      // last chunks still balance the waves' loads
      const int grid = std::min(nf, h->fused_cap[h->precision]);
      const int chunk = std::max(1, (nf + 4 * grid - 1) / (4 * grid));
      hipLaunchKernelGGL(kern, dim3(grid), dim3(64), 0, st, h->dL, (const Tile*)h->dTiles,
                         (const Task*)h->dTasks, (const swc::FusedPanel*)h->dFused + f0, nf, chunk,
                         h->dCounters + l, h->dInfo, h->dLf, h->fp32_mask ? 1 : 0);
    }
    if (ntn > 0) {
      auto kern = h->precision == 0 ? k_tile_narrow<0> : (h->precision == 1 ? k_tile_narrow<1>
                                                                             : k_tile_narrow<2>);
      hipLaunchKernelGGL(kern, dim3(ntn), dim3(64), 0, st, h->dL, (const Tile*)h->dTiles,
                         (const Task*)h->dTasks, t0 + ntf, h->dDiag, (const float*)h->dLf);
    }
    // tiles of the panels this level factors, then (last in the level's list) the
    // others, which run beside the panel factorization unless profiling times the
    // update launches on their own
    const int ntd = profile ? 0 : P.lvl_tiles_deferred[l];
    const int tc0 = t0 + ntf + ntn, ntc = nt - ntf - ntn - ntd;
    if (ntc > 0) {
      auto kern = h->precision == 0 ? k_tile_update<0> : (h->precision == 1 ? k_tile_update<1>
                                                                             : k_tile_update<2>);
      hipLaunchKernelGGL(kern, dim3(ntc), dim3(kThreads), 0, st, h->dL, (const Tile*)h->dTiles,
                         (const Task*)h->dTasks, tc0, h->dDiag, (const float*)h->dLf);
    }
    if (pe) (void)hipEventRecord(pe[1], st);
    {
      // the level's panels by width class: <= 16, <= 32, <= 64 columns
      const int* cls = &P.lvl_trsm_cls[(size_t)4 * l];
      float* lf = h->fp32_mask ? h->dLf : nullptr;
      const Tile* ft = (const Tile*)h->dFtiles;
      const int nblk = ns > 0 ? cls[3] - cls[1] : 0, nreg = ns > 0 ? cls[1] - cls[0] : 0;
      // the narrow panels go into the launch of the wider ones when there is one
      // (four tiles to a workgroup), else into their own launch
      const int nrw = (nreg + 3) / 4;
      if (ntd > 0) {
        auto kern = h->precision == 0
                        ? k_factor_and_update<0>
                        : (h->precision == 1 ? k_factor_and_update<1> : k_factor_and_update<2>);
        hipLaunchKernelGGL(kern, dim3(nrw + nblk + ntd), dim3(kThreads), 0, st, h->dL, ft,
                           ns > 0 ? cls[1] : 0, nblk, (const Tile*)h->dTiles,
                           (const Task*)h->dTasks, tc0 + ntc, h->dDiag, h->dInfo, lf,
                           (const float*)h->dLf, ns > 0 ? cls[0] : 0, nreg);
      } else if (nblk > 0) {
        hipLaunchKernelGGL(k_panel_factor_blk, dim3(nrw + nblk), dim3(256), 0, st, h->dL, ft,
                           cls[1], (const double*)h->dDiag, h->dInfo, lf, cls[0], nreg);
      } else if (nreg > 0) {
        hipLaunchKernelGGL(k_panel_factor_regs<16>, dim3(nreg), dim3(64), 0, st, h->dL, ft,
                           cls[0], (const double*)h->dDiag, h->dInfo, lf);
      }
    }
    if (pe) (void)hipEventRecord(pe[2], st);
  }
  if (dag) {
    auto kern = h->precision == 0 ? k_dag_factor<0> : (h->precision == 1 ? k_dag_factor<1>
                                                                          : k_dag_factor<2>);
    const int nitems = (int)P.dag_items.size();
    const int grid = std::min(nitems, h->dag_cap[h->precision]);
    hipLaunchKernelGGL(kern, dim3(grid), dim3(kThreads), 0, st, h->dL, (const Tile*)h->dTiles,
                       (const Task*)h->dTasks, (const swc::DagItem*)h->dDagItems, nitems,
                       (const swc::DagDep*)h->dDagDeps, h->dDagSync, h->dDiag, h->dInfo, h->dLf,
                       h->fp32_mask ? 1 : 0);
  }
}

// Numeric factorization from values on the device (in upload order).
int factor_device(swc_handle* h, const double* dvals) {
  hipStream_t st = h->stream;
  const Plan& P = h->plan;
  if (dvals != h->dVals)
    SWC_HIP(hipMemcpyAsync(h->dVals, dvals, sizeof(double) * h->nvals, hipMemcpyDeviceToDevice, st));
  if (h->profile && h->pev.size() < (size_t)4 * P.nlevel) {
    for (auto& e : h->pev) (void)hipEventDestroy(e);
    h->pev.assign((size_t)4 * P.nlevel, nullptr);
    for (auto& e : h->pev) SWC_HIP(hipEventCreate(&e));
  }
  const bool graph = h->use_graphs && !h->profile;
  hipGraphExec_t& gx = h->fgraph[h->precision];
  if (graph && !gx) {
    hipGraph_t g = nullptr;
    SWC_HIP(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
    enqueue_factor(h, false);
    SWC_HIP(hipStreamEndCapture(st, &g));
    hipError_t e = hipGraphInstantiate(&gx, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (e != hipSuccess) {
      gx = nullptr;
      return swc_set_error(SWC_EDEVICE, std::string("factor graph: ") + hipGetErrorString(e));
    }
  }
  SWC_HIP(hipEventRecord(h->ev[0], st));
  if (graph) SWC_HIP(hipGraphLaunch(gx, st));
  else enqueue_factor(h, h->profile);
  SWC_HIP(hipEventRecord(h->ev[1], st));
  SWC_HIP(hipGetLastError());
  const int big = 0x7fffffff;
  int info = 0;
  // both copies target host memory of this frame and the handle: the stream is
  // synchronized before any return, also when a copy fails to enqueue
  hipError_t ce = hipMemcpyAsync(&info, h->dInfo, sizeof(int), hipMemcpyDeviceToHost, st);
  if (ce == hipSuccess)
    ce = hipMemcpyAsync(&h->normA, h->dNorm, sizeof(double), hipMemcpyDeviceToHost, st);
  int dag_abort = 0;
  const bool dag = !h->profile && !P.dag_items.empty();
  if (ce == hipSuccess && dag)
    ce = hipMemcpyAsync(&dag_abort, h->dDagSync + 1, sizeof(int), hipMemcpyDeviceToHost, st);
  SWC_HIP(hipStreamSynchronize(st));
  SWC_HIP(ce);
  if (dag_abort) {
    h->factored = false;
    return swc_set_error(SWC_EDEVICE, "factor: a dependency wait of the factor launch timed out");
  }
  float ms = 0;
  SWC_HIP(hipEventElapsedTime(&ms, h->ev[0], h->ev[1]));
  h->t_factor_ms = ms;
  if (h->profile) {
    h->t_tile_ms = h->t_trsm_ms = 0;
    h->n_tile = h->n_trsm = 0;
    h->lvl_tile_ms.assign(P.nlevel, 0.0);
    for (int l = 0; l < P.nlevel; ++l) {
      const int nt = P.lvl_tiles[l + 1] - P.lvl_tiles[l];
      const int ns = P.lvl_trsm[l + 1] - P.lvl_trsm[l];
      hipEvent_t* pe = &h->pev[(size_t)4 * l];
      if (nt > 0) {
        SWC_HIP(hipEventElapsedTime(&ms, pe[0], pe[1]));
        h->t_tile_ms += ms;
        h->lvl_tile_ms[l] = ms;
        h->n_tile++;
      }
      if (ns > 0) {
        SWC_HIP(hipEventElapsedTime(&ms, pe[1], pe[2]));
        h->t_trsm_ms += ms;
        h->n_trsm++;
      }
    }
  }
  h->last_info = info == big ? 0 : info;
  h->factored = true;
  h->factored_precision = h->precision;
  h->zvalid = false;
  return SWC_OK;
}

// Enqueues the forward and backward sweeps on device columns X (factor
// ordering); the panel kernels take the right-hand sides in groups: RF for the
// forward kernel, RP for the backward partials, RD for the backward diagonal.
// Forward sweep L Y = X: X is consumed (it accumulates the pushes), the
// solution goes to Y (ld n), V holds the contribution vectors (ld vsize).
template <int RF>
void enqueue_forward(swc_handle* h, double* dX, int nrhs, int64_t ldx, double* dY, double* dV) {
  const int ngf = (nrhs + RF - 1) / RF;
  hipStream_t st = h->stream;
  const Plan& P = h->plan;
  const int64_t vld = std::max<int64_t>(P.vsize, 1);
  const int64_t ldy = h->n;
  const Panel* sps = (const Panel*)h->dSpanels;
  const swc::SolveBlock* fbl = (const swc::SolveBlock*)h->dFblk;
  hipLaunchKernelGGL(k_fill, dim3(grid_for(vld * nrhs)), dim3(kThreads), 0, st, dV, vld * nrhs, 0.0);
  for (int l = 0; l < P.nslevel; ++l) {
    const int g0 = P.lvl_gcol[l], ng = P.lvl_gcol[l + 1] - g0;
    const int b0 = P.lvl_fblk[l], nb = P.lvl_fblk[l + 1] - b0, nbn = P.lvl_fblk_narrow[l];
    if (ng > 0)
      hipLaunchKernelGGL(k_fwd_gather, dim3((ng + kThreads - 1) / kThreads, nrhs), dim3(kThreads), 0, st,
                         (const int*)h->dGcols, g0, ng, (const int64_t*)h->dGptr,
                         (const int64_t*)h->dGidx, (const double*)dV, vld, dX, nrhs, ldx);
    if (nbn > 0)
      hipLaunchKernelGGL(k_fwd_narrow, dim3(nbn, nrhs), dim3(64), 0, st, (const double*)h->dL, sps,
                         fbl, b0, dV, vld, dX, dY, ldx, ldy);
    if (nb > nbn)
      hipLaunchKernelGGL(k_fwd_panel<RF>, dim3(nb - nbn, ngf), dim3(kThreads), 0, st, (const double*)h->dL,
                         sps, fbl, b0 + nbn, dV, vld, dX, dY, nrhs, ldx, ldy);
  }
}

// Backward sweep L^T X = Y (Y ld n), into X; Pb holds the row blocks' partials.
template <int RP, int RD>
void enqueue_backward(swc_handle* h, double* dX, int nrhs, int64_t ldx, const double* dY,
                      double* dPb) {
  const int ngp = (nrhs + RP - 1) / RP, ngd = (nrhs + RD - 1) / RD;
  hipStream_t st = h->stream;
  const Plan& P = h->plan;
  const int64_t ldy = h->n;
  const Panel* sps = (const Panel*)h->dSpanels;
  const swc::SolveBlock* fbl = (const swc::SolveBlock*)h->dFblk;
  const int64_t pld = (int64_t)std::max(P.max_lvl_fblk, 1) * swc::SPW;
  for (int l = P.nslevel - 1; l >= 0; --l) {
    const int p0 = P.lvl_sp[l], np = P.lvl_sp[l + 1] - p0, npn = P.lvl_sp_narrow[l];
    const int b0 = P.lvl_fblk[l], nb = P.lvl_fblk[l + 1] - b0, nbn = P.lvl_fblk_narrow[l];
    if (nbn > 0)
      hipLaunchKernelGGL(k_bwd_partial_narrow, dim3(nbn, nrhs), dim3(64), 0, st, (const double*)h->dL,
                         (const int*)h->dLs, sps, fbl, b0, b0, dPb, pld, (const double*)dX, ldx);
    if (nb > nbn)
      hipLaunchKernelGGL(k_bwd_partial<RP>, dim3(nb - nbn, ngp), dim3(kThreads), 0, st, (const double*)h->dL,
                         (const int*)h->dLs, sps, fbl, b0 + nbn, b0, dPb, pld, (const double*)dX,
                         nrhs, ldx);
    if (npn > 0)
      hipLaunchKernelGGL(k_bwd_diag_narrow, dim3(npn, nrhs), dim3(64), 0, st, (const double*)h->dL, sps,
                         (const int*)h->dSpOrder, p0, (const int*)h->dFblkPtr, b0,
                         (const double*)dPb, pld, dY, dX, ldx, ldy);
    if (np > npn)
      hipLaunchKernelGGL(k_bwd_diag<RD>, dim3(np - npn, ngd), dim3(kThreads), 0, st, (const double*)h->dL, sps,
                         (const int*)h->dSpOrder, p0 + npn, (const int*)h->dFblkPtr, b0,
                         (const double*)dPb, pld, dY, dX, nrhs, ldx, ldy);
  }
}

template <int RF, int RP, int RD>
void enqueue_solve_rb(swc_handle* h, double* dX, int nrhs, int64_t ldx, double* dY, double* dV,
                      double* dPb) {
  enqueue_forward<RF>(h, dX, nrhs, ldx, dY, dV);
  enqueue_backward<RP, RD>(h, dX, nrhs, ldx, dY, dPb);
}

// Groups of right-hand sides: each block of L is read once per group, while
// the substitution chains of a group run one after another in one wave.  The
// forward kernel takes 4 (at 8 it needs more than 256 registers per lane, one
// wave per SIMD); the backward partials, which have no chain, take 8, or 16
// from 16 right-hand sides on; the backward diagonal, mostly chain, takes one
// per workgroup (measured on the vector 48^3 operator at 8 and 32 RHS).
// dY (n x nrhs), dV (max(vsize, 1) x nrhs) and dPb (max(max_lvl_fblk, 1) SPW x
// nrhs) are the sweeps' workspace.
void enqueue_solve_bufs(swc_handle* h, double* dX, int nrhs, int64_t ldx, double* dY, double* dV,
                        double* dPb) {
  if (nrhs >= 16) enqueue_solve_rb<4, 16, 1>(h, dX, nrhs, ldx, dY, dV, dPb);
  else if (nrhs >= 2) enqueue_solve_rb<4, 8, 1>(h, dX, nrhs, ldx, dY, dV, dPb);
  else enqueue_solve_rb<1, 1, 1>(h, dX, nrhs, ldx, dY, dV, dPb);
}

void enqueue_solve(swc_handle* h, double* dX, int nrhs, int64_t ldx) {
  enqueue_solve_bufs(h, dX, nrhs, ldx, h->dY, h->dV, h->dPb);
}

// In-place solve of L L^T X = B on device columns X (factor ordering); needs
// ensure_work(h, nrhs).
int solve_device(swc_handle* h, double* dX, int nrhs, int64_t ldx) {
  hipStream_t st = h->stream;
  hipGraphExec_t gx = nullptr;
  if (h->use_graphs) {
    for (auto& g : h->sgraphs)
      if (g.x == dX && g.nrhs == nrhs && g.ld == ldx) gx = g.exec;
    if (!gx) {
      if (h->sgraphs.size() >= 8) {
        (void)hipGraphExecDestroy(h->sgraphs.front().exec);
        h->sgraphs.erase(h->sgraphs.begin());
      }
      hipGraph_t g = nullptr;
      SWC_HIP(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
      enqueue_solve(h, dX, nrhs, ldx);
      SWC_HIP(hipStreamEndCapture(st, &g));
      hipError_t e = hipGraphInstantiate(&gx, g, nullptr, nullptr, 0);
      (void)hipGraphDestroy(g);
      if (e != hipSuccess) return swc_set_error(SWC_EDEVICE, std::string("solve graph: ") + hipGetErrorString(e));
      h->sgraphs.push_back({dX, nrhs, ldx, gx});
    }
  }
  SWC_HIP(hipEventRecord(h->ev[2], st));
  if (gx) SWC_HIP(hipGraphLaunch(gx, st));
  else enqueue_solve(h, dX, nrhs, ldx);
  SWC_HIP(hipEventRecord(h->ev[3], st));
  SWC_HIP(hipGetLastError());
  return SWC_OK;
}

// ---- vectors in and out of the factor ordering ----------------------------

// dOut (n x nrhs, ld n, factor ordering) = P B; B host or device with ld ldb.
int to_factor(swc_handle* h, const double* B, int on_device, int64_t ldb, double* dOut, int nrhs) {
  const int n = h->n;
  hipStream_t st = h->stream;
  const double* src = B;
  int64_t lds = ldb;
  if (!on_device) {
    SWC_HIP(hipMemcpy2DAsync(h->dWork, sizeof(double) * n, B, sizeof(double) * ldb,
                             sizeof(double) * n, nrhs, hipMemcpyHostToDevice, st));
    src = h->dWork;
    lds = n;
  }
  if (!h->perm.empty())
    hipLaunchKernelGGL(k_permute, dim3(grid_for((int64_t)n * nrhs)), dim3(kThreads), 0, st, src,
                       dOut, (const int*)h->dPerm, n, nrhs, lds, (int64_t)n, 0);
  else
    SWC_HIP(hipMemcpy2DAsync(dOut, sizeof(double) * n, src, sizeof(double) * lds,
                             sizeof(double) * n, nrhs, hipMemcpyDeviceToDevice, st));
  return SWC_OK;
}

// X (host or device, ld ldx) = P^T dIn.
int from_factor(swc_handle* h, const double* dIn, double* X, int on_device, int64_t ldx, int nrhs) {
  const int n = h->n;
  hipStream_t st = h->stream;
  double* dst = on_device ? X : h->dWork;
  const int64_t ldd = on_device ? ldx : n;
  if (!h->perm.empty())
    hipLaunchKernelGGL(k_permute, dim3(grid_for((int64_t)n * nrhs)), dim3(kThreads), 0, st, dIn,
                       dst, (const int*)h->dPerm, n, nrhs, (int64_t)n, ldd, 1);
  else
    SWC_HIP(hipMemcpy2DAsync(dst, sizeof(double) * ldd, dIn, sizeof(double) * n,
                             sizeof(double) * n, nrhs, hipMemcpyDeviceToDevice, st));
  if (!on_device)
    SWC_HIP(hipMemcpy2DAsync(X, sizeof(double) * ldx, h->dWork, sizeof(double) * n,
                             sizeof(double) * n, nrhs, hipMemcpyDeviceToHost, st));
  SWC_HIP(hipStreamSynchronize(st));
  return SWC_OK;
}

// ---- refinement -------------------------------------------------------------

// Per-column reduction (op 0 max|x|, op 1 x.y) of n x ncol device columns.
int reduce_cols(swc_handle* h, const double* X, const double* Y, int ncol, int op, double* out) {
  hipStream_t st = h->stream;
  hipLaunchKernelGGL(k_reduce_partial, dim3(kRedBlocks, ncol), dim3(kThreads), 0, st, X, Y, h->n,
                     (int64_t)h->n, op, h->dPart);
  hipLaunchKernelGGL(k_reduce_final, dim3(ncol), dim3(kThreads), 0, st, (const double*)h->dPart, op,
                     h->dScal + 64);
  SWC_HIP(hipMemcpyAsync(out, h->dScal + 64, sizeof(double) * ncol, hipMemcpyDeviceToHost, st));
  SWC_HIP(hipStreamSynchronize(st));
  return SWC_OK;
}

// Y = A X, or R = B - A X when B is given, on n x nrhs device columns of
// leading dimension ld (A = the full P A P^T in CSR, uploaded with the pattern).
void launch_csr_mv(swc_handle* h, const double* X, const double* B, double* Y, int n, int64_t ld,
                   int nrhs) {
  const int64_t rows_per_block = kThreads / kMvGroup;
  const int64_t blocks = std::min<int64_t>((n + rows_per_block - 1) / rows_per_block, 2048);
  hipLaunchKernelGGL(k_csr_mv, dim3((unsigned)std::max<int64_t>(blocks, 1)), dim3(kThreads), 0,
                     h->stream, (const int*)h->dRp, (const int*)h->dCi, (const double*)h->dAf, X, B,
                     Y, n, ld, nrhs);
}

int put_coeffs(swc_handle* h, const std::vector<double>& c) {
  SWC_HIP(hipMemcpyAsync(h->dScal + 64 + 2 * c.size(), c.data(), sizeof(double) * c.size(),
                         hipMemcpyHostToDevice, h->stream));
  return SWC_OK;
}

int solve_device(swc_handle* h, double* dX, int nrhs, int64_t ldx);

// Refines X for A X = B (both n x nrhs, factor ordering, ld n) until every
// column's eta = ||b - A x||_inf / (||A||_inf ||x||_inf + ||b||_inf) <= tol.
// method 0: classical refinement x += M^{-1} r; method 1: conjugate gradients
// preconditioned with M = L L^T.  Starts from x = M^{-1} b.
int refine_device(swc_handle* h, const double* Bt, double* Xt, int nrhs, double tol, int maxiter,
                  int method, double* eta, int* iters) {
  const int n = h->n;
  const int64_t ld = n;
  hipStream_t st = h->stream;
  double* R = h->dRef + (int64_t)2 * n * nrhs;
  double* Z = h->dRef + (int64_t)3 * n * nrhs;
  double* Pv = h->dRef + (int64_t)4 * n * nrhs;
  const dim3 g(std::min(grid_for(n), 2048), nrhs);
  std::vector<double> bn(nrhs), xn(nrhs), rn(nrhs), rz(nrhs), pq(nrhs), co(nrhs);
  std::vector<char> conv(nrhs, 0);
  auto copy = [&](double* dst, const double* src) {
    return hipMemcpyAsync(dst, src, sizeof(double) * n * nrhs, hipMemcpyDeviceToDevice, st);
  };
  auto evaluate = [&](double* Rbuf, bool& done) -> int {
    launch_csr_mv(h, (const double*)Xt, Bt, Rbuf, n, ld, nrhs);
    if (int rc = reduce_cols(h, Rbuf, nullptr, nrhs, 0, rn.data())) return rc;
    if (int rc = reduce_cols(h, Xt, nullptr, nrhs, 0, xn.data())) return rc;
    done = true;
    for (int c = 0; c < nrhs; ++c) {
      const double den = h->normA * xn[c] + bn[c];
      eta[c] = den > 0 ? rn[c] / den : 0.0;
      conv[c] = eta[c] <= tol;
      done = done && conv[c];
    }
    return SWC_OK;
  };
  if (int rc = reduce_cols(h, Bt, nullptr, nrhs, 0, bn.data())) return rc;
  SWC_HIP(copy(Xt, Bt));
  if (int rc = solve_device(h, Xt, nrhs, ld)) return rc;
  int it = 0;
  bool done = false;
  if (int rc = evaluate(R, done)) return rc;
  if (method == 0) {
    for (; !done && it < maxiter; ++it) {
      SWC_HIP(copy(Z, R));
      if (int rc = solve_device(h, Z, nrhs, ld)) return rc;
      for (int c = 0; c < nrhs; ++c) co[c] = conv[c] ? 0.0 : 1.0;
      if (int rc = put_coeffs(h, co)) return rc;
      hipLaunchKernelGGL(k_axpy_cols, g, dim3(kThreads), 0, st, Xt, (const double*)Z,
                         (const double*)(h->dScal + 64 + 2 * nrhs), n, ld);
      if (int rc = evaluate(R, done)) return rc;
    }
  } else {
    if (!done) {
      SWC_HIP(copy(Z, R));
      if (int rc = solve_device(h, Z, nrhs, ld)) return rc;
      SWC_HIP(copy(Pv, Z));
      if (int rc = reduce_cols(h, R, Z, nrhs, 1, rz.data())) return rc;
    }
    for (; !done && it < maxiter; ++it) {
      launch_csr_mv(h, (const double*)Pv, nullptr, Z, n, ld, nrhs);
      if (int rc = reduce_cols(h, Pv, Z, nrhs, 1, pq.data())) return rc;
      for (int c = 0; c < nrhs; ++c) {
        if (!conv[c] && !(pq[c] > 0))
          return swc_set_error(SWC_ENOTPOS, "swc_refine: conjugate gradients broke down");
        co[c] = conv[c] ? 0.0 : rz[c] / pq[c];
      }
      if (int rc = put_coeffs(h, co)) return rc;
      hipLaunchKernelGGL(k_axpy_cols, g, dim3(kThreads), 0, st, Xt, (const double*)Pv,
                         (const double*)(h->dScal + 64 + 2 * nrhs), n, ld);
      for (int c = 0; c < nrhs; ++c) co[c] = -co[c];
      if (int rc = put_coeffs(h, co)) return rc;
      hipLaunchKernelGGL(k_axpy_cols, g, dim3(kThreads), 0, st, R, (const double*)Z,
                         (const double*)(h->dScal + 64 + 2 * nrhs), n, ld);
      if (int rc = evaluate(Z, done)) return rc;  // true residual, for eta only
      if (done) {
        ++it;
        break;
      }
      SWC_HIP(copy(Z, R));
      if (int rc = solve_device(h, Z, nrhs, ld)) return rc;
      std::vector<double> rzn(nrhs);
      if (int rc = reduce_cols(h, R, Z, nrhs, 1, rzn.data())) return rc;
      for (int c = 0; c < nrhs; ++c) {
        co[c] = (conv[c] || rz[c] == 0) ? 0.0 : rzn[c] / rz[c];
        rz[c] = rzn[c];
      }
      if (int rc = put_coeffs(h, co)) return rc;
      hipLaunchKernelGGL(k_xpby_cols, g, dim3(kThreads), 0, st, Pv, (const double*)Z,
                         (const double*)(h->dScal + 64 + 2 * nrhs), n, ld);
    }
  }
  SWC_HIP(hipGetLastError());
  if (iters) *iters = it;
  return SWC_OK;
}

// ---- adjoint ---------------------------------------------------------------

int selinv_setup(swc_handle* h) {
  if (h->dNodes) return SWC_OK;
  const int ns = h->nsuper;
  hipStream_t st = h->stream;
  // supernode levels of the descendant DAG; the sweep runs them top-down
  std::vector<int> lev(ns, 0), col2sup(h->n);
  int nlev = 0;
  std::vector<int> pp(h->pruneptr), ps(h->pruneset);
  for (int s = 0; s < ns; ++s) {
    for (int t = pp[s]; t < pp[s + 1]; ++t) lev[s] = std::max(lev[s], lev[ps[t]] + 1);
    nlev = std::max(nlev, lev[s] + 1);
    for (int c = h->super[s]; c < h->super[s + 1]; ++c) col2sup[c] = s;
  }
  std::vector<SelinvNode> nodes;
  h->sel_lvl.assign(nlev + 1, 0);
  h->sel_wmax.clear();
  h->sel_emax.clear();
  int64_t arena = 0;
  for (int l = nlev - 1, k = 0; l >= 0; --l, ++k) {
    int64_t off = 0;
    for (int s = 0; s < ns; ++s) {
      if (lev[s] != l) continue;
      SelinvNode nd;
      nd.vbase = h->px[s];
      nd.w = h->super[s + 1] - h->super[s];
      nd.nrow = (int)(h->pi[s + 1] - h->pi[s]);
      nd.c0 = h->super[s];
      nd.rows_off = (int)h->pi[s];
      nd.scratch = off;
      const int64_t m = nd.nrow - nd.w;
      off += (int64_t)nd.w * nd.w + m * nd.w + m * m;
      nodes.push_back(nd);
    }
    h->sel_lvl[k + 1] = (int)nodes.size();
    arena = std::max(arena, off);
    int wmax = 0;
    int64_t emax = 0;
    for (int q = h->sel_lvl[k]; q < h->sel_lvl[k + 1]; ++q) {
      const int64_t w = nodes[q].w, m = nodes[q].nrow - nodes[q].w;
      wmax = std::max<int>(wmax, (int)w);
      emax = std::max(emax, std::max(w * w, m * w + m * m));
    }
    h->sel_wmax.push_back(wmax);
    h->sel_emax.push_back(emax);
  }
  int rc = SWC_OK;
  if (!rc) rc = upload(&h->dNodes, nodes.data(), nodes.size(), st);
  if (!rc) rc = upload(&h->dCol2sup, col2sup.data(), col2sup.size(), st);
  if (!rc) rc = upload(&h->dSpx, h->px.data(), h->px.size(), st);
  if (!rc) rc = upload(&h->dSsuper, h->super.data(), h->super.size(), st);
  if (!rc) rc = upload(&h->dSpi, h->pi.data(), h->pi.size(), st);
  if (!rc) rc = upload(&h->dArena, (const double*)nullptr, (size_t)std::max<int64_t>(arena, 1), st);
  if (!rc) rc = upload(&h->dZ, (const double*)nullptr, (size_t)h->plan.xsize, st);
  if (rc) return rc;
  SWC_HIP(hipStreamSynchronize(st));
  return SWC_OK;
}

int selinv_device(swc_handle* h) {
  if (h->zvalid) return SWC_OK;
  if (int rc = selinv_setup(h)) return rc;
  hipStream_t st = h->stream;
  const int nl = (int)h->sel_lvl.size() - 1;
  for (int k = 0; k < nl; ++k) {
    const int n0 = h->sel_lvl[k], nn = h->sel_lvl[k + 1] - n0;
    if (nn <= 0) continue;
    const SelinvNode* nd = (const SelinvNode*)h->dNodes;
    // workgroups per supernode: one wave per 64 columns of T; for the other
    // steps about 8 entries per thread of the level's largest supernode
    const int gt = (h->sel_wmax[k] + 63) / 64;
    const int ge = (int)std::min<int64_t>(std::max<int64_t>((h->sel_emax[k] + 8 * kThreads - 1) / (8 * kThreads), 1), 1024);
    hipLaunchKernelGGL(k_selinv_T, dim3(nn, gt), dim3(64), 0, st, (const double*)h->dL, nd, n0,
                       h->dArena);
    hipLaunchKernelGGL(k_selinv_UZ, dim3(nn, ge), dim3(kThreads), 0, st, (const double*)h->dL,
                       (const double*)h->dZ, (const int*)h->dLs, (const int*)h->dCol2sup,
                       (const int64_t*)h->dSpx, (const int*)h->dSsuper, (const int64_t*)h->dSpi, nd,
                       n0, h->dArena);
    hipLaunchKernelGGL(k_selinv_Z21, dim3(nn, ge), dim3(kThreads), 0, st, h->dZ, nd, n0,
                       (const double*)h->dArena);
    hipLaunchKernelGGL(k_selinv_Z11, dim3(nn, ge), dim3(kThreads), 0, st, h->dZ, nd, n0,
                       (const double*)h->dArena);
  }
  SWC_HIP(hipGetLastError());
  SWC_HIP(hipStreamSynchronize(st));
  h->zvalid = true;
  return SWC_OK;
}

int check_handle(const swc_handle* h) {
  if (!h) return swc_set_error(SWC_EINVAL, "null handle");
  return SWC_OK;
}

// The BCSC structure the kernels index: every supernode's row list holds its
// own columns first, then strictly increasing rows below them, all in [0, n);
// its value block (ld = row count) lies at a non-negative offset.
int validate_structure(int n, const int* blockSet, int supNo, const size_t* lC, const int* Ls,
                       const size_t* Li_ptr) {
  if (n <= 0 || supNo <= 0 || !blockSet || !lC || !Ls || !Li_ptr)
    return swc_set_error(SWC_EINVAL, "structure: bad argument");
  if (blockSet[0] != 0 || blockSet[supNo] != n)
    return swc_set_error(SWC_EINVAL, "structure: blockSet must span 0..n");
  for (int s = 0; s < supNo; ++s) {
    const int c0 = blockSet[s], c1 = blockSet[s + 1];
    if (c1 <= c0) return swc_set_error(SWC_EINVAL, "structure: empty supernode");
    const size_t p0 = Li_ptr[c0], p1 = Li_ptr[c1];
    if (p1 < p0 + (size_t)(c1 - c0) || p1 - p0 > 0x7fffffffu)
      return swc_set_error(SWC_EINVAL, "structure: row list shorter than the supernode");
    if (lC[c0] > ((size_t)1 << 62)) return swc_set_error(SWC_EINVAL, "structure: bad value offset");
    for (int c = c0; c < c1; ++c)
      if (Ls[p0 + (c - c0)] != c)
        return swc_set_error(SWC_EINVAL, "structure: a supernode's rows must start with its columns");
    for (size_t i = p0 + (c1 - c0); i < p1; ++i)
      if (Ls[i] <= Ls[i - 1] || Ls[i] >= n)
        return swc_set_error(SWC_EINVAL, "structure: rows must increase and stay below n");
  }
  return SWC_OK;
}

// Structure part of an upload: the plan (solve part only without `factor`), L
// and its FP32 copy, the row lists and the plan arrays on the device.
int upload_structure(swc_handle* h, int n, const int* blockSet, int supNo, const size_t* lC,
                     const int* Ls, const size_t* Li_ptr, const int* prunePtr,
                     const int* pruneSet, bool factor) {
  if (int rc = validate_structure(n, blockSet, supNo, lC, Ls, Li_ptr)) return rc;
  // pruneSet may be null when it is empty (no supernode has descendants)
  if (!prunePtr || (!pruneSet && prunePtr[supNo] != 0) || prunePtr[0] != 0)
    return swc_set_error(SWC_EINVAL, "structure: bad descendant sets");
  for (int s = 0; s < supNo; ++s)
    if (prunePtr[s + 1] < prunePtr[s]) return swc_set_error(SWC_EINVAL, "structure: bad descendant sets");
  h->factored = false;
  h->has_values = false;
  h->drop_graphs();
  h->work_cols = 0;  // the solve buffers are sized by the structure: rebuilt on next use
  // the old structure's L goes first: if this upload fails part way, no buffer
  // sized for the old structure is left beside the new plan, and swc_factor
  // refuses to run until an upload completes
  (void)hipFree(h->dL);
  h->dL = nullptr;
  h->n = n;
  h->nsuper = supNo;
  h->nvals = 0;
  h->super.assign(blockSet, blockSet + supNo + 1);
  h->pi.resize(supNo + 1);
  h->px.resize(supNo);
  for (int s = 0; s <= supNo; ++s) h->pi[s] = (int64_t)Li_ptr[blockSet[s]];
  for (int s = 0; s < supNo; ++s) h->px[s] = (int64_t)lC[blockSet[s]];
  const int64_t ssize = h->pi[supNo];
  h->Ls.assign(Ls, Ls + ssize);
  h->pruneptr.assign(prunePtr, prunePtr + supNo + 1);
  h->pruneset.assign(pruneSet, pruneSet + prunePtr[supNo]);
  try {
    swc::build_plan(n, supNo, h->super.data(), h->pi.data(), h->px.data(), h->Ls.data(),
                    prunePtr, pruneSet, h->plan, factor, h->params,
                    h->hlevel.level_ptr.empty() ? nullptr : &h->hlevel);
  } catch (const std::bad_alloc&) {
    return swc_set_error(SWC_ENOMEM, "upload: out of host memory");
  } catch (const std::exception& e) {
    return swc_set_error(SWC_EINVAL, std::string("upload: ") + e.what());
  }
  const Plan& P = h->plan;
  hipStream_t st = h->stream;
  int rc = SWC_OK;
  if (!rc) rc = upload(&h->dL, (const double*)nullptr, (size_t)P.xsize, st);
  // the FP32 shadow follows the structure's size whatever the current mode: a
  // later switch to a mixed mode finds it sized for this structure
  (void)hipFree(h->dLf);
  h->dLf = nullptr;
  h->lf_cap = 0;
  if (!rc && h->precision > 0 && factor) {
    rc = upload(&h->dLf, (const float*)nullptr, (size_t)P.xsize, st);
    if (!rc) h->lf_cap = P.xsize;
  }
  if (!rc) rc = upload(&h->dLs, h->Ls.data(), h->Ls.size(), st);
  if (!rc) rc = upload(&h->dPanels, P.panels.data(), P.panels.size(), st);
  if (!rc) rc = upload(&h->dTiles, P.tiles.data(), P.tiles.size(), st);
  if (!rc) rc = upload(&h->dTasks, P.tasks.data(), P.tasks.size(), st);
  if (!rc) rc = upload(&h->dGptr, P.gptr.data(), P.gptr.size(), st);
  if (!rc) rc = upload(&h->dGidx, P.gidx.data(), P.gidx.size(), st);
  if (!rc) rc = upload(&h->dGcols, P.gcols.data(), P.gcols.size(), st);
  if (!rc) rc = upload(&h->dFblk, P.fblocks.data(), P.fblocks.size(), st);
  if (!rc) rc = upload(&h->dFblkPtr, P.fblk_ptr.data(), P.fblk_ptr.size(), st);
  if (!rc) rc = upload(&h->dSpanels, P.spanels.data(), P.spanels.size(), st);
  if (!rc) rc = upload(&h->dSpOrder, P.sp_order.data(), P.sp_order.size(), st);
  if (!rc) rc = upload(&h->dOrder, P.panel_order.data(), P.panel_order.size(), st);
  if (!rc) rc = upload(&h->dFtiles, P.ftiles.data(), P.ftiles.size(), st);
  if (!rc) rc = upload(&h->dFused, P.fused.data(), P.fused.size(), st);
  if (!rc) rc = upload(&h->dCounters, (const int*)nullptr, (size_t)std::max(P.nlevel, 1), st);
  if (!rc && h->fused_cap[0] == 0) {
    // a persistent grid: as many workgroups as the device holds at once
    int ncu = 0, per[3] = {0, 0, 0};
    hipError_t e = hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, h->device);
    if (e == hipSuccess) e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per[0], k_narrow_fused<0>, 64, 0);
    if (e == hipSuccess) e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per[1], k_narrow_fused<1>, 64, 0);
    if (e == hipSuccess) e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per[2], k_narrow_fused<2>, 64, 0);
    if (e != hipSuccess)
      rc = swc_set_error(SWC_EDEVICE, std::string("upload: occupancy query: ") + hipGetErrorString(e));
    else
      for (int m = 0; m < 3; ++m) h->fused_cap[m] = std::max(1, per[m]) * std::max(ncu, 1);
  }
  if (!rc) rc = upload(&h->dDagItems, P.dag_items.data(), P.dag_items.size(), st);
  if (!rc) rc = upload(&h->dDagDeps, P.dag_deps.data(), P.dag_deps.size(), st);
  if (!rc)
    rc = upload(&h->dDagSync, (const int*)nullptr, 2 + P.dag_items.size() + (size_t)P.npanel, st);
  if (!rc && h->dag_cap[0] == 0) {
    int ncu = 0, per[3] = {0, 0, 0};
    hipError_t e = hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, h->device);
    if (e == hipSuccess) e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per[0], k_dag_factor<0>, kThreads, 0);
    if (e == hipSuccess) e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per[1], k_dag_factor<1>, kThreads, 0);
    if (e == hipSuccess) e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per[2], k_dag_factor<2>, kThreads, 0);
    if (e != hipSuccess)
      rc = swc_set_error(SWC_EDEVICE, std::string("upload: occupancy query: ") + hipGetErrorString(e));
    else
      for (int m = 0; m < 3; ++m) h->dag_cap[m] = std::max(1, per[m]) * std::max(ncu, 1);
  }
  if (!rc) rc = upload(&h->dDiag, (const double*)nullptr, (size_t)P.diag_scratch, st);
  if (!rc) rc = upload(&h->dNorm, (const double*)nullptr, 8, st);
  if (!rc) rc = upload(&h->dNormPart, (const double*)nullptr, kRedBlocks, st);
  h->zvalid = false;
  (void)hipFree(h->dZ);
  (void)hipFree(h->dArena);
  (void)hipFree(h->dNodes);
  h->dZ = nullptr;
  h->dArena = nullptr;
  h->dNodes = nullptr;
  const hipError_t se = hipStreamSynchronize(st);
  if (rc) {
    (void)hipFree(h->dL);
    h->dL = nullptr;
    return rc;
  }
  SWC_HIP(se);
  return SWC_OK;
}

}  // namespace

extern "C" {

int swc_create(swc_handle** out) {
  if (!out) return swc_set_error(SWC_EINVAL, "swc_create: null argument");
  *out = nullptr;
  auto* h = new (std::nothrow) swc_handle;
  if (!h) return swc_set_error(SWC_ENOMEM, "swc_create: out of host memory");
  hipError_t e = hipGetDevice(&h->device);
  if (e == hipSuccess && bind_scratch(h->device) != SWC_OK) {
    delete h;
    return SWC_EDEVICE;  // bind_scratch recorded why
  }
  if (e == hipSuccess) e = hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking);
  if (e == hipSuccess) e = hipMalloc((void**)&h->dInfo, sizeof(int));
  for (auto& ev : h->ev)
    if (e == hipSuccess) e = hipEventCreate(&ev);
  if (e != hipSuccess) {
    delete h;
    return swc_set_error(SWC_EDEVICE, std::string("swc_create: ") + hipGetErrorString(e));
  }
  *out = h;
  return SWC_OK;
}

int swc_destroy(swc_handle* h) {
  delete h;
  return SWC_OK;
}

int swc_upload(swc_handle* h, int n, const int* blockSet, int supNo, const size_t* lC,
               const int* Ls, const size_t* Li_ptr, const int* prunePtr, const int* pruneSet,
               const int* A2p, const int* A2i, const int64_t* a2src, int64_t nvals) {
  if (int rc = check_handle(h)) return rc;
  if (!A2p || !A2i) return swc_set_error(SWC_EINVAL, "swc_upload: bad argument");
  if (int rc = upload_structure(h, n, blockSet, supNo, lC, Ls, Li_ptr, prunePtr, pruneSet, true))
    return rc;
  if (A2p[0] != 0 || A2p[n] < 0) return swc_set_error(SWC_EINVAL, "swc_upload: bad A2 pointers");
  for (int j = 0; j < n; ++j)
    if (A2p[j + 1] < A2p[j]) return swc_set_error(SWC_EINVAL, "swc_upload: bad A2 pointers");
  h->nvals = a2src ? nvals : (int64_t)A2p[n];
  std::vector<int64_t> amap;
  if (int rc = build_amap(h, A2p, A2i, a2src, nvals, amap)) return rc;
  // full CSR of P A P^T from its lower triangle, each entry mapped to its value
  std::vector<int> frp(n + 1, 0), fci;
  std::vector<int64_t> fvm;
  for (int j = 0; j < n; ++j)
    for (int q = A2p[j]; q < A2p[j + 1]; ++q) {
      frp[A2i[q] + 1]++;
      if (A2i[q] != j) frp[j + 1]++;
    }
  for (int i = 0; i < n; ++i) frp[i + 1] += frp[i];
  fci.resize(frp[n]);
  fvm.resize(frp[n]);
  {
    std::vector<int> fill(frp.begin(), frp.end() - 1);
    // row i: first the lower-triangle entries (columns j <= i, ascending j) ...
    for (int j = 0; j < n; ++j)
      for (int q = A2p[j]; q < A2p[j + 1]; ++q) {
        const int i = A2i[q];
        fci[fill[i]] = j;
        fvm[fill[i]++] = a2src ? a2src[q] : q;
      }
    // ... then the mirrored strictly-upper entries (columns > i, ascending)
    for (int j = 0; j < n; ++j)
      for (int q = A2p[j]; q < A2p[j + 1]; ++q) {
        const int i = A2i[q];
        if (i == j) continue;
        fci[fill[j]] = i;
        fvm[fill[j]++] = a2src ? a2src[q] : q;
      }
  }
  h->nnz_full = frp[n];
  // factor-order coordinates of every input value, and the diagonal positions
  std::vector<int> qi(h->nvals, 0), qj(h->nvals, 0);
  for (int j = 0; j < n; ++j)
    for (int q = A2p[j]; q < A2p[j + 1]; ++q) {
      const int64_t dst = a2src ? a2src[q] : q;
      qi[dst] = A2i[q];
      qj[dst] = j;
    }
  std::vector<int64_t> dpos(n);
  for (int s = 0; s < supNo; ++s) {
    const int64_t R = h->pi[s + 1] - h->pi[s];
    for (int c = blockSet[s]; c < blockSet[s + 1]; ++c)
      dpos[c] = h->px[s] + (int64_t)(c - blockSet[s]) * R + (c - blockSet[s]);
  }
  hipStream_t st = h->stream;
  int rc = SWC_OK;
  if (!rc) rc = upload(&h->dAmap, amap.data(), amap.size(), st);
  if (!rc) rc = upload(&h->dVals, (const double*)nullptr, (size_t)h->nvals, st);
  if (!rc) rc = upload(&h->dRp, frp.data(), frp.size(), st);
  if (!rc) rc = upload(&h->dCi, fci.data(), fci.size(), st);
  if (!rc) rc = upload(&h->dVmap, fvm.data(), fvm.size(), st);
  if (!rc) rc = upload(&h->dAf, (const double*)nullptr, fci.size(), st);
  if (!rc) rc = upload(&h->dQi, qi.data(), qi.size(), st);
  if (!rc) rc = upload(&h->dQj, qj.data(), qj.size(), st);
  if (!rc) rc = upload(&h->dDiagPos, dpos.data(), dpos.size(), st);
  // the copies above read this frame's vectors: wait for them on every path
  const hipError_t se = hipStreamSynchronize(st);
  if (rc) {
    // a partial upload leaves no usable structure: factor calls are refused
    (void)hipFree(h->dL);
    h->dL = nullptr;
    return rc;
  }
  SWC_HIP(se);
  h->has_values = true;
  return SWC_OK;
}

int swc_upload_symbolic(swc_handle* h, const swc_symbolic* sym) {
  if (int rc = check_handle(h)) return rc;
  if (!sym) return swc_set_error(SWC_EINVAL, "swc_upload_symbolic: null symbolic");
  int64_t sz[8];
  double fl;
  if (int rc = swc_symbolic_sizes(sym, sz, &fl)) return rc;
  const int n = (int)sz[0], ns = (int)sz[1];
  std::vector<int> super(ns + 1), Ls(sz[2]), pp(ns + 1), ps(sz[7]), a2p(n + 1), a2i(sz[6]), perm(n);
  std::vector<int64_t> iptr(n + 1), p(n + 1), a2src(sz[6]);
  swc_symbolic_copy(sym, SWC_SYM_SUPER, super.data());
  swc_symbolic_copy(sym, SWC_SYM_S, Ls.data());
  swc_symbolic_copy(sym, SWC_SYM_PRUNEPTR, pp.data());
  swc_symbolic_copy(sym, SWC_SYM_PRUNESET, ps.data());
  swc_symbolic_copy(sym, SWC_SYM_A2P, a2p.data());
  swc_symbolic_copy(sym, SWC_SYM_A2I, a2i.data());
  swc_symbolic_copy(sym, SWC_SYM_A2SRC, a2src.data());
  swc_symbolic_copy(sym, SWC_SYM_I_PTR, iptr.data());
  swc_symbolic_copy(sym, SWC_SYM_P, p.data());
  swc_symbolic_copy(sym, SWC_SYM_PERM, perm.data());
  int rc = swc_upload(h, n, super.data(), ns, (const size_t*)p.data(), Ls.data(),
                      (const size_t*)iptr.data(), pp.data(), ps.data(), a2p.data(), a2i.data(),
                      a2src.data(), (int64_t)a2src.size());
  if (rc) return rc;
  return swc_set_perm(h, perm.data());
}

int swc_set_perm(swc_handle* h, const int* perm) {
  if (int rc = check_handle(h)) return rc;
  if (!perm) {
    h->perm.clear();
    return SWC_OK;
  }
  std::vector<char> seen(h->n, 0);
  for (int k = 0; k < h->n; ++k) {
    if (perm[k] < 0 || perm[k] >= h->n || seen[perm[k]])
      return swc_set_error(SWC_EINVAL, "swc_set_perm: not a permutation");
    seen[perm[k]] = 1;
  }
  h->perm.assign(perm, perm + h->n);
  return upload(&h->dPerm, h->perm.data(), h->perm.size(), h->stream);
}

int swc_factor(swc_handle* h, const double* values, int values_on_device, int* info) {
  if (int rc = check_handle(h)) return rc;
  if (!values) return swc_set_error(SWC_EINVAL, "swc_factor: null values");
  if (!h->dL || !h->has_values) return swc_set_error(SWC_ESTATE, "swc_factor: no structure uploaded");
  const double* dv = values;
  if (!values_on_device) {
    SWC_HIP(hipMemcpyAsync(h->dVals, values, sizeof(double) * h->nvals, hipMemcpyHostToDevice,
                           h->stream));
    dv = h->dVals;
  }
  if (int rc = factor_device(h, dv)) return rc;
  if (info) *info = h->last_info;
  if (h->last_info)
    return swc_set_error(SWC_ENOTPOS, "swc_factor: matrix is not positive definite (column " +
                                          std::to_string(h->last_info - 1) + ")");
  return SWC_OK;
}

int swc_get_factor(swc_handle* h, double* lValues, int on_device) {
  if (int rc = check_handle(h)) return rc;
  if (!h->factored) return swc_set_error(SWC_ESTATE, "swc_get_factor: not factored");
  SWC_HIP(hipMemcpyAsync(lValues, h->dL, sizeof(double) * h->plan.xsize,
                         on_device ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost, h->stream));
  SWC_HIP(hipStreamSynchronize(h->stream));
  return SWC_OK;
}

int swc_solve(swc_handle* h, double* B, int nrhs, int64_t ldb, int on_device) {
  if (int rc = check_handle(h)) return rc;
  if (!h->factored) return swc_set_error(SWC_ESTATE, "swc_solve: not factored");
  if (h->last_info) return swc_set_error(SWC_ENOTPOS, "swc_solve: factorization failed");
  if (!B || nrhs <= 0 || nrhs > 65535 || ldb < h->n)
    return swc_set_error(SWC_EINVAL, "swc_solve: bad argument");
  if (int rc = ensure_work(h, nrhs)) return rc;
  double* dY = h->dRef;
  if (int rc = to_factor(h, B, on_device, ldb, dY, nrhs)) return rc;
  if (int rc = solve_device(h, dY, nrhs, h->n)) return rc;
  if (int rc = from_factor(h, dY, B, on_device, ldb, nrhs)) return rc;
  float ms = 0;
  SWC_HIP(hipEventElapsedTime(&ms, h->ev[2], h->ev[3]));
  h->t_solve_ms = ms;
  return SWC_OK;
}

// swc_solve on device columns with every buffer the sweeps need taken from the
// static `scratch` array (through this device's looked-up address): no device
// allocation, no captured graph (eager launches on the handle's stream).
// Layout, each part starting on a 256-byte boundary: the permuted right-hand
// sides (n x nrhs), the forward solutions (n x nrhs), the contribution vectors
// (max(vsize,1) x nrhs, cleared by the forward sweep's fill kernel) and the
// backward partials.  Calls on one GPU share the array: overlapping calls race.
int swc_solve_noalloc(swc_handle* h, double* B, int nrhs, int64_t ldb) {
  if (int rc = check_handle(h)) return rc;
  if (!h->factored) return swc_set_error(SWC_ESTATE, "swc_solve_noalloc: not factored");
  if (h->last_info) return swc_set_error(SWC_ENOTPOS, "swc_solve_noalloc: factorization failed");
  if (!B || nrhs <= 0 || nrhs > 65535 || ldb < h->n)
    return swc_set_error(SWC_EINVAL, "swc_solve_noalloc: bad argument");
  double* base = scratch_ptr(h->device);
  if (!base) return swc_set_error(SWC_ESTATE, "swc_solve_noalloc: workspace not bound to this device");
  const Plan& P = h->plan;
  auto up = [](int64_t x) { return (x + 31) & ~(int64_t)31; };  // 32 doubles = 256 bytes
  const int64_t n = h->n;
  const int64_t vld = std::max<int64_t>(P.vsize, 1);
  const int64_t pld = (int64_t)std::max(P.max_lvl_fblk, 1) * swc::SPW;
  const int64_t oX = 0, oY = oX + up(n * nrhs), oV = oY + up(n * nrhs), oP = oV + up(vld * nrhs);
  const int64_t need = oP + up(pld * nrhs);
  if (need > kScratchDoubles)
    return swc_set_error(SWC_EINVAL, "swc_solve_noalloc: needs " + std::to_string(need * 8) +
                                         " bytes of workspace for " + std::to_string(nrhs) +
                                         " right-hand sides; the fixed workspace holds " +
                                         std::to_string(kScratchDoubles * 8));
  double* Xt = base + oX;
  if (int rc = to_factor(h, B, 1, ldb, Xt, nrhs)) return rc;
  SWC_HIP(hipEventRecord(h->ev[2], h->stream));
  enqueue_solve_bufs(h, Xt, nrhs, n, base + oY, base + oV, base + oP);
  SWC_HIP(hipEventRecord(h->ev[3], h->stream));
  SWC_HIP(hipGetLastError());
  if (int rc = from_factor(h, Xt, B, 1, ldb, nrhs)) return rc;
  float ms = 0;
  SWC_HIP(hipEventElapsedTime(&ms, h->ev[2], h->ev[3]));
  h->t_solve_ms = ms;
  return SWC_OK;
}

int swc_refine(swc_handle* h, const double* B, double* X, int nrhs, int64_t ld, int on_device,
               double tol, int maxiter, int method, double* eta, int* iters) {
  if (int rc = check_handle(h)) return rc;
  if (!h->factored) return swc_set_error(SWC_ESTATE, "swc_refine: not factored");
  if (h->last_info) return swc_set_error(SWC_ENOTPOS, "swc_refine: factorization failed");
  if (!B || !X || nrhs <= 0 || nrhs > 65535 || ld < h->n || !eta)
    return swc_set_error(SWC_EINVAL, "swc_refine: bad argument");
  if (method < 0) method = h->factored_precision ? 1 : 0;
  if (method > 1) return swc_set_error(SWC_EINVAL, "swc_refine: method must be -1, 0 or 1");
  if (int rc = ensure_work(h, nrhs)) return rc;
  double* Bt = h->dRef;
  double* Xt = h->dRef + (int64_t)h->n * nrhs;
  if (int rc = to_factor(h, B, on_device, ld, Bt, nrhs)) return rc;
  if (int rc = refine_device(h, Bt, Xt, nrhs, tol, maxiter, method, eta, iters)) return rc;
  return from_factor(h, Xt, X, on_device, ld, nrhs);
}

int swc_residual(swc_handle* h, const double* B, const double* X, int nrhs, int64_t ld,
                 int on_device, double* eta) {
  if (int rc = check_handle(h)) return rc;
  if (!h->factored) return swc_set_error(SWC_ESTATE, "swc_residual: no values (factor first)");
  if (!B || !X || nrhs <= 0 || nrhs > 65535 || ld < h->n || !eta)
    return swc_set_error(SWC_EINVAL, "swc_residual: bad argument");
  if (int rc = ensure_work(h, nrhs)) return rc;
  const int n = h->n;
  double* Bt = h->dRef;
  double* Xt = h->dRef + (int64_t)n * nrhs;
  double* R = h->dRef + (int64_t)2 * n * nrhs;
  if (int rc = to_factor(h, B, on_device, ld, Bt, nrhs)) return rc;
  if (int rc = to_factor(h, X, on_device, ld, Xt, nrhs)) return rc;
  launch_csr_mv(h, (const double*)Xt, (const double*)Bt, R, n, (int64_t)n, nrhs);
  std::vector<double> rn(nrhs), xn(nrhs), bn(nrhs);
  if (int rc = reduce_cols(h, R, nullptr, nrhs, 0, rn.data())) return rc;
  if (int rc = reduce_cols(h, Xt, nullptr, nrhs, 0, xn.data())) return rc;
  if (int rc = reduce_cols(h, Bt, nullptr, nrhs, 0, bn.data())) return rc;
  for (int c = 0; c < nrhs; ++c) {
    const double den = h->normA * xn[c] + bn[c];
    eta[c] = den > 0 ? rn[c] / den : 0.0;
  }
  return SWC_OK;
}

namespace {
// Per-column CSC descriptors of the uploaded structure and the factor's entry count.
std::vector<CscColumn> csc_columns(const swc_handle* h, int64_t* nnz) {
  std::vector<CscColumn> cols(h->n);
  int64_t off = 0;
  for (int s = 0; s < h->nsuper; ++s) {
    const int c0 = h->super[s], c1 = h->super[s + 1];
    const int64_t R = h->pi[s + 1] - h->pi[s];
    for (int j = c0; j < c1; ++j) {
      const int64_t d = j - c0;
      cols[j] = CscColumn{h->px[s] + d * R + d, h->pi[s] + d, R - d, off};
      off += R - d;
    }
  }
  *nnz = off;
  return cols;
}
}  // namespace

int swc_factor_nnz(swc_handle* h, int64_t* nnz) {
  if (int rc = check_handle(h)) return rc;
  if (!nnz) return swc_set_error(SWC_EINVAL, "swc_factor_nnz: null argument");
  if (h->n <= 0) return swc_set_error(SWC_ESTATE, "swc_factor_nnz: no structure");
  (void)csc_columns(h, nnz);
  return SWC_OK;
}

int swc_factor_csc(swc_handle* h, int64_t* Cp, int* Ci, double* Cx, int on_device) {
  if (int rc = check_handle(h)) return rc;
  if (!Cp || !Ci || !Cx) return swc_set_error(SWC_EINVAL, "swc_factor_csc: null argument");
  if (!h->factored) return swc_set_error(SWC_ESTATE, "swc_factor_csc: not factored");
  int64_t nnz = 0;
  const std::vector<CscColumn> cols = csc_columns(h, &nnz);
  hipStream_t st = h->stream;
  CscColumn* dcol = nullptr;
  int64_t* dCp = on_device ? Cp : nullptr;
  int* dCi = on_device ? Ci : nullptr;
  double* dCx = on_device ? Cx : nullptr;
  hipError_t e = hipMalloc((void**)&dcol, sizeof(CscColumn) * cols.size());
  if (e == hipSuccess && !on_device) e = hipMalloc((void**)&dCp, sizeof(int64_t) * (h->n + 1));
  if (e == hipSuccess && !on_device) e = hipMalloc((void**)&dCi, sizeof(int) * std::max<int64_t>(nnz, 1));
  if (e == hipSuccess && !on_device) e = hipMalloc((void**)&dCx, sizeof(double) * std::max<int64_t>(nnz, 1));
  if (e == hipSuccess)
    e = hipMemcpyAsync(dcol, cols.data(), sizeof(CscColumn) * cols.size(), hipMemcpyHostToDevice, st);
  if (e == hipSuccess) {
    const int per = kThreads / 64;
    hipLaunchKernelGGL(k_factor_csc, dim3((h->n + per - 1) / per), dim3(kThreads), 0, st,
                       (const double*)h->dL, (const int*)h->dLs, (const CscColumn*)dcol, h->n, dCp,
                       dCi, dCx);
    e = hipGetLastError();
  }
  if (e == hipSuccess && !on_device) {
    e = hipMemcpyAsync(Cp, dCp, sizeof(int64_t) * (h->n + 1), hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipMemcpyAsync(Ci, dCi, sizeof(int) * nnz, hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipMemcpyAsync(Cx, dCx, sizeof(double) * nnz, hipMemcpyDeviceToHost, st);
  }
  const hipError_t se = hipStreamSynchronize(st);
  if (e == hipSuccess) e = se;
  (void)hipFree(dcol);
  if (!on_device) {
    (void)hipFree(dCp);
    (void)hipFree(dCi);
    (void)hipFree(dCx);
  }
  if (e != hipSuccess)
    return swc_set_error(SWC_EDEVICE, std::string("swc_factor_csc: ") + hipGetErrorString(e));
  return SWC_OK;
}

int swc_set_plan_params(swc_handle* h, const int* params, int count) {
  if (int rc = check_handle(h)) return rc;
  swc::PlanParams p;
  if (params && !swc::plan_params_from(params, count, p))
    return swc_set_error(SWC_EINVAL, "swc_set_plan_params: parameter out of range");
  // The dependency-driven launch is switched off: one run of it gave a wrong factor
  // (a non-positive pivot where the per-level launches give none) without a cause
  // found in its waits and fences.
  if (p.dag != 0)
    return swc_set_error(SWC_EINVAL,
                         "swc_set_plan_params: the dependency-driven launch (dag) is disabled");
  h->params = p;  // NULL: the defaults
  return SWC_OK;
}

int swc_set_schedule(swc_handle* h, int nLevels, const int* levelPtr, const int* parPtr,
                     const int* partition) {
  if (int rc = check_handle(h)) return rc;
  if (nLevels <= 0 || !levelPtr) {
    h->hlevel = swc::HLevelSchedule();
    return SWC_OK;
  }
  if (!parPtr || !partition || levelPtr[0] != 0 || levelPtr[nLevels] < 0)
    return swc_set_error(SWC_EINVAL, "swc_set_schedule: bad argument");
  const int npar = levelPtr[nLevels];
  if (parPtr[0] != 0 || parPtr[npar] < 0)
    return swc_set_error(SWC_EINVAL, "swc_set_schedule: bad partition pointers");
  swc::HLevelSchedule hl;
  hl.level_ptr.assign(levelPtr, levelPtr + nLevels + 1);
  hl.par_ptr.assign(parPtr, parPtr + npar + 1);
  hl.partition.assign(partition, partition + parPtr[npar]);
  h->hlevel = std::move(hl);  // checked against the structure at the next upload
  return SWC_OK;
}

int swc_set_precision(swc_handle* h, int mode) {
  if (int rc = check_handle(h)) return rc;
  if (mode < 0 || mode > 2) return swc_set_error(SWC_EINVAL, "swc_set_precision: mode must be 0, 1 or 2");
  if (mode > 0 && h->lf_cap < h->plan.xsize) {
    // a factor graph of a mixed mode captured the old shadow: drop it with the buffer
    for (int m = 1; m < 3; ++m)
      if (h->fgraph[m]) {
        (void)hipGraphExecDestroy(h->fgraph[m]);
        h->fgraph[m] = nullptr;
      }
    (void)hipFree(h->dLf);
    h->dLf = nullptr;
    h->lf_cap = 0;
    SWC_HIP(hipMalloc((void**)&h->dLf, sizeof(float) * (size_t)h->plan.xsize));
    h->lf_cap = h->plan.xsize;
  }
  h->precision = mode;
  h->fp32_mask = mode == 0 ? 0 : (mode == 1 ? 1 : 3);
  return SWC_OK;
}

int swc_selected_inverse(swc_handle* h, double* Z, int on_device) {
  if (int rc = check_handle(h)) return rc;
  if (!h->factored || h->last_info) return swc_set_error(SWC_ESTATE, "swc_selected_inverse: no valid factor");
  if (int rc = selinv_device(h)) return rc;
  if (Z) {
    SWC_HIP(hipMemcpyAsync(Z, h->dZ, sizeof(double) * h->plan.xsize,
                           on_device ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost, h->stream));
    SWC_HIP(hipStreamSynchronize(h->stream));
  }
  return SWC_OK;
}

int swc_logdet(swc_handle* h, double* logdet) {
  if (int rc = check_handle(h)) return rc;
  if (!logdet) return swc_set_error(SWC_EINVAL, "swc_logdet: null argument");
  if (!h->factored || h->last_info) return swc_set_error(SWC_ESTATE, "swc_logdet: no valid factor");
  hipStream_t st = h->stream;
  hipLaunchKernelGGL(k_logdet_partial, dim3(kRedBlocks), dim3(kThreads), 0, st, (const double*)h->dL,
                     (const int64_t*)h->dDiagPos, h->n, h->dNormPart);
  hipLaunchKernelGGL(k_reduce_final, dim3(1), dim3(kThreads), 0, st, (const double*)h->dNormPart, 1,
                     h->dNorm + 1);
  double v = 0;
  SWC_HIP(hipMemcpyAsync(&v, h->dNorm + 1, sizeof(double), hipMemcpyDeviceToHost, st));
  SWC_HIP(hipStreamSynchronize(st));
  *logdet = 2.0 * v;
  return SWC_OK;
}

int swc_gradient(swc_handle* h, const double* x, int x_on_device, double alpha, double beta,
                 double* grad, int grad_on_device) {
  if (int rc = check_handle(h)) return rc;
  if (!grad || (beta != 0.0 && !x)) return swc_set_error(SWC_EINVAL, "swc_gradient: null argument");
  if (!h->factored || h->last_info) return swc_set_error(SWC_ESTATE, "swc_gradient: no valid factor");
  if (alpha != 0.0)
    if (int rc = selinv_device(h)) return rc;
  if (int rc = ensure_work(h, 1)) return rc;
  hipStream_t st = h->stream;
  double* xt = h->dRef;
  if (beta != 0.0)
    if (int rc = to_factor(h, x, x_on_device, h->n, xt, 1)) return rc;
  double* g = grad_on_device ? grad : h->dRef + h->n;  // dRef holds >= 5 n doubles
  double* gbuf = nullptr;
  if (!grad_on_device) {
    SWC_HIP(hipMallocAsync((void**)&gbuf, sizeof(double) * std::max<int64_t>(h->nvals, 1), st));
    g = gbuf;
  }
  hipLaunchKernelGGL(k_grad_values, dim3(grid_for(h->nvals)), dim3(kThreads), 0, st,
                     (const double*)h->dZ, (const int64_t*)h->dAmap, (const int*)h->dQi,
                     (const int*)h->dQj, (const double*)xt, alpha, beta, h->nvals, g);
  SWC_HIP(hipGetLastError());
  if (!grad_on_device) {
    SWC_HIP(hipMemcpyAsync(grad, gbuf, sizeof(double) * h->nvals, hipMemcpyDeviceToHost, st));
    SWC_HIP(hipFreeAsync(gbuf, st));
  }
  SWC_HIP(hipStreamSynchronize(st));
  return SWC_OK;
}

int swc_set_graphs(swc_handle* h, int on) {
  if (int rc = check_handle(h)) return rc;
  h->use_graphs = on != 0;
  return SWC_OK;
}

int swc_set_profiling(swc_handle* h, int on) {
  if (int rc = check_handle(h)) return rc;
  h->profile = on != 0;
  return SWC_OK;
}

int swc_get_timings(const swc_handle* h, double* out) {
  if (int rc = check_handle(h)) return rc;
  if (!out) return swc_set_error(SWC_EINVAL, "swc_get_timings: null argument");
  const Plan& P = h->plan;
  out[0] = h->t_factor_ms;
  out[1] = h->t_solve_ms;
  out[2] = h->t_tile_ms;
  out[3] = h->n_tile;
  out[4] = h->t_trsm_ms;
  out[5] = h->n_trsm;
  out[6] = P.tile_kernel_flops;
  out[7] = P.trsm_kernel_flops;
  return SWC_OK;
}

int swc_level_timings(const swc_handle* h, double* out, int cap, int* nlevel) {
  if (int rc = check_handle(h)) return rc;
  if (!nlevel) return swc_set_error(SWC_EINVAL, "swc_level_timings: null argument");
  const Plan& P = h->plan;
  *nlevel = P.nlevel;
  if (!out) return SWC_OK;
  for (int l = 0; l < std::min(cap, P.nlevel); ++l) {
    out[2 * l] = l < (int)h->lvl_tile_ms.size() ? h->lvl_tile_ms[l] : 0.0;
    out[2 * l + 1] = P.lvl_tile_flops[l];
  }
  return SWC_OK;
}

int swc_plan_stats(const swc_handle* h, int64_t* out, double* flops) {
  if (int rc = check_handle(h)) return rc;
  const Plan& P = h->plan;
  out[0] = P.npanel;
  out[1] = (int64_t)P.tiles.size();
  out[2] = (int64_t)P.tasks.size();
  out[3] = P.nlevel;
  out[4] = P.xsize;
  out[5] = (int64_t)P.dag_items.size();
  out[6] = P.dag_items.empty() ? P.nlevel : P.dag_level0;
  if (flops) {
    flops[0] = P.update_flops;
    flops[1] = P.factor_flops;
    flops[2] = P.dag_flops;
    flops[3] = P.dag_sim_us;
    flops[4] = P.lvl_sim_us;
  }
  return SWC_OK;
}

// Reference entry point (sw_parallel_PB_Cholesky_05.h:52-66, PRUNE variant):
// uploads the structure, factors on the current GPU and writes lValues in place.
// The schedule arguments (levelPtr, parPtr, partition, chunk, threads, ...) are
// accepted and ignored: the GPU schedule is derived from prunePtr/pruneSet.
// Returns 1 on success and 0 on failure (the driver's convention,
// sw_choleskyTest01.cpp:240-244); swc_last_error() says why.
int sw_cholesky_left_par_05(int n, int* c, int* r, double* values, size_t* lC, int* Ls,
                            size_t* Li_ptr, double* lValues, int* blockSet, int supNo,
                            double* timing, int* prunePtr, int* pruneSet, int nLevels,
                            int* levelPtr, int* levelSet, int nPar, int* parPtr, int* partition,
                            int chunk, int threads, int super_max, int col_max, double* nodCost) {
  (void)nLevels; (void)levelPtr; (void)levelSet; (void)nPar; (void)parPtr; (void)partition;
  (void)chunk; (void)threads; (void)super_max; (void)col_max; (void)nodCost;
  swc_handle* h = nullptr;
  if (swc_create(&h)) return 0;
  std::unique_ptr<swc_handle> guard(h);
  int ok = 0;
  if (!swc_upload(h, n, blockSet, supNo, lC, Ls, Li_ptr, prunePtr, pruneSet, c, r, nullptr, 0)) {
    int info = 0;
    const int rc = swc_factor(h, values, 0, &info);
    if (rc == SWC_OK && swc_get_factor(h, lValues, 0) == SWC_OK) ok = 1;
    if (timing) timing[0] = h->t_factor_ms * 1e-3;  // device time of the factorization
  }
  return ok;
}

// Reference forward solve (triangularSolve/Triangular_BCSC.h:14-49, called as
// examples/triangularTest02.cpp:195): L y = x in place on the host vector x, for
// the BCSC factor (Lp = L->p, Li = L->s, Lx = values, Li_ptr = L->i_ptr, sup2col =
// L->super, supNo = L->nsuper), on the current GPU.  NNZ and col2sup are not used
// (the reference does not read them either).  The solve schedule comes from the
// row lists (the call carries no descendant sets); it is built on the first call
// for a structure and kept, with L's device buffer, for later calls with the
// same structure (the arrays are compared, not their addresses).  Returns 1, or
// 0 on failure (the reference's convention; swc_last_error() says why).
namespace {
std::mutex g_lsolve_mu;
swc_handle* g_lsolve = nullptr;  // held until the process ends

bool same_structure(const swc_handle* h, int n, const size_t* Lp, const int* Li,
                    const size_t* Li_ptr, const int* sup2col, int supNo) {
  if (!h || !h->dL || h->n != n || h->nsuper != supNo) return false;
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev != h->device) return false;
  if (std::memcmp(h->super.data(), sup2col, sizeof(int) * (size_t)(supNo + 1)) != 0) return false;
  for (int s = 0; s <= supNo; ++s)
    if ((int64_t)Li_ptr[sup2col[s]] != h->pi[s]) return false;
  for (int s = 0; s < supNo; ++s)
    if ((int64_t)Lp[sup2col[s]] != h->px[s]) return false;
  return std::memcmp(h->Ls.data(), Li, sizeof(int) * h->Ls.size()) == 0;
}
}  // namespace

int blockedLsolve(int n, size_t* Lp, int* Li, double* Lx, int NNZ, size_t* Li_ptr, int* col2sup,
                  int* sup2col, int supNo, double* x) {
  (void)NNZ; (void)col2sup;
  if (!Lp || !Li || !x) return 0;  // the reference's own check
  if (!Lx) {
    swc_set_error(SWC_EINVAL, "blockedLsolve: null values");
    return 0;
  }
  std::lock_guard<std::mutex> lock(g_lsolve_mu);
  if (!same_structure(g_lsolve, n, Lp, Li, Li_ptr, sup2col, supNo)) {
    if (validate_structure(n, sup2col, supNo, Lp, Li, Li_ptr)) return 0;
    int dev = 0;
    if (g_lsolve && (hipGetDevice(&dev) != hipSuccess || dev != g_lsolve->device)) {
      delete g_lsolve;
      g_lsolve = nullptr;
    }
    if (!g_lsolve && swc_create(&g_lsolve)) return 0;
    std::vector<int> pp, ps;
    std::vector<int64_t> pi(supNo + 1);
    for (int s = 0; s <= supNo; ++s) pi[s] = (int64_t)Li_ptr[sup2col[s]];
    try {
      swc::descendant_sets(n, supNo, sup2col, pi.data(), Li, pp, ps);
    } catch (const std::exception& e) {
      swc_set_error(SWC_EINVAL, std::string("blockedLsolve: ") + e.what());
      return 0;
    }
    if (upload_structure(g_lsolve, n, sup2col, supNo, Lp, Li, Li_ptr, pp.data(), ps.data(), false))
      return 0;
  }
  swc_handle* h = g_lsolve;
  if (ensure_work(h, 1)) return 0;
  hipStream_t st = h->stream;
  hipError_t e = hipMemcpyAsync(h->dL, Lx, sizeof(double) * (size_t)h->plan.xsize,
                                hipMemcpyHostToDevice, st);
  if (e == hipSuccess)
    e = hipMemcpyAsync(h->dWork, x, sizeof(double) * (size_t)n, hipMemcpyHostToDevice, st);
  if (e == hipSuccess) {
    enqueue_forward<1>(h, h->dWork, 1, n, h->dY, h->dV);
    e = hipGetLastError();
  }
  if (e == hipSuccess)
    e = hipMemcpyAsync(x, h->dY, sizeof(double) * (size_t)n, hipMemcpyDeviceToHost, st);
  const hipError_t se = hipStreamSynchronize(st);
  if (e == hipSuccess) e = se;
  if (e != hipSuccess) {
    swc_set_error(SWC_EDEVICE, std::string("blockedLsolve: ") + hipGetErrorString(e));
    return 0;
  }
  return 1;
}

}  // extern "C"
