// Device work plan of the numeric phase, built once per symbolic structure.
//
// Every supernode is cut into panels of at most PW consecutive columns and its
// rows into tiles of TM rows; a tile owns its accumulator, so each entry of L is
// updated by one workgroup at a time, in a fixed source order, without atomics.
//
// A supernode of one panel is one left-looking target: its tiles apply the
// updates of the reference's pruned descendants (pruneSet order), the diagonal
// tile then factors the panel, and the other tiles are solved against it.
// A wider supernode first applies all descendant updates to every tile of its
// lower trapezoid in a single level (fully 2D-parallel), then factors itself
// right-looking, one level per panel, over blocks of rlb panels: within a block
// each factored panel updates the block's later panels; after the block's last
// panel the block updates all later panels with one wide inner dimension.
#pragma once
#include <cstdint>
#include <vector>

namespace swc {

constexpr int PW = 64;  // panel width (columns)
constexpr int TM = 64;  // tile height (rows)
constexpr int SPW = 128;       // widest solve panel the solve kernels take (columns)
constexpr int SPW_NARROW = 16;  // solve panels this narrow run one wave each

// Shape parameters of a plan, chosen per structure class by the autotuner's
// table (swcholesky_amd/tuning.json, written by swcholesky_amd/tune.py).
struct PlanParams {
  int rlb = 8;                 // panels per block of a wide supernode's right-looking steps
  int spw = SPW;               // solve panel width: 64 or 128 columns
  int narrow_min_tiles = 1024; // a level runs its narrow tiles one wave each (own launch)
  int narrow_max_k = 256;      // ... if it has this many; narrow = panel <= 16 columns and
                               // total inner dimension <= narrow_max_k
  int fuse_narrow = 1;         // 1: a level's panels whose tiles are all narrow are updated and
                               // factored together, one wave per panel from a device worklist
  int dag = 0;                 // levels after the last one with narrow or fused tiles run as one
                               // dependency-driven launch: 0 off, 1 in the order of a list
                               // schedule simulated on the host, 2 in level order
};
// entries of the int[] form {rlb, spw, narrow_min_tiles, narrow_max_k, fuse_narrow, dag}
constexpr int kPlanParams = 6;
// Checks and copies the int[] form; returns false if a value is out of range.
bool plan_params_from(const int* v, int count, PlanParams& out);

struct Panel {
  int64_t vbase;  // value offset of (panel row 0, panel column 0)
  int64_t voff;   // offset of the supernode's contribution vector (solves)
  int ld;         // leading dimension (row count of the supernode)
  int ncol;       // columns in the panel (<= PW)
  int nrow;       // rows from the panel's first column down
  int c0;         // first global column
  int rows_off;   // offset into Ls of the panel's row 0
  int sn;         // owning supernode
  int sc0;        // first column of the supernode
  int sw;         // width of the supernode
};

struct Tile {
  int64_t vbase;   // the panel's value offset of (row 0, column 0)
  int panel;
  int r0;          // first panel row of the tile
  int nr;          // rows in the tile (<= TM)
  int task_begin;  // [task_begin, task_end) into the task array
  int task_end;
  int kind;        // 0: the panel's diagonal tile, 1: below it
  int slot;        // factor level: offset of the panel's diagonal block copy in the
                   // per-level scratch (updated, not yet factored), -1 otherwise
  int ld;          // the panel's leading dimension, columns and first column (copied
  int ncol;        // from the panel, so that a kernel reads one descriptor)
  int c0;
};

// One source contribution to one tile: rows [ra,rb) x columns [ca,cb) of the
// source's row list, inner dimension k.  Where they land is two bit masks: the
// e-th set bit of rmask is the tile row that source row ra+e lands on, the e-th
// set bit of cmask the panel column of source row ca+e (both lists are sorted,
// so the landing positions are increasing).
struct Task {
  int64_t src_base;  // value offset of (source row 0, source column 0)
  int ld;            // source leading dimension
  int k;             // source column count
  int rows_off;      // offset into Ls of the source's row 0
  int ra, rb;        // source rows landing in the tile's rows
  int ca, cb;        // source rows landing in the panel's columns
  int kind;          // 0: descendant supernode, 1: earlier panel of the same supernode
  uint64_t rmask;    // tile rows receiving source rows ra..rb-1
  uint64_t cmask;    // panel columns receiving source rows ca..cb-1
};

// Solve row block of a solve panel: rows [r0, r0+nr) below its diagonal block
// (nr = 0: the panel has no rows below it).  In the forward sweep the block with
// writer = 1 also stores the panel's solution.
struct SolveBlock {
  int panel;
  int r0;
  int nr;
  int writer;
};

// A panel updated and factored by one wave of the fused narrow-panel kernel:
// its tiles are tiles[tile0 .. tile0 + ntile), the diagonal tile first.
struct FusedPanel {
  int tile0;
  int ntile;
};

// One work item of the dependency-driven factor kernel: the updates one level
// applies to one tile and, when that level completes the tile's panel, the
// factorization of the tile's rows.  Items are claimed in array order, which is
// a topological order: every item an item waits for comes before it.
struct DagItem {
  int tile;        // index into tiles
  int fclass;      // -1: updates only; 0: factor, panel <= 16 columns; 1: 17..64 columns
  int prev;        // the item that last wrote this tile (-1: written before the launch)
  int diag;        // a factoring tile below the diagonal: its panel's diagonal-tile item
  int dep0, dep1;  // [dep0, dep1) into dag_deps: panels the updates read
};

// Wait until `count` tiles of `panel` are factored (all of them: the panel is final).
struct DagDep {
  int panel;
  int count;
};

static_assert(sizeof(Panel) == 48, "Panel layout is part of the plan export");
static_assert(sizeof(Tile) == 48, "Tile layout is part of the plan export");
static_assert(sizeof(Task) == 56, "Task layout is part of the plan export");
static_assert(sizeof(SolveBlock) == 16, "SolveBlock layout is part of the plan export");

struct Plan {
  int n = 0, nsuper = 0, npanel = 0;
  int64_t xsize = 0;
  std::vector<Panel> panels;
  std::vector<Tile> tiles;          // grouped by level (one update launch per level)
  std::vector<int> lvl_tiles;       // nlevel+1 offsets into tiles
  std::vector<int> lvl_tiles_fused;   // nlevel: tiles of fused narrow panels, first in their level
  std::vector<int> lvl_tiles_narrow;  // nlevel: one-wave tiles (panels <= 16 wide, short
                                      // total inner dimension), next in their level
  std::vector<int> lvl_tiles_deferred;  // nlevel: tiles of panels the level does not factor
                                        // (later panels of a wide supernode), last in their
                                        // level: they run beside the level's panel factor
  std::vector<FusedPanel> fused;      // fused narrow panels, by level
  std::vector<int> lvl_fused;         // nlevel+1 offsets into fused
  std::vector<int> trsm_tiles;      // tiles of the panels each level completes (factor kernel)
  std::vector<Tile> ftiles;         // copies of those tiles in the same order
  std::vector<int> lvl_trsm;        // nlevel+1 offsets into trsm_tiles
  std::vector<int> lvl_trsm_cls;    // 4 per level: starts of the panel-width classes
                                    // (<= 16, <= 32, <= 64 columns) in trsm_tiles, and the end
  std::vector<Task> tasks;
  std::vector<int> panel_level;     // per panel
  std::vector<int> lvl_panels;      // nlevel+1 offsets into panel_order
  std::vector<int> panel_order;     // panels grouped by level
  // solves: solve panels of up to SPW columns of one supernode, with their own
  // level structure (nslevel levels)
  std::vector<Panel> spanels;
  std::vector<int> sp_order;        // solve panels grouped by level
  std::vector<int> lvl_sp;          // nslevel+1
  std::vector<int> lvl_sp_narrow;   // nslevel: narrow solve panels at the start of each level
  std::vector<int> lvl_fblk_narrow; // nslevel: their row blocks at the start of the level's
  int nslevel = 0;
  // forward solve: column c (factor ordering) takes the contributions
  // V[gidx[gptr[c] .. gptr[c+1])] of its descendants, in increasing supernode
  // order; the columns of the supernodes whose first panel is at level l are
  // gathered at that level (gcols[lvl_gcol[l] ..])
  std::vector<int64_t> gptr;        // n+1
  std::vector<int64_t> gidx;        // vsize
  std::vector<int> gcols;           // n
  std::vector<int> lvl_gcol;        // nslevel+1
  std::vector<SolveBlock> fblocks;  // row blocks of each level's solve panels (both sweeps)
  std::vector<int> lvl_fblk;        // nslevel+1
  std::vector<int> fblk_ptr;        // nsp+1: blocks of sp_order[i]
  int max_lvl_fblk = 0;             // most row blocks in one level
  int64_t diag_scratch = 0;         // doubles of diagonal-block copies one level needs (the
                                    // levels of the dependency-driven launch hold theirs together)
  // dependency-driven factor launch over levels dag_level0 .. nlevel-1 (none when
  // dag_items is empty)
  int dag_level0 = 0;
  std::vector<DagItem> dag_items;
  std::vector<DagDep> dag_deps;
  double dag_flops = 0;             // useful flops of its updates and factorizations
  double dag_sim_us = 0;            // makespan of the simulated list schedule (cost model)
  double lvl_sim_us = 0;            // the same levels as per-level launches, same cost model
  int64_t vsize = 0;                // sum over supernodes of (rows - width)
  double update_flops = 0, factor_flops = 0;
  // algorithmic flops of each kernel class: lower-triangle update products (tile
  // kernel); POTRF + TRSM of every panel, counted once (panel factor kernel)
  double tile_kernel_flops = 0, trsm_kernel_flops = 0;
  std::vector<double> lvl_tile_flops;  // the tile kernel's useful flops per level
  int nlevel = 0;
};

// super[nsuper+1], pi[nsuper+1] (row-list pointers), px[nsuper] value offsets,
// Ls row indices, prunePtr/pruneSet as the reference's analysis produces them.
// With factor = false only the solve part (solve panels, their levels, row
// blocks and forward gathers) is built; the factor lists stay empty.
// The reference's H-level schedule (getCoarseLevelSet_6, InspectionLevel_06.h:
// 18-391): clusters [level_ptr[k], level_ptr[k+1]) of partitions, partition q the
// supernodes partition[par_ptr[q] .. par_ptr[q+1]) in execution order.
struct HLevelSchedule {
  std::vector<int> level_ptr, par_ptr, partition;
};

// hl == nullptr: the wavefront schedule (each supernode at the first level its
// descendants allow); otherwise the levels follow hl (for comparison).
void build_plan(int n, int nsuper, const int* super, const int64_t* pi, const int64_t* px,
                const int* Ls, const int* prunePtr, const int* pruneSet, Plan& P,
                bool factor = true, const PlanParams& params = PlanParams(),
                const HLevelSchedule* hl = nullptr);

// Descendant sets for a structure given without them: supernode s depends on
// every d < s whose rows below its own columns meet s's columns, listed in
// increasing d (the set the reference's pruneSet holds, in another order).
void descendant_sets(int n, int nsuper, const int* super, const int64_t* pi, const int* Ls,
                     std::vector<int>& prunePtr, std::vector<int>& pruneSet);

}  // namespace swc
