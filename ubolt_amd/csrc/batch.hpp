// Many independent transport problems driven as one: every operator apply,
// block inverse and sweep level of all problems is one launch over a work
// list of (problem, tile) items, and the Krylov solver keeps one Hessenberg
// per problem on the device so a whole restart cycle replays as one captured
// graph. Problems may differ in mesh size, quadrature and group count; they
// share the discretisation type (dimension, basis, faces per cell).
#pragma once

#include "fgmres.hpp"
#include "solver.hpp"

namespace ubt {

template <class T> struct BatchItem {
  OpArgs<T> g;  // the problem's current group
  int64_t off;  // first row of the problem in the batch vector
  BoxArgs box;  // the box-tile layout (box body only)
};

// apply bodies the batch autotuner chooses from
constexpr int kBatchBodies = 4;  // 0 tile, 1 four-angle tile, 2 box tile, 3 strip march (uniform DG1 boxes)

// the direction groups of every active problem, one workgroup each
struct BatchSweep {
  int n_wg = 0, threads = 256;
  DevBuf prob, angle, cells, upwind, lvl, base;
  // the forward sweep records of every item in this schedule's precision
  // (see Sweep), built by Batch::sweep_records for the current group
  int slots = 2;
  int64_t n_items = 0, widest = 0;
  mutable DevBuf rec, rnb;
  mutable uint64_t rec_tag = ~0ull;
  std::vector<char> lists_for;  // the active set the lists were built for (empty: none built)
};
// The batched sweep of uniform all-vacuum 2D DG1 boxes (no records): one
// workgroup per problem, quadrant and group of kBoxSweepAngles of its angles
constexpr int kBoxSweepThreads = 256;
template <class T> constexpr int kBoxSweepAngles = sizeof(T) == 8 ? 4 : 8;
// diagonals whose r and sigma_t are in flight while one is solved
template <class T> constexpr int kBoxSweepDepth = sizeof(T) == 8 ? 2 : 1;
struct BoxSweepGroup {
  int item;      // the problem's entry in the strip body's work list (-1: an idle slot)
  int q, k0;     // the quadrant, the group's first position among its angles
  int angle[8];  // the quadrant's angles (-1: none)
  int ipt;       // cells per thread on the longest diagonal (1, 2, 4 or 8)
  int64_t rbase; // the group's first value in the quadrant-major copies
};
// r and psi pass through a quadrant-major copy, row (q, cell, basis) of the
// quadrant's A / 4 angles contiguous. The groups of one (problem, quadrant)
// read neighbouring parts of the same lines, so they are placed on one XCD
// (workgroup i runs on XCD i mod 8) and share its L2.
struct BoxSweepPlan {
  bool ok = false;   // every active problem is such a box (else the record sweep runs)
  int n_wg = 0, nx_max = 0;
  DevBuf groups;
  DevBuf qpos;       // int per (problem, angle): q * A / 4 + position, problems concatenated
  DevBuf qoff;       // int64 per problem + 1: prefix of the problems' angle counts
  mutable DevBuf rq, pq;  // the quadrant-major copies of r and psi
};

// directions of one octant per batched sweep workgroup, per precision
// [FP64, FP32] (measured on the 64-problem ensemble: 2 and 4)
constexpr int kBatchSweepGroup[2] = {2, 4};

struct Batch {
  std::vector<Operator *> ops;
  std::vector<std::unique_ptr<Sweep>> sweeps;  // per problem, host schedule kept
  int dim = 0, nb = 0, nf = 0, group = 0;
  int64_t n_rows = 0;
  std::vector<int64_t> off;  // n + 1 row offsets
  std::vector<char> active;  // the problem has the current group
  // per precision, per body: work list of the active problems
  DevBuf items[2][kBatchBodies], tiles[2][kBatchBodies];
  int n_items[2][kBatchBodies] = {};
  int64_t n_tiles[2][kBatchBodies] = {};
  size_t lds[2][kBatchBodies] = {};
  bool body_ok[2][kBatchBodies] = {};
  int body[2] = {0, 0};
  int persist[2] = {8, 8};     // workgroups per CU of the persistent grid
  float tuned_us[2][kBatchBodies] = {};
  int n_cu = 256;
  // box body: prefix of boundary-cell tiles over the items, per precision
  DevBuf edge_off[2][kBatchBodies];  // box bodies: prefix of boundary-cell tiles per problem
  int64_t n_edge[2][kBatchBodies] = {};
  int gen = 0;                 // bumped whenever the work lists are rebuilt
  // block inverse and sweep work lists
  DevBuf bblk;                 // blocks: prefix of cells x angles over the active problems
  DevBuf block_cnt;            // int per problem: singular blocks found by blocks_setup
  int64_t n_blocks = 0;
  // per precision: its own grouping and records (built when the record
  // sweep first runs on an active set)
  mutable BatchSweep sweep[2];
  mutable std::vector<SweepLists> sched[2];  // per precision, per problem: the regrouped schedule, once computed
  BoxSweepPlan box_sweep[2];   // per precision: the record-free sweep of box batches

  void create(const std::vector<Operator *> &o);
  // problems without group g go inactive, and with `want` (n flags) those it
  // leaves out
  void set_group(int g, const std::vector<char> &want = {});
  // quad: x and y in the box sweep's quadrant-major order (box batches)
  void mult(int prec, const void *x, void *y, bool scatter, hipStream_t s, bool quad = false) const;
  void blocks_setup(hipStream_t s);  // singular check of every active problem's blocks
  void blocks_apply(int prec, const void *x, void *y, hipStream_t s) const;
  void sweep_apply(int prec, const void *r, void *psi, hipStream_t s, bool staged = false) const;
  // the solver may run in the box sweep's quadrant-major order: every active
  // problem a box of the box sweep and the box apply body applicable
  bool quad_ready(int prec) const { const int pi = prec == UBT_FP64 ? 0 : 1; return box_sweep[pi].ok && body_ok[pi][2]; }
  void quad_order(int prec, const void *src, void *dst, bool to_q, hipStream_t s) const;
  // the sweep inverts every active problem's L exactly (the box sweep, or
  // record sweeps with no lagged or reflective coupling)
  bool sweep_exact(int prec) const;
  // (re)build the sweep records when the group or any cross sections changed
  void sweep_records(hipStream_t s) const;
  uint64_t xs_tag() const;
  void rhs(int prec, void *b, hipStream_t s) const;
  void set_scalar_flux(int prec, const void *psi, hipStream_t s) const;
  void tune();
  void build_lists();
  void build_box_sweep();
  void record_lists(int pi) const;  // the record sweep's work list for the active set
  std::vector<char> lists_for;  // the active set the block prefix, sweep lists and box plan were built for
  bool box_sweep_enabled = true;  // false: the record sweep on every batch
};

// inflow slots of the sweep record layout for nf faces and at most nfu inflow faces
int sweep_slots(int nf, int nfu);
void batch_records_launch(const Batch &b, int pi, hipStream_t s);

// launchers (dev_operator.hip, dev_sweep.hip)
// quad: x and y in the quadrant-major order of the box sweep (box body only)
template <class T>
void batch_apply_launch(const Batch &b, int body, const T *x, T *y, bool scatter, hipStream_t s, bool quad = false);
template <class T> void batch_blocks_launch(const Batch &b, const T *x, T *y, hipStream_t s);
template <class T> void batch_sweep_launch(const Batch &b, const T *r, T *psi, hipStream_t s);
// staged: r and psi already in the quadrant-major order (no reorder either side)
template <class T> void batch_box_sweep_launch(const Batch &b, const T *r, T *psi, hipStream_t s, bool staged = false);
// reference order <-> quadrant-major order of every active problem's rows
template <class T> void batch_quad_order(const Batch &b, const T *src, T *dst, bool to_q, hipStream_t s);
// per-problem tile size, LDS and tile count of a body (the box body also
// fills its layout); false where it does not apply
bool batch_body_layout(const Operator &op, int prec, int body, int *cpb, int *red_lanes, size_t *lds,
                       int64_t *n_tiles, BoxArgs *box);
constexpr int kBatchBoxRows = 4;  // rows per thread of the box body

// The batch's solver: the device FGMRES engine over the batch's problems,
// the batch's persistent apply as the operator and the composite
// preconditioner (element blocks, then the sweep on the L-residual - the
// sweep alone where it inverts L exactly)
struct BatchSolver : KrylovSystem {
  Batch *batch = nullptr;
  SolverOptions opt;
  Fgmres eng;
  DevBuf t64, u64, t32, u32;  // the composite preconditioner's work vectors
  // on a box-sweep batch the Krylov vectors live in the sweep's quadrant-major
  // order (the box apply reads and writes it, the sweep needs no reorder):
  // b and x pass through bq, xq once per solve
  bool quad = false;
  DevBuf bq, xq;
  int gen = -1;               // the batch work lists the solver was refreshed on
  uint64_t xs_tag = 0;        // the members' cross-section generations the records follow
  std::vector<GroupReport> groups;  // per problem: the last group sweep (solve_groups)

  void create(Batch *b, const SolverOptions &o);
  void refresh(hipStream_t s);
  // guess: start from x as given (else from zero); rtol (n values, may be
  // null): per-problem tolerances in place of opt.rtol
  void solve(const double *b, double *x, hipStream_t s, bool guess = false, const double *rtol = nullptr);
  template <class T> void apply_pc(const T *r, T *y, hipStream_t s);
  void mult(int prec, const void *x, void *y, hipStream_t s) override;
  void pc(int prec, const void *r, void *y, hipStream_t s) override;
};

}  // namespace ubt
