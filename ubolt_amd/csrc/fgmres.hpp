// Right-preconditioned flexible GMRES on the device over np independent
// problems laid end to end in one vector (np = 1 for a single solve).
//
// Every reduction is segmented by problem: the vector is cut into chunks
// that never straddle two problems, a block reduces one chunk, and a second
// kernel sums each problem's chunks in a fixed order (deterministic). The
// Hessenberg matrix, the Givens rotations, the residual estimates and the
// convergence flags of every problem live in device arrays, so a restart
// cycle - the start (residual, norm), blocks of Arnoldi steps (preconditioner
// apply, operator apply, classical Gram-Schmidt, rotation) and the end (the
// least-squares update) - has no host round trip and each part is captured
// once as a HIP graph and replayed; the host reads the done flags once per
// block of steps. A problem that converges mid-cycle freezes its rotation
// state; the rest of the cycle still runs over its (inert) span.
//
// The stored basis is unnormalised, V_i = s_i W_i with W_0 the residual and
// W_{j+1} the orthogonalised vector itself (of norm h_{j+1,j}, so the scales
// never compound), and s folded into the projections and the update: no pass
// over a vector only to normalise it.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <vector>

#include "ubt_internal.hpp"

namespace ubt {

struct SolverOptions {
  double rtol = 1e-8;       // relative true residual ||b - A x|| / ||b||
  int max_it = 500;         // total Krylov iterations (inner ones in mixed mode)
  int restart = 30;
  int mixed = 0;            // 1: FP32 inner FGMRES inside FP64 iterative refinement
  double inner_rtol = 1e-4; // inner solve tolerance (mixed)
  int max_outer = 20;       // refinement steps (mixed)
  int pc_blocks = 1;        // stage 0: element block inverse
  int pc_sweep = 1;         // stage 1: transport sweep (the streaming stage)
  int pc_dsa = 1;           // stage 2: DSA (if a DSA object is given)
  int use_amat = 1;         // residual updates between stages with A (else L) when DSA is on
  int gs_passes = 2;        // classical Gram-Schmidt passes per Arnoldi step (1 or 2)
};

// What the Krylov engine asks of the system it solves: y = A x and
// y = M^{-1} r over the whole vector, in precision prec (UBT_FP64 or
// UBT_FP32), enqueued on stream s (which may be capturing)
struct KrylovSystem {
  virtual ~KrylovSystem() = default;
  virtual void mult(int prec, const void *x, void *y, hipStream_t s) = 0;
  virtual void pc(int prec, const void *r, void *y, hipStream_t s) = 0;
};

constexpr int kKrylovStride = 32;  // per-problem stride of the coefficient arrays (restart + 1 at most)

struct Fgmres {
  KrylovSystem *sys = nullptr;
  SolverOptions opt;
  int64_t n = 0;
  int np = 0;
  // chunking of the vector for the per-problem reductions
  int64_t chunk = 0;
  int n_chunks = 0;
  DevBuf chunk_start, chunk_len, chunk_prob;
  DevBuf cfirst;              // per problem: its first chunk (prefix, np + 1)
  DevBuf V64, Z64, V32, Z32, d32, b32, r64;
  DevBuf part, H, cs, sn, gv, hc, yv, nrm, tol, res, bnorm, onrm, vsc;
  DevBuf done;                // int per problem x 4: done, steps this cycle, iterations, live
  DevBuf hist;                // per problem: the residual estimate after each iteration (hist_cap of them)
  int hist_cap = 0;
  std::vector<int> live;
  // results per problem
  std::vector<int> iterations, reason;
  std::vector<double> rel_residual;
  std::vector<double> history;  // problem 0: relative residual estimate per iteration (FP64) or per refinement step (mixed)
  int outer = 0, cycles = 0;
  bool use_graph = true;
  uint64_t tag = 0;           // what the graphs were captured for
  static constexpr int kStepsPerCheck = 4;       // Arnoldi steps per graph between flag reads
  std::vector<hipGraphExec_t> graphs[2];         // per precision: start, step blocks, end
  hipStream_t cap = nullptr;
  Fgmres() = default;
  Fgmres(const Fgmres &) = delete;
  Fgmres &operator=(const Fgmres &) = delete;
  ~Fgmres();
  void drop_graphs();

  // n rows in np problems; hist_cap residual estimates kept per problem
  void create(KrylovSystem *system, int64_t rows, int problems, const SolverOptions &o, int hist_cap);
  // the problems' row spans (np + 1 offsets) and which take part; drops the graphs
  void set_problems(const std::vector<int64_t> &off, const std::vector<char> &active, hipStream_t s);
  // x = A^{-1} b per problem, b and x FP64, from x = 0 or (guess) from x as
  // given; rtol (np values, may be null) replaces opt.rtol per problem.
  // Graphs captured under another tag are dropped first (the system's state
  // they replay changed)
  void solve(const double *b, double *x, uint64_t state_tag, hipStream_t s, bool guess = false,
             const double *rtol = nullptr);
  template <class T> void cycle_start(const T *b, const T *x, hipStream_t s);
  template <class T> void cycle_steps(int j0, int j1, hipStream_t s);
  template <class T> void cycle_end(T *x, hipStream_t s);
  // restart cycles until every problem is done; true_stop: done is decided
  // on the true residual at a cycle's start (so the last action is a start),
  // else on the estimate as well (an inner solve)
  template <class T> void fgmres(const T *b, T *x, const double *tol_abs, int max_cycles, bool true_stop, hipStream_t s);
};

// vector helpers on a stream: y = a x + b y (y not read when b = 0), y = v
template <class T> void vec_axpby(int64_t n, double a, const T *x, double b, T *y, hipStream_t s);
template <class T> void vec_fill(int64_t n, T v, T *y, hipStream_t s);
template <class T> constexpr int prec_id() { return sizeof(T) == 8 ? UBT_FP64 : UBT_FP32; }

}  // namespace ubt
