// The GPU solver around the operator: right-preconditioned flexible GMRES
// with the reference's multiplicative composite preconditioner, in FP64 or
// as FP32-apply inner solves inside FP64 iterative refinement.
#pragma once

#include "dsa.hpp"
#include "sweep.hpp"

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

struct Solver {
  Operator *op = nullptr;
  DSA *dsa = nullptr;
  SolverOptions opt;
  Sweep sweep;
  BlockInverse blocks;
  int64_t n = 0;
  // Krylov storage per precision: V (restart+1 vectors), Z (restart)
  DevBuf V64, Z64, V32, Z32;
  DevBuf w64, t64, u64, r64, d32, w32, t32, u32, b32;  // work vectors
  DevBuf hcol, partial;   // device GS coefficients (restart+2), multi-dot partials
  // pinned host copy of the Hessenberg column (both passes; an asynchronous
  // copy, where pageable memory would be staged) and the event recorded
  // after it
  double *h_pin = nullptr;
  hipEvent_t h_ev = nullptr;
  Solver() = default;
  Solver(const Solver &) = delete;
  Solver &operator=(const Solver &) = delete;
  ~Solver();
  // results
  int iterations = 0, outer = 0, reason = 0;
  double rel_residual = 0.0;
  std::vector<double> history;  // relative residual estimate per Krylov iteration

  void create(Operator *o, DSA *d, const SolverOptions &opts);
  void refresh(hipStream_t s);  // per group: block-inverse check
  bool transpose = false;  // the current solve is with A^T (the adjoint)
  // x = A^{-1} b, or (adj) x = A^{-T} b with every preconditioner stage transposed
  void solve(const double *b, double *x, hipStream_t s, bool adj = false);
  template <class T> void apply_pc(const T *r, T *y, hipStream_t s);
  template <class T> void apply_op(const T *x, T *y, bool scatter, hipStream_t s);
  template <class T> double fgmres(const T *b, T *x, double tol_abs, int max_it, int *its, hipStream_t s);
};

}  // namespace ubt
