// The GPU solver around the operator: right-preconditioned flexible GMRES
// (the device engine, one problem) with the reference's multiplicative
// composite preconditioner, in FP64 or as FP32-apply inner solves inside FP64
// iterative refinement.
#pragma once

#include "dsa.hpp"
#include "fgmres.hpp"
#include "groups.hpp"
#include "sweep.hpp"

namespace ubt {

struct Solver : KrylovSystem {
  Operator *op = nullptr;
  DSA *dsa = nullptr;
  SolverOptions opt;
  Sweep sweep;
  BlockInverse blocks;
  int64_t n = 0;
  Fgmres eng;
  DevBuf t64, u64, t32, u32;  // the composite preconditioner's work vectors
  Solver() = default;
  Solver(const Solver &) = delete;
  Solver &operator=(const Solver &) = delete;
  // results
  int iterations = 0, outer = 0, reason = 0;
  double rel_residual = 0.0;
  std::vector<double> history;  // relative residual estimate per Krylov iteration (per refinement step when mixed)
  GroupReport groups;           // the last group sweep (solve_groups)

  void create(Operator *o, DSA *d, const SolverOptions &opts);
  void refresh(hipStream_t s);  // per group: block-inverse check
  bool transpose = false;  // the current solve is with A^T (the adjoint)
  // x = A^{-1} b, or (adj) x = A^{-T} b with every preconditioner stage transposed
  // guess: start from x as given (else from zero)
  void solve(const double *b, double *x, hipStream_t s, bool adj = false, bool guess = false);
  void set_rtol(double rtol) { opt.rtol = eng.opt.rtol = rtol; }
  // the preconditioner's work vectors exist in this precision
  bool has_precision(int prec) const { return prec == UBT_FP64 ? t64.ptr != nullptr : t32.ptr != nullptr; }
  // no_dsa: the removal and streaming stages alone (the L preconditioner)
  template <class T> void apply_pc(const T *r, T *y, hipStream_t s, bool no_dsa = false);
  template <class T> void apply_op(const T *x, T *y, bool scatter, hipStream_t s);
  void mult(int prec, const void *x, void *y, hipStream_t s) override;
  void pc(int prec, const void *r, void *y, hipStream_t s) override;
};

}  // namespace ubt
