// Krylov solver on gfx950 for one problem: the device FGMRES engine
// (right-preconditioned flexible GMRES(m), classical Gram-Schmidt once or
// twice, the Hessenberg matrix and the Givens rotations on the device, each
// part of a restart cycle captured as a HIP graph and replayed; the host reads
// the convergence flag once per block of Arnoldi steps) with the operator's
// apply and the composite preconditioner below.
//
// Preconditioner: the reference's multiplicative composite
// (src/transport_solverk.kokkos.cxx:196-318): stage 0 the element block
// inverse, stage 1 the streaming stage (here the transport sweep), stage 2
// DSA; each later stage corrects the residual r - A_pc y left by the ones
// before it, A_pc the full operator when DSA is on (PCSetUseAmat, :289) and
// the unscattered L otherwise. Right preconditioning, so the stopping test is
// on the true residual (:292-309).
//
// Mixed precision: the FP64 true residual r = b - A x drives an iterative
// refinement whose correction solves A d = r with FP32 vectors, FP32 applies
// and sweeps and FP32 block solves, FP64 orthogonalisation arithmetic and the
// FP64 DSA inner solve.
#include <hip/hip_runtime.h>

#include <cmath>

#include "solver.hpp"

namespace ubt {

void Solver::create(Operator *o, DSA *d, const SolverOptions &opts) {
  op = o;
  dsa = d;
  opt = opts;
  n = o->disc->n_rows();
  // the engine checks the options; its residual history holds every iteration
  eng.create(this, n, 1, opt, opt.max_it);
  eng.set_problems({0, n}, {1}, nullptr);
  if (opt.pc_sweep) sweep.create(o);
  blocks.op = o;
  if (!opt.mixed) {
    t64.alloc(sizeof(double) * n);
    u64.alloc(sizeof(double) * n);
  } else {
    t32.alloc(sizeof(float) * n);
    u32.alloc(sizeof(float) * n);
  }
}

void Solver::refresh(hipStream_t s) {
  if (opt.pc_blocks) blocks.setup(s);
}

template <class T> void Solver::apply_op(const T *x, T *y, bool scatter, hipStream_t s) {
  if (transpose)
    op->mult_transpose(prec_id<T>(), x, y, scatter, s);
  else
    op->mult(prec_id<T>(), x, y, scatter, s);
}

void Solver::mult(int prec, const void *x, void *y, hipStream_t s) {
  if (prec == UBT_FP64)
    apply_op<double>((const double *)x, (double *)y, true, s);
  else
    apply_op<float>((const float *)x, (float *)y, true, s);
}

void Solver::pc(int prec, const void *r, void *y, hipStream_t s) {
  if (prec == UBT_FP64)
    apply_pc<double>((const double *)r, (double *)y, s);
  else
    apply_pc<float>((const float *)r, (float *)y, s);
}

template <class T> void Solver::apply_pc(const T *r, T *y, hipStream_t s, bool no_dsa) {
  T *t = sizeof(T) == 8 ? (T *)t64.ptr : (T *)t32.ptr;
  T *u = sizeof(T) == 8 ? (T *)u64.ptr : (T *)u32.ptr;
  const bool with_dsa = dsa && opt.pc_dsa && !no_dsa;
  const bool amat = with_dsa && opt.use_amat;
  const int prec = prec_id<T>();
  bool first = true;
  auto stage = [&](auto &&fn) {
    if (first) {
      fn(r, y);
      first = false;
      return;
    }
    apply_op<T>(y, t, amat, s);
    vec_axpby<T>(n, 1.0, r, -1.0, t, s);
    fn((const T *)t, u);
    vec_axpby<T>(n, 1.0, (const T *)u, 1.0, y, s);
  };
  // where the sweep is L's exact inverse S and the stages correct with L, the
  // first two stages are the sweep alone: y + S (r - L y) = S r for any y
  const bool fold = opt.pc_sweep && sweep.exact && !amat;
  if (opt.pc_blocks && !fold) stage([&](const T *a, T *b) { blocks.apply(prec, a, b, s, transpose); });
  if (opt.pc_sweep) stage([&](const T *a, T *b) { sweep.apply(prec, a, b, s, transpose); });
  if (with_dsa) stage([&](const T *a, T *b) { dsa->apply(prec, a, b, s, transpose); });
  if (first) vec_axpby<T>(n, 1.0, r, 0.0, y, s);
}

void Solver::solve(const double *b, double *x, hipStream_t s, bool adj, bool guess) {
  UBT_CHECK(b && x, UBT_ERR_ARG_NULL, "b and x are required");
  UBT_CHECK(b != x, UBT_ERR_ARG_IDN, "the solve needs distinct b and x");
  UBT_CHECK(op->xs_set, UBT_ERR_ARG_WRONGSTATE, "set the cross sections first");
  if (opt.pc_blocks) UBT_CHECK(blocks.group_ready == op->group, UBT_ERR_ARG_WRONGSTATE,
                               "refresh the solver after switching the operator to group %d", op->group);
  const bool with_dsa = dsa && opt.pc_dsa;
  if (with_dsa) UBT_CHECK(dsa->group == op->group, UBT_ERR_ARG_WRONGSTATE,
                          "set the DSA group to the operator's group %d", op->group);
  transpose = adj;
  // the DSA inner solve launches all its iterations during the solve instead
  // of polling its convergence flag from the host (the CG kernels do nothing
  // once the flag is set, so the result is the same), so every cycle can be
  // captured; the caller's setting comes back afterwards
  struct KeepCheck {
    DSA *d;
    int every;
    ~KeepCheck() {
      if (d) d->opt.check_every = every;
    }
  } keep{with_dsa ? dsa : nullptr, with_dsa ? dsa->opt.check_every : 0};
  if (with_dsa) dsa->opt.check_every = 0;
  // what the captured cycles replay: the operator's group and cross
  // sections, the direction, DSA's group
  uint64_t tag = op->xs_gen;
  tag = tag * 0x100000001B3ull + (uint64_t)(op->group + 1);
  tag = tag * 0x100000001B3ull + (adj ? 2u : 1u);
  tag = tag * 0x100000001B3ull + (uint64_t)(with_dsa ? dsa->group + 2 : 0);
  eng.solve(b, x, tag, s, guess);
  iterations = eng.iterations[0];
  outer = eng.outer;
  rel_residual = eng.rel_residual[0];
  reason = eng.reason[0];
  history = eng.history;
}

template void Solver::apply_pc<double>(const double *, double *, hipStream_t, bool);
template void Solver::apply_pc<float>(const float *, float *, hipStream_t, bool);

}  // namespace ubt
