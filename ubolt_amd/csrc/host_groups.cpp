// The multigroup driver (reference tests/transportk.kokkos.cxx:925-1208): the
// Gauss-Seidel group sweep and the upscatter outer iteration, for one problem
// (Solver) and for a batch (BatchSolver, every problem's groups swept
// together, each problem with its own outer test).
#include "groups.hpp"

#include <algorithm>
#include <cmath>
#include <limits>

#include "batch.hpp"
#include "solver.hpp"

namespace ubt {

void check_group_options(const GroupOptions &o) {
  UBT_CHECK(o.recon_steps >= 0, UBT_ERR_ARG_OUTOFRANGE, "recon_steps must not be negative, was given %d", o.recon_steps);
  UBT_CHECK(o.outer_rtol > 0.0, UBT_ERR_ARG_OUTOFRANGE, "outer_rtol must be positive, was given %g", o.outer_rtol);
  UBT_CHECK(o.max_outer >= 1, UBT_ERR_ARG_OUTOFRANGE, "max_outer must be at least 1, was given %d", o.max_outer);
  UBT_CHECK(o.inner_gamma >= 0.0 && o.inner_gamma < 1.0, UBT_ERR_ARG_OUTOFRANGE,
            "inner_gamma must be in [0, 1), was given %g", o.inner_gamma);
  UBT_CHECK(o.inner_rtol >= 0.0, UBT_ERR_ARG_OUTOFRANGE, "inner_rtol must not be negative, was given %g", o.inner_rtol);
}

// the sweep's thermal rtol (reference :1113-1120): the first sweep's change
// is 1 by definition, a later one capped at 1
void OuterTest::begin_sweep(int outer) {
  if (upscatter && o.inner_gamma > 0.0)
    sweep_rtol = final_sweep ? inner : std::max(inner, o.inner_gamma * (outer == 0 ? 1.0 : std::min(outer_change, 1.0)));
  else
    sweep_rtol = inner;
}

// the outer test (reference :1159-1207)
int OuterTest::end_sweep(int outer, double change_sq, double norm_sq, GroupReport &rep) {
  const double change = norm_sq > 0.0 ? std::sqrt(change_sq / norm_sq) : std::sqrt(change_sq);
  const bool have_rho = outer >= 2 && outer_change > 0.0;
  const double rho = have_rho ? change / outer_change : 0.0;
  outer_change = change;
  const double est = (have_rho && rho < 1.0) ? change * rho / (1.0 - rho) : std::numeric_limits<double>::infinity();
  rep.change = change;
  rep.rho = rho;
  rep.err_est = est;
  if (change == 0.0 || est < o.outer_rtol) {
    if (sweep_rtol == inner) return 1;
    final_sweep = true;
  }
  if (outer + 1 >= o.max_outer) return -2;
  return 0;
}

static OuterTest make_test(const GroupOptions &o, bool upscatter, double solver_rtol) {
  OuterTest t;
  t.o = o;
  t.upscatter = upscatter;
  t.inner = upscatter ? (o.inner_rtol > 0.0 ? o.inner_rtol : kUpscatterInnerRtol) : solver_rtol;
  // a warm-started solve whose rhs moves by less than its rtol takes no
  // iterations, so the flux stops moving and the test would pass falsely
  UBT_CHECK(!upscatter || o.outer_rtol >= t.inner, UBT_ERR_ARG_INCOMP,
            "outer_rtol %g is below the inner rtol %g: a warm-started inner solve whose rhs moves by less than its "
            "rtol takes no iterations, so the scalar flux stops moving and the outer iteration would converge "
            "falsely",
            o.outer_rtol, t.inner);
  return t;
}

static void reset_report(GroupReport &r, int G, int g_up) {
  r = GroupReport();
  r.g_up = g_up;
  r.group_iterations.assign(G, 0);
  r.group_reason.assign(G, 0);
  r.group_residual.assign(G, 0.0);
}

// puts the solver's own rtol back however the sweep ends
struct KeepRtol {
  double *a, *b, v;
  ~KeepRtol() { *a = *b = v; }
};

// The moments store's warm start of group g (reference Reconstruct,
// tests/transportk.kokkos.cxx:300-345): psi = the P1 expansion of the cached
// scalar flux and the stored current, then Richardson steps on
// L psi = b + S_gg phi_g (the within-group scatter of the stored flux),
// psi += P_L r with P_L the removal and streaming stages, each kept only if it
// lowers ||r||, until ||r|| <= eta ||b + S_gg phi_g|| or max_steps. Where the
// sweep is L's exact inverse one step solves it. Returns the steps taken.
static int reconstruct(Solver &sv, int g, const double *b, double *psi, int max_steps, double eta, double *r,
                       double *t, double *dz, hipStream_t s) {
  Operator &op = *sv.op;
  const int64_t n = op.disc->n_rows();
  op.expand_p1(g, psi, s);
  if (max_steps == 0) return 0;
  // r = b + S_gg phi_g - L psi
  vec_axpby<double>(n, 1.0, b, 0.0, r, s);
  op.scatter_source(g, g, UBT_FP64, r, s);
  const double c_norm = op.norm2(r, s);
  op.mult(UBT_FP64, psi, t, false, s);
  vec_axpby<double>(n, -1.0, (const double *)t, 1.0, r, s);
  double r_norm = op.norm2(r, s);
  const double target = eta * c_norm;
  int steps = 0;
  while (steps < max_steps && r_norm > target) {
    sv.apply_pc<double>(r, dz, s, true);
    steps++;
    // the trial iterate in t, so psi stays the best one if the step is refused
    vec_axpby<double>(n, 1.0, (const double *)psi, 0.0, t, s);
    vec_axpby<double>(n, 1.0, (const double *)dz, 1.0, t, s);
    op.mult(UBT_FP64, t, r, false, s);
    vec_axpby<double>(n, 1.0, b, -1.0, r, s);
    op.scatter_source(g, g, UBT_FP64, r, s);
    const double new_norm = op.norm2(r, s);
    if (!(new_norm < r_norm)) break;
    vec_axpby<double>(n, 1.0, (const double *)t, 0.0, psi, s);
    r_norm = new_norm;
  }
  return steps;
}

void solve_groups(Solver &sv, const GroupOptions &o, double *psi, hipStream_t s, GroupReport &rep) {
  check_group_options(o);
  UBT_CHECK(psi, UBT_ERR_ARG_NULL, "psi is required");
  Operator &op = *sv.op;
  UBT_CHECK(op.xs_set, UBT_ERR_ARG_WRONGSTATE, "set the cross sections first");
  const int G = op.n_groups, g_up = op.g_up;
  const int64_t n = op.disc->n_rows();
  const bool up = g_up < G;
  OuterTest t = make_test(o, up, sv.opt.rtol);
  reset_report(rep, G, g_up);
  const bool with_dsa = sv.dsa && sv.opt.pc_dsa;
  // the moments store keeps one psi; without upscatter nothing is
  // reconstructed, so it is the plain sweep over that one vector
  const bool store = o.store_moments != 0;
  KeepRtol keep{&sv.opt.rtol, &sv.eng.opt.rtol, sv.opt.rtol};
  DevBuf b, sums, rw;
  b.alloc(sizeof(double) * (size_t)n);
  sums.alloc(sizeof(double) * 2);
  if (store && up) rw.alloc(sizeof(double) * 3 * (size_t)n);  // the reconstruction's r, t and dz
  UBT_HIP(hipMemsetAsync(psi, 0, sizeof(double) * (size_t)(store ? 1 : G) * n, s));
  // the first sweep reads every thermal group below the one it solves before
  // that group is solved: a zero flux cached for each (reference :964-976)
  if (up)
    for (int g = g_up; g < G; g++) {
      op.cache_zero_flux(g, s);
      if (store) op.zero_moments(g, s);
    }
  for (int outer = 0;; outer++) {
    t.begin_sweep(outer);
    UBT_HIP(hipMemsetAsync(sums.ptr, 0, sizeof(double) * 2, s));
    bool failed = false;
    for (int g = outer == 0 ? 0 : g_up; g < G && !failed; g++) {
      op.group = g;
      if (with_dsa) sv.dsa->set_group(g, s);
      sv.refresh(s);
      op.rhs(g, UBT_FP64, b.ptr, s, g >= g_up);
      const double rtol = t.rtol(g, g_up);
      sv.set_rtol(rtol);
      double *x = store ? psi : psi + (int64_t)g * n;
      // from the second sweep every solve starts from the group's last psi,
      // or under the moments store from its reconstruction
      if (store && outer >= 1) {
        double *rr = rw.as<double>();
        rep.recon_steps += reconstruct(sv, g, b.as<double>(), x, o.recon_steps, rtol, rr, rr + n, rr + 2 * n, s);
      } else if (store) {
        UBT_HIP(hipMemsetAsync(x, 0, sizeof(double) * (size_t)n, s));
      }
      sv.solve(b.as<double>(), x, s, false, outer >= 1);
      rep.group_iterations[g] += sv.iterations;
      rep.iterations += sv.iterations;
      rep.group_residual[g] = sv.rel_residual;
      rep.group_reason[g] = sv.reason;
      if (sv.reason <= 0) {
        failed = true;
        break;
      }
      op.set_scalar_flux(g, UBT_FP64, x, s, g >= g_up ? sums.as<double>() : nullptr);
      if (store && up && g >= g_up) op.store_moments(g, x, s);
    }
    rep.sweeps = outer + 1;
    if (failed) {
      rep.status = -1;
      break;
    }
    if (!up) {
      rep.status = 1;
      break;
    }
    double h[2];
    UBT_HIP(hipMemcpyAsync(h, sums.ptr, sizeof(h), hipMemcpyDeviceToHost, s));
    UBT_HIP(hipStreamSynchronize(s));
    const int st = t.end_sweep(outer, h[0], h[1], rep);
    if (st) {
      rep.status = st;
      break;
    }
  }
  UBT_HIP(hipStreamSynchronize(s));
}

void solve_groups(BatchSolver &bs, const GroupOptions &o, double *psi, hipStream_t s, std::vector<GroupReport> &rep) {
  check_group_options(o);
  UBT_CHECK(!o.store_moments, UBT_ERR_SUP, "the moments store is offered for a single problem's group sweep only");
  UBT_CHECK(psi, UBT_ERR_ARG_NULL, "psi is required");
  Batch &B = *bs.batch;
  const int np = (int)B.ops.size();
  const int64_t n = B.n_rows;
  int G = 0;
  std::vector<OuterTest> t(np);
  std::vector<char> up(np), going(np, 1), want(np);
  rep.assign(np, GroupReport());
  for (int p = 0; p < np; p++) {
    const Operator &op = *B.ops[p];
    G = std::max(G, op.n_groups);
    up[p] = op.g_up < op.n_groups;
    t[p] = make_test(o, up[p], bs.opt.rtol);
    reset_report(rep[p], op.n_groups, op.g_up);
  }
  DevBuf b, sums;
  b.alloc(sizeof(double) * (size_t)n);
  sums.alloc(sizeof(double) * 2 * np);
  UBT_HIP(hipMemsetAsync(b.ptr, 0, sizeof(double) * (size_t)n, s));
  UBT_HIP(hipMemsetAsync(psi, 0, sizeof(double) * (size_t)G * n, s));
  for (int p = 0; p < np; p++)
    if (up[p])
      for (int g = B.ops[p]->g_up; g < B.ops[p]->n_groups; g++) B.ops[p]->cache_zero_flux(g, s);
  std::vector<double> rt(np), h(2 * np);
  for (int outer = 0;; outer++) {
    int g0 = G;
    for (int p = 0; p < np; p++)
      if (going[p]) {
        t[p].begin_sweep(outer);
        g0 = std::min(g0, outer == 0 ? 0 : B.ops[p]->g_up);
      }
    if (g0 == G) break;
    UBT_HIP(hipMemsetAsync(sums.ptr, 0, sizeof(double) * 2 * np, s));
    for (int g = g0; g < G; g++) {
      bool any = false;
      for (int p = 0; p < np; p++) {
        const Operator &op = *B.ops[p];
        want[p] = going[p] && g < op.n_groups && (outer == 0 || g >= op.g_up);
        any = any || want[p];
      }
      if (!any) continue;
      B.set_group(g, want);
      bs.refresh(s);
      // the rhs of every active problem (an idle span holds zeros or an
      // earlier group's rhs: no solve reads it)
      for (int p = 0; p < np; p++) {
        rt[p] = bs.opt.rtol;
        if (!B.active[p]) continue;
        const Operator &op = *B.ops[p];
        op.rhs(g, UBT_FP64, b.as<double>() + B.off[p], s, g >= op.g_up);
        rt[p] = t[p].rtol(g, op.g_up);
      }
      // always from x as given: psi is zero before a group's first solve, and
      // the spans of the problems sitting out keep their values
      double *x = psi + (int64_t)g * n;
      bs.solve(b.as<double>(), x, s, true, rt.data());
      for (int p = 0; p < np; p++) {
        if (!B.active[p]) continue;
        Operator &op = *B.ops[p];
        GroupReport &r = rep[p];
        r.group_iterations[g] += bs.eng.iterations[p];
        r.iterations += bs.eng.iterations[p];
        r.group_residual[g] = bs.eng.rel_residual[p];
        r.group_reason[g] = bs.eng.reason[p];
        if (bs.eng.reason[p] <= 0) {
          going[p] = 0;  // the problem's sweep stops here
          r.status = -1;
          r.sweeps = outer + 1;
          continue;
        }
        op.set_scalar_flux(g, UBT_FP64, x + B.off[p], s, g >= op.g_up ? sums.as<double>() + 2 * p : nullptr);
      }
    }
    UBT_HIP(hipMemcpyAsync(h.data(), sums.ptr, sizeof(double) * 2 * np, hipMemcpyDeviceToHost, s));
    UBT_HIP(hipStreamSynchronize(s));
    for (int p = 0; p < np; p++) {
      if (!going[p]) continue;
      GroupReport &r = rep[p];
      r.sweeps = outer + 1;
      const int st = up[p] ? t[p].end_sweep(outer, h[2 * p], h[2 * p + 1], r) : 1;
      if (st) {
        r.status = st;
        going[p] = 0;
      }
    }
  }
}

}  // namespace ubt
