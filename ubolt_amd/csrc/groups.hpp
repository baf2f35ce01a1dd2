// The multigroup driver: the Gauss-Seidel sweep over the energy groups and,
// with upscatter, the outer iteration over the thermal block (reference
// tests/transportk.kokkos.cxx:925-1208).
//
// Downscatter only: the group system is block lower triangular, so one
// forward sweep is exact - per group the rhs (inflow, source, the transfers
// from the groups above, out of their cached scalar flux), the solve, and the
// scalar flux cached for the groups below. With upscatter, g_up is the first
// group anything upscatters into; the groups above it are solved once, in the
// first sweep, and the thermal block g_up .. G-1 is swept again until the
// estimated error change * rho / (1 - rho) of its scalar flux is below
// outer_rtol (change = ||phi_new - phi_old|| / ||phi_new|| over the block,
// rho the ratio of successive changes). A thermal group's rhs takes every
// other group's latest flux (zero before it is first solved); from the second
// sweep on its solve starts from its last angular flux, at the adaptive rtol
// max(inner rtol, gamma * the last change), and the iteration only stops on a
// sweep whose thermal solves all ran at the inner rtol.
//
// Everything between two solves stays on the device (rhs, transfers, scalar
// flux, the change sums); the host reads the Krylov convergence flags inside
// a solve and the two change sums once per sweep.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <vector>

namespace ubt {

struct Solver;
struct BatchSolver;

// the inner rtol the reference driver defaults to under upscatter
constexpr double kUpscatterInnerRtol = 1e-10;

struct GroupOptions {
  double outer_rtol = 1e-8;   // the outer test on the error estimate
  int max_outer = 1000;       // sweeps over the thermal block, the first included
  double inner_gamma = 0.01;  // adaptive inner rtol factor, [0, 1); 0 runs every solve at the inner rtol
  double inner_rtol = 0.0;    // every solve's rtol under upscatter; 0: kUpscatterInnerRtol
  // the moments store (single problem): psi is one vector every group's
  // solve reuses; a thermal group's scalar flux and current are kept, and
  // after the first sweep its solve starts from their P1 expansion refined by
  // at most recon_steps safeguarded Richardson steps on L_g (0: the expansion
  // alone)
  int store_moments = 0;
  int recon_steps = 10;
};

struct GroupReport {
  int g_up = 0;               // first thermal group (n_groups: no upscatter)
  int sweeps = 0;             // the first included
  int status = 0;             // 1 converged, -1 an inner solve failed, -2 out of sweeps
  double change = 0.0, rho = 0.0, err_est = 0.0;  // of the last sweep
  int64_t iterations = 0;     // over every solve
  int64_t recon_steps = 0;    // Richardson steps of the moments store's reconstructions
  std::vector<int> group_iterations, group_reason;  // per group: summed over sweeps, of the last solve
  std::vector<double> group_residual;               // per group: the last solve's relative residual
};

// One problem's outer-iteration bookkeeping on the host
struct OuterTest {
  GroupOptions o;
  bool upscatter = false;
  double inner = 0.0;         // the effective inner rtol
  double outer_change = 0.0;  // the last sweep's change
  double sweep_rtol = 0.0;    // this sweep's thermal rtol
  bool final_sweep = false;   // the test passed on a looser sweep: sweep once more at the inner rtol

  void begin_sweep(int outer);
  double rtol(int g, int g_up) const { return g >= g_up ? sweep_rtol : inner; }
  // after a sweep: 1 converged, -2 out of sweeps, 0 sweep again
  int end_sweep(int outer, double change_sq, double norm_sq, GroupReport &rep);
};

void check_group_options(const GroupOptions &o);

// psi: n_groups vectors of the operator's rows (group g at g * rows)
void solve_groups(Solver &sv, const GroupOptions &o, double *psi, hipStream_t s, GroupReport &rep);
// psi: G_max batch vectors (group g of problem p at g * n_rows + off[p]);
// a problem with fewer groups leaves the later vectors' span zero
void solve_groups(BatchSolver &bs, const GroupOptions &o, double *psi, hipStream_t s, std::vector<GroupReport> &rep);

}  // namespace ubt
