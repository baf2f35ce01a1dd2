// Diffusion synthetic acceleration on the device: the diffusion operator of
// each backend assembled per group by a kernel, its FP64 matvec, a block-Jacobi
// preconditioned CG inner solve with device-resident scalars (no host round
// trip, so a stream can be captured), and the restriction / prolongation.
#pragma once

#include "operator.hpp"

namespace ubt {

// the most diffusion nodes the allocation-free inner solve takes (2^21: a
// 128^3 structured box; its workspace is 5 node vectors, 80 MiB per GPU)
constexpr int64_t kNoAllocNodes = int64_t(1) << 21;

struct DSAOptions {
  int consistent_d = 1;      // blend the upwind numerical diffusion m h into face D
  double consistent_power = 1.5;
  double mip_penalty = 4.0;  // C in kappa = max(C/2 (D_c/h_c + D_n/h_n), 1/4)
  int max_it = 200;          // inner CG iterations at most
  double rtol = 1e-6;        // inner CG relative residual (preconditioned norm)
  int check_every = 8;       // read the converged flag on the host every this many
                             // iterations and stop launching (0: never - no host
                             // synchronisation, so the solve can be graph-captured)
  // the void policy (-dsa_void_bridge, -dsa_void_sigma_t, -dsa_void_d)
  int void_bridge = 1;       // bridge voids where the operator allows it; 0 masks them
  double void_sigma_t = 0.0; // a cell is void where sigma_t <= this
  double void_d = 0.0;       // a bridged void's D; 0: a third of the free flight
};

struct DSA {
  Operator *op = nullptr;
  DSAOptions opt;
  int group = -1;
  int nb = 1, ns = 0;        // block size, neighbour slots per cell
  int64_t n_cells = 0, n_nodes = 0;
  double half_range[3] = {0, 0, 0};
  int vacuum[3][2] = {{0, 0}, {0, 0}, {0, 0}};
  bool any_vacuum = false;
  // the group's void policy (void: sigma_t <= opt.void_sigma_t). Voids are
  // bridged - kept in the operator with D = 1 / (3 (max(sigma_t, 0) + 1 / L)),
  // L = 4 V / S the voids' mean chord (V their volume, S their faces onto a
  // non-void cell or a vacuum face), or opt.void_d - when bridging is on,
  // something is not void and a vacuum face or some absorption holds the
  // operator; otherwise masked: identity rows coupled to nothing, a face into
  // them a vacuum face, nothing restricted or corrected there. A group void
  // everywhere is masked whole (a zero correction)
  int n_void = 0;
  double inv_chord = 0.0;    // 1 / L, 0 without bridged voids
  bool bridged = false, masked = false, all_void = false;
  DevBuf vsum;               // double: per-block (V, S) partials, then the two sums
  DevBuf wmask;              // double, nodes: weight, zero on masked void nodes
  DevBuf nbr;                // int, cells * ns: neighbour cell per slot, -1 none
  DevBuf own, off, jinv;     // double: cells*nb*nb, cells*ns*nb*nb, cells*nb*nb
  DevBuf weight;             // double, nodes: V (plex) or 1 (structured)
  DevBuf slot_hr;            // double, cells*ns: half-range current per face (DG0)
  DevBuf rhs, sol, r, z, p, p2, q, partial, scal;
  double *flag_host = nullptr;  // pinned: the converged flag read back every check_every iterations
  ~DSA();
  int last_iterations = 0;

  void create(Operator *o);
  void set_group(int g, hipStream_t s);
  // y = P D^{-1} W R x, or (transpose) y = R^T W D^{-1} P^T x: the diffusion
  // operators are symmetric, so the adjoint restricts with the unweighted angle
  // sum / sum_weights and prolongs with w_a times the node weight
  void apply(int prec, const void *x, void *y, hipStream_t s, bool transpose = false);
  // y = D_diff x on node vectors (FP64)
  void matvec(const double *x, double *y, hipStream_t s) const;
  // sol = D^{-1} rhs by the inner CG (x0 = 0)
  void solve(const double *b, double *x, hipStream_t s);
  // the same solve with its workspace in a file-scope device array (at most
  // kNoAllocNodes nodes): allocates nothing
  void solve_noalloc(const double *b, double *x, hipStream_t s);
};

}  // namespace ubt
