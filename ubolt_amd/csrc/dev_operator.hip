// The transport operator's device kernels for gfx950 (CDNA4): the fused
// streaming + removal + within-group scatter apply for structured FD and
// unstructured DG0/DG1, the composed diagonal, the right-hand side, the group
// transfer, the angular moments on the matrix cores and the element block
// inverse. One template per discretisation family, instantiated for FP32 and
// FP64; every kernel is stream-ordered and allocates nothing.
//
// Work mapping of the applies: a workgroup owns cpb whole cells and every
// angle of them (thread = (cell, angle), cpb = blockDim / n_angles), so the
// scalar flux the scatter needs is reduced inside the workgroup from values
// already in registers - psi is read from HBM once and y written once. The
// rows of one cell are contiguous (angle fastest), so the own-cell loads and
// the stores are coalesced, and an upwind gather reads a neighbour's whole
// contiguous run for the angles that see it upwind.
#include <hip/hip_runtime.h>

#include "operator.hpp"

namespace ubt {

namespace {

constexpr int kReflect = UBT_BC_REFLECT;

// Sum nseg contiguous segments of len values in LDS, S lanes per segment
// (S a power of two dividing the wavefront): strided partial sums, then a
// butterfly inside each S-lane group. Deterministic order. Every thread of the
// block must call it.
template <class T>
__device__ inline void seg_sum(const T *red, int nseg, int len, T *out, int S) {
  const int t = threadIdx.x, per = blockDim.x / S;
  for (int base = 0; base < nseg; base += per) {
    const int seg = base + t / S, l = t & (S - 1);
    T p = T(0);
    if (seg < nseg)
      for (int k = l; k < len; k += S) p += red[seg * len + k];
    for (int o = S >> 1; o > 0; o >>= 1) p += __shfl_xor(p, o, S);
    if (l == 0 && seg < nseg) out[seg] = p;
  }
}

// ---------------------------------------------------------------------------
// Workgroup -> XCD-contiguous tile. The dispatcher deals consecutive
// workgroup ids round-robin over the 8 XCDs; remapping id b to tile
// (b % 8) * per + b / 8 hands each XCD one contiguous run of cells, so the
// upwind neighbours a tile gathers were mostly just read through the same
// XCD's L2. The grid is launched with 8 * per workgroups; ids past the last
// tile return at once.
// ---------------------------------------------------------------------------
__device__ inline int64_t xcd_tile(int64_t n_tiles) {
  const int64_t per = (n_tiles + 7) / 8;
  return (int64_t)(blockIdx.x % 8) * per + blockIdx.x / 8;
}

// ---------------------------------------------------------------------------
// Structured FD apply (reference src/termsk.kokkos.cxx:70-98 streaming,
// :480-509 removal, :584-649 scatter; upwind/BC classification
// src/structured_fd_commonk.hpp:213-397 under the ghost-flux treatment)
// ---------------------------------------------------------------------------
// ADJ: the adjoint A^T - direction reversed (the upwind neighbour becomes the
// downwind one), the scatter transposed (unweighted moment, w_a on output).
// On a grid every cell has the same volume, so no volume ratio appears.
template <class T, int DIM, bool SCATTER, bool ADJ = false>
__global__ void __launch_bounds__(1024) fd_apply(const T *__restrict__ x, T *__restrict__ y, OpArgs<T> g) {
  extern __shared__ __align__(16) unsigned char smem[];
  T *red = reinterpret_cast<T *>(smem);
  T *phi = red + g.cpb * g.A;
  const int A = g.A;
  const int64_t tile = xcd_tile(((int64_t)g.n_cells + g.cpb - 1) / g.cpb);
  if (tile * g.cpb >= g.n_cells) return;
  const int lc = threadIdx.x / A, a = threadIdx.x - lc * A;
  const int64_t c = tile * g.cpb + lc;
  const bool live = lc < g.cpb && c < g.n_cells;
  T psi = T(0), acc = T(0);
  if (live) {
    psi = x[c * A + a];
    acc = g.sigt[c] * psi;
    int idx[3];
    int64_t cc = c;
    idx[0] = (int)(cc % g.n[0]);
    cc /= g.n[0];
    idx[1] = (int)(cc % g.n[1]);
    idx[2] = (int)(cc / g.n[1]);
    const int64_t stride[3] = {1, g.n[0], (int64_t)g.n[0] * g.n[1]};
#pragma unroll
    for (int d = 0; d < DIM; d++) {
      const T cs = ADJ ? -g.omega[3 * a + d] : g.omega[3 * a + d];
      const T cd = fabs(cs) * g.invh[d];
      acc += cd * psi;
      if (cs != T(0)) {
        const int up = cs > T(0) ? idx[d] - 1 : idx[d] + 1;
        if (up >= 0 && up < g.n[d])
          acc -= cd * x[(c + (cs > T(0) ? -stride[d] : stride[d])) * A + a];
        else if (g.face_type[d][cs > T(0) ? 0 : 1] == kReflect)
          acc -= cd * x[c * A + g.reflect[3 * a + d]];
      }
    }
  }
  if (SCATTER) {
    if (lc < g.cpb) red[lc * A + a] = live ? (ADJ ? psi : g.w[a] * psi) : T(0);
    __syncthreads();
    seg_sum(red, g.cpb, A, phi, g.red_lanes);
    __syncthreads();
    if (live) acc -= g.sigs[c] * g.inv_sum_w * (ADJ ? g.w[a] : T(1)) * phi[lc];
  }
  if (live) y[c * A + a] = acc;
}

// ---------------------------------------------------------------------------
// Unstructured DG0 / DG1 apply (reference src/termsk.kokkos.cxx:197-235 DG0,
// :328-385 DG1, removal and scatter as above). NB basis functions and NF faces
// per cell are compile-time; a thread owns (cell, angle) and all NB rows. The
// tile's geometry (contiguous per-cell records in HBM) is staged into LDS with
// coalesced loads once, then every angle of a cell reads it from LDS.
// ---------------------------------------------------------------------------
// ADJ: the adjoint A^T = V P L P V^{-1} - S^T (P reverses every angle): the
// same kernel with Omega -> -Omega, each neighbour's trace scaled by V_c / V_n,
// and the scatter transposed.
template <class T, int DIM, int NB, int NF, bool SCATTER, bool ADJ = false>
__global__ void __launch_bounds__(1024) dg_apply(const T *__restrict__ x, T *__restrict__ y, OpArgs<T> g) {
  extern __shared__ __align__(16) unsigned char smem[];
  constexpr int FM = NB * NB;
  const int cpb = g.cpb, A = g.A;
  const int64_t n_tiles = ((int64_t)g.n_cells + cpb - 1) / cpb;
  const int64_t tile = xcd_tile(n_tiles);
  if (tile >= n_tiles) return;
  const int64_t c0 = tile * cpb;
  const int ncell = (int)((int64_t)g.n_cells - c0 < cpb ? (int64_t)g.n_cells - c0 : cpb);
  // LDS: F_own, F_up, face s (DIM per slot), grad (NB-1 x DIM per cell),
  // sigma_t, sigma_s, then the reduction buffers; face_nb after the reals
  T *s_own = reinterpret_cast<T *>(smem);
  T *s_up = s_own + (NB > 1 ? cpb * NF * FM : 0);
  T *s_fs = s_up + (NB > 1 ? cpb * NF * FM : 0);
  T *s_gr = s_fs + cpb * NF * DIM;
  T *s_st = s_gr + (NB > 1 ? cpb * (NB - 1) * DIM : 0);
  T *s_ss = s_st + cpb;
  T *red = s_ss + cpb;
  T *phi = red + (SCATTER ? cpb * NB * A : 0);
  int *s_nb = reinterpret_cast<int *>(phi + (SCATTER ? cpb * NB : 0));
  const int tid = threadIdx.x, nt = blockDim.x;
  if (NB > 1) {
    const T *go = g.face_own + c0 * NF * FM, *gu = g.face_up + c0 * NF * FM;
    for (int e = tid; e < ncell * NF * FM; e += nt) {
      s_own[e] = go[e];
      s_up[e] = gu[e];
    }
    for (int e = tid; e < ncell * (NB - 1) * DIM; e += nt) {
      const int lc = e / ((NB - 1) * DIM), r = e - lc * ((NB - 1) * DIM);
      const int i = 1 + r / DIM, d = r - (i - 1) * DIM;
      s_gr[e] = g.grad[((c0 + lc) * NB + i) * 3 + d];
    }
  }
  for (int e = tid; e < ncell * NF; e += nt) {
    s_nb[e] = g.face_nb[c0 * NF + e];
#pragma unroll
    for (int d = 0; d < DIM; d++) s_fs[e * DIM + d] = g.face_s[(c0 * NF + e) * 3 + d];
  }
  for (int e = tid; e < ncell; e += nt) {
    s_st[e] = g.sigt[c0 + e];
    s_ss[e] = g.sigs[c0 + e];
  }
  __syncthreads();

  const int lc = tid / A, a = tid - lc * A;
  const int64_t c = c0 + lc;
  const bool live = lc < ncell;
  T psi[NB], acc[NB];
#pragma unroll
  for (int j = 0; j < NB; j++) psi[j] = acc[j] = T(0);
  if (live) {
    const int64_t base = c * NB * A + a;
#pragma unroll
    for (int j = 0; j < NB; j++) psi[j] = x[base + (int64_t)j * A];
    const T sig = s_st[lc];
#pragma unroll
    for (int i = 0; i < NB; i++) acc[i] = sig * psi[i];
    T om[DIM];
#pragma unroll
    for (int d = 0; d < DIM; d++) om[d] = ADJ ? -g.omega[3 * a + d] : g.omega[3 * a + d];
    if (NB > 1) {
      // volume term -(Omega . grad phi_i) psi_0; grad phi_0 = 0
#pragma unroll
      for (int i = 1; i < NB; i++) {
        const T *gr = s_gr + (lc * (NB - 1) + i - 1) * DIM;
        T od = T(0);
#pragma unroll
        for (int d = 0; d < DIM; d++) od += om[d] * gr[d];
        acc[i] -= od * psi[0];
      }
    }
#pragma unroll
    for (int f = 0; f < NF; f++) {
      const int kl = lc * NF + f;
      T s = T(0);
#pragma unroll
      for (int d = 0; d < DIM; d++) s += om[d] * s_fs[kl * DIM + d];
      if (s > T(0)) {
        if (NB == 1) {
          acc[0] += s * psi[0];
        } else {
          const T *F = s_own + kl * FM;
#pragma unroll
          for (int i = 0; i < NB; i++) {
            T v = T(0);
#pragma unroll
            for (int j = 0; j < NB; j++) v += F[i * NB + j] * psi[j];
            acc[i] += s * v;
          }
        }
      } else {
        const int nbk = s_nb[kl];
        if (nbk == -1) continue;  // vacuum: the inflow lives in the rhs
        int64_t ub;
        int ua;
        if (nbk >= 0) {
          ub = (int64_t)nbk * NB * A;
          ua = a;
        } else {  // reflective: this cell, mirrored angle
          ub = c * NB * A;
          ua = g.reflect[3 * a + (-2 - nbk)];
        }
        T pu[NB];
        const T vr = (ADJ && nbk >= 0) ? g.vol[c] / g.vol[nbk] : T(1);
#pragma unroll
        for (int j = 0; j < NB; j++) pu[j] = vr * x[ub + (int64_t)j * A + ua];
        if (NB == 1) {
          acc[0] += s * pu[0];
        } else {
          const T *F = s_up + kl * FM;
#pragma unroll
          for (int i = 0; i < NB; i++) {
            T v = T(0);
#pragma unroll
            for (int j = 0; j < NB; j++) v += F[i * NB + j] * pu[j];
            acc[i] += s * v;
          }
        }
      }
    }
  }
  if (SCATTER) {
    if (lc < cpb) {
      const T wa = live ? (ADJ ? T(1) : g.w[a]) : T(0);
#pragma unroll
      for (int j = 0; j < NB; j++) red[(lc * NB + j) * A + a] = wa * psi[j];
    }
    __syncthreads();
    seg_sum(red, cpb * NB, A, phi, g.red_lanes);
    __syncthreads();
    if (live) {
      const T ss = s_ss[lc] * g.inv_sum_w * (ADJ ? g.w[a] : T(1));
#pragma unroll
      for (int i = 0; i < NB; i++) acc[i] -= ss * phi[lc * NB + i];
    }
  }
  if (live) {
    const int64_t base = c * NB * A + a;
#pragma unroll
    for (int i = 0; i < NB; i++) y[base + (int64_t)i * A] = acc[i];
  }
}

// ---------------------------------------------------------------------------
// Wavefront-per-cell applies. A wavefront owns one cell and all its angles
// (lane l takes angles l, l + 64, ...: APL per lane), so everything per cell -
// cross sections, face normals and matrices, neighbour ids - sits at a
// wave-uniform address and is read through the scalar cache, and the
// scatter's moment is a butterfly across the wave. No LDS, no barrier: the
// latency chain of a launch is the own-psi load, then the upwind gather.
// Used where the angle count keeps most lanes busy (A > 32); the tile kernels
// above serve the small sets. The autotuner picks per operator.
// ---------------------------------------------------------------------------
template <class T>
__device__ inline T wave_sum(T v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

template <class T, int DIM, int NB, int NF, int APL, bool SCATTER>
__global__ void __launch_bounds__(256) dg_apply_wave(const T *__restrict__ x, T *__restrict__ y, OpArgs<T> g) {
  constexpr int FM = NB * NB;
  const int64_t n_tiles = ((int64_t)g.n_cells + 3) / 4;
  const int64_t tile = xcd_tile(n_tiles);
  const int lane = threadIdx.x & 63;
  const int64_t c = __builtin_amdgcn_readfirstlane((int)(tile * 4 + (threadIdx.x >> 6)));
  if (tile >= n_tiles || c >= g.n_cells) return;
  const int A = g.A;
  T psi[APL][NB], acc[APL][NB], om[APL][3];
  bool on[APL];
  const T sig = g.sigt[c];
#pragma unroll
  for (int q = 0; q < APL; q++) {
    const int a = lane + 64 * q;
    on[q] = a < A;
    const int ac = on[q] ? a : 0;
#pragma unroll
    for (int d = 0; d < 3; d++) om[q][d] = (d < DIM) ? g.omega[3 * ac + d] : T(0);
#pragma unroll
    for (int j = 0; j < NB; j++) psi[q][j] = on[q] ? x[(c * NB + j) * A + a] : T(0);
#pragma unroll
    for (int i = 0; i < NB; i++) acc[q][i] = sig * psi[q][i];
  }
  if (NB > 1) {
#pragma unroll
    for (int i = 1; i < NB; i++) {
      const T *gr = g.grad + (c * NB + i) * 3;
#pragma unroll
      for (int q = 0; q < APL; q++) {
        T od = T(0);
#pragma unroll
        for (int d = 0; d < DIM; d++) od += om[q][d] * gr[d];
        acc[q][i] -= od * psi[q][0];
      }
    }
  }
#pragma unroll
  for (int f = 0; f < NF; f++) {
    const int64_t k = c * NF + f;
    const int nbk = g.face_nb[k];
    const T *sv = g.face_s + 3 * k;
    const T *Fo = g.face_own + k * FM, *Fu = g.face_up + k * FM;
    int64_t ub = 0;
    int axis = 0;
    if (nbk >= 0) ub = (int64_t)nbk * NB * A;
    else if (nbk <= -2) { ub = c * NB * A; axis = -2 - nbk; }
#pragma unroll
    for (int q = 0; q < APL; q++) {
      if (!on[q]) continue;
      T s = T(0);
#pragma unroll
      for (int d = 0; d < DIM; d++) s += om[q][d] * sv[d];
      if (s > T(0)) {
        if (NB == 1) {
          acc[q][0] += s * psi[q][0];
        } else {
#pragma unroll
          for (int i = 0; i < NB; i++) {
            T v = T(0);
#pragma unroll
            for (int j = 0; j < NB; j++) v += Fo[i * NB + j] * psi[q][j];
            acc[q][i] += s * v;
          }
        }
      } else if (nbk != -1) {
        const int a = lane + 64 * q;
        const int ua = nbk >= 0 ? a : g.reflect[3 * a + axis];
        T pu[NB];
#pragma unroll
        for (int j = 0; j < NB; j++) pu[j] = x[ub + (int64_t)j * A + ua];
        if (NB == 1) {
          acc[q][0] += s * pu[0];
        } else {
#pragma unroll
          for (int i = 0; i < NB; i++) {
            T v = T(0);
#pragma unroll
            for (int j = 0; j < NB; j++) v += Fu[i * NB + j] * pu[j];
            acc[q][i] += s * v;
          }
        }
      }
    }
  }
  if (SCATTER) {
    const T ss = g.sigs[c] * g.inv_sum_w;
#pragma unroll
    for (int j = 0; j < NB; j++) {
      T part = T(0);
#pragma unroll
      for (int q = 0; q < APL; q++)
        if (on[q]) part += g.w[lane + 64 * q] * psi[q][j];
      const T phi = wave_sum(part);
#pragma unroll
      for (int q = 0; q < APL; q++) acc[q][j] -= ss * phi;
    }
  }
#pragma unroll
  for (int q = 0; q < APL; q++) {
    if (!on[q]) continue;
    const int a = lane + 64 * q;
#pragma unroll
    for (int i = 0; i < NB; i++) y[(c * NB + i) * A + a] = acc[q][i];
  }
}

template <class T, int DIM, int APL, bool SCATTER>
__global__ void __launch_bounds__(256) fd_apply_wave(const T *__restrict__ x, T *__restrict__ y, OpArgs<T> g) {
  const int64_t n_tiles = ((int64_t)g.n_cells + 3) / 4;
  const int64_t tile = xcd_tile(n_tiles);
  const int lane = threadIdx.x & 63;
  const int64_t c = __builtin_amdgcn_readfirstlane((int)(tile * 4 + (threadIdx.x >> 6)));
  if (tile >= n_tiles || c >= g.n_cells) return;
  const int A = g.A;
  int idx[3];
  {
    int64_t cc = c;
    idx[0] = (int)(cc % g.n[0]);
    cc /= g.n[0];
    idx[1] = (int)(cc % g.n[1]);
    idx[2] = (int)(cc / g.n[1]);
  }
  const int64_t stride[3] = {1, g.n[0], (int64_t)g.n[0] * g.n[1]};
  const T sig = g.sigt[c];
  T psi[APL], acc[APL];
#pragma unroll
  for (int q = 0; q < APL; q++) {
    const int a = lane + 64 * q;
    psi[q] = T(0);
    acc[q] = T(0);
    if (a >= A) continue;
    psi[q] = x[c * A + a];
    T v = sig * psi[q];
#pragma unroll
    for (int d = 0; d < DIM; d++) {
      const T cs = g.omega[3 * a + d];
      const T cd = fabs(cs) * g.invh[d];
      v += cd * psi[q];
      if (cs == T(0)) continue;
      const int up = cs > T(0) ? idx[d] - 1 : idx[d] + 1;
      if (up >= 0 && up < g.n[d])
        v -= cd * x[(c + (cs > T(0) ? -stride[d] : stride[d])) * A + a];
      else if (g.face_type[d][cs > T(0) ? 0 : 1] == kReflect)
        v -= cd * x[c * A + g.reflect[3 * a + d]];
    }
    acc[q] = v;
  }
  if (SCATTER) {
    T part = T(0);
#pragma unroll
    for (int q = 0; q < APL; q++)
      if (lane + 64 * q < A) part += g.w[lane + 64 * q] * psi[q];
    const T phi = wave_sum(part);
    const T ss = g.sigs[c] * g.inv_sum_w;
#pragma unroll
    for (int q = 0; q < APL; q++) acc[q] -= ss * phi;
  }
#pragma unroll
  for (int q = 0; q < APL; q++) {
    const int a = lane + 64 * q;
    if (a < A) y[c * A + a] = acc[q];
  }
}

// ---------------------------------------------------------------------------
// Composed diagonal (reference src/transport_operatork.kokkos.cxx:240-256 and
// the add_diagonal of each term, src/termsk.kokkos.cxx:242-270,391-419,517-539)
// ---------------------------------------------------------------------------
template <class T>
__global__ void diag_kernel(T *__restrict__ d, OpArgs<T> g, int is_fd) {
  const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t rows = (int64_t)g.n_cells * g.nb * g.A;
  if (r >= rows) return;
  const int a = (int)(r % g.A);
  const int64_t node = r / g.A;
  const int64_t c = node / g.nb;
  const int i = (int)(node % g.nb);
  T v = T(0);
  if (is_fd) {
    for (int dd = 0; dd < g.dim; dd++) v += fabs(g.omega[3 * a + dd]) * g.invh[dd];
  } else {
    for (int f = 0; f < g.nf; f++) {
      const int64_t k = c * g.nf + f;
      T s = T(0);
      for (int dd = 0; dd < g.dim; dd++) s += g.omega[3 * a + dd] * g.face_s[3 * k + dd];
      if (s > T(0)) v += g.nb == 1 ? s : s * g.face_own[(k * g.nb + i) * g.nb + i];
    }
  }
  d[r] = v + g.sigt[c];
}

// ---------------------------------------------------------------------------
// Right-hand side for group g: ghost inflow + isotropic source on basis 0 +
// transfers from every group above (reference tests/transportk.kokkos.cxx:
// 1058-1065; src/discretisationk.kokkos.cxx:187-206;
// src/material_speck.kokkos.cxx:164-187; src/multigroupk.kokkos.cxx:332-359).
// All FP64 arithmetic, stored in T.
// ---------------------------------------------------------------------------
struct RhsArgs {
  int fd, dim, nb, nf, A, n_cells, g, n_groups;
  int gf_lo, gf_hi, skip;     // transfers from groups [gf_lo, gf_hi), skipping group `skip`
  double inv_sum_w;
  const double *omega;
  const double *source;       // G * cells or null
  const double *sigs;         // G * G * cells ([g_from][g_to][cell])
  const double *phi_cache;    // G * nodes
  // structured
  int n[3];
  double h[3];
  double face_inflow[3][2];   // inflow / W on a vacuum face, 0 otherwise
  int face_npairs[3][2];
  double face_window[3][2][4];
  // unstructured
  const int *face_nb;
  const double *face_nA;      // raw outward area normals, 3 per slot
  const double *inv_volume;
  const double *face_inflow_slot;
  const double *face_basis_value;
};

template <class T>
__global__ void rhs_kernel(T *__restrict__ b, RhsArgs r, int accumulate_only_transfers) {
  const int64_t row = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t rows = (int64_t)r.n_cells * r.nb * r.A;
  if (row >= rows) return;
  const int a = (int)(row % r.A);
  const int64_t node = row / r.A;
  const int64_t c = node / r.nb;
  const int i = (int)(node % r.nb);
  double v = 0.0;
  if (!accumulate_only_transfers) {
    if (r.fd) {
      int idx[3];
      int64_t cc = c;
      idx[0] = (int)(cc % r.n[0]);
      cc /= r.n[0];
      idx[1] = (int)(cc % r.n[1]);
      idx[2] = (int)(cc / r.n[1]);
      for (int d = 0; d < r.dim; d++) {
        const double cs = r.omega[3 * a + d];
        if (cs == 0.0) continue;
        const int up = cs > 0.0 ? idx[d] - 1 : idx[d] + 1;
        if (up >= 0 && up < r.n[d]) continue;
        const int side = cs > 0.0 ? 0 : 1;
        const double inflow = r.face_inflow[d][side];
        if (inflow == 0.0) continue;
        bool inside = true;
        int p = 0;
        for (int e = 0; e < r.dim && inside; e++) {
          if (e == d) continue;
          if (p < r.face_npairs[d][side]) {
            const double t = (idx[e] + 0.5) * r.h[e];
            inside = t >= r.face_window[d][side][2 * p] && t <= r.face_window[d][side][2 * p + 1];
          }
          p++;
        }
        if (inside) v += fabs(cs) / r.h[d] * inflow;
      }
    } else {
      for (int f = 0; f < r.nf; f++) {
        const int64_t k = c * r.nf + f;
        if (r.face_nb[k] != -1) continue;
        const double fi = r.face_inflow_slot[k];
        if (fi == 0.0) continue;
        double s = 0.0;
        for (int d = 0; d < r.dim; d++) s += r.omega[3 * a + d] * r.face_nA[3 * k + d];
        if (s < 0.0) v += -s * r.inv_volume[c] * r.face_basis_value[k * r.nb + i] * fi;
      }
    }
    if (i == 0 && r.source) v += r.source[(int64_t)r.g * r.n_cells + c] * r.inv_sum_w;
  }
  const int64_t nodes = (int64_t)r.n_cells * r.nb;
  for (int gf = r.gf_lo; gf < r.gf_hi; gf++) {
    if (gf == r.skip) continue;
    const double s = r.sigs[((int64_t)gf * r.n_groups + r.g) * r.n_cells + c];
    if (s != 0.0) v += s * r.inv_sum_w * r.phi_cache[(int64_t)gf * nodes + node];
  }
  if (accumulate_only_transfers)
    b[row] += (T)v;
  else
    b[row] = (T)v;
}

// ---------------------------------------------------------------------------
// Angular moments on the matrix cores (reference UboltAngularIntegral,
// src/sn_quadraturek.kokkos.cxx:400-427): phi[N x M] = psi[N x A] . W[A x M],
// W[a][m] = w_a {1, Omega_x, Omega_y, Omega_z}_m. A wavefront owns 16 nodes;
// the block stages 64 nodes x A (one contiguous run of psi) through LDS, then
// each wavefront issues Apad/4 MFMA 16x16x4 steps with W zero-padded to 16
// columns. Operand layout of the 16x16x4 MFMAs: lane l supplies A[l%16][l/16]
// and B[l/16][l%16]; accumulator item r = 0..3 of lane l is D[row][l%16] with
// row = 4*(l/16) + r for f32 and row = 4*r + l/16 for f64 (measured on gfx950
// by the moments parity test).
// ---------------------------------------------------------------------------
typedef float v4f __attribute__((ext_vector_type(4)));
typedef double v4d __attribute__((ext_vector_type(4)));

__device__ inline v4f mfma16x16x4(float a, float b, v4f c) {
  return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
}
__device__ inline v4d mfma16x16x4(double a, double b, v4d c) {
  return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}

template <class T, class OutT>
__global__ void __launch_bounds__(256) moments_mfma(const T *__restrict__ psi, OutT *__restrict__ phi,
                                                    const T *__restrict__ wext, int64_t n_nodes, int A,
                                                    int Apad, int M, int phi_stride) {
  using V4 = typename std::conditional<sizeof(T) == 4, v4f, v4d>::type;
  extern __shared__ __align__(16) unsigned char smem[];
  T *tile = reinterpret_cast<T *>(smem);  // 64 x Apad
  const int64_t node0 = (int64_t)blockIdx.x * 64;
  const int64_t nn = n_nodes - node0 < 64 ? n_nodes - node0 : 64;
  // stage: contiguous psi[node0 .. node0+nn) x A, zero-padded columns
  for (int e = threadIdx.x; e < 64 * Apad; e += blockDim.x) {
    const int row = e / Apad, col = e - row * Apad;
    tile[e] = (row < nn && col < A) ? psi[(node0 + row) * A + col] : T(0);
  }
  __syncthreads();
  const int wave = threadIdx.x >> 6, l = threadIdx.x & 63;
  const int i = l & 15, kq = l >> 4;
  V4 acc = {T(0), T(0), T(0), T(0)};
  const T *trow = tile + (wave * 16 + i) * Apad;
  for (int k0 = 0; k0 < Apad; k0 += 4) acc = mfma16x16x4(trow[k0 + kq], wext[(k0 + kq) * 16 + i], acc);
  const int col = i;
  if (col < M) {
#pragma unroll
    for (int r = 0; r < 4; r++) {
      const int row = wave * 16 + (sizeof(T) == 8 ? 4 * r + kq : 4 * kq + r);
      if (row < nn) phi[(node0 + row) * phi_stride + col] = (OutT)acc[r];
    }
  }
}

// ---------------------------------------------------------------------------
// Element block inverse (reference src/block_inversek.kokkos.cxx:83-150,
// 206-222). The (cell, angle) block - own-cell outflow face couplings, the
// volume term and sigma_t; reflective couplings excluded - is rebuilt from the
// geometry on the fly and solved by Gauss-Jordan with partial pivoting, which
// reads NB values per row instead of NB*NB stored inverse entries.
// ---------------------------------------------------------------------------
template <class T, int NB>
__device__ inline bool build_block(const OpArgs<T> &g, int64_t c, int a, int is_fd, T B[NB][NB], T dir = T(1)) {
#pragma unroll
  for (int i = 0; i < NB; i++)
#pragma unroll
    for (int j = 0; j < NB; j++) B[i][j] = T(0);
  const T sig = g.sigt[c];
  if (is_fd) {
    T v = T(0);
    for (int d = 0; d < g.dim; d++) v += fabs(g.omega[3 * a + d]) * g.invh[d];
    B[0][0] = v + sig;
    return true;
  }
  const T ox = dir * g.omega[3 * a], oy = dir * g.omega[3 * a + 1], oz = dir * g.omega[3 * a + 2];
  for (int f = 0; f < g.nf; f++) {
    const int64_t k = c * g.nf + f;
    const T s = ox * g.face_s[3 * k] + oy * g.face_s[3 * k + 1] + oz * g.face_s[3 * k + 2];
    if (s > T(0)) {
      if (NB == 1) {
        B[0][0] += s;
      } else {
        const T *F = g.face_own + k * NB * NB;
#pragma unroll
        for (int i = 0; i < NB; i++)
#pragma unroll
          for (int j = 0; j < NB; j++) B[i][j] += s * F[i * NB + j];
      }
    }
  }
  if (NB > 1) {
#pragma unroll
    for (int i = 1; i < NB; i++) {
      const T *gr = g.grad + (c * NB + i) * 3;
      B[i][0] -= ox * gr[0] + oy * gr[1] + oz * gr[2];
    }
  }
#pragma unroll
  for (int i = 0; i < NB; i++) B[i][i] += sig;
  return true;
}

// Solve B z = v in place by Gauss-Jordan with partial pivoting; false if singular
template <class T, int NB>
__device__ inline bool gj_solve(T B[NB][NB], T v[NB]) {
#pragma unroll
  for (int k = 0; k < NB; k++) {
    int piv = k;
#pragma unroll
    for (int i = k + 1; i < NB; i++)
      if (fabs(B[i][k]) > fabs(B[piv][k])) piv = i;
    if (B[piv][k] == T(0)) return false;
    if (piv != k) {
#pragma unroll
      for (int j = 0; j < NB; j++) {
        T t = B[k][j];
        B[k][j] = B[piv][j];
        B[piv][j] = t;
      }
      T t = v[k];
      v[k] = v[piv];
      v[piv] = t;
    }
    const T inv = T(1) / B[k][k];
#pragma unroll
    for (int j = 0; j < NB; j++) B[k][j] *= inv;
    v[k] *= inv;
#pragma unroll
    for (int i = 0; i < NB; i++) {
      if (i == k) continue;
      const T f = B[i][k];
#pragma unroll
      for (int j = 0; j < NB; j++) B[i][j] -= f * B[k][j];
      v[i] -= f * v[k];
    }
  }
  return true;
}

// transpose: the blocks of A^T are the forward blocks at the reversed direction
template <class T, int NB>
__global__ void blockinv_apply(const T *__restrict__ x, T *__restrict__ y, OpArgs<T> g, int is_fd, int transpose) {
  const int64_t b = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= (int64_t)g.n_cells * g.A) return;
  const int64_t c = b / g.A;
  const int a = (int)(b - c * g.A);
  T B[NB][NB], v[NB];
  build_block<T, NB>(g, c, a, is_fd, B, transpose ? T(-1) : T(1));
  const int64_t base = c * NB * g.A + a;
#pragma unroll
  for (int j = 0; j < NB; j++) v[j] = x[base + (int64_t)j * g.A];
  if (!gj_solve<T, NB>(B, v)) {
#pragma unroll
    for (int j = 0; j < NB; j++) v[j] = T(0);
  }
#pragma unroll
  for (int i = 0; i < NB; i++) y[base + (int64_t)i * g.A] = v[i];
}

template <int NB>
__global__ void blockinv_check(OpArgs<double> g, int is_fd, int *n_singular) {
  const int64_t b = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= (int64_t)g.n_cells * g.A) return;
  const int64_t c = b / g.A;
  const int a = (int)(b - c * g.A);
  double B[NB][NB], v[NB];
  build_block<double, NB>(g, c, a, is_fd, B);
#pragma unroll
  for (int j = 0; j < NB; j++) v[j] = 1.0;
  if (!gj_solve<double, NB>(B, v)) atomicAdd(n_singular, 1);
}

// ---------------------------------------------------------------------------
// host-side launch helpers
// ---------------------------------------------------------------------------
struct Launch {
  int threads, cpb, lanes;
  size_t lds;
};

int pow2_floor(int v) {
  int p = 1;
  while (p * 2 <= v) p *= 2;
  return p;
}

template <class T> Launch apply_launch(int A, int nb) {
  Launch L;
  L.threads = A <= 256 ? 256 : ((A + 63) / 64) * 64;
  UBT_CHECK(L.threads <= 1024, UBT_ERR_SUP, "%d angles per cell exceed one workgroup", A);
  L.cpb = L.threads / A;
  const int nseg = L.cpb * nb;
  int S = pow2_floor(std::max(1, L.threads / nseg));
  S = std::min(S, 64);
  S = std::min(S, pow2_floor(A));
  L.lanes = std::max(1, S);
  L.lds = (size_t)(L.cpb * nb * A + L.cpb * nb) * sizeof(T);
  UBT_CHECK(L.lds <= 65536, UBT_ERR_SUP, "%d angles x %d basis need %zu bytes of LDS per workgroup", A, nb, L.lds);
  return L;
}

// The wavefront kernels: one cell per wavefront, 4 per workgroup, APL angles per lane
template <class T>
bool launch_apply_wave(const Operator &op, const OpArgs<T> &a, const T *x, T *y, bool scatter, hipStream_t s) {
  const Disc &D = *op.disc;
  const int apl = (D.n_angles + 63) / 64;
  if (apl > 4) return false;
  const int64_t n_tiles = (D.n_cells + 3) / 4;
  const unsigned grid = (unsigned)(8 * ((n_tiles + 7) / 8));
  bool launched = false;
#define W_FD(DIM, APL)                                                                                         \
  if (!launched && D.kind == DISC_FD && D.dim == DIM && apl == APL) {                                         \
    if (scatter)                                                                                               \
      hipLaunchKernelGGL((fd_apply_wave<T, DIM, APL, true>), dim3(grid), dim3(256), 0, s, x, y, a);            \
    else                                                                                                       \
      hipLaunchKernelGGL((fd_apply_wave<T, DIM, APL, false>), dim3(grid), dim3(256), 0, s, x, y, a);           \
    launched = true;                                                                                           \
  }
#define W_DG(DIM, NB, NF, APL)                                                                                 \
  if (!launched && D.kind == DISC_DG && D.dim == DIM && D.nb == NB && D.nf == NF && apl == APL) {             \
    if (scatter)                                                                                               \
      hipLaunchKernelGGL((dg_apply_wave<T, DIM, NB, NF, APL, true>), dim3(grid), dim3(256), 0, s, x, y, a);    \
    else                                                                                                       \
      hipLaunchKernelGGL((dg_apply_wave<T, DIM, NB, NF, APL, false>), dim3(grid), dim3(256), 0, s, x, y, a);   \
    launched = true;                                                                                           \
  }
#define W_DG_ALL(APL) W_DG(2, 1, 3, APL) W_DG(2, 1, 4, APL) W_DG(3, 1, 4, APL) W_DG(3, 1, 6, APL) \
  W_DG(2, 3, 3, APL) W_DG(2, 3, 4, APL) W_DG(3, 4, 4, APL) W_DG(3, 4, 6, APL)
  W_FD(2, 1) W_FD(2, 2) W_FD(3, 1) W_FD(3, 2) W_FD(3, 3) W_FD(3, 4)
  W_DG_ALL(1) W_DG_ALL(2)
#undef W_DG_ALL
#undef W_DG
#undef W_FD
  return launched;
}

template <class T>
void launch_apply(const Operator &op, const OpArgs<T> &a0, const T *x, T *y, bool scatter, hipStream_t s,
                  bool adj = false) {
  const Disc &D = *op.disc;
  OpArgs<T> a = a0;
  if (!adj && op.variant[sizeof(T) == 8 ? 0 : 1] == 1 && launch_apply_wave<T>(op, a, x, y, scatter, s)) {
    UBT_HIP(hipGetLastError());
    return;
  }
  Launch L = apply_launch<T>(D.n_angles, D.nb);
  a.cpb = L.cpb;
  a.red_lanes = L.lanes;
  const int64_t n_tiles = (D.n_cells + L.cpb - 1) / L.cpb;
  // XCD-contiguous tiles: 8 * ceil(tiles / 8) workgroups
  const unsigned grid = (unsigned)(8 * ((n_tiles + 7) / 8));
  if (D.kind == DISC_DG) {
    // the staged geometry in LDS
    const int nb = D.nb, nf = D.nf, dim = D.dim;
    const size_t reals = (size_t)L.cpb * ((nb > 1 ? 2 * nf * nb * nb : 0) + nf * dim + (nb > 1 ? (nb - 1) * dim : 0) + 2) +
                         (size_t)L.cpb * nb * D.n_angles + (size_t)L.cpb * nb;
    L.lds = reals * sizeof(T) + sizeof(int) * (size_t)L.cpb * nf;
    UBT_CHECK(L.lds <= 65536, UBT_ERR_SUP, "the DG tile needs %zu bytes of LDS", L.lds);
  }
  if (D.kind == DISC_FD) {
#define FD_LAUNCH(DIM)                                                                                       \
  if (adj) {                                                                                                 \
    if (scatter)                                                                                             \
      hipLaunchKernelGGL((fd_apply<T, DIM, true, true>), dim3(grid), dim3(L.threads), L.lds, s, x, y, a);     \
    else                                                                                                     \
      hipLaunchKernelGGL((fd_apply<T, DIM, false, true>), dim3(grid), dim3(L.threads), L.lds, s, x, y, a);    \
  } else if (scatter)                                                                                        \
    hipLaunchKernelGGL((fd_apply<T, DIM, true>), dim3(grid), dim3(L.threads), L.lds, s, x, y, a);             \
  else                                                                                                       \
    hipLaunchKernelGGL((fd_apply<T, DIM, false>), dim3(grid), dim3(L.threads), L.lds, s, x, y, a);
    if (D.dim == 2) {
      FD_LAUNCH(2)
    } else {
      FD_LAUNCH(3)
    }
#undef FD_LAUNCH
  } else {
#define DG_LAUNCH(DIM, NB, NF)                                                                                    \
  if (D.dim == DIM && D.nb == NB && D.nf == NF) {                                                                 \
    if (adj) {                                                                                                    \
      if (scatter)                                                                                                \
        hipLaunchKernelGGL((dg_apply<T, DIM, NB, NF, true, true>), dim3(grid), dim3(L.threads), L.lds, s, x, y, a); \
      else                                                                                                        \
        hipLaunchKernelGGL((dg_apply<T, DIM, NB, NF, false, true>), dim3(grid), dim3(L.threads), L.lds, s, x, y, a); \
    } else if (scatter)                                                                                           \
      hipLaunchKernelGGL((dg_apply<T, DIM, NB, NF, true>), dim3(grid), dim3(L.threads), L.lds, s, x, y, a);      \
    else                                                                                                          \
      hipLaunchKernelGGL((dg_apply<T, DIM, NB, NF, false>), dim3(grid), dim3(L.threads), L.lds, s, x, y, a);     \
    launched = true;                                                                                              \
  }
    bool launched = false;
    DG_LAUNCH(2, 1, 3)
    DG_LAUNCH(2, 1, 4)
    DG_LAUNCH(3, 1, 4)
    DG_LAUNCH(3, 1, 6)
    DG_LAUNCH(2, 3, 3)
    DG_LAUNCH(2, 3, 4)
    DG_LAUNCH(3, 4, 4)
    DG_LAUNCH(3, 4, 6)
#undef DG_LAUNCH
    UBT_CHECK(launched, UBT_ERR_SUP, "no DG kernel for dim %d, %d basis, %d faces", D.dim, D.nb, D.nf);
  }
  UBT_HIP(hipGetLastError());
}

template <class T> OpArgs<T> make_args(const Operator &op, const PrecTables &t) {
  const Disc &D = *op.disc;
  OpArgs<T> a{};
  a.dim = D.dim;
  a.nb = D.nb;
  a.nf = D.nf;
  a.A = D.n_angles;
  a.n_cells = D.n_cells;
  a.inv_sum_w = (T)(1.0 / D.quad.sum_weights);
  a.omega = t.omega.as<T>();
  a.w = t.w.as<T>();
  a.reflect = op.reflect.as<int>();
  a.neg = op.neg.as<int>();
  a.vol = t.vol.as<T>();
  a.sigt = t.sigt.as<T>() + (int64_t)op.group * D.n_cells;
  a.sigs = t.sigs.as<T>() + ((int64_t)op.group * op.n_groups + op.group) * D.n_cells;
  for (int d = 0; d < 3; d++) {
    a.n[d] = D.n[d];
    a.invh[d] = (T)(1.0 / D.h[d]);
    a.face_type[d][0] = D.face_type[d][0];
    a.face_type[d][1] = D.face_type[d][1];
  }
  a.face_nb = op.face_nb.as<int>();
  a.face_s = t.face_s.as<T>();
  a.face_own = t.face_own.as<T>();
  a.face_up = t.face_up.as<T>();
  a.grad = t.grad.as<T>();
  return a;
}

}  // namespace

template <> OpArgs<double> Operator::args<double>() const { return make_args<double>(*this, t64); }
template <> OpArgs<float> Operator::args<float>() const { return make_args<float>(*this, t32); }

// ---------------------------------------------------------------------------
// Operator construction: upload every table once
// ---------------------------------------------------------------------------
static void upload_prec(PrecTables &t, const Disc &D, bool fp32) {
  std::vector<double> omega(D.quad.omega), w(D.quad.w);
  std::vector<double> face_s;
  if (D.kind == DISC_DG) {
    const size_t slots = (size_t)D.n_cells * D.nf;
    face_s.assign(3 * slots, 0.0);
    for (size_t k = 0; k < slots; k++) {
      const double sc = D.order == 0 ? 1.0 / D.volume[k / D.nf] : 1.0;
      for (int d = 0; d < 3; d++) face_s[3 * k + d] = D.face_nA[3 * k + d] * sc;
    }
  }
  const int A = D.n_angles, Apad = (A + 3) / 4 * 4;
  std::vector<double> wext((size_t)Apad * 16, 0.0), wunit((size_t)Apad * 16, 0.0);
  for (int a = 0; a < A; a++) {
    wunit[(size_t)a * 16] = 1.0;
    wext[(size_t)a * 16] = w[a];
    for (int d = 0; d < 3; d++) wext[(size_t)a * 16 + 1 + d] = w[a] * omega[3 * a + d];
  }
  if (fp32) {
    t.wext.upload_as<float>(wext);
    t.wunit.upload_as<float>(wunit);
    t.omega.upload_as<float>(omega);
    t.w.upload_as<float>(w);
    t.face_s.upload_as<float>(face_s);
    t.face_own.upload_as<float>(D.face_own);
    t.face_up.upload_as<float>(D.face_up);
    t.grad.upload_as<float>(D.basis_grad);
    t.vol.upload_as<float>(D.volume);
  } else {
    t.wext.upload(wext);
    t.wunit.upload(wunit);
    t.omega.upload(omega);
    t.w.upload(w);
    t.face_s.upload(face_s);
    t.face_own.upload(D.face_own);
    t.face_up.upload(D.face_up);
    t.grad.upload(D.basis_grad);
    t.vol.upload(D.volume);
  }
}

std::unique_ptr<Operator> make_operator(std::shared_ptr<Disc> disc, int n_groups) {
  UBT_CHECK(n_groups >= 1, UBT_ERR_ARG_OUTOFRANGE, "n_groups must be positive, got %d", n_groups);
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) fail(UBT_ERR_GPU, "no HIP device is visible");
  auto op = std::make_unique<Operator>();
  op->disc = disc;
  op->n_groups = n_groups;
  UBT_HIP(hipGetDevice(&op->device));
  const Disc &D = *disc;
  upload_prec(op->t64, D, false);
  upload_prec(op->t32, D, true);
  std::vector<int> refl(3 * (size_t)D.n_angles, -1);
  for (int d = 0; d < D.dim; d++)
    for (int a = 0; a < D.n_angles; a++) refl[3 * a + d] = D.quad.reflect[d][a];
  op->reflect.upload(refl);
  std::vector<int> neg(D.n_angles);
  for (int a = 0; a < D.n_angles; a++) {
    int b = a;
    for (int d = 0; d < D.dim; d++) b = D.quad.reflect[d][b];
    neg[a] = b;
  }
  op->neg.upload(neg);
  if (D.kind == DISC_DG) {
    op->face_nb.upload(D.face_nb);
    op->face_nA_raw.upload(D.face_nA);
    op->volume_d.upload(D.volume);
    op->face_area_d.upload(D.face_area);
    op->face_dist_d.upload(D.face_distance);
    std::vector<double> iv(D.n_cells);
    for (int c = 0; c < D.n_cells; c++) iv[c] = 1.0 / D.volume[c];
    op->inv_volume.upload(iv);
    op->face_inflow.upload(D.face_inflow);
    op->face_basis_value.upload(D.face_basis_value);
  }
  op->phi_cache.alloc(sizeof(double) * (size_t)n_groups * D.n_nodes());
  op->phi_set.assign(n_groups, 0);
  op->flag.alloc(sizeof(int) * 4);
  return op;
}

void operator_set_xs(Operator &op, const double *sigma_t, const double *sigma_s, const double *source) {
  const Disc &D = *op.disc;
  const size_t G = op.n_groups, nc = D.n_cells;
  UBT_CHECK(sigma_t && sigma_s, UBT_ERR_ARG_NULL, "sigma_t and sigma_s are required");
  std::vector<double> st(sigma_t, sigma_t + G * nc), ss(sigma_s, sigma_s + G * G * nc);
  for (double v : st) UBT_CHECK(std::isfinite(v), UBT_ERR_ARG_WRONG, "sigma_t holds a non-finite value");
  op.t64.sigt.upload(st);
  op.t64.sigs.upload(ss);
  op.t32.sigt.upload_as<float>(st);
  op.t32.sigs.upload_as<float>(ss);
  op.sigs_d.upload(ss);
  const bool first = !op.xs_set;
  if (source) {
    std::vector<double> q(source, source + G * nc);
    op.source.upload(q);
  } else {
    op.source.release();
  }
  op.xs_set = true;
  if (first) op.tune();
}

// ---------------------------------------------------------------------------
// Operator methods
// ---------------------------------------------------------------------------
void Operator::mult(int prec, const void *x, void *y, bool scatter, hipStream_t s) const {
  UBT_CHECK(xs_set, UBT_ERR_ARG_WRONGSTATE, "set the cross sections before applying the operator");
  UBT_CHECK(x && y, UBT_ERR_ARG_NULL, "x and y are required");
  UBT_CHECK(x != y, UBT_ERR_ARG_IDN, "the operator apply needs distinct x and y");
  if (prec == UBT_FP64)
    launch_apply<double>(*this, args<double>(), (const double *)x, (double *)y, scatter, s);
  else if (prec == UBT_FP32)
    launch_apply<float>(*this, args<float>(), (const float *)x, (float *)y, scatter, s);
  else
    fail(UBT_ERR_ARG_OUTOFRANGE, "unknown precision %d", prec);
}

// ---------------------------------------------------------------------------
// Autotuner: time each apply variant (with scatter) on zero-filled scratch
// vectors - 2 warm-up and 5 timed launches, HIP events - and keep the fastest
// per precision. Runs when the cross sections are first set; the scratch is
// freed before returning, so applies still allocate nothing.
// ---------------------------------------------------------------------------
void Operator::tune() {
  const Disc &D = *disc;
  const int64_t n = D.n_rows();
  const bool wave_ok = (D.n_angles + 63) / 64 <= (D.kind == DISC_FD ? (D.dim == 3 ? 4 : 2) : 2);
  hipStream_t s = nullptr;
  UBT_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
  hipEvent_t e0, e1;
  UBT_HIP(hipEventCreate(&e0));
  UBT_HIP(hipEventCreate(&e1));
  DevBuf xb, yb;
  xb.alloc(sizeof(double) * n);
  yb.alloc(sizeof(double) * n);
  UBT_HIP(hipMemsetAsync(xb.ptr, 0, sizeof(double) * n, s));
  for (int p = 0; p < 2; p++) {
    float best = 1e30f;
    int best_v = 0;
    for (int v = 0; v < (wave_ok ? 2 : 1); v++) {
      variant[p] = v;
      for (int it = 0; it < 2; it++) mult(p == 0 ? UBT_FP64 : UBT_FP32, xb.ptr, yb.ptr, true, s);
      UBT_HIP(hipEventRecord(e0, s));
      for (int it = 0; it < 5; it++) mult(p == 0 ? UBT_FP64 : UBT_FP32, xb.ptr, yb.ptr, true, s);
      UBT_HIP(hipEventRecord(e1, s));
      UBT_HIP(hipEventSynchronize(e1));
      float ms = 0.f;
      UBT_HIP(hipEventElapsedTime(&ms, e0, e1));
      tuned_us[p][v] = ms * 1000.f / 5.f;
      if (tuned_us[p][v] < best) {
        best = tuned_us[p][v];
        best_v = v;
      }
    }
    variant[p] = best_v;
  }
  UBT_HIP(hipStreamSynchronize(s));
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  (void)hipStreamDestroy(s);
}

void Operator::mult_transpose(int prec, const void *x, void *y, bool scatter, hipStream_t s) const {
  UBT_CHECK(xs_set, UBT_ERR_ARG_WRONGSTATE, "set the cross sections before applying the operator");
  UBT_CHECK(x && y, UBT_ERR_ARG_NULL, "x and y are required");
  UBT_CHECK(x != y, UBT_ERR_ARG_IDN, "the operator apply needs distinct x and y");
  if (prec == UBT_FP64)
    launch_apply<double>(*this, args<double>(), (const double *)x, (double *)y, scatter, s, true);
  else if (prec == UBT_FP32)
    launch_apply<float>(*this, args<float>(), (const float *)x, (float *)y, scatter, s, true);
  else
    fail(UBT_ERR_ARG_OUTOFRANGE, "unknown precision %d", prec);
}

void Operator::diagonal(int prec, void *d, hipStream_t s) const {
  UBT_CHECK(xs_set, UBT_ERR_ARG_WRONGSTATE, "set the cross sections first");
  const int64_t rows = disc->n_rows();
  const unsigned grid = (unsigned)((rows + 255) / 256);
  const int fd = disc->kind == DISC_FD;
  if (prec == UBT_FP64)
    hipLaunchKernelGGL(diag_kernel<double>, dim3(grid), dim3(256), 0, s, (double *)d, args<double>(), fd);
  else
    hipLaunchKernelGGL(diag_kernel<float>, dim3(grid), dim3(256), 0, s, (float *)d, args<float>(), fd);
  UBT_HIP(hipGetLastError());
}

namespace {
template <class T, class OutT>
void launch_moments(const Operator &op, const PrecTables &t, const T *psi, OutT *phi, int M, int stride,
                    hipStream_t s, bool unit = false) {
  const Disc &D = *op.disc;
  const int A = D.n_angles, Apad = (A + 3) / 4 * 4;
  const int64_t N = D.n_nodes();
  const unsigned grid = (unsigned)((N + 63) / 64);
  const size_t lds = (size_t)64 * Apad * sizeof(T);
  UBT_CHECK(lds <= 65536, UBT_ERR_SUP, "the moments tile of %d angles needs %zu bytes of LDS", A, lds);
  hipLaunchKernelGGL((moments_mfma<T, OutT>), dim3(grid), dim3(256), lds, s, psi, phi,
                     unit ? t.wunit.as<T>() : t.wext.as<T>(), N, A, Apad, M, stride);
  UBT_HIP(hipGetLastError());
}

RhsArgs make_rhs_args(const Operator &op, int g, int gf_lo, int gf_hi, int skip) {
  const Disc &D = *op.disc;
  RhsArgs r{};
  r.fd = D.kind == DISC_FD;
  r.dim = D.dim;
  r.nb = D.nb;
  r.nf = D.nf;
  r.A = D.n_angles;
  r.n_cells = D.n_cells;
  r.g = g;
  r.n_groups = op.n_groups;
  r.gf_lo = gf_lo;
  r.gf_hi = gf_hi;
  r.skip = skip;
  r.inv_sum_w = 1.0 / D.quad.sum_weights;
  r.omega = op.t64.omega.as<double>();
  r.source = op.source.as<double>();
  r.sigs = op.sigs_d.as<double>();
  r.phi_cache = op.phi_cache.as<double>();
  for (int d = 0; d < 3; d++) {
    r.n[d] = D.n[d];
    r.h[d] = D.h[d];
    for (int side = 0; side < 2; side++) {
      const BCFace &f = D.face_bc[d][side];
      r.face_inflow[d][side] = (D.face_type[d][side] == UBT_BC_VACUUM) ? f.inflow / D.quad.sum_weights : 0.0;
      r.face_npairs[d][side] = f.n_window_pairs;
      for (int q = 0; q < 4; q++) r.face_window[d][side][q] = f.window[q];
    }
  }
  r.face_nb = op.face_nb.as<int>();
  r.face_nA = op.face_nA_raw.as<double>();
  r.inv_volume = op.inv_volume.as<double>();
  r.face_inflow_slot = op.face_inflow.as<double>();
  r.face_basis_value = op.face_basis_value.as<double>();
  return r;
}
}  // namespace

void Operator::moments(int prec, const void *psi, void *phi, int M, hipStream_t s) const {
  UBT_CHECK(M == 1 || M == 1 + disc->dim, UBT_ERR_ARG_OUTOFRANGE, "n_moments must be 1 or 1 + dim, got %d", M);
  UBT_CHECK(psi && phi, UBT_ERR_ARG_NULL, "psi and phi are required");
  if (prec == UBT_FP64)
    launch_moments<double, double>(*this, t64, (const double *)psi, (double *)phi, M, M, s);
  else
    launch_moments<float, float>(*this, t32, (const float *)psi, (float *)phi, M, M, s);
}

void Operator::moments_f64(int prec, const void *psi, double *phi, hipStream_t s, bool unit) const {
  UBT_CHECK(psi && phi, UBT_ERR_ARG_NULL, "psi and phi are required");
  if (prec == UBT_FP64)
    launch_moments<double, double>(*this, t64, (const double *)psi, phi, 1, 1, s, unit);
  else
    launch_moments<float, double>(*this, t32, (const float *)psi, phi, 1, 1, s, unit);
}

void Operator::set_scalar_flux(int g, int prec, const void *psi, hipStream_t s) {
  UBT_CHECK(g >= 0 && g < n_groups, UBT_ERR_ARG_OUTOFRANGE, "group %d out of range, n_groups is %d", g, n_groups);
  double *dst = phi_cache.as<double>() + (int64_t)g * disc->n_nodes();
  if (prec == UBT_FP64)
    launch_moments<double, double>(*this, t64, (const double *)psi, dst, 1, 1, s);
  else
    launch_moments<float, double>(*this, t32, (const float *)psi, dst, 1, 1, s);
  phi_set[g] = 1;
}

void Operator::scatter_source(int gf, int gt, int prec, void *b, hipStream_t s) const {
  UBT_CHECK(xs_set, UBT_ERR_ARG_WRONGSTATE, "set the cross sections first");
  UBT_CHECK(gf >= 0 && gf < n_groups && gt >= 0 && gt < n_groups, UBT_ERR_ARG_OUTOFRANGE,
            "group pair (%d, %d) out of range, n_groups is %d", gf, gt, n_groups);
  UBT_CHECK(phi_set[gf], UBT_ERR_ARG_WRONGSTATE,
            "no scalar flux cached for group %d - set it when that group is solved", gf);
  // one transfer (the self-transfer g -> g is legal): the rhs kernel in accumulate mode
  RhsArgs r = make_rhs_args(*this, gt, gf, gf + 1, -1);
  const int64_t rows = disc->n_rows();
  const unsigned grid = (unsigned)((rows + 255) / 256);
  // the kernel indexes sigs[(gf' * G + g) * cells] with gf' = 0 after the shift and skips gf' == g;
  // a transfer into group 0 from gf > 0 would be skipped, so pass g relative to the shift
  if (prec == UBT_FP64)
    hipLaunchKernelGGL(rhs_kernel<double>, dim3(grid), dim3(256), 0, s, (double *)b, r, 1);
  else
    hipLaunchKernelGGL(rhs_kernel<float>, dim3(grid), dim3(256), 0, s, (float *)b, r, 1);
  UBT_HIP(hipGetLastError());
}

void Operator::rhs(int g, int prec, void *b, hipStream_t s) const {
  UBT_CHECK(xs_set, UBT_ERR_ARG_WRONGSTATE, "set the cross sections first");
  UBT_CHECK(g >= 0 && g < n_groups, UBT_ERR_ARG_OUTOFRANGE, "group %d out of range, n_groups is %d", g, n_groups);
  for (int gf = 0; gf < g; gf++)
    UBT_CHECK(phi_set[gf], UBT_ERR_ARG_WRONGSTATE,
              "no scalar flux cached for group %d - set it when that group is solved", gf);
  RhsArgs r = make_rhs_args(*this, g, 0, g, g);
  const int64_t rows = disc->n_rows();
  const unsigned grid = (unsigned)((rows + 255) / 256);
  if (prec == UBT_FP64)
    hipLaunchKernelGGL(rhs_kernel<double>, dim3(grid), dim3(256), 0, s, (double *)b, r, 0);
  else
    hipLaunchKernelGGL(rhs_kernel<float>, dim3(grid), dim3(256), 0, s, (float *)b, r, 0);
  UBT_HIP(hipGetLastError());
}

// ---------------------------------------------------------------------------
// Block inverse
// ---------------------------------------------------------------------------
void BlockInverse::setup(hipStream_t s) {
  const Operator &o = *op;
  UBT_CHECK(o.xs_set, UBT_ERR_ARG_WRONGSTATE, "set the cross sections first");
  const Disc &D = *o.disc;
  const int64_t nblk = (int64_t)D.n_cells * D.n_angles;
  const unsigned grid = (unsigned)((nblk + 255) / 256);
  int *cnt = o.flag.as<int>();
  UBT_HIP(hipMemsetAsync(cnt, 0, sizeof(int), s));
  const int fd = D.kind == DISC_FD;
  switch (D.nb) {
    case 1: hipLaunchKernelGGL(blockinv_check<1>, dim3(grid), dim3(256), 0, s, o.args<double>(), fd, cnt); break;
    case 3: hipLaunchKernelGGL(blockinv_check<3>, dim3(grid), dim3(256), 0, s, o.args<double>(), fd, cnt); break;
    case 4: hipLaunchKernelGGL(blockinv_check<4>, dim3(grid), dim3(256), 0, s, o.args<double>(), fd, cnt); break;
    default: fail(UBT_ERR_SUP, "element blocks of %d basis functions are not supported", D.nb);
  }
  UBT_HIP(hipGetLastError());
  int h = 0;
  UBT_HIP(hipMemcpyAsync(&h, cnt, sizeof(int), hipMemcpyDeviceToHost, s));
  UBT_HIP(hipStreamSynchronize(s));
  UBT_CHECK(h == 0, UBT_ERR_MAT_LU_ZRPVT, "%d of %lld element blocks are singular", h, (long long)nblk);
  group_ready = o.group;
}

void BlockInverse::apply(int prec, const void *x, void *y, hipStream_t s, bool transpose) const {
  const Operator &o = *op;
  UBT_CHECK(group_ready == o.group, UBT_ERR_ARG_WRONGSTATE,
            "the block inverse was not set up for the operator's current group %d", o.group);
  UBT_CHECK(x && y, UBT_ERR_ARG_NULL, "x and y are required");
  UBT_CHECK(x != y, UBT_ERR_ARG_IDN, "the block inverse apply needs distinct x and y");
  const Disc &D = *o.disc;
  const int64_t nblk = (int64_t)D.n_cells * D.n_angles;
  const unsigned grid = (unsigned)((nblk + 255) / 256);
  const int fd = D.kind == DISC_FD;
#define BI(NB)                                                                                               \
  if (prec == UBT_FP64)                                                                                      \
    hipLaunchKernelGGL((blockinv_apply<double, NB>), dim3(grid), dim3(256), 0, s, (const double *)x, (double *)y, \
                       o.args<double>(), fd, transpose ? 1 : 0);                                             \
  else                                                                                                       \
    hipLaunchKernelGGL((blockinv_apply<float, NB>), dim3(grid), dim3(256), 0, s, (const float *)x, (float *)y,   \
                       o.args<float>(), fd, transpose ? 1 : 0);
  switch (D.nb) {
    case 1: BI(1) break;
    case 3: BI(3) break;
    case 4: BI(4) break;
    default: fail(UBT_ERR_SUP, "element blocks of %d basis functions are not supported", D.nb);
  }
#undef BI
  UBT_HIP(hipGetLastError());
}

}  // namespace ubt
