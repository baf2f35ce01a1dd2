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

#include <algorithm>
#include <cmath>
#include <type_traits>
#include <unordered_map>

#include "batch.hpp"
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

// A load through a pointer the compiler cannot place (one read out of device
// memory, as a batch's work items hold them) is a flat load, which also
// counts against the LDS counter: every barrier's wait for the LDS then waits
// for it too, and no such load stays in flight across a barrier. The tables
// behind the operator's pointers are global memory; ldg loads through the
// global address space.
template <class T> __device__ __forceinline__ T ldg(const T *p) {
  return *(const __attribute__((address_space(1))) T *)p;
}

// the geometry class of cell c: with IDC, the cell itself when every cell is
// its own class (one dependent load fewer on a mesh with no repeated
// geometry). The test is a branch around the load, and a load under a branch
// makes the compiler's waits for the other loads in flight conservative, so
// kernels for box meshes (whose geometry always repeats) leave IDC off: in the
// batched box tile the branch cost 30%.
template <bool IDC = true, class T> __device__ __forceinline__ int64_t cls_of(const OpArgs<T> &g, int64_t c) {
  return IDC && g.gcls_identity ? c : (int64_t)g.gcls[c];
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
// The grid indices of cell c (a grid holds fewer than 2^31 cells, so 32-bit
// division)
template <int DIM>
__device__ __forceinline__ void fd_index(const int n[3], int64_t c, int idx[3]) {
  const unsigned cu = (unsigned)c, n0 = (unsigned)n[0], n1 = (unsigned)n[1];
  const unsigned q = cu / n0;
  idx[0] = (int)(cu - q * n0);
  idx[1] = DIM == 3 ? (int)(q % n1) : (int)q;
  idx[2] = DIM == 3 ? (int)(q / n1) : 0;
}

// Streaming + removal of row (c, a): every load is issued before any is
// used - the upwind neighbour's index is clamped to the cell itself where
// there is none and its weight is then zero - so the own trace and the DIM
// upwind traces are in flight together (a load under a branch, or used right
// after it, would serialise them). A reflective boundary face, whose trace is
// the cell's own reflected angle, takes one more load on a branch only
// boundary lanes enter.
template <class T, int DIM, bool ADJ>
__device__ __forceinline__ T fd_row(const T *__restrict__ x, const OpArgs<T> &g, int64_t c, const int idx[3], int a,
                                    T &psi) {
  const int A = g.A;
  const int64_t stride[3] = {1, g.n[0], (int64_t)g.n[0] * g.n[1]};
  T cd[DIM], nbv[DIM];
  bool refl[DIM];
  psi = x[c * A + a];
  const T sig = g.sigt[c];
#pragma unroll
  for (int d = 0; d < DIM; d++) {
    const T cs = ADJ ? -g.omega[3 * a + d] : g.omega[3 * a + d];
    const int up = cs > T(0) ? idx[d] - 1 : idx[d] + 1;
    const bool in = cs != T(0) && up >= 0 && up < g.n[d];
    refl[d] = cs != T(0) && !in && g.face_type[d][cs > T(0) ? 0 : 1] == kReflect;
    cd[d] = fabs(cs) * g.invh[d];
    nbv[d] = x[(in ? c + (cs > T(0) ? -stride[d] : stride[d]) : c) * A + a];
    if (!in) nbv[d] = T(0);
  }
  T acc = sig * psi;
#pragma unroll
  for (int d = 0; d < DIM; d++) acc += cd[d] * psi - cd[d] * nbv[d];
#pragma unroll
  for (int d = 0; d < DIM; d++)
    if (refl[d]) acc -= cd[d] * x[c * A + g.reflect[3 * a + d]];
  return acc;
}

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
  // every thread computes on a valid cell; the dead ones store nothing
  const int64_t cl = live ? c : tile * g.cpb;
  int idx[3];
  fd_index<DIM>(g.n, cl, idx);
  // the scatter's cross section too is loaded before the barriers
  const T ssg = SCATTER ? g.sigs[cl] : T(0);
  T psi;
  T acc = fd_row<T, DIM, ADJ>(x, g, cl, idx, a, psi);
  if (SCATTER) {
    if (lc < g.cpb) red[lc * A + a] = live ? (ADJ ? psi : g.w[a] * psi) : T(0);
    __syncthreads();
    seg_sum(red, g.cpb, A, phi, g.red_lanes);
    __syncthreads();
    if (live) acc -= ssg * g.inv_sum_w * (ADJ ? g.w[a] : T(1)) * phi[lc];
  }
  if (live) y[c * A + a] = acc;
}

// Multi-cell structured apply: a lane column owns CPT consecutive cells (one
// angle per thread), so each workgroup moves CPT times the data of the tile
// kernel per barrier pair, and every load - the CPT own traces, their y (and
// z) upwind traces, the one x halo beyond the run - is in flight together.
// The x upwind trace of every cell but the run's first (last, for a negative
// cosine) is the neighbouring cell's own trace, already in registers.
template <class T, int DIM, int CPT, bool SCATTER>
__global__ void __launch_bounds__(256) fd_apply_mc(const T *__restrict__ x, T *__restrict__ y, OpArgs<T> g) {
  extern __shared__ __align__(16) unsigned char smem[];
  const int A = g.A, R = g.cpb;
  const int64_t per = (int64_t)R * CPT;
  T *red = reinterpret_cast<T *>(smem);
  T *phi = red + per * A;
  const int64_t tile = xcd_tile(((int64_t)g.n_cells + per - 1) / per);
  if (tile * per >= g.n_cells) return;
  const int lc = threadIdx.x / A, a = threadIdx.x - lc * A;
  const bool lane = lc < R;
  const int64_t c0 = tile * per + (int64_t)(lane ? lc : 0) * CPT;
  const int64_t last = g.n_cells - 1;
  const int64_t stride[3] = {1, g.n[0], (int64_t)g.n[0] * g.n[1]};
  T cs[DIM], cd[DIM];
#pragma unroll
  for (int d = 0; d < DIM; d++) {
    cs[d] = g.omega[3 * a + d];
    cd[d] = fabs(cs[d]) * g.invh[d];
  }
  int idx0[3];
  fd_index<DIM>(g.n, c0 < last ? c0 : last, idx0);
  // the loads
  T p[CPT], st[CPT], ss[CPT], nb[DIM > 1 ? DIM - 1 : 1][CPT];
  bool live[CPT];
  int ix[CPT];
#pragma unroll
  for (int k = 0; k < CPT; k++) {
    const int64_t c = c0 + k < last ? c0 + k : last;
    live[k] = lane && c0 + k <= last;
    int idx[3];
    fd_index<DIM>(g.n, c, idx);
    ix[k] = idx[0];
    p[k] = x[c * A + a];
    st[k] = g.sigt[c];
    ss[k] = SCATTER ? g.sigs[c] : T(0);
#pragma unroll
    for (int d = 1; d < DIM; d++) {
      const int up = cs[d] > T(0) ? idx[d] - 1 : idx[d] + 1;
      const bool in = cs[d] != T(0) && up >= 0 && up < g.n[d];
      const T v = x[(in ? c + (cs[d] > T(0) ? -stride[d] : stride[d]) : c) * A + a];
      nb[d - 1][k] = in ? v : T(0);
    }
  }
  // the x halo: the cell before the run (positive cosine) or after it
  const bool pos = cs[0] > T(0);
  const int64_t ch = pos ? c0 - 1 : c0 + CPT;
  const T hx = x[(ch < 0 ? 0 : (ch > last ? last : ch)) * A + a];
  T acc[CPT];
#pragma unroll
  for (int k = 0; k < CPT; k++) {
    T v = st[k] * p[k];
#pragma unroll
    for (int d = 0; d < DIM; d++) v += cd[d] * p[k];
#pragma unroll
    for (int d = 1; d < DIM; d++) v -= cd[d] * nb[d - 1][k];
    // x upwind: inside the box unless the cell is on the inflow x face
    const bool xin = cs[0] != T(0) && (pos ? ix[k] > 0 : ix[k] < g.n[0] - 1);
    const T ux = pos ? (k > 0 ? p[k > 0 ? k - 1 : 0] : hx) : (k < CPT - 1 ? p[k < CPT - 1 ? k + 1 : 0] : hx);
    if (xin) v -= cd[0] * ux;
    acc[k] = v;
  }
  // reflective faces: the cell's own reflected angle, on a branch only
  // boundary lanes enter
#pragma unroll
  for (int k = 0; k < CPT; k++) {
    if (!live[k]) continue;
    int idx[3];
    fd_index<DIM>(g.n, c0 + k, idx);
#pragma unroll
    for (int d = 0; d < DIM; d++) {
      if (cs[d] == T(0)) continue;
      const int up = cs[d] > T(0) ? idx[d] - 1 : idx[d] + 1;
      if ((up < 0 || up >= g.n[d]) && g.face_type[d][cs[d] > T(0) ? 0 : 1] == kReflect)
        acc[k] -= cd[d] * x[(c0 + k) * A + g.reflect[3 * a + d]];
    }
  }
  if (SCATTER) {
    if (lane) {
      const T wa = g.w[a];
#pragma unroll
      for (int k = 0; k < CPT; k++) red[(lc * CPT + k) * A + a] = live[k] ? wa * p[k] : T(0);
    }
    __syncthreads();
    seg_sum(red, (int)per, A, phi, g.red_lanes);
    __syncthreads();
    if (lane) {
#pragma unroll
      for (int k = 0; k < CPT; k++) acc[k] -= ss[k] * g.inv_sum_w * phi[lc * CPT + k];
    }
  }
#pragma unroll
  for (int k = 0; k < CPT; k++)
    if (live[k]) y[(c0 + k) * A + a] = acc[k];
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
// LIST: the tile's cells are entries tile * cpb ... of g.cells (n_list of
// them) instead of a contiguous run of cell ids. IDC: see cls_of.
template <class T, int DIM, int NB, int NF, bool SCATTER, bool ADJ, bool LIST = false, bool IDC = true>
__device__ __forceinline__ void dg_tile_body(const T *__restrict__ x, T *__restrict__ y, const OpArgs<T> &g,
                                             int64_t tile) {
  extern __shared__ __align__(16) unsigned char smem[];
  constexpr int FM = NB * NB;
  const int cpb = g.cpb, A = g.A;
  const int64_t c0 = tile * cpb;
  const int64_t n_here = LIST ? (int64_t)g.n_list : (int64_t)g.n_cells;
  const int ncell = (int)(n_here - c0 < cpb ? n_here - c0 : cpb);
  auto cell_of = [&](int l) -> int64_t { return LIST ? (int64_t)g.cells[c0 + l] : c0 + l; };
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
    for (int e = tid; e < ncell * NF * FM; e += nt) {
      const int lc = e / (NF * FM);
      const int64_t ge = (int64_t)cls_of<IDC>(g, cell_of(lc)) * NF * FM + (e - lc * NF * FM);
      s_own[e] = g.face_own[ge];
      s_up[e] = g.face_up[ge];
    }
    for (int e = tid; e < ncell * (NB - 1) * DIM; e += nt) {
      const int lc = e / ((NB - 1) * DIM), r = e - lc * ((NB - 1) * DIM);
      const int i = 1 + r / DIM, d = r - (i - 1) * DIM;
      s_gr[e] = g.grad[((int64_t)cls_of<IDC>(g, cell_of(lc)) * NB + i) * 3 + d];
    }
  }
  for (int e = tid; e < ncell * NF; e += nt) {
    const int lc = e / NF;
    const int64_t ce = cell_of(lc);
    s_nb[e] = g.face_nb[ce * NF + (e - lc * NF)];
    const int64_t gk = (int64_t)cls_of<IDC>(g, ce) * NF + (e - lc * NF);
#pragma unroll
    for (int d = 0; d < DIM; d++) s_fs[e * DIM + d] = g.face_s[gk * 3 + d];
  }
  for (int e = tid; e < ncell; e += nt) {
    const int64_t ce = cell_of(e);
    s_st[e] = g.sigt[ce];
    s_ss[e] = g.sigs[ce];
  }
  __syncthreads();

  const int lc = tid / A, a = tid - lc * A;
  const bool live = lc < ncell;
  const int64_t c = live ? cell_of(lc) : 0;
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
        const T vr = (ADJ && nbk >= 0) ? g.vol[cls_of<IDC>(g, c)] / g.vol[cls_of<IDC>(g, nbk)] : T(1);
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

template <class T, int DIM, int NB, int NF, bool SCATTER, bool ADJ = false>
__global__ void __launch_bounds__(1024) dg_apply(const T *__restrict__ x, T *__restrict__ y, OpArgs<T> g) {
  const int64_t n_tiles = ((int64_t)g.n_cells + g.cpb - 1) / g.cpb;
  const int64_t tile = xcd_tile(n_tiles);
  if (tile >= n_tiles) return;
  dg_tile_body<T, DIM, NB, NF, SCATTER, ADJ>(x, y, g, tile);
}

// ---------------------------------------------------------------------------
// Multi-cell tile apply: the tile above with CPT cells per thread (thread =
// angle a of cells r, r + R, ..., r + (CPT-1) R, R = blockDim / A), so one
// workgroup carries CPT times the work through the same latency chain, and
// every global load is issued before the first barrier: the own psi, and the
// upwind traces chosen by the inflow test on the class tables (face_nb, the
// geometry class and its normals are small and cache-resident). Only the face
// matrices and gradients go through LDS. Same arithmetic as dg_apply.
// ---------------------------------------------------------------------------
template <class T, int DIM, int NB, int NF, int CPT, bool SCATTER>
__global__ void __launch_bounds__(256) dg_apply_mc(const T *__restrict__ x, T *__restrict__ y, OpArgs<T> g) {
  extern __shared__ __align__(16) unsigned char smem[];
  constexpr int FM = NB * NB;
  const int R = g.cpb, A = g.A, tcells = R * CPT;
  const int64_t n_tiles = ((int64_t)g.n_cells + tcells - 1) / tcells;
  const int64_t tile = xcd_tile(n_tiles);
  if (tile >= n_tiles) return;
  const int64_t c0 = tile * tcells;
  const int ncell = (int)((int64_t)g.n_cells - c0 < tcells ? (int64_t)g.n_cells - c0 : tcells);
  T *s_own = reinterpret_cast<T *>(smem);
  T *s_up = s_own + (NB > 1 ? tcells * NF * FM : 0);
  T *s_gr = s_up + (NB > 1 ? tcells * NF * FM : 0);
  T *red = s_gr + (NB > 1 ? tcells * (NB - 1) * DIM : 0);
  T *phi = red + (SCATTER ? tcells * NB * A : 0);
  const int tid = threadIdx.x, nt = blockDim.x;
  const int r = tid / A, a = tid - r * A;
  const bool lane = r < R;
  const int ac = lane ? a : 0;
  T om[DIM];
#pragma unroll
  for (int d = 0; d < DIM; d++) om[d] = g.omega[3 * ac + d];

  // 1. global loads: own psi, cross sections, inflow test, upwind traces
  T psi[CPT][NB], pu[CPT][NF][NB], sf[CPT][NF], sig[CPT], ssg[CPT];
#pragma unroll
  for (int k = 0; k < CPT; k++) {
    const int lc = r + k * R;
    const bool live = lane && lc < ncell;
    const int64_t c = c0 + (live ? lc : 0);
    const int64_t base = c * NB * A + ac;
#pragma unroll
    for (int j = 0; j < NB; j++) psi[k][j] = live ? x[base + (int64_t)j * A] : T(0);
    sig[k] = g.sigt[c];
    ssg[k] = SCATTER ? g.sigs[c] : T(0);
    const int64_t gc = cls_of(g, c);
#pragma unroll
    for (int f = 0; f < NF; f++) {
      const int nbk = g.face_nb[c * NF + f];
      T sv = T(0);
#pragma unroll
      for (int d = 0; d < DIM; d++) sv += om[d] * g.face_s[(gc * NF + f) * 3 + d];
      sf[k][f] = sv;
      const bool up = live && !(sv > T(0)) && nbk != -1;
      int64_t ub = base;
      if (up) ub = nbk >= 0 ? (int64_t)nbk * NB * A + ac : c * NB * A + g.reflect[3 * ac + (-2 - nbk)];
#pragma unroll
      for (int j = 0; j < NB; j++) pu[k][f][j] = up ? x[ub + (int64_t)j * A] : T(0);
    }
  }
  // 2. stage the face matrices and gradients of the tile's cells
  if (NB > 1) {
    for (int e = tid; e < ncell * NF * FM; e += nt) {
      const int lc = e / (NF * FM);
      const int64_t ge = (int64_t)cls_of(g, c0 + lc) * NF * FM + (e - lc * NF * FM);
      s_own[e] = g.face_own[ge];
      s_up[e] = g.face_up[ge];
    }
    for (int e = tid; e < ncell * (NB - 1) * DIM; e += nt) {
      const int lc = e / ((NB - 1) * DIM), rr = e - lc * ((NB - 1) * DIM);
      const int i = 1 + rr / DIM, d = rr - (i - 1) * DIM;
      s_gr[e] = g.grad[((int64_t)cls_of(g, c0 + lc) * NB + i) * 3 + d];
    }
  }
  if (SCATTER) {
#pragma unroll
    for (int k = 0; k < CPT; k++) {
      const int lc = r + k * R;
      if (lane) {
        const T wa = lc < ncell ? g.w[a] : T(0);
#pragma unroll
        for (int j = 0; j < NB; j++) red[(lc * NB + j) * A + a] = wa * psi[k][j];
      }
    }
  }
  __syncthreads();
  if (SCATTER) {
    seg_sum(red, tcells * NB, A, phi, g.red_lanes);
    __syncthreads();
  }
  // 3. compute and store
#pragma unroll
  for (int k = 0; k < CPT; k++) {
    const int lc = r + k * R;
    if (!lane || lc >= ncell) continue;
    const int64_t c = c0 + lc;
    T acc[NB];
#pragma unroll
    for (int i = 0; i < NB; i++) acc[i] = sig[k] * psi[k][i];
    if (NB > 1) {
#pragma unroll
      for (int i = 1; i < NB; i++) {
        const T *gr = s_gr + (lc * (NB - 1) + i - 1) * DIM;
        T od = T(0);
#pragma unroll
        for (int d = 0; d < DIM; d++) od += om[d] * gr[d];
        acc[i] -= od * psi[k][0];
      }
    }
#pragma unroll
    for (int f = 0; f < NF; f++) {
      const T sv = sf[k][f];
      const bool out = sv > T(0);
      if (NB == 1) {
        acc[0] += sv * (out ? psi[k][0] : pu[k][f][0]);
      } else {
        const T *F = (out ? s_own : s_up) + (lc * NF + f) * FM;
#pragma unroll
        for (int i = 0; i < NB; i++) {
          T v = T(0);
#pragma unroll
          for (int j = 0; j < NB; j++) v += F[i * NB + j] * (out ? psi[k][j] : pu[k][f][j]);
          acc[i] += sv * v;
        }
      }
    }
    if (SCATTER) {
      const T ss = ssg[k] * g.inv_sum_w;
#pragma unroll
      for (int i = 0; i < NB; i++) acc[i] -= ss * phi[lc * NB + i];
    }
    const int64_t base = c * NB * A + a;
#pragma unroll
    for (int i = 0; i < NB; i++) y[base + (int64_t)i * A] = acc[i];
  }
}

// ---------------------------------------------------------------------------
// Four-angle tile apply: thread = (cell, four consecutive angles), so a
// thread moves 16-byte vectors (FP32; two 16-byte halves in FP64) and every
// per-cell cost - neighbour ids, normals, face matrices from LDS, address
// arithmetic - is paid once per four angles. The upwind choice is per angle
// and branch-free: both the outflow product F_own psi and the inflow product
// F_up psi_nb are formed and the face flux picks one. Needs A % 4 == 0 (every
// level-symmetric set). Same arithmetic as dg_apply, geometry staged the same
// way.
// ---------------------------------------------------------------------------
template <class T> struct Vec4 { T v[4]; };
template <class T> __device__ inline Vec4<T> ld4(const T *__restrict__ p) {
  Vec4<T> r;
  if constexpr (sizeof(T) == 4) {
    const float4 q = *reinterpret_cast<const float4 *>(p);
    r.v[0] = q.x; r.v[1] = q.y; r.v[2] = q.z; r.v[3] = q.w;
  } else {
    const double2 q0 = *reinterpret_cast<const double2 *>(p), q1 = *reinterpret_cast<const double2 *>(p + 2);
    r.v[0] = q0.x; r.v[1] = q0.y; r.v[2] = q1.x; r.v[3] = q1.y;
  }
  return r;
}
template <class T> __device__ inline void st4(T *__restrict__ p, const T v[4]) {
  if constexpr (sizeof(T) == 4) {
    *reinterpret_cast<float4 *>(p) = make_float4(v[0], v[1], v[2], v[3]);
  } else {
    *reinterpret_cast<double2 *>(p) = make_double2(v[0], v[1]);
    *reinterpret_cast<double2 *>(p + 2) = make_double2(v[2], v[3]);
  }
}

template <class T, int DIM, int NB, int NF, bool SCATTER>
__device__ __forceinline__ void dg_q4_body(const T *__restrict__ x, T *__restrict__ y, const OpArgs<T> &g,
                                           int64_t tile) {
  extern __shared__ __align__(16) unsigned char smem[];
  constexpr int FM = NB * NB;
  const int cpb = g.cpb, A = g.A, TPC = A >> 2;
  const int64_t c0 = tile * cpb;
  const int ncell = (int)((int64_t)g.n_cells - c0 < cpb ? (int64_t)g.n_cells - c0 : cpb);
  T *s_own = reinterpret_cast<T *>(smem);
  T *s_up = s_own + (NB > 1 ? cpb * NF * FM : 0);
  T *s_fs = s_up + (NB > 1 ? cpb * NF * FM : 0);
  T *s_gr = s_fs + cpb * NF * DIM;
  T *s_st = s_gr + (NB > 1 ? cpb * (NB - 1) * DIM : 0);
  T *s_ss = s_st + cpb;
  T *red = s_ss + cpb;
  T *phi = red + (SCATTER ? cpb * NB * TPC : 0);
  int *s_nb = reinterpret_cast<int *>(phi + (SCATTER ? cpb * NB : 0));
  const int tid = threadIdx.x, nt = blockDim.x;
  const int lc = tid / TPC, q = tid - lc * TPC, a0 = 4 * q;
  const bool live = lc < ncell;
  const int64_t c = c0 + (live ? lc : 0);
  // own psi first: its loads overlap the staging below
  Vec4<T> psi[NB];
#pragma unroll
  for (int j = 0; j < NB; j++) {
    if (live) psi[j] = ld4(x + (c * NB + j) * A + a0);
    else
#pragma unroll
      for (int t = 0; t < 4; t++) psi[j].v[t] = T(0);
  }
  if (NB > 1) {
    for (int e = tid; e < ncell * NF * FM; e += nt) {
      const int l = e / (NF * FM);
      const int64_t ge = (int64_t)cls_of(g, c0 + l) * NF * FM + (e - l * NF * FM);
      s_own[e] = g.face_own[ge];
      s_up[e] = g.face_up[ge];
    }
    for (int e = tid; e < ncell * (NB - 1) * DIM; e += nt) {
      const int l = e / ((NB - 1) * DIM), r = e - l * ((NB - 1) * DIM);
      const int i = 1 + r / DIM, d = r - (i - 1) * DIM;
      s_gr[e] = g.grad[((int64_t)cls_of(g, c0 + l) * NB + i) * 3 + d];
    }
  }
  for (int e = tid; e < ncell * NF; e += nt) {
    s_nb[e] = g.face_nb[c0 * NF + e];
    const int l = e / NF;
    const int64_t gk = (int64_t)cls_of(g, c0 + l) * NF + (e - l * NF);
#pragma unroll
    for (int d = 0; d < DIM; d++) s_fs[e * DIM + d] = g.face_s[gk * 3 + d];
  }
  for (int e = tid; e < ncell; e += nt) {
    s_st[e] = g.sigt[c0 + e];
    s_ss[e] = g.sigs[c0 + e];
  }
  T om[4][DIM];
#pragma unroll
  for (int t = 0; t < 4; t++)
#pragma unroll
    for (int d = 0; d < DIM; d++) om[t][d] = live ? g.omega[3 * (a0 + t) + d] : T(0);
  if (SCATTER && live) {
#pragma unroll
    for (int j = 0; j < NB; j++) {
      T p = T(0);
#pragma unroll
      for (int t = 0; t < 4; t++) p += g.w[a0 + t] * psi[j].v[t];
      red[(lc * NB + j) * TPC + q] = p;
    }
  }
  __syncthreads();
  T acc[NB][4];
  if (live) {
    const T sig = s_st[lc];
#pragma unroll
    for (int j = 0; j < NB; j++)
#pragma unroll
      for (int t = 0; t < 4; t++) acc[j][t] = sig * psi[j].v[t];
    if (NB > 1) {
#pragma unroll
      for (int i = 1; i < NB; i++) {
        const T *gr = s_gr + (lc * (NB - 1) + i - 1) * DIM;
#pragma unroll
        for (int t = 0; t < 4; t++) {
          T od = T(0);
#pragma unroll
          for (int d = 0; d < DIM; d++) od += om[t][d] * gr[d];
          acc[i][t] -= od * psi[0].v[t];
        }
      }
    }
#pragma unroll
    for (int f = 0; f < NF; f++) {
      const int kl = lc * NF + f;
      const int nbk = s_nb[kl];
      T sv[4];
#pragma unroll
      for (int t = 0; t < 4; t++) {
        sv[t] = T(0);
#pragma unroll
        for (int d = 0; d < DIM; d++) sv[t] += om[t][d] * s_fs[kl * DIM + d];
      }
      Vec4<T> nbv[NB];
      if (nbk >= 0) {
#pragma unroll
        for (int j = 0; j < NB; j++) nbv[j] = ld4(x + ((int64_t)nbk * NB + j) * A + a0);
      } else if (nbk <= -2) {  // reflective: this cell, mirrored angles
#pragma unroll
        for (int t = 0; t < 4; t++) {
          const int ra = g.reflect[3 * (a0 + t) + (-2 - nbk)];
#pragma unroll
          for (int j = 0; j < NB; j++) nbv[j].v[t] = x[(c * NB + j) * A + ra];
        }
      } else {  // vacuum: the inflow lives in the rhs
#pragma unroll
        for (int j = 0; j < NB; j++)
#pragma unroll
          for (int t = 0; t < 4; t++) nbv[j].v[t] = T(0);
      }
      if (NB == 1) {
#pragma unroll
        for (int t = 0; t < 4; t++) acc[0][t] += sv[t] * (sv[t] > T(0) ? psi[0].v[t] : nbv[0].v[t]);
      } else {
        const T *Fo = s_own + kl * FM, *Fu = s_up + kl * FM;
#pragma unroll
        for (int i = 0; i < NB; i++) {
          T fo[NB], fu[NB];
#pragma unroll
          for (int j = 0; j < NB; j++) {
            fo[j] = Fo[i * NB + j];
            fu[j] = Fu[i * NB + j];
          }
#pragma unroll
          for (int t = 0; t < 4; t++) {
            T vo = T(0), vu = T(0);
#pragma unroll
            for (int j = 0; j < NB; j++) {
              vo += fo[j] * psi[j].v[t];
              vu += fu[j] * nbv[j].v[t];
            }
            acc[i][t] += sv[t] * (sv[t] > T(0) ? vo : vu);
          }
        }
      }
    }
  }
  if (SCATTER) {
    // the moment of each (cell, basis): sum over the cell's TPC partials
    for (int e = tid; e < ncell * NB; e += nt) {
      T p = T(0);
      for (int k = 0; k < TPC; k++) p += red[e * TPC + k];
      phi[e] = p;
    }
    __syncthreads();
    if (live) {
      const T ss = s_ss[lc] * g.inv_sum_w;
#pragma unroll
      for (int j = 0; j < NB; j++)
#pragma unroll
        for (int t = 0; t < 4; t++) acc[j][t] -= ss * phi[lc * NB + j];
    }
  }
  if (live) {
#pragma unroll
    for (int j = 0; j < NB; j++) st4(y + (c * NB + j) * A + a0, acc[j]);
  }
}

template <class T, int DIM, int NB, int NF, bool SCATTER>
__global__ void __launch_bounds__(256) dg_apply_q4(const T *__restrict__ x, T *__restrict__ y, OpArgs<T> g) {
  const int64_t n_tiles = ((int64_t)g.n_cells + g.cpb - 1) / g.cpb;
  const int64_t tile = xcd_tile(n_tiles);
  if (tile >= n_tiles) return;
  dg_q4_body<T, DIM, NB, NF, SCATTER>(x, y, g, tile);
}

// ---------------------------------------------------------------------------
// Pencil apply for DG on a structured box (cells numbered lexicographically,
// x fastest, every cell's local faces in one axis/side order; detected on the
// host). A workgroup owns W consecutive cells along x (thread = (cell, angle),
// all NB rows) and marches along the last axis (y in 2D, z in 3D) over H
// cells. The upwind trace along the march axis comes from the thread's own
// register window (rows m-1, m, m+1 resident, m+2 in flight), along x from
// the row buffer in LDS (segment ends: a halo load issued one step ahead),
// and in 3D along y from one global load per step, also issued one step
// ahead. Every geometry class is staged in LDS once. So a DoF costs one
// global read (two in 3D) instead of one per face, and the loads of the next
// step overlap the current step's barriers.
// ---------------------------------------------------------------------------
struct PencilArgs {
  int n[3];        // cells per axis
  int W, H;        // cells per segment along x, steps along the march axis
  int fmap[3][2];  // local face of (axis, side)
  int ncls;        // geometry classes (all staged)
  int segs[3];     // workgroup grid: x segments, y rows (3D), march chunks
};

template <class T, int DIM, int NB>
__global__ void __launch_bounds__(256) dg_apply_pencil(const T *__restrict__ x, T *__restrict__ y, OpArgs<T> g,
                                                       PencilArgs pa, int scatter) {
  extern __shared__ __align__(16) unsigned char smem[];
  constexpr int NF = 2 * DIM, FM = NB * NB, MA = DIM - 1;  // march axis
  const int A = g.A, W = pa.W, ncls = pa.ncls;
  const int nx = pa.n[0], ny = pa.n[1], nm = pa.n[MA];
  const int64_t nwg = (int64_t)pa.segs[0] * pa.segs[1] * pa.segs[2];
  const int64_t wg = xcd_tile(nwg);
  if (wg >= nwg) return;
  const int sx = (int)(wg % pa.segs[0]);
  const int64_t rest = wg / pa.segs[0];
  const int jy = DIM == 3 ? (int)(rest % pa.segs[1]) : 0;  // the fixed y row in 3D
  const int sm = (int)(DIM == 3 ? rest / pa.segs[1] : rest);
  const int i0 = sx * W, m0 = sm * pa.H, m1 = min(nm, m0 + pa.H);
  const int64_t sstride = DIM == 2 ? (int64_t)nx : (int64_t)nx * ny;  // march stride
  // LDS: class tables, two row buffers, the weighted row, the moments
  T *c_fs = reinterpret_cast<T *>(smem);
  T *c_own = c_fs + ncls * NF * DIM;
  T *c_up = c_own + (NB > 1 ? ncls * NF * FM : 0);
  T *c_gr = c_up + (NB > 1 ? ncls * NF * FM : 0);
  T *rowb = c_gr + (NB > 1 ? ncls * (NB - 1) * DIM : 0);
  T *red = rowb + 2 * W * NB * A;
  T *phi = red + W * NB * A;
  const int tid = threadIdx.x, nt = blockDim.x;
  for (int e = tid; e < ncls * NF * DIM; e += nt) {
    const int cf = e / DIM, d = e - cf * DIM;
    c_fs[e] = g.face_s[(int64_t)cf * 3 + d];
  }
  if (NB > 1) {
    for (int e = tid; e < ncls * NF * FM; e += nt) {
      c_own[e] = g.face_own[e];
      c_up[e] = g.face_up[e];
    }
    for (int e = tid; e < ncls * (NB - 1) * DIM; e += nt) {
      const int cl = e / ((NB - 1) * DIM), r = e - cl * ((NB - 1) * DIM);
      const int i = 1 + r / DIM, d = r - (i - 1) * DIM;
      c_gr[e] = g.grad[((int64_t)cl * NB + i) * 3 + d];
    }
  }
  const int lc = tid / A, a = tid - lc * A;
  const bool lane = lc < W;
  const int ac = lane ? a : 0;
  const int i = i0 + lc;
  const bool col = lane && i < nx;
  T om[DIM];
#pragma unroll
  for (int d = 0; d < DIM; d++) om[d] = g.omega[3 * ac + d];
  const T wa = col ? g.w[ac] : T(0);
  // the x halo this lane reads from global (segment end, inflow side), else -1
  const bool hx_lo = col && lc == 0 && i > 0 && om[0] > T(0);
  const bool hx_hi = col && (lc == W - 1 || i == nx - 1) && i + 1 < nx && om[0] < T(0);
  const int64_t colbase = (int64_t)(DIM == 3 ? (int64_t)jy * nx : 0) + i;  // cell at march index 0
  auto cell_at = [&](int m) { return colbase + (int64_t)m * sstride; };
  auto load_row = [&](int m, T v[NB]) {
#pragma unroll
    for (int k = 0; k < NB; k++) v[k] = T(0);
    if (col && m >= 0 && m < nm) {
      const int64_t b = cell_at(m) * NB * A + ac;
#pragma unroll
      for (int k = 0; k < NB; k++) v[k] = x[b + (int64_t)k * A];
    }
  };
  // 3D: the y-upwind neighbour (one side per lane by the sign of Omega_y)
  const int ydir = DIM == 3 ? (om[1] > T(0) ? -1 : 1) : 0;
  const bool yin = DIM == 3 && col && jy + ydir >= 0 && jy + ydir < ny;
  auto load_y = [&](int m, T v[NB]) {
#pragma unroll
    for (int k = 0; k < NB; k++) v[k] = T(0);
    if (yin && m < nm) {
      const int64_t b = (cell_at(m) + (int64_t)ydir * nx) * NB * A + ac;
#pragma unroll
      for (int k = 0; k < NB; k++) v[k] = x[b + (int64_t)k * A];
    }
  };
  auto load_hx = [&](int m, T v[NB]) {
#pragma unroll
    for (int k = 0; k < NB; k++) v[k] = T(0);
    if ((hx_lo || hx_hi) && m < nm) {
      const int64_t b = (cell_at(m) + (hx_lo ? -1 : 1)) * NB * A + ac;
#pragma unroll
      for (int k = 0; k < NB; k++) v[k] = x[b + (int64_t)k * A];
    }
  };
  T pm[NB], p0[NB], pp[NB], pn[NB], hy[NB], hyn[NB], hx[NB], hxn[NB];
  load_row(m0 - 1, pm);
  load_row(m0, p0);
  load_row(m0 + 1, pp);
  load_y(m0, hy);
  load_hx(m0, hx);
  int gc = col && m0 < nm ? cls_of<false>(g, cell_at(m0)) : 0;
  T st = col && m0 < nm ? g.sigt[cell_at(m0)] : T(0), ssg = col && m0 < nm ? g.sigs[cell_at(m0)] : T(0);
  __syncthreads();  // class tables staged
  for (int m = m0, par = 0; m < m1; m++, par ^= 1) {
    // next step's loads first
    load_row(m + 2, pn);
    load_y(m + 1, hyn);
    load_hx(m + 1, hxn);
    const bool nxt = col && m + 1 < nm;
    const int gcn = nxt ? cls_of<false>(g, cell_at(m + 1)) : 0;
    const T stn = nxt ? g.sigt[cell_at(m + 1)] : T(0), ssn = nxt ? g.sigs[cell_at(m + 1)] : T(0);
    const int64_t c = cell_at(m);
    int fnb[NF];
#pragma unroll
    for (int f = 0; f < NF; f++) fnb[f] = col ? g.face_nb[c * NF + f] : -1;
    T *rb = rowb + par * W * NB * A;
    if (lane) {
#pragma unroll
      for (int k = 0; k < NB; k++) {
        rb[(lc * NB + k) * A + a] = p0[k];
        red[(lc * NB + k) * A + a] = wa * p0[k];
      }
    }
    __syncthreads();
    if (scatter) seg_sum(red, W * NB, A, phi, g.red_lanes);
    __syncthreads();
    if (col) {
      T acc[NB];
#pragma unroll
      for (int k = 0; k < NB; k++) acc[k] = st * p0[k];
      if (NB > 1) {
#pragma unroll
        for (int k = 1; k < NB; k++) {
          const T *gr = c_gr + (gc * (NB - 1) + k - 1) * DIM;
          T od = T(0);
#pragma unroll
          for (int d = 0; d < DIM; d++) od += om[d] * gr[d];
          acc[k] -= od * p0[0];
        }
      }
#pragma unroll
      for (int ax = 0; ax < DIM; ax++) {
#pragma unroll
        for (int side = 0; side < 2; side++) {
          const int f = pa.fmap[ax][side];
          const T *fs = c_fs + (gc * NF + f) * DIM;
          T sv = T(0);
#pragma unroll
          for (int d = 0; d < DIM; d++) sv += om[d] * fs[d];
          T pu[NB];
          bool have = true;
          if (sv > T(0)) {
#pragma unroll
            for (int k = 0; k < NB; k++) pu[k] = p0[k];
          } else if (fnb[f] == -1) {
            have = false;  // vacuum: the inflow lives in the rhs
          } else if (fnb[f] <= -2) {  // reflective: this cell, mirrored angle (row buffer)
            const int ra = g.reflect[3 * a + (-2 - fnb[f])];
#pragma unroll
            for (int k = 0; k < NB; k++) pu[k] = rb[(lc * NB + k) * A + ra];
          } else if (ax == MA) {
#pragma unroll
            for (int k = 0; k < NB; k++) pu[k] = side == 0 ? pm[k] : pp[k];
          } else if (ax == 0) {
            const int l2 = side == 0 ? lc - 1 : lc + 1;
            if (l2 >= 0 && l2 < W) {
#pragma unroll
              for (int k = 0; k < NB; k++) pu[k] = rb[(l2 * NB + k) * A + a];
            } else {
#pragma unroll
              for (int k = 0; k < NB; k++) pu[k] = hx[k];
            }
          } else {
#pragma unroll
            for (int k = 0; k < NB; k++) pu[k] = hy[k];
          }
          if (!have) continue;
          if (NB == 1) {
            acc[0] += sv * pu[0];
          } else {
            const T *F = (sv > T(0) ? c_own : c_up) + (gc * NF + f) * FM;
#pragma unroll
            for (int r = 0; r < NB; r++) {
              T v = T(0);
#pragma unroll
              for (int k = 0; k < NB; k++) v += F[r * NB + k] * pu[k];
              acc[r] += sv * v;
            }
          }
        }
      }
      if (scatter) {
        const T ss = ssg * g.inv_sum_w;
#pragma unroll
        for (int k = 0; k < NB; k++) acc[k] -= ss * phi[lc * NB + k];
      }
      const int64_t b = c * NB * A + a;
#pragma unroll
      for (int k = 0; k < NB; k++) y[b + (int64_t)k * A] = acc[k];
    }
#pragma unroll
    for (int k = 0; k < NB; k++) {
      pm[k] = p0[k];
      p0[k] = pp[k];
      pp[k] = pn[k];
      hy[k] = hyn[k];
      hx[k] = hxn[k];
    }
    gc = gcn;
    st = stn;
    ssg = ssn;
  }
}

// ---------------------------------------------------------------------------
// Box-tile apply for DG1 on a uniform structured box (reference
// src/termsk.kokkos.cxx streaming/removal/scatter terms, as the tile kernel).
// Every interior cell of such a box has one geometry class, and with face
// normals exactly on the axes the upwind side of each axis is fixed by the
// sign of the ordinate's cosine alone. So a thread (column, angle) folds the
// interior class once into three small matrices for its angle:
//   M   = sum over outflow faces of s_f F_own,f  -  the gradient column,
//   U_x = s F_up of the x inflow face, U_y (and U_z) likewise,
// and an interior cell costs sigma psi + M psi + sum_axis U_axis psi_upwind
// minus the scatter: about NB (1 + DIM + 1) NB multiply-adds per angle, no
// branch on the face orientation, no class-table reads. Cells on the box
// boundary (or of another class) are skipped here and computed afterwards in
// the same launch by the tile kernel's body over their list, with vacuum and
// reflective faces; keeping that general path out of this loop keeps the
// register count, and so the waves per SIMD, where the interior needs it.
//
// A thread holds K rows of its column along the last axis in registers, so
// the upwind row is a register (the first or last row of the tile a halo
// load); the x neighbour is a neighbouring column through LDS (tile edges: a
// halo load), the y neighbour in 3D a global load. One LDS write, two
// barriers and one segmented sum per tile of W x K cells serve both the
// x exchange and the scatter's angular moments.
// ---------------------------------------------------------------------------
template <class T, int DIM, int NB> struct BoxConst {
  T om[DIM];
  T M[NB][NB];
  T U[DIM][NB][NB];
  int in_side[DIM];  // the inflow face of each axis: 0 (the - side) when the cosine is positive
};

template <class T, int DIM, int NB>
__device__ __forceinline__ void box_consts(const OpArgs<T> &g, const BoxArgs &bx, int a, BoxConst<T, DIM, NB> &k) {
  constexpr int NF = 2 * DIM, FM = NB * NB;
#pragma unroll
  for (int d = 0; d < DIM; d++) k.om[d] = g.omega[3 * a + d];
#pragma unroll
  for (int ax = 0; ax < DIM; ax++) k.in_side[ax] = k.om[ax] > T(0) ? 0 : 1;
#pragma unroll
  for (int r = 0; r < NB; r++)
#pragma unroll
    for (int c = 0; c < NB; c++) {
      k.M[r][c] = T(0);
#pragma unroll
      for (int ax = 0; ax < DIM; ax++) k.U[ax][r][c] = T(0);
    }
  const int64_t cls = bx.cls;
#pragma unroll
  for (int ax = 0; ax < DIM; ax++) {
#pragma unroll
    for (int side = 0; side < 2; side++) {
      const int f = bx.fmap[ax][side];
      const T *fs = g.face_s + (cls * NF + f) * 3;
      T s = T(0);
#pragma unroll
      for (int d = 0; d < DIM; d++) s += k.om[d] * fs[d];
      if (s > T(0)) {
        const T *F = g.face_own + (cls * NF + f) * FM;
#pragma unroll
        for (int r = 0; r < NB; r++)
#pragma unroll
          for (int c = 0; c < NB; c++) k.M[r][c] += s * F[r * NB + c];
      } else if (side == k.in_side[ax]) {
        const T *F = g.face_up + (cls * NF + f) * FM;
#pragma unroll
        for (int r = 0; r < NB; r++)
#pragma unroll
          for (int c = 0; c < NB; c++) k.U[ax][r][c] = s * F[r * NB + c];
      }
    }
  }
#pragma unroll
  for (int r = 1; r < NB; r++) {
    T od = T(0);
#pragma unroll
    for (int d = 0; d < DIM; d++) od += k.om[d] * g.grad[(cls * NB + r) * 3 + d];
    k.M[r][0] -= od;
  }
}

// the segmented sum of seg_sum with the angle weights ws applied
template <class T>
__device__ inline void seg_wsum(const T *red, const T *ws, int nseg, int len, T *out, int S) {
  const int t = threadIdx.x, per = blockDim.x / S;
  for (int base = 0; base < nseg; base += per) {
    const int seg = base + t / S, l = t & (S - 1);
    T p = T(0);
    if (seg < nseg)
      for (int k = l; k < len; k += S) p += ws[k] * red[seg * len + k];
    for (int o = S >> 1; o > 0; o >>= 1) p += __shfl_xor(p, o, S);
    if (l == 0 && seg < nseg) out[seg] = p;
  }
}

// ---------------------------------------------------------------------------
// Z-march structured apply (3D FD; reference src/termsk.kokkos.cxx:70-98
// streaming, :480-509 removal, :584-649 scatter). A workgroup owns a run of
// XR = R * CPT cells along x at one y index and walks a chunk of zc planes
// along z; a thread (lane column, angle) holds CPT consecutive cells of each
// plane. Planes are loaded two ahead, so the z upwind trace - the previous
// plane's own trace for a positive cosine, the next plane's for a negative
// one - is already in registers; the x upwind trace is a neighbouring cell of
// the lane's run, or of the neighbouring lane column through the LDS copy the
// moments are summed from (one global halo load per plane at the run's
// inflow edge); the y upwind trace is a load of the neighbouring row, which
// that row's workgroup reads as its own trace at about the same time, so it
// comes through L2. psi is read from HBM about once (plus one plane per chunk
// on each side). Every load is clamped into the grid, so none sits under a
// branch; a missing neighbour's weight is zero.
// ---------------------------------------------------------------------------
template <class T, int CPT> struct ZPlane {
  T p[CPT], yu[CPT];
  T hx;
  T sg;  // one cross section of the plane's run (threads below 2 XR), staged through LDS
};

template <class T, int CPT, bool SCATTER>
__global__ void __launch_bounds__(256) fd_apply_zm(const T *__restrict__ x, T *__restrict__ y, OpArgs<T> g, int zc) {
  extern __shared__ __align__(16) unsigned char smem[];
  const int A = g.A, R = g.cpb, XR = R * CPT;
  const int nx = g.n[0], ny = g.n[1], nz = g.n[2];
  const int tiles_x = (nx + XR - 1) / XR, chunks = (nz + zc - 1) / zc;
  const int64_t n_tiles = (int64_t)tiles_x * ny * chunks;
  const int64_t t = xcd_tile(n_tiles);
  if (t >= n_tiles) return;
  T *ws = reinterpret_cast<T *>(smem);
  T *buf = ws + ((A + 1) & ~1);
  T *phi = buf + (int64_t)XR * A;
  T *sgl = phi + XR;  // [2 planes][sigma_t, sigma_s][XR], by the plane's parity
  for (int e = threadIdx.x; e < A; e += blockDim.x) ws[e] = ldg(g.w + e);
  const int tx = (int)(t % tiles_x);
  const int64_t rest = t / tiles_x;
  const int j = (int)(rest % ny), ch = (int)(rest / ny);
  const int z0 = ch * zc, z1 = z0 + zc < nz ? z0 + zc : nz;
  const int lc = (int)threadIdx.x / A;
  const bool lane = lc < R;
  const int a = lane ? (int)threadIdx.x - lc * A : 0;
  const int xb = tx * XR + (lane ? lc : 0) * CPT;
  T cs[3], cd[3];
#pragma unroll
  for (int d = 0; d < 3; d++) {
    cs[d] = ldg(g.omega + 3 * a + d);
    cd[d] = fabs(cs[d]) * g.invh[d];
  }
  const bool xp = cs[0] > T(0), yp = cs[1] > T(0), zp = cs[2] > T(0);
  const int ju = yp ? j - 1 : j + 1;
  const bool yin = cs[1] != T(0) && ju >= 0 && ju < ny;
  const int jl = yin ? ju : j;
  int xc[CPT];
  bool xv[CPT];
#pragma unroll
  for (int k = 0; k < CPT; k++) {
    xv[k] = lane && xb + k < nx;
    xc[k] = xb + k < nx ? xb + k : nx - 1;
  }
  // the cell beyond the run on its x inflow side: a global load where it lies
  // outside the workgroup, else the neighbouring lane column's cell in LDS
  const int xh = xp ? xb - 1 : xb + CPT;
  const bool xhin = cs[0] != T(0) && xh >= 0 && xh < nx;
  const bool xhalo = xhin && (xp ? lc == 0 : lc == R - 1);
  const int xhc = xhalo ? xh : xc[0];
  const int nbl = xp ? lc - 1 : lc + 1;
  const int nbk = xp ? CPT - 1 : 0;
  // a plane's rows start at the (uniform) plane base; the lane's offsets
  // inside a plane are the same on every plane and fit 32 bits (the
  // launcher checks), so every load is a scalar base plus a 32-bit offset
  const int64_t prow = (int64_t)nx * ny * A, pcell = (int64_t)nx * ny;
  constexpr unsigned TS = sizeof(T);
  unsigned op[CPT], oy[CPT];  // byte offsets in a plane
#pragma unroll
  for (int k = 0; k < CPT; k++) {
    op[k] = (unsigned)((((int64_t)j * nx + xc[k]) * A + a) * TS);
    oy[k] = (unsigned)((((int64_t)jl * nx + xc[k]) * A + a) * TS);
  }
  const unsigned oh = (unsigned)((((int64_t)j * nx + xhc) * A + a) * TS);
  // the cross sections of the run: thread i < XR loads sigma_t of cell i,
  // XR <= i < 2 XR sigma_s (threads past that load sigma_t of the run's
  // first cell, unused)
  const int si = (int)threadIdx.x < 2 * XR ? (int)threadIdx.x : 0;
  const int sx0 = tx * XR + (si < XR ? si : si - XR);
  const unsigned os = (unsigned)(((int64_t)j * nx + (sx0 < nx ? sx0 : nx - 1)) * TS);
  const T *sgp = SCATTER && si >= XR ? g.sigs : g.sigt;
  auto at = [](const T *base, unsigned off) -> T {
    return ldg(reinterpret_cast<const T *>(reinterpret_cast<const char *>(base) + off));
  };
  auto load = [&](ZPlane<T, CPT> &s, int z) {
    const T *xz = x + (int64_t)z * prow;
#pragma unroll
    for (int k = 0; k < CPT; k++) {
      s.p[k] = at(xz, op[k]);
      s.yu[k] = at(xz, oy[k]);
    }
    s.hx = at(xz, oh);
    s.sg = at(sgp + (int64_t)z * pcell, os);
  };
  T prv[CPT];
  {
    const T *xq = x + (int64_t)(z0 > 0 ? z0 - 1 : 0) * prow;
#pragma unroll
    for (int k = 0; k < CPT; k++) prv[k] = at(xq, op[k]);
  }
  ZPlane<T, CPT> s0, s1, s2;
  load(s0, z0);
  load(s1, z0 + 1 < nz ? z0 + 1 : nz - 1);
  const T diag = cd[0] + cd[1] + cd[2];
  for (int z = z0; z < z1; z++) {
    load(s2, z + 2 < nz ? z + 2 : nz - 1);
    if (lane) {
#pragma unroll
      for (int k = 0; k < CPT; k++) buf[(lc * CPT + k) * A + a] = s0.p[k];
    }
    // the plane's cross sections, in the half of the staging its parity owns:
    // the half written two planes ago is no longer read (every thread has
    // passed this plane's first barrier)
    T *sgz = sgl + (z & 1) * 2 * XR;
    if ((int)threadIdx.x < 2 * XR) sgz[threadIdx.x] = s0.sg;
    __syncthreads();
    T xl = s0.hx;
    if (lane && !xhalo && nbl >= 0 && nbl < R) xl = buf[(nbl * CPT + nbk) * A + a];
    if (SCATTER) seg_wsum(buf, ws, XR, A, phi, g.red_lanes);
    __syncthreads();
    const bool zin = cs[2] != T(0) && (zp ? z > 0 : z < nz - 1);
#pragma unroll
    for (int k = 0; k < CPT; k++) {
      if (!xv[k]) continue;
      const int ck = lc * CPT + k;
      T v = (sgz[ck] + diag) * s0.p[k];
      const bool xi = cs[0] != T(0) && (xp ? xc[k] > 0 : xc[k] < nx - 1);
      const T ux = xp ? (k > 0 ? s0.p[k > 0 ? k - 1 : 0] : xl) : (k < CPT - 1 ? s0.p[k < CPT - 1 ? k + 1 : 0] : xl);
      if (xi) v -= cd[0] * ux;
      if (yin) v -= cd[1] * s0.yu[k];
      if (zin) v -= cd[2] * (zp ? prv[k] : s1.p[k]);
      if (SCATTER) v -= sgz[XR + ck] * g.inv_sum_w * phi[ck];
      // reflective faces: the cell's own reflected angle, on a branch only
      // boundary cells enter
      const T *xc0 = x + (int64_t)z * prow + (op[k] / TS - a);
      const bool bx = cs[0] != T(0) && !xi, by = cs[1] != T(0) && !yin, bz = cs[2] != T(0) && !zin;
      if (bx && g.face_type[0][xp ? 0 : 1] == kReflect) v -= cd[0] * xc0[g.reflect[3 * a]];
      if (by && g.face_type[1][yp ? 0 : 1] == kReflect) v -= cd[1] * xc0[g.reflect[3 * a + 1]];
      if (bz && g.face_type[2][zp ? 0 : 1] == kReflect) v -= cd[2] * xc0[g.reflect[3 * a + 2]];
      *reinterpret_cast<T *>(reinterpret_cast<char *>(y + (int64_t)z * prow) + op[k]) = v;
    }
#pragma unroll
    for (int k = 0; k < CPT; k++) prv[k] = s0.p[k];
    s0 = s1;
    s1 = s2;
  }
}

template <class T, int DIM, int NB, int K, bool SCATTER>
__device__ __forceinline__ void box_tile(const T *__restrict__ x, T *__restrict__ y, const OpArgs<T> &g,
                                         const BoxArgs &bx, const BoxConst<T, DIM, NB> &kc, int64_t t, T *buf,
                                         T *phi, const T *ws, int lc, int a, int64_t RS, int BS, int64_t aoff) {
  constexpr int R = DIM - 1;
  const int A = g.A, W = bx.W;
  const int nx = bx.n[0], ny = bx.n[1], nR = bx.n[R];
  // tiles walk bands of kBoxBand row blocks along the register axis, the
  // blocks of a band one after another at each x (and y) position: the row
  // two blocks share is read by the second while the first's read of it is
  // still in L2, so only the band's two outer halo rows come from HBM
  const int64_t per_band = (int64_t)bx.tiles[0] * bx.tiles[1] * kBoxBand;
  const int band = (int)(t / per_band);
  int64_t idx = t - (int64_t)band * per_band;
  const int nbh = bx.tiles[2] - band * kBoxBand < kBoxBand ? bx.tiles[2] - band * kBoxBand : kBoxBand;
  const int rbi = (int)(idx % nbh);
  idx /= nbh;
  const int tx = (int)(idx % bx.tiles[0]);
  const int j = DIM == 3 ? (int)(idx / bx.tiles[0]) : 0;
  const int r0 = (band * kBoxBand + rbi) * K;
  const int i = tx * W + lc;
  const bool lane = lc < W;
  const bool col = lane && i < nx;
  auto cell = [&](int ii, int jj, int rr) -> int64_t {
    return DIM == 2 ? (int64_t)rr * nx + ii : ((int64_t)rr * ny + jj) * nx + ii;
  };
  // value (cell c, basis b) of this thread's angle at row_of(c) + b BS: the
  // reference order (RS = NB A, BS = A, aoff = a) or the quadrant-major one
  auto row_of = [&](int64_t c) { return c * RS + aoff; };
  T p[K][NB], hx[K][NB], hy[K][NB], hr[NB], st[K], ss[K];
  bool okc[K];
#pragma unroll
  for (int k = 0; k < K; k++) {
    const bool v = col && r0 + k < nR;
    const int64_t c = v ? cell(i, j, r0 + k) : 0;
#pragma unroll
    for (int b = 0; b < NB; b++) p[k][b] = v ? x[row_of(c) + (int64_t)b * BS] : T(0);
    st[k] = v ? ldg(g.sigt + c) : T(0);
    ss[k] = v ? ldg(g.sigs + c) : T(0);
    okc[k] = v && (bx.ok == nullptr || ldg(bx.ok + c));
  }
  // the upwind row beyond the tile along the register axis
  const int rh = kc.in_side[R] == 0 ? r0 - 1 : r0 + K;
  {
    const bool v = col && rh >= 0 && rh < nR;
    const int64_t c = v ? cell(i, j, rh) : 0;
#pragma unroll
    for (int b = 0; b < NB; b++) hr[b] = v ? x[row_of(c) + (int64_t)b * BS] : T(0);
  }
  // the x upwind column: another lane of the tile, or a halo load at its edge
  const int ix = kc.in_side[0] == 0 ? i - 1 : i + 1;
  const bool xin = col && ix >= 0 && ix < nx;
  const bool xhalo = xin && (kc.in_side[0] == 0 ? lc == 0 : lc == W - 1);
#pragma unroll
  for (int k = 0; k < K; k++) {
    const bool v = xhalo && r0 + k < nR;
    const int64_t c = v ? cell(ix, j, r0 + k) : 0;
#pragma unroll
    for (int b = 0; b < NB; b++) hx[k][b] = v ? x[row_of(c) + (int64_t)b * BS] : T(0);
  }
  // 3D: the y upwind row, a global load
  const int jn = DIM == 3 ? (kc.in_side[DIM == 3 ? 1 : 0] == 0 ? j - 1 : j + 1) : 0;
  const bool yin = DIM == 3 && col && jn >= 0 && jn < ny;
#pragma unroll
  for (int k = 0; k < K; k++) {
    const bool v = yin && r0 + k < nR;
    const int64_t c = v ? cell(i, jn, r0 + k) : 0;
#pragma unroll
    for (int b = 0; b < NB; b++) hy[k][b] = v ? x[row_of(c) + (int64_t)b * BS] : T(0);
  }
  if (lane) {
#pragma unroll
    for (int k = 0; k < K; k++)
#pragma unroll
      for (int b = 0; b < NB; b++) buf[((lc * K + k) * NB + b) * A + a] = p[k][b];
  }
  __syncthreads();
  if (xin && !xhalo) {
    const int l2 = kc.in_side[0] == 0 ? lc - 1 : lc + 1;
#pragma unroll
    for (int k = 0; k < K; k++)
#pragma unroll
      for (int b = 0; b < NB; b++) hx[k][b] = buf[((l2 * K + k) * NB + b) * A + a];
  }
  if (SCATTER) seg_wsum(buf, ws, W * K * NB, A, phi, bx.lanes);
  __syncthreads();
  if (!col) return;
#pragma unroll
  for (int k = 0; k < K; k++) {
    const int r = r0 + k;
    if (r >= nR) break;
    const int64_t c = cell(i, j, r);
    T pr[NB];
#pragma unroll
    for (int b = 0; b < NB; b++) {
      const T lo = k > 0 ? p[k > 0 ? k - 1 : 0][b] : hr[b];
      const T hi = k < K - 1 ? p[k < K - 1 ? k + 1 : 0][b] : hr[b];
      pr[b] = kc.in_side[R] == 0 ? lo : hi;
    }
    // cells the box formula does not cover belong to the list phase (a
    // vacuum boundary face reads the zero halo loaded for it above)
    if (!okc[k]) continue;
    T acc[NB];
#pragma unroll
    for (int b = 0; b < NB; b++) {
      T v = st[k] * p[k][b];
#pragma unroll
      for (int q = 0; q < NB; q++) v += kc.M[b][q] * p[k][q];
#pragma unroll
      for (int q = 0; q < NB; q++) v += kc.U[0][b][q] * hx[k][q];
      if (DIM == 3) {
#pragma unroll
        for (int q = 0; q < NB; q++) v += kc.U[DIM == 3 ? 1 : 0][b][q] * hy[k][q];
      }
#pragma unroll
      for (int q = 0; q < NB; q++) v += kc.U[R][b][q] * pr[q];
      acc[b] = v;
    }
    if (SCATTER) {
      const T sc = ss[k] * g.inv_sum_w;
#pragma unroll
      for (int b = 0; b < NB; b++) acc[b] -= sc * phi[(lc * K + k) * NB + b];
    }
#pragma unroll
    for (int b = 0; b < NB; b++) y[row_of(c) + (int64_t)b * BS] = acc[b];
  }
}

// LDS of the box tile: the angle weights, the moments, the psi tile. One
// buffer each serves consecutive tiles: a tile's LDS reads all come before
// its second barrier, and the next tile's writes after it.
__host__ __device__ inline size_t box_lds_reals(int A, int W, int K, int nb) {
  const size_t wpad = ((size_t)A + 1) & ~(size_t)1;
  return wpad + (size_t)W * K * nb + (size_t)W * K * nb * A;
}

// Runs tiles [t0, t1) of one problem; every thread of the workgroup calls it.
template <class T, int DIM, int NB, int K, bool SCATTER>
__device__ __forceinline__ void box_run(const T *__restrict__ x, T *__restrict__ y, const OpArgs<T> &g,
                                        const BoxArgs &bx, int64_t t0, int64_t t1, const int *qp = nullptr) {
  extern __shared__ __align__(16) unsigned char smem[];
  const int A = g.A, W = bx.W;
  const int lc = threadIdx.x / A;
  const int a = lc < W ? (int)threadIdx.x - lc * A : 0;
  T *ws = reinterpret_cast<T *>(smem);
  T *phib = ws + ((A + 1) & ~1);
  T *bufb = phib + W * K * NB;
  for (int e = threadIdx.x; e < A; e += blockDim.x) ws[e] = ldg(g.w + e);
  BoxConst<T, DIM, NB> kc;
  box_consts<T, DIM, NB>(g, bx, a, kc);
  // vectors in the reference order, or (qp: the problem's angle positions
  // q A / 4 + kq) in the quadrant-major order (q, cell, basis, kq)
  int64_t RS = (int64_t)NB * A, aoff = a;
  int BS = A;
  if (qp) {
    const int Aq = A / 4, qk = ldg(qp + a), q = qk / Aq;
    RS = (int64_t)NB * Aq;
    BS = Aq;
    aoff = (int64_t)q * g.n_cells * NB * Aq + (qk - q * Aq);
  }
  for (int64_t t = t0; t < t1; t++)
    box_tile<T, DIM, NB, K, SCATTER>(x, y, g, bx, kc, t, bufb, phib, ws, lc, a, RS, BS, aoff);
}

// One problem: a persistent grid (a multiple of 8 workgroups), each walking one
// contiguous run of tiles, ranked XCD-major so a run's halos stay in one L2,
// then its share of the boundary-cell list through the tile kernel's body.
template <class T, int DIM, int NB, int K, bool SCATTER>
__global__ void __launch_bounds__((kBoxTileThreads<T, DIM>)) dg_apply_box(const T *__restrict__ x, T *__restrict__ y, OpArgs<T> g,
                                                    BoxArgs bx, int64_t total) {
  const int64_t G = gridDim.x;
  const int64_t rank = (int64_t)(blockIdx.x % 8) * (G / 8) + blockIdx.x / 8;
  const int64_t chunk = (total + G - 1) / G;
  const int64_t t0 = rank * chunk, t1 = t0 + chunk < total ? t0 + chunk : total;
  if (t0 < t1) box_run<T, DIM, NB, K, SCATTER>(x, y, g, bx, t0, t1);
  const int64_t E = bx.edge_tiles, ce = (E + G - 1) / G;
  const int64_t e0 = rank * ce, e1 = e0 + ce < E ? e0 + ce : E;
  for (int64_t e = e0; e < e1; e++) {
    __syncthreads();  // the LDS changes layout
    dg_tile_body<T, DIM, NB, 2 * DIM, SCATTER, false, true, false>(x, y, g, e);
  }
}

// ---------------------------------------------------------------------------
// Strip march over a uniform DG1 box. A unit is W columns along x (thread =
// column, angle; at one y index in 3D) by a band of R cells along the last
// axis, which the workgroup walks one row at a time. The rows either side of
// the current one stay in registers, so the upwind trace along the march axis
// is never loaded twice, and the next rows are requested two steps ahead. Per
// row one LDS write of the row, one barrier (the LDS is double-buffered by
// row parity): the x upwind trace comes from the neighbouring lane's slot
// (the strip's edge lane loads it: the neighbouring strip, marching at the
// same time on the same XCD, has it in L2), the scatter's moments are a
// segmented sum over the row, and a row's result is stored one step later,
// once its moments have passed the next barrier. Boundary cells whose box
// faces are vacuum run the same formula with a zero trace; cells the formula
// does not cover are left to the list phase.
// ---------------------------------------------------------------------------
// rows requested ahead by the strip march, per precision
template <class T> constexpr int kStripAhead = sizeof(T) == 8 ? 3 : 4;

__host__ __device__ inline size_t strip_lds_reals(int A, int W, int nb) {
  const size_t wpad = ((size_t)A + 1) & ~(size_t)1;
  return wpad + 2 * (size_t)W * nb + 2 * (size_t)(W + 2) * nb * A;
}

// Calls f(integral_constant<K>) for K = 0 .. N - 1 while f returns true: a
// loop body unrolled with a compile-time step index.
template <int K, int N, class F> __device__ __forceinline__ bool unroll_while(F &&f) {
  if constexpr (K < N) {
    if (!f(std::integral_constant<int, K>{})) return false;
    return unroll_while<K + 1, N>(f);
  }
  return true;
}

// PF: rows requested ahead of the current one. The rows live in a ring of
// L = PF + 3 register slots (rows r - 1 .. r + PF + 1), the row loop is
// unrolled by L so every slot is a fixed register, and every load is
// unconditional (clamped to a valid cell, the value selected at its use):
// copying a register whose load is still in flight, or a load under a
// branch, makes the compiler wait for every load in flight.
template <class T, int DIM, int NB, int PF, bool SCATTER>
__device__ __forceinline__ void strip_unit(const T *__restrict__ x, T *__restrict__ y, const OpArgs<T> &g,
                                           const BoxArgs &bx, const BoxConst<T, DIM, NB> &kc, int64_t u, T *buf,
                                           T *phib, const T *ws, int lc, int a) {
  constexpr int R = DIM - 1, L = PF + 3;
  const int A = g.A, W = bx.W, WN = W * NB, NA = NB * A;
  const int nx = bx.n[0], ny = bx.n[1], nR = bx.n[R];
  const int sx = (int)(u % bx.units[0]);
  const int64_t rest = u / bx.units[0];
  const int j = DIM == 3 ? (int)(rest % bx.units[1]) : 0;
  const int r0 = (int)(DIM == 3 ? rest / bx.units[1] : rest) * bx.R;
  const int r1 = r0 + bx.R < nR ? r0 + bx.R : nR;
  const int i0 = sx * W;
  const int i = i0 + lc;
  const bool lane = lc < W;
  const bool col = lane && i < nx;
  const int ic = col ? i : i0;
  auto cell = [&](int ii, int jj, int rr) -> int64_t {
    return DIM == 2 ? (int64_t)rr * nx + ii : ((int64_t)rr * ny + jj) * nx + ii;
  };
  auto clampr = [&](int rr) { return rr < 0 ? 0 : (rr < nR ? rr : nR - 1); };
  // the march axis' upwind row is r - 1 where the cosine is positive
  const bool lo = kc.in_side[R] == 0;
  // x: the neighbour column in the LDS row (columns 0 and W + 1 hold the
  // strip's halo cells); the halo element e = threadIdx.x of the row, taken
  // from the left or the right halo cell by its angle's inflow side
  const int xcol = kc.in_side[0] == 0 ? lc : lc + 2;
  const int e = threadIdx.x;
  const bool he = e < NA;
  const int ea = he ? e % A : 0;
  const bool hleft = g.omega[3 * ea] > T(0);
  const int ih = hleft ? i0 - 1 : i0 + W;
  const bool hv = he && ih >= 0 && ih < nx;
  const int ihc = hv ? ih : i0;
  const int hslot = (hleft ? 0 : W + 1) * NA + e;
  // y (3D): the neighbour row j -+ 1 by the angle's inflow side
  const int jn = DIM == 3 ? (kc.in_side[DIM == 3 ? 1 : 0] == 0 ? j - 1 : j + 1) : 0;
  const bool yin = DIM == 3 && col && jn >= 0 && jn < ny;
  const int jnc = yin ? jn : j;
  T q[L][NB], hy[DIM == 3 ? L : 1][NB], hx[L], st[L], ss[L], acc[NB];
  unsigned char ok[L];
  // the loads of row rr into ring slot s
  auto fetch = [&](int s, int rr) {
    const int rc = clampr(rr);
    const int64_t c = cell(ic, j, rc);
#pragma unroll
    for (int b = 0; b < NB; b++) q[s][b] = ldg(x + c * NA + b * A + a);
    st[s] = ldg(g.sigt + c);
    ss[s] = ldg(g.sigs + c);
    ok[s] = ldg(bx.ok + c);
    hx[s] = ldg(x + cell(ihc, j, rc) * NA + (he ? e : 0));
    if (DIM == 3) {
      const int64_t cy = cell(ic, jnc, rc);
#pragma unroll
      for (int b = 0; b < NB; b++) hy[DIM == 3 ? s : 0][b] = ldg(x + cy * NA + b * A + a);
    }
  };
  // rows r0 - 1 .. r0 + PF into slots 0 .. PF + 1
#pragma unroll
  for (int d = -1; d <= PF; d++) fetch(d + 1, r0 + d);
  T ssp = T(0);
  bool okp = false;
  int64_t cp = 0;
  int t = 0;
  auto step = [&](auto kk) -> bool {
    constexpr int k = decltype(kk)::value;
    constexpr int s0 = (k + 1) % L, sm = k % L, sp = (k + 2) % L, sa = (k + PF + 2) % L;
    const int r = r0 + t + k;
    if (r >= r1) return false;
    fetch(sa, r + PF + 1);
    const int64_t c = cell(ic, j, r);
    T *bb = buf + (size_t)((t + k) & 1) * (W + 2) * NA;
    T *ph = phib + ((t + k) & 1) * WN;
    if (lane) {
#pragma unroll
      for (int b = 0; b < NB; b++) bb[(lc + 1) * NA + b * A + a] = col ? q[s0][b] : T(0);
    }
    if (he) bb[hslot] = hv ? hx[s0] : T(0);
    __syncthreads();
    T hn[NB];
#pragma unroll
    for (int b = 0; b < NB; b++) hn[b] = lane ? bb[xcol * NA + b * A + a] : T(0);
    if (SCATTER) {
      seg_wsum(bb + NA, ws, WN, A, ph, bx.lanes);
      if (okp) {  // the previous row: its moments passed this step's barrier
        const T sc = ssp * g.inv_sum_w;
        const T *pp = phib + ((t + k + 1) & 1) * WN + lc * NB;
#pragma unroll
        for (int b = 0; b < NB; b++) y[cp * NA + b * A + a] = acc[b] - sc * pp[b];
      }
    }
    const bool upv = lo ? r > 0 : r + 1 < nR;
#pragma unroll
    for (int b = 0; b < NB; b++) {
      T v = st[s0] * q[s0][b];
#pragma unroll
      for (int qq = 0; qq < NB; qq++) v += kc.M[b][qq] * q[s0][qq];
#pragma unroll
      for (int qq = 0; qq < NB; qq++) v += kc.U[0][b][qq] * hn[qq];
      if (DIM == 3) {
#pragma unroll
        for (int qq = 0; qq < NB; qq++) v += kc.U[DIM == 3 ? 1 : 0][b][qq] * (yin ? hy[DIM == 3 ? s0 : 0][qq] : T(0));
      }
#pragma unroll
      for (int qq = 0; qq < NB; qq++) v += kc.U[R][b][qq] * (upv ? (lo ? q[sm][qq] : q[sp][qq]) : T(0));
      acc[b] = v;
    }
    const bool okc = col && ok[s0];
    if (!SCATTER && okc) {
#pragma unroll
      for (int b = 0; b < NB; b++) y[c * NA + b * A + a] = acc[b];
    }
    okp = okc;
    cp = c;
    ssp = ss[s0];
    return true;
  };
  while (unroll_while<0, L>(step)) t += L;
  if (SCATTER) {
    __syncthreads();
    if (okp) {
      const T sc = ssp * g.inv_sum_w;
      const T *pp = phib + ((r1 - 1 - r0) & 1) * WN + lc * NB;
#pragma unroll
      for (int b = 0; b < NB; b++) y[cp * NA + b * A + a] = acc[b] - sc * pp[b];
    }
  }
}

// LDS carve of the strip march: angle weights, two rows of moments, two rows
// of W + 2 columns (the strip and its two halo cells)
template <class T, int NB>
__device__ __forceinline__ void strip_carve(int A, int W, T *&ws, T *&phib, T *&buf) {
  extern __shared__ __align__(16) unsigned char smem[];
  ws = reinterpret_cast<T *>(smem);
  phib = ws + ((A + 1) & ~1);
  buf = phib + 2 * W * NB;
}

// One problem: a persistent grid, unit u = rank + k * grid, ranked XCD-major
// so the units one XCD runs at a time are neighbouring strips; then the
// boundary-cell list through the tile kernel's body.
template <class T, int DIM, int NB, int PF, bool SCATTER>
__global__ void __launch_bounds__(kBoxThreads) dg_apply_strip(const T *__restrict__ x, T *__restrict__ y,
                                                      OpArgs<T> g, BoxArgs bx, int64_t total) {
  const int64_t G = gridDim.x;
  const int64_t rank = (int64_t)(blockIdx.x % 8) * (G / 8) + blockIdx.x / 8;
  const int A = g.A, W = bx.W;
  const int lc = threadIdx.x / A;
  const int a = lc < W ? (int)threadIdx.x - lc * A : 0;
  T *ws, *phib, *buf;
  strip_carve<T, NB>(A, W, ws, phib, buf);
  for (int e = threadIdx.x; e < A; e += blockDim.x) ws[e] = ldg(g.w + e);
  BoxConst<T, DIM, NB> kc;
  box_consts<T, DIM, NB>(g, bx, a, kc);
  for (int64_t u = rank; u < total; u += G) {
    __syncthreads();  // the previous unit's last reads of the row buffers
    strip_unit<T, DIM, NB, PF, SCATTER>(x, y, g, bx, kc, u, buf, phib, ws, lc, a);
  }
  const int64_t E = bx.edge_tiles, ce = (E + G - 1) / G;
  const int64_t e0 = rank * ce, e1 = e0 + ce < E ? e0 + ce : E;
  for (int64_t e = e0; e < e1; e++) {
    __syncthreads();  // the LDS changes layout
    dg_tile_body<T, DIM, NB, 2 * DIM, SCATTER, false, true, false>(x, y, g, e);
  }
}

// ---------------------------------------------------------------------------
// Batched apply: one persistent launch over the tiles of many problems. The
// work list holds each active problem's operator arguments and row offset;
// tile_off is the prefix of tiles per problem. Workgroups are ranked
// XCD-major and each walks one contiguous run of tiles, so a run stays within
// one or two problems and its upwind gathers in one XCD's L2.
// ---------------------------------------------------------------------------
template <class T, int DIM, int NB, int NF, int BODY, bool SCATTER>
__global__ void __launch_bounds__(256) dg_apply_batch(const T *__restrict__ x, T *__restrict__ y,
                                                      const BatchItem<T> *__restrict__ items,
                                                      const int64_t *__restrict__ tile_off, int64_t total) {
  const int64_t G = gridDim.x;
  const int64_t rank = (int64_t)(blockIdx.x % 8) * (G / 8) + blockIdx.x / 8;
  const int64_t chunk = (total + G - 1) / G;
  const int64_t t0 = rank * chunk, t1 = t0 + chunk < total ? t0 + chunk : total;
  int p = 0;
  for (int64_t t = t0; t < t1; t++) {
    while (t >= tile_off[p + 1]) p++;
    const BatchItem<T> &it = items[p];
    const OpArgs<T> &g = it.g;
    if (BODY == 0)
      dg_tile_body<T, DIM, NB, NF, SCATTER, false, false, false>(x + it.off, y + it.off, g, t - tile_off[p]);
    else
      dg_q4_body<T, DIM, NB, NF, SCATTER>(x + it.off, y + it.off, g, t - tile_off[p]);
    __syncthreads();  // the next tile restages the LDS
  }
}

// Batched box tile: a workgroup's contiguous run of tiles, one problem at a
// time (the per-angle matrices are refolded at each problem change), then
// its contiguous share of every problem's boundary-cell tiles.
template <class T, int DIM, int NB, int K, bool SCATTER>
__global__ void __launch_bounds__((kBoxTileThreads<T, DIM>)) dg_apply_batch_box(const T *__restrict__ x, T *__restrict__ y,
                                                          const BatchItem<T> *__restrict__ items,
                                                          const int64_t *__restrict__ tile_off, int64_t total,
                                                          const int64_t *__restrict__ edge_off, int64_t n_edge,
                                                          const int *__restrict__ qpos,
                                                          const int64_t *__restrict__ qoff) {
  const int64_t G = gridDim.x;
  const int64_t rank = (int64_t)(blockIdx.x % 8) * (G / 8) + blockIdx.x / 8;
  const int64_t chunk = (total + G - 1) / G;
  int64_t t = rank * chunk;
  const int64_t t1 = t + chunk < total ? t + chunk : total;
  int p = 0;
  bool first = true;
  while (t < t1) {
    while (t >= tile_off[p + 1]) p++;
    const int64_t e = t1 < tile_off[p + 1] ? t1 : tile_off[p + 1];
    if (!first) __syncthreads();  // the next problem lays the LDS out anew
    first = false;
    const BatchItem<T> &it = items[p];
    box_run<T, DIM, NB, K, SCATTER>(x + it.off, y + it.off, it.g, it.box, t - tile_off[p],
                                    e - tile_off[p], qpos ? qpos + qoff[p] : nullptr);
    t = e;
  }
  const int64_t ce = (n_edge + G - 1) / G;
  const int64_t e1 = rank * ce + ce < n_edge ? rank * ce + ce : n_edge;
  p = 0;
  for (int64_t e = rank * ce; e < e1; e++) {
    while (e >= edge_off[p + 1]) p++;
    __syncthreads();  // the LDS changes layout
    const BatchItem<T> &it = items[p];
    dg_tile_body<T, DIM, NB, 2 * DIM, SCATTER, false, true, false>(x + it.off, y + it.off, it.g,
                                                                    e - edge_off[p]);
  }
}

// Batched strip march: unit u = rank + k * grid over every problem's units
// (the problem found by bisection of the unit prefix), so the units one XCD
// runs at a time are neighbouring strips of one problem; the per-angle
// matrices are refolded at each problem change. Then the workgroup's
// contiguous share of every problem's boundary-cell tiles.
template <class T, int DIM, int NB, int PF, bool SCATTER>
__global__ void __launch_bounds__(kBoxThreads) dg_apply_batch_strip(const T *__restrict__ x, T *__restrict__ y,
                                                            const BatchItem<T> *__restrict__ items,
                                                            const int64_t *__restrict__ unit_off, int n_items,
                                                            int64_t total, const int64_t *__restrict__ edge_off,
                                                            int64_t n_edge) {
  const int64_t G = gridDim.x;
  const int64_t rank = (int64_t)(blockIdx.x % 8) * (G / 8) + blockIdx.x / 8;
  int p = -1;
  int lc = 0, a = 0;
  T *ws = nullptr, *phib = nullptr, *buf = nullptr;
  BoxConst<T, DIM, NB> kc;
  for (int64_t u = rank; u < total; u += G) {
    int lo = 0, hi = n_items;  // unit_off[lo] <= u < unit_off[hi]
    while (hi - lo > 1) {
      const int mid = (lo + hi) >> 1;
      if (u >= unit_off[mid]) lo = mid; else hi = mid;
    }
    __syncthreads();  // the previous unit's last reads of the LDS
    const BatchItem<T> &it = items[lo];
    const OpArgs<T> &g = it.g;
    const BoxArgs &bx = it.box;
    if (lo != p) {
      p = lo;
      const int A = g.A, W = bx.W;
      lc = threadIdx.x / A;
      a = lc < W ? (int)threadIdx.x - lc * A : 0;
      strip_carve<T, NB>(A, W, ws, phib, buf);
      for (int e = threadIdx.x; e < A; e += blockDim.x) ws[e] = ldg(g.w + e);
      box_consts<T, DIM, NB>(g, bx, a, kc);
      __syncthreads();  // the weights
    }
    strip_unit<T, DIM, NB, PF, SCATTER>(x + it.off, y + it.off, g, bx, kc, u - unit_off[lo], buf, phib, ws, lc, a);
  }
  const int64_t ce = (n_edge + G - 1) / G;
  const int64_t e1 = rank * ce + ce < n_edge ? rank * ce + ce : n_edge;
  p = 0;
  for (int64_t e = rank * ce; e < e1; e++) {
    while (e >= edge_off[p + 1]) p++;
    __syncthreads();  // the LDS changes layout
    const BatchItem<T> &it = items[p];
    dg_tile_body<T, DIM, NB, 2 * DIM, SCATTER, false, true, false>(x + it.off, y + it.off, it.g,
                                                                    e - edge_off[p]);
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
  const int gc = cls_of(g, c);
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
      const T *gr = g.grad + ((int64_t)gc * NB + i) * 3;
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
    const int64_t k = c * NF + f, kg = (int64_t)gc * NF + f;
    const int nbk = g.face_nb[k];
    const T *sv = g.face_s + 3 * kg;
    const T *Fo = g.face_own + kg * FM, *Fu = g.face_up + kg * FM;
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
  fd_index<DIM>(g.n, c, idx);
  T psi[APL], acc[APL];
#pragma unroll
  for (int q = 0; q < APL; q++) {
    // lanes past the last angle compute on the last one and store nothing
    const int a = lane + 64 * q < A ? lane + 64 * q : A - 1;
    acc[q] = fd_row<T, DIM, false>(x, g, c, idx, a, psi[q]);
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
      const int64_t kg = (int64_t)cls_of(g, c) * g.nf + f;
      T s = T(0);
      for (int dd = 0; dd < g.dim; dd++) s += g.omega[3 * a + dd] * g.face_s[3 * kg + dd];
      if (s > T(0)) v += g.nb == 1 ? s : s * g.face_own[(kg * g.nb + i) * g.nb + i];
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
                                                    int Apad, int M, int phi_stride, int w_in_lds) {
  using V4 = typename std::conditional<sizeof(T) == 4, v4f, v4d>::type;
  extern __shared__ __align__(16) unsigned char smem[];
  T *tile = reinterpret_cast<T *>(smem);  // 64 x Apad
  // the weights (Apad x 16) beside the tile when they fit: each MFMA step
  // then reads both operands from LDS instead of waiting on a cached load
  T *wl = tile + 64 * Apad;
  if (w_in_lds)
    for (int e = threadIdx.x; e < Apad * 16; e += blockDim.x) wl[e] = wext[e];
  const int64_t node0 = (int64_t)blockIdx.x * 64;
  const int64_t nn = n_nodes - node0 < 64 ? n_nodes - node0 : 64;
  // stage: contiguous psi[node0 .. node0+nn) x A, zero-padded columns. With
  // A a multiple of 4 (every level-symmetric set) the tile is psi's own run,
  // copied in 16-byte words; otherwise a row index from a float reciprocal
  // (exact here: e < 2^14, Apad < 2^9, so (e + 1/2) / Apad stays more than
  // 2^-10 inside its integer interval)
  if (Apad == A && ((reinterpret_cast<uintptr_t>(psi) & 15) == 0)) {
    // rounds of U 16-byte loads per thread issued together (indices past the
    // tile's values clamped to its first word, their values dropped), then
    // written to LDS: a round trip to HBM per U words, not per word
    constexpr int VW = 16 / sizeof(T), U = 8;
    struct alignas(16) W16 {
      T v[VW];
    };
    const int total = (int)(nn * A);  // a multiple of 4 values
    const T *src = psi + node0 * A;
    const int step = (int)blockDim.x * VW;
    for (int base = (int)threadIdx.x * VW; base < 64 * A; base += U * step) {
      W16 w[U];
#pragma unroll
      for (int u = 0; u < U; u++) {
        const int e = base + u * step;
        w[u] = *reinterpret_cast<const W16 *>(src + (e < total ? e : 0));
      }
#pragma unroll
      for (int u = 0; u < U; u++) {
        const int e = base + u * step;
        if (e >= 64 * A) break;
        if (e >= total)
#pragma unroll
          for (int q = 0; q < VW; q++) w[u].v[q] = T(0);
        *reinterpret_cast<W16 *>(tile + e) = w[u];
      }
    }
  } else {
    const float inv = 1.0f / (float)Apad;
    for (int e = threadIdx.x; e < 64 * Apad; e += blockDim.x) {
      const int row = (int)(((float)e + 0.5f) * inv), col = e - row * Apad;
      tile[e] = (row < nn && col < A) ? psi[(node0 + row) * A + col] : T(0);
    }
  }
  __syncthreads();
  const int wave = threadIdx.x >> 6, l = threadIdx.x & 63;
  const int i = l & 15, kq = l >> 4;
  V4 acc = {T(0), T(0), T(0), T(0)};
  const T *trow = tile + (wave * 16 + i) * Apad;
  const T *wsrc = w_in_lds ? wl : wext;
  for (int k0 = 0; k0 < Apad; k0 += 4) acc = mfma16x16x4(trow[k0 + kq], wsrc[(k0 + kq) * 16 + i], acc);
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
  const int64_t gc = cls_of(g, c);
  for (int f = 0; f < g.nf; f++) {
    const int64_t k = gc * g.nf + f;
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
      const T *gr = g.grad + (gc * NB + i) * 3;
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

// element blocks of many problems: block index -> problem by bisection of the
// block prefix
template <class T, int NB>
__global__ void blockinv_batch(const T *__restrict__ x, T *__restrict__ y, const BatchItem<T> *__restrict__ items,
                               const int64_t *__restrict__ blk_off, int n_items, int64_t total) {
  const int64_t b = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= total) return;
  int lo = 0, hi = n_items;
  while (hi - lo > 1) {
    const int mid = (lo + hi) >> 1;
    if (blk_off[mid] <= b) lo = mid;
    else hi = mid;
  }
  const BatchItem<T> &it = items[lo];
  const OpArgs<T> &g = it.g;
  const int A = g.A;
  const int64_t bb = b - blk_off[lo];
  const int64_t c = bb / A;
  const int a = (int)(bb - c * A);
  T B[NB][NB], v[NB];
  build_block<T, NB>(g, c, a, 0, B);
  const int64_t base = it.off + c * NB * A + a;
#pragma unroll
  for (int j = 0; j < NB; j++) v[j] = x[base + (int64_t)j * A];
  if (!gj_solve<T, NB>(B, v)) {
#pragma unroll
    for (int j = 0; j < NB; j++) v[j] = T(0);
  }
#pragma unroll
  for (int i = 0; i < NB; i++) y[base + (int64_t)i * A] = v[i];
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

// geo_bytes: LDS per cell for staged geometry (DG); the cells per workgroup
// shrink until the tile fits the 64 KB a workgroup may allocate
template <class T> Launch apply_launch(int A, int nb, size_t geo_bytes = 0) {
  Launch L;
  L.threads = A <= 256 ? 256 : ((A + 63) / 64) * 64;
  UBT_CHECK(L.threads <= 1024, UBT_ERR_SUP, "%d angles per cell exceed one workgroup", A);
  L.cpb = L.threads / A;
  auto lds_for = [&](int cpb) { return (size_t)(cpb * nb * A + cpb * nb) * sizeof(T) + (size_t)cpb * geo_bytes; };
  while (L.cpb > 1 && lds_for(L.cpb) > 65536) L.cpb--;
  L.threads = std::max(64, ((L.cpb * A + 63) / 64) * 64);
  const int nseg = L.cpb * nb;
  int S = pow2_floor(std::max(1, L.threads / nseg));
  S = std::min(S, 64);
  S = std::min(S, pow2_floor(A));
  L.lanes = std::max(1, S);
  L.lds = lds_for(L.cpb);
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

// The four-angle tile kernel; false where it does not apply (FD, A not a
// multiple of 4, or a tile that overflows the LDS)
template <class T>
bool launch_apply_q4(const Operator &op, OpArgs<T> a, const T *x, T *y, bool scatter, hipStream_t s,
                     bool dry = false) {
  const Disc &D = *op.disc;
  const int A = D.n_angles, nb = D.nb, nf = D.nf, dim = D.dim;
  if (D.kind != DISC_DG || A % 4 != 0 || A / 4 > 256) return false;
  const int TPC = A / 4;
  int cpb = 256 / TPC;
  auto lds_for = [&](int cp) {
    const size_t reals = (size_t)cp * ((nb > 1 ? 2 * nf * nb * nb + (nb - 1) * dim : 0) + nf * dim + 2) +
                         (size_t)cp * nb * TPC + (size_t)cp * nb;
    return reals * sizeof(T) + sizeof(int) * (size_t)cp * nf;
  };
  while (cpb > 1 && lds_for(cpb) > 65536) cpb--;
  if (lds_for(cpb) > 65536) return false;
  if (dry) return true;
  a.cpb = cpb;
  const int threads = ((cpb * TPC + 63) / 64) * 64;
  const int64_t n_tiles = (D.n_cells + cpb - 1) / cpb;
  const unsigned grid = (unsigned)(8 * ((n_tiles + 7) / 8));
  const size_t lds = lds_for(cpb);
  bool launched = false;
#define Q4(DIM, NB, NF)                                                                                           \
  if (!launched && dim == DIM && nb == NB && nf == NF) {                                                          \
    if (scatter)                                                                                                  \
      hipLaunchKernelGGL((dg_apply_q4<T, DIM, NB, NF, true>), dim3(grid), dim3(threads), lds, s, x, y, a);         \
    else                                                                                                          \
      hipLaunchKernelGGL((dg_apply_q4<T, DIM, NB, NF, false>), dim3(grid), dim3(threads), lds, s, x, y, a);        \
    launched = true;                                                                                              \
  }
  Q4(2, 1, 3) Q4(2, 1, 4) Q4(3, 1, 4) Q4(3, 1, 6) Q4(2, 3, 3) Q4(2, 3, 4) Q4(3, 4, 4) Q4(3, 4, 6)
#undef Q4
  return launched;
}

// The pencil kernel on a structured box marching H cells per workgroup; false
// where it does not apply (not a box, more than 256 angles, too many geometry
// classes to stage, or a row that overflows the LDS)
template <class T>
bool launch_apply_pencil(const Operator &op, OpArgs<T> a, const T *x, T *y, bool scatter, hipStream_t s, int H,
                         bool dry = false) {
  const Disc &D = *op.disc;
  if (!op.box || D.n_angles > 256 || op.n_geo_classes > kPencilMaxClasses) return false;
  const int A = D.n_angles, nb = D.nb, dim = D.dim, nf = D.nf, ncls = op.n_geo_classes, FM = nb * nb;
  PencilArgs pa;
  pa.W = 256 / A;
  pa.H = H;
  pa.ncls = ncls;
  for (int d = 0; d < 3; d++) {
    pa.n[d] = op.box_n[d];
    pa.fmap[d][0] = op.box_fmap[d][0];
    pa.fmap[d][1] = op.box_fmap[d][1];
  }
  const size_t reals = (size_t)ncls * nf * dim + (nb > 1 ? (size_t)2 * ncls * nf * FM + (size_t)ncls * (nb - 1) * dim : 0) +
                       (size_t)3 * pa.W * nb * A + (size_t)pa.W * nb;
  const size_t lds = reals * sizeof(T);
  if (lds > 65536) return false;
  if (dry) return true;
  pa.segs[0] = (pa.n[0] + pa.W - 1) / pa.W;
  pa.segs[1] = dim == 3 ? pa.n[1] : 1;
  pa.segs[2] = (pa.n[dim - 1] + H - 1) / H;
  const int64_t nwg = (int64_t)pa.segs[0] * pa.segs[1] * pa.segs[2];
  const unsigned grid = (unsigned)(8 * ((nwg + 7) / 8));
  const int threads = ((pa.W * A + 63) / 64) * 64;
  int S = pow2_floor(std::max(1, threads / (pa.W * nb)));
  S = std::min(std::min(S, 64), pow2_floor(A));
  a.red_lanes = std::max(1, S);
  const int sc = scatter ? 1 : 0;
  if (dim == 2 && nb == 1)
    hipLaunchKernelGGL((dg_apply_pencil<T, 2, 1>), dim3(grid), dim3(threads), lds, s, x, y, a, pa, sc);
  else if (dim == 2 && nb == 3)
    hipLaunchKernelGGL((dg_apply_pencil<T, 2, 3>), dim3(grid), dim3(threads), lds, s, x, y, a, pa, sc);
  else if (dim == 3 && nb == 1)
    hipLaunchKernelGGL((dg_apply_pencil<T, 3, 1>), dim3(grid), dim3(threads), lds, s, x, y, a, pa, sc);
  else if (dim == 3 && nb == 4)
    hipLaunchKernelGGL((dg_apply_pencil<T, 3, 4>), dim3(grid), dim3(threads), lds, s, x, y, a, pa, sc);
  else
    return false;
  return true;
}

// The multi-cell tile kernel with CPT cells per thread; false where it does
// not apply (FD, more than 256 angles, or a tile that overflows the LDS)
static int device_cus(int dev) {
  static int cache[64] = {};
  if (dev < 0 || dev >= 64) return 256;
  if (cache[dev] == 0) {
    int n = 0;
    UBT_HIP(hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, dev));
    cache[dev] = n > 0 ? n : 256;
  }
  return cache[dev];
}

// The box-tile layout of an operator for K rows per thread: columns per tile
// (as many as 256 threads hold, fewer if the LDS overflows), tiles, threads,
// the segmented-sum lanes and the LDS; false where the kernel does not apply
// (not a uniform DG1 box, more than 256 angles, or no column fits the LDS)
bool box_layout(const Operator &op, size_t ts, int K, BoxArgs &bx, int &threads, int &cpb, int &lanes,
                size_t &lds) {
  const Disc &D = *op.disc;
  const int A = D.n_angles, nb = D.nb, nf = D.nf, dim = D.dim;
  if (op.box_cls < 0 || A > 256) return false;
  int W = box_tile_threads(ts, dim) / A;
  while (W > 1 && box_lds_reals(A, W, K, nb) * ts > 65536) W--;
  lds = box_lds_reals(A, W, K, nb) * ts;
  if (lds > 65536) return false;
  threads = ((W * A + 63) / 64) * 64;
  // the boundary-cell phase: the tile body, cpb cells per tile
  const size_t geo = (size_t)(2 * nf * nb * nb + nf * dim + (nb - 1) * dim + 2) * ts + sizeof(int) * (size_t)nf;
  auto tile_lds = [&](int cp) { return (size_t)(cp * nb * A + cp * nb) * ts + (size_t)cp * geo; };
  cpb = threads / A;
  while (cpb > 1 && tile_lds(cpb) > 65536) cpb--;
  if (tile_lds(cpb) > 65536) return false;
  lds = std::max(lds, tile_lds(cpb));
  int Se = pow2_floor(std::max(1, threads / (cpb * nb)));
  Se = std::min(std::min(Se, 64), pow2_floor(A));
  lanes = std::max(1, Se);
  bx.edge_tiles = (op.n_box_edge + cpb - 1) / cpb;
  for (int d = 0; d < 3; d++) {
    bx.n[d] = op.box_n[d];
    bx.fmap[d][0] = op.box_fmap[d][0];
    bx.fmap[d][1] = op.box_fmap[d][1];
  }
  bx.cls = op.box_cls;
  bx.W = W;
  bx.tiles[0] = (op.box_n[0] + W - 1) / W;
  bx.tiles[1] = dim == 3 ? op.box_n[1] : 1;
  bx.tiles[2] = (op.box_n[dim - 1] + K - 1) / K;
  int S = pow2_floor(std::max(1, threads / (W * K * nb)));
  S = std::min(std::min(S, 64), pow2_floor(A));
  bx.lanes = std::max(1, S);
  bx.R = 0;
  bx.units[0] = bx.units[1] = bx.units[2] = 0;
  bx.ok = op.n_box_edge > 0 ? op.box_ok.as<unsigned char>() : nullptr;
  return true;
}

template <class T>
bool launch_apply_box(const Operator &op, OpArgs<T> a, const T *x, T *y, bool scatter, hipStream_t s, int K,
                      bool dry = false) {
  const Disc &D = *op.disc;
  BoxArgs bx;
  int threads = 0, cpb = 1, lanes = 1;
  size_t lds = 0;
  if (!box_layout(op, sizeof(T), K, bx, threads, cpb, lanes, lds)) return false;
  if (dry) return true;
  a.cpb = cpb;
  a.red_lanes = lanes;
  const int64_t total = (int64_t)bx.tiles[0] * bx.tiles[1] * bx.tiles[2];
  int64_t g = std::min<int64_t>(total, (int64_t)device_cus(op.device) * 8);
  g = std::max<int64_t>(8, (g + 7) / 8 * 8);
  const unsigned grid = (unsigned)g;
  bool launched = false;
#define BOX(DIM, NB, KK)                                                                                          \
  if (!launched && D.dim == DIM && D.nb == NB && K == KK) {                                                       \
    if (scatter)                                                                                                  \
      hipLaunchKernelGGL((dg_apply_box<T, DIM, NB, KK, true>), dim3(grid), dim3(threads), lds, s, x, y, a, bx,     \
                         total);                                                                                  \
    else                                                                                                          \
      hipLaunchKernelGGL((dg_apply_box<T, DIM, NB, KK, false>), dim3(grid), dim3(threads), lds, s, x, y, a, bx,    \
                         total);                                                                                  \
    launched = true;                                                                                              \
  }
  BOX(2, 3, 2) BOX(2, 3, 4) BOX(3, 4, 2) BOX(3, 4, 4)
#undef BOX
  return launched;
}

// The strip-march layout of an operator: columns per strip (as many as 256
// threads hold), units, threads, the list phase's tile and the LDS; false
// where the kernel does not apply (as box_layout)
bool strip_layout(const Operator &op, size_t ts, BoxArgs &bx, int &threads, int &cpb, int &lanes, size_t &lds) {
  const Disc &D = *op.disc;
  const int A = D.n_angles, nb = D.nb, nf = D.nf, dim = D.dim;
  if (op.box_cls < 0 || A > 256) return false;
  const int W = kBoxThreads / A;
  threads = ((W * A + 63) / 64) * 64;
  if (nb * A > threads) return false;  // one halo element per thread
  lds = strip_lds_reals(A, W, nb) * ts;
  // the boundary-cell phase: the tile body, cpb cells per tile
  const size_t geo = (size_t)(2 * nf * nb * nb + nf * dim + (nb - 1) * dim + 2) * ts + sizeof(int) * (size_t)nf;
  auto tile_lds = [&](int cp) { return (size_t)(cp * nb * A + cp * nb) * ts + (size_t)cp * geo; };
  cpb = threads / A;
  while (cpb > 1 && tile_lds(cpb) > 65536) cpb--;
  if (tile_lds(cpb) > 65536 || lds > 65536) return false;
  lds = std::max(lds, tile_lds(cpb));
  int Se = pow2_floor(std::max(1, threads / (cpb * nb)));
  Se = std::min(std::min(Se, 64), pow2_floor(A));
  lanes = std::max(1, Se);
  bx.edge_tiles = (op.n_box_edge + cpb - 1) / cpb;
  for (int d = 0; d < 3; d++) {
    bx.n[d] = op.box_n[d];
    bx.fmap[d][0] = op.box_fmap[d][0];
    bx.fmap[d][1] = op.box_fmap[d][1];
  }
  bx.cls = op.box_cls;
  bx.W = W;
  bx.tiles[0] = bx.tiles[1] = bx.tiles[2] = 0;
  bx.R = kStripRows;
  bx.units[0] = (op.box_n[0] + W - 1) / W;
  bx.units[1] = dim == 3 ? op.box_n[1] : 1;
  bx.units[2] = (op.box_n[dim - 1] + bx.R - 1) / bx.R;
  int S = pow2_floor(std::max(1, threads / (W * nb)));
  S = std::min(std::min(S, 64), pow2_floor(A));
  bx.lanes = std::max(1, S);
  bx.ok = op.box_ok.as<unsigned char>();  // read per cell by the strip march
  return true;
}

template <class T>
bool launch_apply_strip(const Operator &op, OpArgs<T> a, const T *x, T *y, bool scatter, hipStream_t s,
                        bool dry = false) {
  const Disc &D = *op.disc;
  BoxArgs bx;
  int threads = 0, cpb = 1, lanes = 1;
  size_t lds = 0;
  if (!strip_layout(op, sizeof(T), bx, threads, cpb, lanes, lds)) return false;
  if (D.nb != (D.dim == 2 ? 3 : 4)) return false;
  if (dry) return true;
  a.cpb = cpb;
  a.red_lanes = lanes;
  const int64_t total = (int64_t)bx.units[0] * bx.units[1] * bx.units[2];
  int64_t g = std::min<int64_t>(total, (int64_t)device_cus(op.device) * 8);
  g = std::max<int64_t>(8, (g + 7) / 8 * 8);
  const unsigned grid = (unsigned)g;
#define STRIP(DIM, NB)                                                                                      \
  if (D.dim == DIM) {                                                                                       \
    if (scatter)                                                                                            \
      hipLaunchKernelGGL((dg_apply_strip<T, DIM, NB, kStripAhead<T>, true>), dim3(grid), dim3(threads), lds, s, x, y, a, bx, \
                         total);                                                                            \
    else                                                                                                    \
      hipLaunchKernelGGL((dg_apply_strip<T, DIM, NB, kStripAhead<T>, false>), dim3(grid), dim3(threads), lds, s, x, y, a,    \
                         bx, total);                                                                        \
    return true;                                                                                            \
  }
  STRIP(2, 3) STRIP(3, 4)
#undef STRIP
  return false;
}

// The structured multi-cell apply, CPT cells per lane column
template <class T>
bool launch_apply_fd_mc(const Operator &op, OpArgs<T> a, const T *x, T *y, bool scatter, hipStream_t s, int cpt,
                        bool dry = false) {
  const Disc &D = *op.disc;
  const int A = D.n_angles;
  if (D.kind != DISC_FD || A > 256) return false;
  const int R = 256 / A;
  const int64_t per = (int64_t)R * cpt;
  const size_t lds = (size_t)(per * A + per) * sizeof(T);
  if (lds > 65536) return false;
  if (dry) return true;
  const int threads = ((R * A + 63) / 64) * 64;
  int S = pow2_floor(std::max<int64_t>(1, threads / per));
  S = std::min(std::min(S, 64), pow2_floor(A));
  a.cpb = R;
  a.red_lanes = std::max(1, S);
  const int64_t n_tiles = (D.n_cells + per - 1) / per;
  const unsigned grid = (unsigned)(8 * ((n_tiles + 7) / 8));
#define FDMC(DIM, CPT)                                                                                        \
  if (D.dim == DIM && cpt == CPT) {                                                                           \
    if (scatter)                                                                                              \
      hipLaunchKernelGGL((fd_apply_mc<T, DIM, CPT, true>), dim3(grid), dim3(threads), lds, s, x, y, a);       \
    else                                                                                                      \
      hipLaunchKernelGGL((fd_apply_mc<T, DIM, CPT, false>), dim3(grid), dim3(threads), lds, s, x, y, a);      \
    return true;                                                                                              \
  }
  FDMC(2, 4) FDMC(2, 8) FDMC(3, 4) FDMC(3, 8)
#undef FDMC
  return false;
}

// The z-march apply (3D FD): planes walked in chunks of 32 per workgroup,
// 4 (cpt8 false) or 8 cells per lane column
template <class T>
bool launch_apply_fd_zm(const Operator &op, OpArgs<T> a, const T *x, T *y, bool scatter, hipStream_t s, bool cpt8,
                        bool dry = false) {
  const Disc &D = *op.disc;
  const int A = D.n_angles;
  if (D.kind != DISC_FD || D.dim != 3 || A > 256) return false;
  const int CPT = cpt8 ? 8 : 4;
  const int R = 256 / A, XR = R * CPT;
  const int threads = ((R * A + 63) / 64) * 64;
  const size_t lds = (size_t)(((A + 1) & ~1) + XR * A + XR + 4 * XR) * sizeof(T);
  // a plane's rows are addressed with 32-bit offsets
  const int64_t prow = (int64_t)D.n[0] * D.n[1] * A;
  if (lds > 65536 || 2 * XR > threads || prow * (int64_t)sizeof(T) >= ((int64_t)1 << 31)) return false;
  if (dry) return true;
  int S = pow2_floor(std::max(1, threads / XR));
  S = std::min(std::min(S, 64), pow2_floor(A));
  a.cpb = R;
  a.red_lanes = std::max(1, S);
  const int nz = D.n[2];
  const int zc = std::min(nz, 32);
  const int64_t n_tiles = (int64_t)((D.n[0] + XR - 1) / XR) * D.n[1] * ((nz + zc - 1) / zc);
  const unsigned grid = (unsigned)(8 * ((n_tiles + 7) / 8));
#define ZM(C)                                                                                                \
  if (CPT == C) {                                                                                            \
    if (scatter)                                                                                             \
      hipLaunchKernelGGL((fd_apply_zm<T, C, true>), dim3(grid), dim3(threads), lds, s, x, y, a, zc);         \
    else                                                                                                     \
      hipLaunchKernelGGL((fd_apply_zm<T, C, false>), dim3(grid), dim3(threads), lds, s, x, y, a, zc);        \
  }
  ZM(4) ZM(8)
#undef ZM
  return true;
}

template <class T>
bool launch_apply_mc(const Operator &op, OpArgs<T> a, const T *x, T *y, bool scatter, hipStream_t s, int cpt) {
  const Disc &D = *op.disc;
  const int A = D.n_angles, nb = D.nb, nf = D.nf, dim = D.dim;
  if (D.kind == DISC_FD) return launch_apply_fd_mc<T>(op, a, x, y, scatter, s, cpt == 2 ? 4 : 8);
  if (A > 256) return false;
  const int R = 256 / A, tcells = R * cpt;
  const size_t reals = (size_t)tcells * ((nb > 1 ? 2 * nf * nb * nb + (nb - 1) * dim : 0) + nb * A + nb);
  const size_t lds = reals * sizeof(T);
  if (lds > 65536) return false;
  const int threads = ((R * A + 63) / 64) * 64;
  int S = pow2_floor(std::max(1, threads / (tcells * nb)));
  S = std::min(std::min(S, 64), pow2_floor(A));
  a.cpb = R;
  a.red_lanes = std::max(1, S);
  const int64_t n_tiles = (D.n_cells + tcells - 1) / tcells;
  const unsigned grid = (unsigned)(8 * ((n_tiles + 7) / 8));
  bool launched = false;
#define MC(DIM, NB, NF, CPT)                                                                                      \
  if (!launched && dim == DIM && nb == NB && nf == NF && cpt == CPT) {                                          \
    if (scatter)                                                                                                \
      hipLaunchKernelGGL((dg_apply_mc<T, DIM, NB, NF, CPT, true>), dim3(grid), dim3(threads), lds, s, x, y, a);  \
    else                                                                                                        \
      hipLaunchKernelGGL((dg_apply_mc<T, DIM, NB, NF, CPT, false>), dim3(grid), dim3(threads), lds, s, x, y, a); \
    launched = true;                                                                                            \
  }
#define MC_ALL(CPT) MC(2, 1, 3, CPT) MC(2, 1, 4, CPT) MC(3, 1, 4, CPT) MC(3, 1, 6, CPT) \
                    MC(2, 3, 3, CPT) MC(2, 3, 4, CPT) MC(3, 4, 4, CPT) MC(3, 4, 6, CPT)
  MC_ALL(2)
  MC_ALL(4)
#undef MC_ALL
#undef MC
  return launched;
}

template <class T>
void launch_apply(const Operator &op, const OpArgs<T> &a0, const T *x, T *y, bool scatter, hipStream_t s,
                  bool adj = false) {
  const Disc &D = *op.disc;
  OpArgs<T> a = a0;
  const int var = op.variant[sizeof(T) == 8 ? 0 : 1];
  if (!adj && (var == 8 || var == 9) && launch_apply_box<T>(op, a, x, y, scatter, s, var == 8 ? 2 : 4)) {
    UBT_HIP(hipGetLastError());
    return;
  }
  if (!adj && var == 10 && launch_apply_strip<T>(op, a, x, y, scatter, s)) {
    UBT_HIP(hipGetLastError());
    return;
  }
  if (!adj && var == 7 && launch_apply_q4<T>(op, a, x, y, scatter, s)) {
    UBT_HIP(hipGetLastError());
    return;
  }
  if (!adj && var >= 4 && var <= 6 && launch_apply_pencil<T>(op, a, x, y, scatter, s, 8 << (var - 4))) {
    UBT_HIP(hipGetLastError());
    return;
  }
  if (!adj && (var == 2 || var == 3) && launch_apply_mc<T>(op, a, x, y, scatter, s, var == 2 ? 2 : 4)) {
    UBT_HIP(hipGetLastError());
    return;
  }
  if (!adj && (var == 11 || var == 12) && launch_apply_fd_zm<T>(op, a, x, y, scatter, s, var == 12)) {
    UBT_HIP(hipGetLastError());
    return;
  }
  if (!adj && var == 1 && launch_apply_wave<T>(op, a, x, y, scatter, s)) {
    UBT_HIP(hipGetLastError());
    return;
  }
  size_t geo = 0;
  if (D.kind == DISC_DG) {
    // the staged geometry in LDS, per cell
    const int nb = D.nb, nf = D.nf, dim = D.dim;
    geo = (size_t)((nb > 1 ? 2 * nf * nb * nb : 0) + nf * dim + (nb > 1 ? (nb - 1) * dim : 0) + 2) * sizeof(T) +
          sizeof(int) * (size_t)nf;
  }
  Launch L = apply_launch<T>(D.n_angles, D.nb, geo);
  a.cpb = L.cpb;
  a.red_lanes = L.lanes;
  const int64_t n_tiles = (D.n_cells + L.cpb - 1) / L.cpb;
  // XCD-contiguous tiles: 8 * ceil(tiles / 8) workgroups
  const unsigned grid = (unsigned)(8 * ((n_tiles + 7) / 8));
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
  a.gcls = op.gcls.as<int>();
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
  a.gcls_identity = op.gcls_identity ? 1 : 0;
  a.cells = op.box_edge.as<int>();
  a.n_list = op.n_box_edge;
  return a;
}

}  // namespace

template <> OpArgs<double> Operator::args<double>() const { return make_args<double>(*this, t64); }
template <> OpArgs<float> Operator::args<float>() const { return make_args<float>(*this, t32); }

// ---------------------------------------------------------------------------
// Operator construction: upload every table once
// ---------------------------------------------------------------------------
// Geometry classes: cells whose DG geometry record (face normals, face
// matrices, basis gradients, volume) agrees within 1e-13 of each quantity's
// largest magnitude share a class. Records are bucketed by a hash of their
// values rounded to 1e-9 of that scale, then compared within the bucket, so a
// record that straddles a rounding boundary only costs an extra class.
static void geometry_classes(const Disc &D, std::vector<int> &cls, std::vector<int> &rep) {
  const int nc = D.n_cells, nf = D.nf, nb = D.nb, FM = nb * nb;
  cls.assign(nc, 0);
  rep.clear();
  double scale[4] = {0, 0, 0, 0};  // nA, face matrices, gradients, volume
  for (double v : D.face_nA) scale[0] = std::max(scale[0], std::fabs(v));
  for (double v : D.face_own) scale[1] = std::max(scale[1], std::fabs(v));
  for (double v : D.face_up) scale[1] = std::max(scale[1], std::fabs(v));
  for (double v : D.basis_grad) scale[2] = std::max(scale[2], std::fabs(v));
  for (double v : D.volume) scale[3] = std::max(scale[3], std::fabs(v));
  for (double &s : scale) s = s > 0.0 ? s : 1.0;
  auto visit = [&](int c, auto &&fn) {
    for (int e = 0; e < 3 * nf; e++) fn(D.face_nA[(size_t)c * 3 * nf + e], 0);
    if (nb > 1) {
      for (int e = 0; e < nf * FM; e++) fn(D.face_own[(size_t)c * nf * FM + e], 1);
      for (int e = 0; e < nf * FM; e++) fn(D.face_up[(size_t)c * nf * FM + e], 1);
      for (int e = 0; e < 3 * nb; e++) fn(D.basis_grad[(size_t)c * 3 * nb + e], 2);
    }
    fn(D.volume[c], 3);
  };
  std::unordered_map<uint64_t, std::vector<int>> buckets;
  std::vector<double> rc, rr;
  for (int c = 0; c < nc; c++) {
    uint64_t h = 0x9E3779B97F4A7C15ull;
    visit(c, [&](double v, int kind) {
      const int64_t q = (int64_t)std::llround(v / scale[kind] * 1e9);
      h ^= (uint64_t)q + 0x9E3779B97F4A7C15ull + (h << 6) + (h >> 2);
    });
    auto &list = buckets[h];
    int found = -1;
    rc.clear();
    visit(c, [&](double v, int kind) { rc.push_back(v / scale[kind]); });
    for (int k : list) {
      rr.clear();
      visit(rep[k], [&](double v, int kind) { rr.push_back(v / scale[kind]); });
      double worst = 0.0;
      for (size_t e = 0; e < rc.size(); e++) worst = std::max(worst, std::fabs(rc[e] - rr[e]));
      if (worst <= 1e-13) {
        found = k;
        break;
      }
    }
    if (found < 0) {
      found = (int)rep.size();
      rep.push_back(c);
      list.push_back(found);
    }
    cls[c] = found;
  }
}

// A DG mesh is a lexicographic box when every cell's faces are axis-aligned in
// one local (axis, side) order and its neighbours are the grid neighbours of
// cell c = (k * ny + j) * nx + i (x fastest), boundary faces exactly on the
// grid's edges. The pencil kernel needs nothing else.
static bool detect_box(const Disc &D, int n[3], int fmap[3][2]) {
  if (D.kind != DISC_DG || D.nf != 2 * D.dim) return false;
  const int dim = D.dim, nf = D.nf, nc = D.n_cells;
  auto classify = [&](int64_t c, int f, int &ax, int &side) {
    const double *v = &D.face_nA[3 * ((size_t)c * nf + f)];
    ax = 0;
    for (int d = 1; d < dim; d++)
      if (std::fabs(v[d]) > std::fabs(v[ax])) ax = d;
    for (int d = 0; d < 3; d++)
      if (d != ax && std::fabs(v[d]) > 1e-12 * std::fabs(v[ax])) return false;
    side = v[ax] > 0.0 ? 1 : 0;
    return v[ax] != 0.0;
  };
  for (int d = 0; d < 3; d++) fmap[d][0] = fmap[d][1] = -1;
  for (int f = 0; f < nf; f++) {
    int ax, side;
    if (!classify(0, f, ax, side) || fmap[ax][side] >= 0) return false;
    fmap[ax][side] = f;
  }
  auto nbr = [&](int64_t c, int ax, int side) { return D.face_nb[(size_t)c * nf + fmap[ax][side]]; };
  int64_t c = 0;
  while (c + 1 < nc && nbr(c, 0, 1) == c + 1) c++;
  n[0] = (int)(c + 1);
  n[1] = n[2] = 1;
  if (dim == 2) {
    if (nc % n[0]) return false;
    n[1] = nc / n[0];
  } else {
    c = 0;
    while (c + n[0] < nc && nbr(c, 1, 1) == c + n[0]) c += n[0];
    n[1] = (int)(c / n[0] + 1);
    if (nc % ((int64_t)n[0] * n[1])) return false;
    n[2] = (int)(nc / ((int64_t)n[0] * n[1]));
  }
  const int64_t stride[3] = {1, n[0], (int64_t)n[0] * n[1]};
  for (c = 0; c < nc; c++) {
    const int idx[3] = {(int)(c % n[0]), (int)((c / n[0]) % n[1]), (int)(c / stride[2])};
    for (int f = 0; f < nf; f++) {
      int ax, side;
      if (!classify(c, f, ax, side) || fmap[ax][side] != f) return false;
      const int o = idx[ax] + (side ? 1 : -1);
      const int got = D.face_nb[(size_t)c * nf + f];
      if (o >= 0 && o < n[ax]) {
        if (got != c + (side ? stride[ax] : -stride[ax])) return false;
      } else if (got >= 0) {
        return false;
      }
    }
  }
  return true;
}

// The geometry class of a box's interior cells, or -1 where the box-tile
// kernel does not apply: a box thinner than 3 cells along an axis, a face
// normal with an off-axis component (the upwind side must follow from the
// sign of one cosine, exactly, in either precision), or an ordinate with a
// zero cosine along an axis of the box.
static int box_interior_class(const Disc &D, const std::vector<int> &cls, const int n[3]) {
  const int dim = D.dim, nf = D.nf;
  for (int d = 0; d < dim; d++)
    if (n[d] < 3) return -1;
  for (size_t e = 0; e < (size_t)D.n_cells * nf; e++) {
    int nonzero = 0;
    for (int d = 0; d < 3; d++) nonzero += D.face_nA[3 * e + d] != 0.0;
    if (nonzero != 1) return -1;
  }
  for (int a = 0; a < D.n_angles; a++)
    for (int d = 0; d < dim; d++)
      if (D.quad.omega[3 * (size_t)a + d] == 0.0) return -1;
  const int64_t c = dim == 2 ? (int64_t)n[0] + 1 : ((int64_t)n[1] + 1) * n[0] + 1;
  return cls[c];
}

// Per cell whether the box kernels' formula holds: a cell of the interior
// class, or a cell whose faces on the box boundary are all vacuum and whose
// geometry record matches the interior class's within 2e-13 of each
// quantity's scale except for the upwind face matrices of those faces, which
// the ghost-flux treatment leaves zero: the kernels read a zero trace there,
// so the product is the same. (Twice the class tolerance: a box's cell widths
// differ in the last bits of its coordinates, by up to 1.0e-13 of the face
// matrices' scale on a 253 x 252 box, and a cell that misses by that much
// would cost a list phase; the apply then differs from the cell's own
// geometry by at most that fraction, well inside the 1e-12 parity bar.)
static void box_formula_cells(const Disc &D, const std::vector<int> &cls, int rep, int box_cls,
                              std::vector<unsigned char> &ok) {
  const int nc = D.n_cells, nf = D.nf, nb = D.nb, FM = nb * nb;
  double scale[4] = {0, 0, 0, 0};  // nA, face matrices, gradients, volume
  for (double v : D.face_nA) scale[0] = std::max(scale[0], std::fabs(v));
  for (double v : D.face_own) scale[1] = std::max(scale[1], std::fabs(v));
  for (double v : D.face_up) scale[1] = std::max(scale[1], std::fabs(v));
  for (double v : D.basis_grad) scale[2] = std::max(scale[2], std::fabs(v));
  for (double v : D.volume) scale[3] = std::max(scale[3], std::fabs(v));
  for (double &v : scale) v = (v > 0.0 ? v : 1.0) * 2e-13;
  auto same = [&](double a, double b, int k) { return std::fabs(a - b) <= scale[k]; };
  ok.assign(nc, 0);
  for (int c = 0; c < nc; c++) {
    if (cls[c] == box_cls) {
      ok[c] = 1;
      continue;
    }
    bool m = same(D.volume[c], D.volume[rep], 3);
    for (int e = 0; e < 3 * nb && m; e++) m = same(D.basis_grad[(size_t)c * 3 * nb + e], D.basis_grad[(size_t)rep * 3 * nb + e], 2);
    for (int f = 0; f < nf && m; f++) {
      const int nbr = D.face_nb[(size_t)c * nf + f];
      if (nbr < -1) m = false;  // reflective: the mirrored angle of the same cell
      for (int e = 0; e < 3 && m; e++) m = same(D.face_nA[((size_t)c * nf + f) * 3 + e], D.face_nA[((size_t)rep * nf + f) * 3 + e], 0);
      for (int e = 0; e < FM && m; e++) {
        const size_t kc = ((size_t)c * nf + f) * FM + e, kr = ((size_t)rep * nf + f) * FM + e;
        m = same(D.face_own[kc], D.face_own[kr], 1) && (nbr == -1 || same(D.face_up[kc], D.face_up[kr], 1));
      }
    }
    ok[c] = m ? 1 : 0;
  }
}

static void upload_prec(PrecTables &t, const Disc &D, const std::vector<int> &rep, bool fp32) {
  std::vector<double> omega(D.quad.omega), w(D.quad.w);
  std::vector<double> face_s, face_own, face_up, grad, vol;
  if (D.kind == DISC_DG) {
    const int nf = D.nf, FM = D.nb * D.nb, ncls = (int)rep.size();
    face_s.assign(3 * (size_t)ncls * nf, 0.0);
    for (int k = 0; k < ncls; k++) {
      const int r = rep[k];
      const double sc = D.order == 0 ? 1.0 / D.volume[r] : 1.0;
      for (int e = 0; e < 3 * nf; e++) face_s[(size_t)k * 3 * nf + e] = D.face_nA[(size_t)r * 3 * nf + e] * sc;
      vol.push_back(D.volume[r]);
      if (D.nb == 1) continue;  // DG0 has no basis gradients or face matrices
      for (int e = 0; e < 3 * D.nb; e++) grad.push_back(D.basis_grad[(size_t)r * 3 * D.nb + e]);
      for (int e = 0; e < nf * FM; e++) {
        face_own.push_back(D.face_own[(size_t)r * nf * FM + e]);
        face_up.push_back(D.face_up[(size_t)r * nf * FM + e]);
      }
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
    t.face_own.upload_as<float>(face_own);
    t.face_up.upload_as<float>(face_up);
    t.grad.upload_as<float>(grad);
    t.vol.upload_as<float>(vol);
  } else {
    t.wext.upload(wext);
    t.wunit.upload(wunit);
    t.omega.upload(omega);
    t.w.upload(w);
    t.face_s.upload(face_s);
    t.face_own.upload(face_own);
    t.face_up.upload(face_up);
    t.grad.upload(grad);
    t.vol.upload(vol);
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
  if (D.kind == DISC_DG) {
    std::vector<int> cls;
    geometry_classes(D, cls, op->geo_rep);
    op->n_geo_classes = (int)op->geo_rep.size();
    op->gcls_identity = op->n_geo_classes == D.n_cells;
    for (int c = 0; c < D.n_cells && op->gcls_identity; c++) op->gcls_identity = cls[c] == c;
    op->gcls.upload(cls);
    op->box = detect_box(D, op->box_n, op->box_fmap);
    if (op->box && D.nb > 1) op->box_cls = box_interior_class(D, cls, op->box_n);
    if (op->box_cls >= 0) {
      // the cells the box kernels leave to their list phase: those the box
      // formula does not cover (a reflective face, another geometry)
      std::vector<unsigned char> ok;
      box_formula_cells(D, cls, op->geo_rep[op->box_cls], op->box_cls, ok);
      std::vector<int> edge;
      for (int64_t c = 0; c < D.n_cells; c++)
        if (!ok[c]) edge.push_back((int)c);
      op->box_edge.upload(edge);
      op->n_box_edge = (int)edge.size();
      op->box_ok.upload(ok);
    }
  }
  upload_prec(op->t64, D, op->geo_rep, false);
  upload_prec(op->t32, D, op->geo_rep, true);
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
  // the first group anything upscatters into (reference
  // tests/transportk.kokkos.cxx:953-962, over cells instead of materials)
  op.g_up = (int)G;
  for (size_t gf = 1; gf < G; gf++)
    for (size_t gt = 0; gt < gf && (int)gt < op.g_up; gt++) {
      const double *v = ss.data() + (gf * G + gt) * nc;
      if (std::any_of(v, v + nc, [](double x) { return x != 0.0; })) op.g_up = (int)gt;
    }
  const bool first = !op.xs_set;
  if (source) {
    std::vector<double> q(source, source + G * nc);
    op.source.upload(q);
  } else {
    op.source.release();
  }
  op.xs_set = true;
  op.xs_gen++;
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
  auto applicable = [&](int v, int p) {
    if (v == 0) return true;
    if (v == 1) return wave_ok;
    if (v == 8 || v == 9)
      return p == 0 ? launch_apply_box<double>(*this, args<double>(), nullptr, nullptr, true, nullptr, 2, true)
                    : launch_apply_box<float>(*this, args<float>(), nullptr, nullptr, true, nullptr, 2, true);
    if (v == 10)
      return p == 0 ? launch_apply_strip<double>(*this, args<double>(), nullptr, nullptr, true, nullptr, true)
                    : launch_apply_strip<float>(*this, args<float>(), nullptr, nullptr, true, nullptr, true);
    if (v == 11 || v == 12)
      return p == 0 ? launch_apply_fd_zm<double>(*this, args<double>(), nullptr, nullptr, true, nullptr, v == 12, true)
                    : launch_apply_fd_zm<float>(*this, args<float>(), nullptr, nullptr, true, nullptr, v == 12, true);
    if (v == 7)
      return p == 0 ? launch_apply_q4<double>(*this, args<double>(), nullptr, nullptr, true, nullptr, true)
                    : launch_apply_q4<float>(*this, args<float>(), nullptr, nullptr, true, nullptr, true);
    if (v >= 4)
      return p == 0 ? launch_apply_pencil<double>(*this, args<double>(), nullptr, nullptr, true, nullptr, 8, true)
                    : launch_apply_pencil<float>(*this, args<float>(), nullptr, nullptr, true, nullptr, 8, true);
    if (D.kind == DISC_FD)
      return p == 0 ? launch_apply_fd_mc<double>(*this, args<double>(), nullptr, nullptr, true, nullptr, v == 2 ? 4 : 8,
                                                 true)
                    : launch_apply_fd_mc<float>(*this, args<float>(), nullptr, nullptr, true, nullptr, v == 2 ? 4 : 8,
                                                true);
    if (D.n_angles > 256) return false;
    const int R = 256 / D.n_angles, tcells = R * (v == 2 ? 2 : 4);
    const size_t reals = (size_t)tcells * ((D.nb > 1 ? 2 * D.nf * D.nb * D.nb + (D.nb - 1) * D.dim : 0) +
                                           D.nb * D.n_angles + D.nb);
    return reals * (p == 0 ? 8 : 4) <= 65536;
  };
  hipStream_t s = nullptr;
  UBT_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
  hipEvent_t e0, e1;
  UBT_HIP(hipEventCreate(&e0));
  UBT_HIP(hipEventCreate(&e1));
  DevBuf xb, yb;
  xb.alloc(sizeof(double) * n);
  yb.alloc(sizeof(double) * n);
  vec_fill<double>(n, 1.0, xb.as<double>(), s);  // tuned on a non-zero input
  for (int p = 0; p < 2; p++) {
    float best = 1e30f;
    int best_v = 0;
    for (int v = 0; v < kApplyVariants; v++) {
      tuned_us[p][v] = 0.f;
      if (!applicable(v, p)) continue;
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
  size_t lds = (size_t)64 * Apad * sizeof(T);
  UBT_CHECK(lds <= 65536, UBT_ERR_SUP, "the moments tile of %d angles needs %zu bytes of LDS", A, lds);
  const int w_in_lds = lds + (size_t)16 * Apad * sizeof(T) <= 65536 ? 1 : 0;
  if (w_in_lds) lds += (size_t)16 * Apad * sizeof(T);
  hipLaunchKernelGGL((moments_mfma<T, OutT>), dim3(grid), dim3(256), lds, s, psi, phi,
                     unit ? t.wunit.as<T>() : t.wext.as<T>(), N, A, Apad, M, stride, w_in_lds);
  UBT_HIP(hipGetLastError());
}

// The outer iteration's change measure: cache = fresh while summing
// (fresh - cache)^2 and fresh^2, per block, then one block adds the partials
// in block order (deterministic) to sums[0..1]
__global__ void __launch_bounds__(256) measure_overwrite(const double *__restrict__ fresh, double *__restrict__ cache,
                                                         int64_t n, double *__restrict__ part) {
  __shared__ double sd[256], sq[256];
  double d = 0.0, q = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256) {
    const double v = fresh[i], o = cache[i];
    d += (v - o) * (v - o);
    q += v * v;
    cache[i] = v;
  }
  sd[threadIdx.x] = d;
  sq[threadIdx.x] = q;
  __syncthreads();
  for (int k = 128; k > 0; k >>= 1) {
    if ((int)threadIdx.x < k) {
      sd[threadIdx.x] += sd[threadIdx.x + k];
      sq[threadIdx.x] += sq[threadIdx.x + k];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    part[2 * blockIdx.x] = sd[0];
    part[2 * blockIdx.x + 1] = sq[0];
  }
}

__global__ void __launch_bounds__(256) measure_final(const double *__restrict__ part, int nblk, double *sums) {
  __shared__ double sd[256], sq[256];
  double d = 0.0, q = 0.0;
  for (int i = threadIdx.x; i < nblk; i += 256) {
    d += part[2 * i];
    q += part[2 * i + 1];
  }
  sd[threadIdx.x] = d;
  sq[threadIdx.x] = q;
  __syncthreads();
  for (int k = 128; k > 0; k >>= 1) {
    if ((int)threadIdx.x < k) {
      sd[threadIdx.x] += sd[threadIdx.x + k];
      sq[threadIdx.x] += sq[threadIdx.x + k];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    sums[0] += sd[0];
    sums[1] += sq[0];
  }
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

void Operator::set_scalar_flux(int g, int prec, const void *psi, hipStream_t s, double *sums) {
  UBT_CHECK(g >= 0 && g < n_groups, UBT_ERR_ARG_OUTOFRANGE, "group %d out of range, n_groups is %d", g, n_groups);
  const int64_t nodes = disc->n_nodes();
  double *cache = phi_cache.as<double>() + (int64_t)g * nodes;
  // measured: integrate into the scratch, then overwrite the cache from it
  // while summing the change (reference GroupTransfer::set_scalar_flux,
  // src/multigroupk.kokkos.cxx:266-299)
  if (sums && phi_tmp.bytes < sizeof(double) * (size_t)nodes) phi_tmp.alloc(sizeof(double) * (size_t)nodes);
  double *dst = sums ? phi_tmp.as<double>() : cache;
  if (prec == UBT_FP64)
    launch_moments<double, double>(*this, t64, (const double *)psi, dst, 1, 1, s);
  else
    launch_moments<float, double>(*this, t32, (const float *)psi, dst, 1, 1, s);
  if (sums) {
    if (!measure_part.bytes) measure_part.alloc(sizeof(double) * 2 * kMeasureBlocks);
    const unsigned nblk = (unsigned)std::min<int64_t>(kMeasureBlocks, (nodes + 255) / 256);
    hipLaunchKernelGGL(measure_overwrite, dim3(nblk), dim3(256), 0, s, (const double *)dst, cache, nodes,
                       measure_part.as<double>());
    hipLaunchKernelGGL(measure_final, dim3(1), dim3(256), 0, s, (const double *)measure_part.as<double>(), (int)nblk, sums);
    UBT_HIP(hipGetLastError());
  }
  phi_set[g] = 1;
}

namespace {
// psi(node, a) = phi(node) / W + sum_d p1(a, d) J(node, d), J the columns
// 1 .. dim of the stored moments
__global__ void expand_p1_kernel(double *__restrict__ psi, const double *__restrict__ phi,
                                 const double *__restrict__ mom, const double *__restrict__ p1, int A, int dim,
                                 double inv_w, int64_t rows) {
  const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= rows) return;
  const int64_t node = r / A;
  const int a = (int)(r - node * A);
  double v = phi[node] * inv_w;
  for (int d = 0; d < dim; d++) v += p1[3 * a + d] * mom[node * (1 + dim) + 1 + d];
  psi[r] = v;
}

// per block: the sum of x^2 (and a zero, the layout measure_final reads)
__global__ void __launch_bounds__(256) sumsq_part(const double *__restrict__ x, int64_t n, double *__restrict__ part) {
  __shared__ double sq[256];
  double q = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256) q += x[i] * x[i];
  sq[threadIdx.x] = q;
  __syncthreads();
  for (int k = 128; k > 0; k >>= 1) {
    if ((int)threadIdx.x < k) sq[threadIdx.x] += sq[threadIdx.x + k];
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    part[2 * blockIdx.x] = sq[0];
    part[2 * blockIdx.x + 1] = 0.0;
  }
}
}  // namespace

void Operator::store_moments(int g, const double *psi, hipStream_t s) {
  UBT_CHECK(g >= 0 && g < n_groups, UBT_ERR_ARG_OUTOFRANGE, "group %d out of range, n_groups is %d", g, n_groups);
  const int M = 1 + disc->dim;
  const int64_t nodes = disc->n_nodes();
  if (!mom_store.bytes) {
    mom_store.alloc(sizeof(double) * (size_t)n_groups * nodes * M);
    UBT_HIP(hipMemsetAsync(mom_store.ptr, 0, mom_store.bytes, s));
  }
  launch_moments<double, double>(*this, t64, psi, mom_store.as<double>() + (int64_t)g * nodes * M, M, M, s);
}

void Operator::zero_moments(int g, hipStream_t s) {
  UBT_CHECK(g >= 0 && g < n_groups, UBT_ERR_ARG_OUTOFRANGE, "group %d out of range, n_groups is %d", g, n_groups);
  const int M = 1 + disc->dim;
  const int64_t nodes = disc->n_nodes();
  if (!mom_store.bytes) mom_store.alloc(sizeof(double) * (size_t)n_groups * nodes * M);
  UBT_HIP(hipMemsetAsync(mom_store.as<double>() + (int64_t)g * nodes * M, 0, sizeof(double) * (size_t)nodes * M, s));
}

void Operator::expand_p1(int g, double *psi, hipStream_t s) {
  UBT_CHECK(g >= 0 && g < n_groups, UBT_ERR_ARG_OUTOFRANGE, "group %d out of range, n_groups is %d", g, n_groups);
  UBT_CHECK(phi_set[g] && mom_store.bytes, UBT_ERR_ARG_WRONGSTATE,
            "no scalar flux and current stored for group %d to expand", g);
  const Disc &D = *disc;
  const int A = D.n_angles, dim = D.dim;
  if (!p1.bytes) {
    // T = sum_a w_a Omega_a Omega_a^T; the first moment must vanish for the
    // expansion's scalar flux to be phi (reference multigroupk:384-433)
    const Quadrature &q = D.quad;
    double t[3][3] = {}, first[3] = {};
    for (int a = 0; a < A; a++)
      for (int d = 0; d < dim; d++) {
        first[d] += q.w[a] * q.omega[3 * a + d];
        for (int e = 0; e < dim; e++) t[d][e] += q.w[a] * q.omega[3 * a + d] * q.omega[3 * a + e];
      }
    for (int d = 0; d < dim; d++)
      UBT_CHECK(std::fabs(first[d]) <= 1e-12 * std::fabs(q.sum_weights), UBT_ERR_ARG_INCOMP,
                "the P1 expansion needs a quadrature whose weighted directions sum to zero: sum_a w_a Omega_a[%d] is %g",
                d, first[d]);
    // T^{-1} by Gauss-Jordan with partial pivoting
    double inv[3][3] = {};
    for (int d = 0; d < dim; d++) inv[d][d] = 1.0;
    for (int k = 0; k < dim; k++) {
      int piv = k;
      for (int i = k + 1; i < dim; i++)
        if (std::fabs(t[i][k]) > std::fabs(t[piv][k])) piv = i;
      UBT_CHECK(t[piv][k] != 0.0, UBT_ERR_ARG_INCOMP, "the quadrature's second moment sum_a w_a Omega_a Omega_a^T is singular");
      std::swap(t[k], t[piv]);
      std::swap(inv[k], inv[piv]);
      const double f = 1.0 / t[k][k];
      for (int j = 0; j < dim; j++) {
        t[k][j] *= f;
        inv[k][j] *= f;
      }
      for (int i = 0; i < dim; i++) {
        if (i == k) continue;
        const double c = t[i][k];
        for (int j = 0; j < dim; j++) {
          t[i][j] -= c * t[k][j];
          inv[i][j] -= c * inv[k][j];
        }
      }
    }
    std::vector<double> h(3 * (size_t)A, 0.0);
    for (int a = 0; a < A; a++)
      for (int d = 0; d < dim; d++)
        for (int e = 0; e < dim; e++) h[3 * a + d] += inv[d][e] * q.omega[3 * a + e];
    p1.upload(h);
  }
  const int64_t rows = D.n_rows(), nodes = D.n_nodes();
  hipLaunchKernelGGL(expand_p1_kernel, dim3((unsigned)((rows + 255) / 256)), dim3(256), 0, s, psi,
                     (const double *)phi_cache.as<double>() + (int64_t)g * nodes,
                     (const double *)mom_store.as<double>() + (int64_t)g * nodes * (1 + dim),
                     (const double *)p1.as<double>(), A, dim, 1.0 / D.quad.sum_weights, rows);
  UBT_HIP(hipGetLastError());
}

double Operator::norm2(const double *x, hipStream_t s) {
  const int64_t n = disc->n_rows();
  if (!measure_part.bytes) measure_part.alloc(sizeof(double) * 2 * kMeasureBlocks);
  if (!norm_sums.bytes) norm_sums.alloc(2 * sizeof(double));
  double *sums = norm_sums.as<double>();
  UBT_HIP(hipMemsetAsync(sums, 0, 2 * sizeof(double), s));
  const unsigned nblk = (unsigned)std::min<int64_t>(kMeasureBlocks, (n + 255) / 256);
  hipLaunchKernelGGL(sumsq_part, dim3(nblk), dim3(256), 0, s, x, n, measure_part.as<double>());
  hipLaunchKernelGGL(measure_final, dim3(1), dim3(256), 0, s, (const double *)measure_part.as<double>(), (int)nblk, sums);
  UBT_HIP(hipGetLastError());
  double h[2];
  UBT_HIP(hipMemcpyAsync(h, sums, sizeof(h), hipMemcpyDeviceToHost, s));
  UBT_HIP(hipStreamSynchronize(s));
  return std::sqrt(h[0]);
}

void Operator::cache_zero_flux(int g, hipStream_t s) {
  UBT_CHECK(g >= 0 && g < n_groups, UBT_ERR_ARG_OUTOFRANGE, "group %d out of range, n_groups is %d", g, n_groups);
  const int64_t nodes = disc->n_nodes();
  UBT_HIP(hipMemsetAsync(phi_cache.as<double>() + (int64_t)g * nodes, 0, sizeof(double) * (size_t)nodes, s));
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

void Operator::rhs(int g, int prec, void *b, hipStream_t s, bool all_groups) const {
  UBT_CHECK(xs_set, UBT_ERR_ARG_WRONGSTATE, "set the cross sections first");
  UBT_CHECK(g >= 0 && g < n_groups, UBT_ERR_ARG_OUTOFRANGE, "group %d out of range, n_groups is %d", g, n_groups);
  const int g_end = all_groups ? n_groups : g;
  for (int gf = 0; gf < g_end; gf++)
    UBT_CHECK(gf == g || phi_set[gf], UBT_ERR_ARG_WRONGSTATE,
              "no scalar flux cached for group %d - set it when that group is solved", gf);
  RhsArgs r = make_rhs_args(*this, g, 0, g_end, g);
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
// counts o's singular element blocks into *cnt (which the caller zeroed)
void blockinv_check_launch(const Operator &o, int *cnt, hipStream_t s) {
  UBT_CHECK(o.xs_set, UBT_ERR_ARG_WRONGSTATE, "set the cross sections first");
  const Disc &D = *o.disc;
  const int64_t nblk = (int64_t)D.n_cells * D.n_angles;
  const unsigned grid = (unsigned)((nblk + 255) / 256);
  const int fd = D.kind == DISC_FD;
  switch (D.nb) {
    case 1: hipLaunchKernelGGL(blockinv_check<1>, dim3(grid), dim3(256), 0, s, o.args<double>(), fd, cnt); break;
    case 3: hipLaunchKernelGGL(blockinv_check<3>, dim3(grid), dim3(256), 0, s, o.args<double>(), fd, cnt); break;
    case 4: hipLaunchKernelGGL(blockinv_check<4>, dim3(grid), dim3(256), 0, s, o.args<double>(), fd, cnt); break;
    default: fail(UBT_ERR_SUP, "element blocks of %d basis functions are not supported", D.nb);
  }
  UBT_HIP(hipGetLastError());
}

void BlockInverse::setup(hipStream_t s) {
  const Operator &o = *op;
  UBT_CHECK(o.xs_set, UBT_ERR_ARG_WRONGSTATE, "set the cross sections first");
  const int64_t nblk = (int64_t)o.disc->n_cells * o.disc->n_angles;
  int *cnt = o.flag.as<int>();
  UBT_HIP(hipMemsetAsync(cnt, 0, sizeof(int), s));
  blockinv_check_launch(o, cnt, s);
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

// ---------------------------------------------------------------------------
// Batched launchers
// ---------------------------------------------------------------------------
bool batch_body_layout(const Operator &op, int prec, int body, int *cpb_out, int *lanes_out, size_t *lds_out,
                       int64_t *n_tiles, BoxArgs *box) {
  const Disc &D = *op.disc;
  if (D.kind != DISC_DG) return false;
  if (body == 2) {
    int threads = 0;
    if (!box_layout(op, prec == UBT_FP64 ? 8 : 4, kBatchBoxRows, *box, threads, *cpb_out, *lanes_out, *lds_out))
      return false;
    *n_tiles = (int64_t)box->tiles[0] * box->tiles[1] * box->tiles[2];
    return true;
  }
  if (body == 3) {
    int threads = 0;
    if (D.nb != (D.dim == 2 ? 3 : 4) ||
        !strip_layout(op, prec == UBT_FP64 ? 8 : 4, *box, threads, *cpb_out, *lanes_out, *lds_out))
      return false;
    *n_tiles = (int64_t)box->units[0] * box->units[1] * box->units[2];
    return true;
  }
  const int A = D.n_angles, nb = D.nb, nf = D.nf, dim = D.dim;
  const size_t ts = prec == UBT_FP64 ? 8 : 4;
  const size_t geo = (size_t)((nb > 1 ? 2 * nf * nb * nb : 0) + nf * dim + (nb > 1 ? (nb - 1) * dim : 0) + 2) * ts +
                     sizeof(int) * (size_t)nf;
  int cpb, lanes = 1;
  size_t lds;
  if (body == 0) {
    if (A > 256) return false;
    cpb = 256 / A;
    auto lds_for = [&](int cp) { return (size_t)(cp * nb * A + cp * nb) * ts + (size_t)cp * geo; };
    while (cpb > 1 && lds_for(cpb) > 65536) cpb--;
    lds = lds_for(cpb);
    int S = pow2_floor(std::max(1, 256 / (cpb * nb)));
    S = std::min(std::min(S, 64), pow2_floor(A));
    lanes = std::max(1, S);
  } else {
    if (A % 4 != 0 || A / 4 > 256) return false;
    const int TPC = A / 4;
    cpb = 256 / TPC;
    auto lds_for = [&](int cp) { return (size_t)cp * geo + (size_t)(cp * nb * TPC + cp * nb) * ts; };
    while (cpb > 1 && lds_for(cpb) > 65536) cpb--;
    lds = lds_for(cpb);
  }
  if (lds > 65536) return false;
  *cpb_out = cpb;
  *lanes_out = lanes;
  *lds_out = lds;
  *n_tiles = (D.n_cells + cpb - 1) / cpb;
  return true;
}

template <class T>
void batch_apply_launch(const Batch &b, int body, const T *x, T *y, bool scatter, hipStream_t s, bool quad) {
  const int pi = sizeof(T) == 8 ? 0 : 1;
  UBT_CHECK(!quad || (body == 2 && b.box_sweep[pi].ok), UBT_ERR_PLIB,
            "the quadrant-major apply needs the box body on a box-sweep batch");
  const int *qpos = quad ? b.box_sweep[pi].qpos.as<int>() : nullptr;
  const int64_t *qoff = quad ? b.box_sweep[pi].qoff.as<int64_t>() : nullptr;
  const int64_t total = b.n_tiles[pi][body];
  if (total == 0) return;
  int64_t g = std::min<int64_t>(total, (int64_t)b.n_cu * b.persist[pi]);
  g = std::max<int64_t>(8, (g + 7) / 8 * 8);
  const unsigned grid = (unsigned)g;
  const BatchItem<T> *items = b.items[pi][body].as<BatchItem<T>>();
  const int64_t *toff = b.tiles[pi][body].as<int64_t>();
  const size_t lds = b.lds[pi][body];
  bool launched = false;
#define BT(DIM, NB, NF, BODY)                                                                                       \
  if (!launched && b.dim == DIM && b.nb == NB && b.nf == NF && body == BODY) {                                     \
    if (scatter)                                                                                                    \
      hipLaunchKernelGGL((dg_apply_batch<T, DIM, NB, NF, BODY, true>), dim3(grid), dim3(256), lds, s, x, y, items, \
                         toff, total);                                                                              \
    else                                                                                                            \
      hipLaunchKernelGGL((dg_apply_batch<T, DIM, NB, NF, BODY, false>), dim3(grid), dim3(256), lds, s, x, y,       \
                         items, toff, total);                                                                       \
    launched = true;                                                                                                \
  }
#define BT_ALL(BODY) BT(2, 1, 3, BODY) BT(2, 1, 4, BODY) BT(3, 1, 4, BODY) BT(3, 1, 6, BODY) \
                     BT(2, 3, 3, BODY) BT(2, 3, 4, BODY) BT(3, 4, 4, BODY) BT(3, 4, 6, BODY)
  BT_ALL(0)
  BT_ALL(1)
#undef BT_ALL
#undef BT
  const int64_t *eoff = body >= 2 ? b.edge_off[pi][body].as<int64_t>() : nullptr;
  const int64_t n_edge = body >= 2 ? b.n_edge[pi][body] : 0;
#define BB(DIM, NB)                                                                                    \
  if (!launched && b.dim == DIM && b.nb == NB && body == 2) {                                          \
    if (scatter)                                                                                       \
      hipLaunchKernelGGL((dg_apply_batch_box<T, DIM, NB, kBatchBoxRows, true>), dim3(grid),             \
                         dim3(kBoxTileThreads<T, DIM>), lds, s, x, y, items, toff, total, eoff, n_edge, qpos, qoff); \
    else                                                                                               \
      hipLaunchKernelGGL((dg_apply_batch_box<T, DIM, NB, kBatchBoxRows, false>), dim3(grid),            \
                         dim3(kBoxTileThreads<T, DIM>), lds, s, x, y, items, toff, total, eoff, n_edge, qpos, qoff); \
    launched = true;                                                                                   \
  }
  BB(2, 3) BB(3, 4)
#undef BB
  const int nit = b.n_items[pi][body];
#define BS(DIM, NB)                                                                                    \
  if (!launched && b.dim == DIM && b.nb == NB && body == 3) {                                          \
    if (scatter)                                                                                       \
      hipLaunchKernelGGL((dg_apply_batch_strip<T, DIM, NB, kStripAhead<T>, true>), dim3(grid), dim3(kBoxThreads), lds, \
                         s, x, y, items, toff, nit, total, eoff, n_edge);                              \
    else                                                                                               \
      hipLaunchKernelGGL((dg_apply_batch_strip<T, DIM, NB, kStripAhead<T>, false>), dim3(grid), dim3(kBoxThreads), lds, \
                         s, x, y, items, toff, nit, total, eoff, n_edge);                              \
    launched = true;                                                                                   \
  }
  BS(2, 3) BS(3, 4)
#undef BS
  UBT_CHECK(launched, UBT_ERR_SUP, "no batched apply for dim %d, %d basis, %d faces", b.dim, b.nb, b.nf);
  UBT_HIP(hipGetLastError());
}

template <class T> void batch_blocks_launch(const Batch &b, const T *x, T *y, hipStream_t s) {
  const int pi = sizeof(T) == 8 ? 0 : 1;
  if (b.n_blocks == 0) return;
  const unsigned grid = (unsigned)((b.n_blocks + 255) / 256);
  const BatchItem<T> *items = b.items[pi][0].as<BatchItem<T>>();
  const int64_t *boff = b.bblk.as<int64_t>();
  const int n = b.n_items[pi][0];
  switch (b.nb) {
    case 1: hipLaunchKernelGGL((blockinv_batch<T, 1>), dim3(grid), dim3(256), 0, s, x, y, items, boff, n, b.n_blocks); break;
    case 3: hipLaunchKernelGGL((blockinv_batch<T, 3>), dim3(grid), dim3(256), 0, s, x, y, items, boff, n, b.n_blocks); break;
    case 4: hipLaunchKernelGGL((blockinv_batch<T, 4>), dim3(grid), dim3(256), 0, s, x, y, items, boff, n, b.n_blocks); break;
    default: fail(UBT_ERR_SUP, "element blocks of %d basis functions are not supported", b.nb);
  }
  UBT_HIP(hipGetLastError());
}

namespace {
// ---------------------------------------------------------------------------
// Batched sweep of uniform all-vacuum 2D DG1 boxes. On such a box every
// item's block is sigma_t(c) I + M(a) and its upwind couplings U_x(a),
// U_y(a) - the box apply's folded per-angle matrices - so no per-item record
// is stored or read: a workgroup owns one problem, one quadrant and up to AG
// of its angles, keeps those matrices in registers, and marches the
// quadrant's anti-diagonals i' + j' = d (i', j' counted from the quadrant's
// inflow corner). Cells of a diagonal are independent; their upwind traces
// are the previous diagonal's results, held in LDS by i' (x upwind at
// i' - 1, y upwind at i'), so the only global traffic is r in, psi out and
// sigma_t, r and psi in the quadrant-major copies (addressed by 32-bit
// offsets from the group's base). A thread solves up to IPT cells of a
// diagonal (AG angle lanes per cell); the next DEPTH diagonals' r and
// sigma_t are requested before the current one is solved, into register
// sets in rotation (the loop is unrolled by the rotation so no set is
// copied). Every load is unconditional - past the last diagonal the fetch
// repeats it, a cell past a diagonal's end reads its first cell - so at
// least one diagonal's loads always follow a set and the compiler's wait
// for that set never covers the prefetches behind it.
// ---------------------------------------------------------------------------
template <class T, int AG, int DEPTH, int IPT>
__device__ __forceinline__ void box_march(const T *__restrict__ rg, T *__restrict__ pg, const OpArgs<T> &g,
                                          const BoxArgs &bx, const BoxSweepGroup &gw, T *ring, int nx_max) {
  constexpr int NB = 3, SL = kBoxSweepThreads / AG;
  const int k = threadIdx.x % AG, slot = threadIdx.x / AG;
  const int a0 = gw.angle[k];
  const bool va = a0 >= 0;
  const int a = va ? a0 : gw.angle[0];
  const int nx = bx.n[0], ny = bx.n[1];
  const int Aq = g.A / 4, NA = NB * Aq;
  // a lane past the quadrant's angles reads lane 0's values (and stores
  // nothing)
  const int lane = va ? k : 0;
  // q = 2 sx + sy, sx = 0 for omega_x > 0 (the march runs up x from i = 0),
  // sy likewise for y
  const bool px = (gw.q >> 1) == 0, py = (gw.q & 1) == 0;
  auto cell_of = [&](int ip, int jp) -> unsigned {
    const int i = px ? ip : nx - 1 - ip, j = py ? jp : ny - 1 - jp;
    return (unsigned)(j * nx + i);
  };
  BoxConst<T, 2, NB> kc;
  box_consts<T, 2, NB>(g, bx, a, kc);
  // the block sigma_t I + M solved in closed form (Cayley-Hamilton for 3 x 3):
  // adj(s I + M) = s^2 I + s (tr M I - M) + adj M,
  // det(s I + M) = s^3 + tr M s^2 + c2 s + det M (c2: the principal 2 x 2
  // minors of M), so per item a few multiply-adds, one division, no pivoting
  const T (&M)[NB][NB] = kc.M;
  const T trM = M[0][0] + M[1][1] + M[2][2];
  T adjM[NB][NB];
  adjM[0][0] = M[1][1] * M[2][2] - M[1][2] * M[2][1];
  adjM[0][1] = M[0][2] * M[2][1] - M[0][1] * M[2][2];
  adjM[0][2] = M[0][1] * M[1][2] - M[0][2] * M[1][1];
  adjM[1][0] = M[1][2] * M[2][0] - M[1][0] * M[2][2];
  adjM[1][1] = M[0][0] * M[2][2] - M[0][2] * M[2][0];
  adjM[1][2] = M[0][2] * M[1][0] - M[0][0] * M[1][2];
  adjM[2][0] = M[1][0] * M[2][1] - M[1][1] * M[2][0];
  adjM[2][1] = M[0][1] * M[2][0] - M[0][0] * M[2][1];
  adjM[2][2] = M[0][0] * M[1][1] - M[0][1] * M[1][0];
  const T c2 = adjM[0][0] + adjM[1][1] + adjM[2][2];
  const T detM = M[0][0] * adjM[0][0] + M[0][1] * adjM[1][0] + M[0][2] * adjM[2][0];
  const int n_diag = nx + ny - 1;
  // the cells of diagonal d this thread solves: i' = lo + e, e = slot + m SL
  struct Set {
    T r[IPT][NB], st[IPT];
  };
  auto fetch = [&](int d, Set &s) {
    d = d < n_diag ? d : n_diag - 1;
    const int lo = d - (ny - 1) > 0 ? d - (ny - 1) : 0, hi = d < nx - 1 ? d : nx - 1;
    const int n = hi - lo;
#pragma unroll
    for (int m = 0; m < IPT; m++) {
      int e = slot + m * SL;
      e = e <= n ? e : 0;  // clamped: a valid cell, its values discarded
      const unsigned c = cell_of(lo + e, d - lo - e);
#pragma unroll
      for (int b = 0; b < NB; b++) s.r[m][b] = ldg(rg + (c * NA + b * Aq + lane));
      s.st[m] = ldg(g.sigt + c);
    }
  };
  // every loaded value is read once at the start of its solve (into sink,
  // stored on a path never taken), so no load of a set is still in flight
  // when its registers are reused - the clamped cells past a diagonal's end
  // would otherwise leave loads whose completion is first known at the
  // reuse, a wait for every load in flight
  T sink = T(0);
  auto solve = [&](int d, const Set &s) {
    const int lo = d - (ny - 1) > 0 ? d - (ny - 1) : 0, hi = d < nx - 1 ? d : nx - 1;
    const int n = hi - lo;
    const T *prev = ring + (size_t)((d + 1) & 1) * nx_max * AG * NB;
    T *cur = ring + (size_t)(d & 1) * nx_max * AG * NB;
#pragma unroll
    for (int m = 0; m < IPT; m++) sink += s.r[m][0] + s.r[m][1] + s.r[m][2] + s.st[m];
#pragma unroll
    for (int m = 0; m < IPT; m++) {
      const int e = slot + m * SL;
      if (e > n) break;
      const int ip = lo + e, jp = d - ip;
      T hx[NB], hy[NB], v[NB];
#pragma unroll
      for (int b = 0; b < NB; b++) {
        hx[b] = ip > 0 ? prev[((ip - 1) * AG + k) * NB + b] : T(0);
        hy[b] = jp > 0 ? prev[(ip * AG + k) * NB + b] : T(0);
      }
#pragma unroll
      for (int b = 0; b < NB; b++) {
        T w = s.r[m][b];
#pragma unroll
        for (int q = 0; q < NB; q++) w -= kc.U[0][b][q] * hx[q] + kc.U[1][b][q] * hy[q];
        v[b] = w;
      }
      const T sgm = s.st[m];
      const T inv_det = T(1) / (((sgm + trM) * sgm + c2) * sgm + detM);
      T x[NB];
#pragma unroll
      for (int p = 0; p < NB; p++) {
        T w = T(0);
#pragma unroll
        for (int q = 0; q < NB; q++) {
          const T adj = sgm * ((p == q ? sgm + trM : T(0)) - M[p][q]) + adjM[p][q];
          w += adj * v[q];
        }
        x[p] = w * inv_det;
      }
#pragma unroll
      for (int b = 0; b < NB; b++) cur[(ip * AG + k) * NB + b] = x[b];
      if (va) {
        const unsigned c = cell_of(ip, jp);
#pragma unroll
        for (int b = 0; b < NB; b++) pg[c * NA + b * Aq + k] = x[b];
      }
    }
    __syncthreads();
  };
  // diagonals d + 1 .. d + DEPTH requested while d is solved: DEPTH + 1
  // register sets in rotation, the loop unrolled by DEPTH + 1
  if (DEPTH == 1) {
    Set s0, s1;
    fetch(0, s0);
    for (int d = 0; d < n_diag; d += 2) {
      fetch(d + 1, s1);
      solve(d, s0);
      if (d + 1 >= n_diag) break;
      fetch(d + 2, s0);
      solve(d + 1, s1);
    }
  } else {
    Set s0, s1, s2;
    fetch(0, s0);
    fetch(1, s1);
    for (int d = 0; d < n_diag; d += 3) {
      fetch(d + 2, s2);
      solve(d, s0);
      if (d + 1 >= n_diag) break;
      fetch(d + 3, s0);
      solve(d + 1, s1);
      if (d + 2 >= n_diag) break;
      fetch(d + 4, s1);
      solve(d + 2, s2);
    }
  }
  if (gw.item == -2) ring[threadIdx.x] = sink;  // never: item is -1 (idle) or an index
}

template <class T, int AG, int DEPTH>
__global__ void __launch_bounds__(kBoxSweepThreads) sweep_box_batch(const T *__restrict__ r, T *__restrict__ psi,
                                                            const BatchItem<T> *__restrict__ items,
                                                            const BoxSweepGroup *__restrict__ groups,
                                                            int nx_max) {
  extern __shared__ __align__(16) unsigned char smem[];
  T *ring = reinterpret_cast<T *>(smem);
  const BoxSweepGroup &gw = groups[blockIdx.x];
  if (gw.item < 0) return;  // an idle slot of a shorter XCD list
  const BatchItem<T> &it = items[gw.item];
  const T *rg = r + gw.rbase;
  T *pg = psi + gw.rbase;
  // the cells per thread the group's box needs (uniform over the workgroup)
  switch (gw.ipt) {
    case 1: box_march<T, AG, DEPTH, 1>(rg, pg, it.g, it.box, gw, ring, nx_max); break;
    case 2: box_march<T, AG, DEPTH, 2>(rg, pg, it.g, it.box, gw, ring, nx_max); break;
    case 4: box_march<T, AG, DEPTH, (AG < 4 ? AG : 4)>(rg, pg, it.g, it.box, gw, ring, nx_max); break;
    default: box_march<T, AG, DEPTH, AG>(rg, pg, it.g, it.box, gw, ring, nx_max); break;
  }
}

}  // namespace

// Reference order <-> quadrant-major order of every active problem's rows:
// (cell, basis, angle a) <-> (q, cell, basis, position), q * A / 4 +
// position = qpos[a]. A grid-stride loop per problem; TO_Q reads the
// reference order in sequence, the other direction writes it in sequence.
template <class T, bool TO_Q>
__global__ void __launch_bounds__(256) quad_reorder(const T *__restrict__ src, T *__restrict__ dst,
                                                   const BatchItem<T> *__restrict__ items, int n_items,
                                                   const int *__restrict__ qpos, const int64_t *__restrict__ qoff) {
  constexpr int NB = 3;
  for (int p = 0; p < n_items; p++) {
    const BatchItem<T> &it = items[p];
    const int A = it.g.A, Aq = A / 4;
    const unsigned nc = (unsigned)it.g.n_cells;
    const unsigned rows = nc * NB * (unsigned)A;
    const int *qp = qpos + qoff[p];
    for (unsigned lr = blockIdx.x * blockDim.x + threadIdx.x; lr < rows; lr += gridDim.x * blockDim.x) {
      const unsigned node = lr / (unsigned)A, a = lr - node * (unsigned)A;
      const unsigned c = node / NB, b = node - c * NB;
      const unsigned qk = (unsigned)qp[a], q = qk / (unsigned)Aq, kq = qk - q * (unsigned)Aq;
      const int64_t qi = (((int64_t)q * nc + c) * NB + b) * Aq + kq;
      if (TO_Q)
        dst[it.off + qi] = src[it.off + lr];
      else
        dst[it.off + lr] = src[it.off + qi];
    }
  }
}

template <class T>
void batch_box_sweep_launch(const Batch &b, const T *r, T *psi, hipStream_t s, bool staged) {
  const int pi = sizeof(T) == 8 ? 0 : 1;
  const BoxSweepPlan &w = b.box_sweep[pi];
  constexpr int AG = kBoxSweepAngles<T>;
  const size_t lds = (size_t)2 * w.nx_max * AG * 3 * sizeof(T);
  const BatchItem<T> *items = b.items[pi][3].as<BatchItem<T>>();
  if (staged) {  // r and psi are already in the quadrant-major order
    hipLaunchKernelGGL((sweep_box_batch<T, AG, kBoxSweepDepth<T>>), dim3(w.n_wg), dim3(kBoxSweepThreads), lds, s, r,
                       psi, items, w.groups.as<BoxSweepGroup>(), w.nx_max);
    UBT_HIP(hipGetLastError());
    return;
  }
  T *rq = w.rq.as<T>(), *pq = w.pq.as<T>();
  batch_quad_order<T>(b, r, rq, true, s);
  hipLaunchKernelGGL((sweep_box_batch<T, AG, kBoxSweepDepth<T>>), dim3(w.n_wg), dim3(kBoxSweepThreads), lds, s,
                     (const T *)rq, pq, items, w.groups.as<BoxSweepGroup>(), w.nx_max);
  batch_quad_order<T>(b, (const T *)pq, psi, false, s);
  UBT_HIP(hipGetLastError());
}

template <class T> void batch_quad_order(const Batch &b, const T *src, T *dst, bool to_q, hipStream_t s) {
  const int pi = sizeof(T) == 8 ? 0 : 1;
  const BoxSweepPlan &w = b.box_sweep[pi];
  const BatchItem<T> *items = b.items[pi][3].as<BatchItem<T>>();
  const int n_items = b.n_items[pi][3];
  const unsigned grid = (unsigned)(b.n_cu * 8);
  if (to_q)
    hipLaunchKernelGGL((quad_reorder<T, true>), dim3(grid), dim3(256), 0, s, src, dst, items, n_items,
                       w.qpos.as<int>(), w.qoff.as<int64_t>());
  else
    hipLaunchKernelGGL((quad_reorder<T, false>), dim3(grid), dim3(256), 0, s, src, dst, items, n_items,
                       w.qpos.as<int>(), w.qoff.as<int64_t>());
  UBT_HIP(hipGetLastError());
}

template void batch_box_sweep_launch<double>(const Batch &, const double *, double *, hipStream_t, bool);
template void batch_box_sweep_launch<float>(const Batch &, const float *, float *, hipStream_t, bool);
template void batch_quad_order<double>(const Batch &, const double *, double *, bool, hipStream_t);
template void batch_quad_order<float>(const Batch &, const float *, float *, bool, hipStream_t);

template void batch_apply_launch<double>(const Batch &, int, const double *, double *, bool, hipStream_t, bool);
template void batch_apply_launch<float>(const Batch &, int, const float *, float *, bool, hipStream_t, bool);
template void batch_blocks_launch<double>(const Batch &, const double *, double *, hipStream_t);
template void batch_blocks_launch<float>(const Batch &, const float *, float *, hipStream_t);

}  // namespace ubt
