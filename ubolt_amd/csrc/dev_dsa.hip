// DSA on gfx950: per-group assembly of the diffusion operator by a kernel (no
// host copy of the cross sections), its block-sparse FP64 matvec, and the
// block-Jacobi preconditioned CG inner solve whose scalars never leave the
// device. Restriction = the matrix-core moments kernel times the node weight;
// prolongation = the isotropic broadcast / sum_weights.
//
// Operators (reference src/dsa_structuredk.kokkos.cxx:183-285,
// src/dsa_plexk.kokkos.cxx:288-475, helpers src/dsa_operatork.hpp:110-147):
//  * structured: cell-centred star per unit volume, harmonic face D, Marshak
//    vacuum faces 1/(h (2 + h/(2 D_b))), zero Neumann reflective faces;
//  * DG0 plex: two-point flux A/(d_c/D_c + d_n/D_n), Marshak A/(2 + d_c/D_b),
//    volume-weighted;
//  * DG1 plex: the MIP interior-penalty form on the modal basis,
//    volume-weighted;
// with D = 1/(3 sigma_t), sigma_a = sigma_t - sigma_s(g,g) and the consistent-D
// blend (D^p + (m h)^p)^(1/p), p = 1.5, on FD and DG0 faces.
// Voids (sigma_t <= a threshold, 0 by default; reference
// src/dsak.kokkos.cxx:233-345) are bridged where the operator allows it: D =
// 1/(3 (max(sigma_t, 0) + 1/L)) with L = 4 V / S the voids' mean chord (or a
// given D), a DG1 face touching a void with the harmonic face D on both
// sides. Otherwise, or when bridging is off, they are masked: identity rows
// (V times identity blocks on the plex) coupled to nothing, a face into one a
// vacuum face, and a zero restriction / correction weight on their nodes. A
// group void everywhere is masked whole (zero correction).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <mutex>
#include <utility>
#include <vector>

#include <cmath>

#include "dsa.hpp"

namespace ubt {

namespace {

struct AsmArgs {
  int kind, dim, nb, nf, ns, n_cells, cons;
  double p, C;
  // the void policy: a cell is void where sigma_t <= void_st; bridged voids
  // take D = void_d, or 1/(3 (max(sigma_t, 0) + inv_chord)); masked voids are
  // identity rows and a face into one is a vacuum face
  double void_st, void_d, inv_chord;
  int bridged, masked;
  const double *sigt, *sigs;  // current group
  const int *nbr;             // cells * ns
  // structured
  double h[3], hr[3];
  int vac[3][2];
  // unstructured
  const double *volume, *area, *dist, *nA, *slot_hr;
  const double *face_own, *face_up, *grad;  // by geometry class
  const int *face_nb, *gcls;
  double *own, *off;
};

__device__ inline double blend(double d, double m, double p) {
  return p == 1.0 ? d + m : pow(pow(d, p) + pow(m, p), 1.0 / p);
}

__device__ inline bool is_void(const AsmArgs &a, int64_t c) { return !(a.sigt[c] > a.void_st); }
__device__ inline bool is_masked(const AsmArgs &a, int64_t c) { return a.masked && is_void(a, c); }

// D of a cell: 1 / (3 sigma_t); in a bridged void the given D or a third of
// the free flight, 1 / (3 (max(sigma_t, 0) + 1 / L)) (Wigner's rational form)
__device__ inline double dcoef(const AsmArgs &a, int64_t c) {
  const double st = a.sigt[c];
  if (!a.bridged || st > a.void_st) return 1.0 / (3.0 * st);
  return a.void_d > 0.0 ? a.void_d : 1.0 / (3.0 * (fmax(st, 0.0) + a.inv_chord));
}

__global__ void asm_structured(AsmArgs a) {
  const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= a.n_cells) return;
  if (is_masked(a, c)) {  // an identity row, decoupled both ways
    for (int k = 0; k < 2 * a.dim; k++) a.off[c * a.ns + k] = 0.0;
    a.own[c] = 1.0;
    return;
  }
  const double st = a.sigt[c];
  const double dc = dcoef(a, c);
  double diag = st - a.sigs[c];
  for (int d = 0; d < a.dim; d++) {
    const double h = a.h[d];
    for (int side = 0; side < 2; side++) {
      const int s = 2 * d + side;
      const int n = a.nbr[c * a.ns + s];
      double o = 0.0;
      if (n >= 0 && !is_masked(a, n)) {
        const double dn = dcoef(a, n);
        double df = 2.0 * dc * dn / (dc + dn);
        if (a.cons) df = blend(df, a.hr[d] * h, a.p);
        o = -df / (h * h);
        diag += df / (h * h);
      } else if (n >= 0 || a.vac[d][side]) {  // a vacuum face, or one into a masked void
        const double db = a.cons ? blend(dc, a.hr[d] * h, a.p) : dc;
        diag += 1.0 / (h * (2.0 + h / (2.0 * db)));
      }
      a.off[c * a.ns + s] = o;
    }
  }
  a.own[c] = diag;
}

__global__ void asm_dg0(AsmArgs a) {
  const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= a.n_cells) return;
  if (is_masked(a, c)) {  // V times an identity row, decoupled both ways
    for (int f = 0; f < a.nf; f++) a.off[c * a.nf + f] = 0.0;
    a.own[c] = a.volume[c];
    return;
  }
  const double st = a.sigt[c];
  const double dc = dcoef(a, c);
  double diag = a.volume[c] * (st - a.sigs[c]);
  for (int f = 0; f < a.nf; f++) {
    const int64_t k = c * a.nf + f;
    const int n = a.face_nb[k];
    const double A = a.area[k], d_own = a.dist[2 * k];
    double o = 0.0;
    if (n >= 0 && !is_masked(a, n)) {
      const double dn = dcoef(a, n), d_nb = a.dist[2 * k + 1];
      double tf = A / (d_own / dc + d_nb / dn);
      if (a.cons) {
        const double len = d_own + d_nb;
        tf = A / len * blend(tf * len / A, a.slot_hr[k] * len, a.p);
      }
      o = -tf;
      diag += tf;
    } else if (n >= 0 || n == -1) {  // a vacuum face, or one into a masked void
      const double db = a.cons ? blend(dc, 2.0 * d_own * a.slot_hr[k], a.p) : dc;
      diag += A / (2.0 + d_own / db);
    }
    a.off[k] = o;
  }
  a.own[c] = diag;
}

template <int NB>
__global__ void asm_dg1(AsmArgs a) {
  const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= a.n_cells) return;
  const double V = a.volume[c];
  if (is_masked(a, c)) {  // V times the identity block, every coupling zeroed
    for (int e = 0; e < a.nf * NB * NB; e++) a.off[c * a.nf * NB * NB + e] = 0.0;
    for (int i = 0; i < NB; i++)
      for (int j = 0; j < NB; j++) a.own[(c * NB + i) * NB + j] = i == j ? V : 0.0;
    return;
  }
  const double st = a.sigt[c];
  const double dc = dcoef(a, c), sa = st - a.sigs[c];
  const int64_t gc = a.gcls[c];
  double own[NB][NB];
  for (int i = 0; i < NB; i++)
    for (int j = 0; j < NB; j++) {
      double gd = 0.0;
      for (int d = 0; d < 3; d++) gd += a.grad[(gc * NB + i) * 3 + d] * a.grad[(gc * NB + j) * 3 + d];
      own[i][j] = V * (dc * gd + (i == j ? sa : 0.0));
    }
  for (int f = 0; f < a.nf; f++) {
    const int64_t k = c * a.nf + f;
    const int n = a.face_nb[k];
    double *nbk = a.off + k * NB * NB;
    for (int e = 0; e < NB * NB; e++) nbk[e] = 0.0;
    if (n < -1) continue;  // reflective: zero Neumann
    const double A = a.area[k];
    const double scale = V * A;
    const int64_t kg = gc * a.nf + f;
    const double *Fo = a.face_own + kg * NB * NB, *Fu = a.face_up + kg * NB * NB;
    double nrm[3];
    for (int d = 0; d < 3; d++) nrm[d] = a.nA[3 * k + d] / A;
    double gcv[NB], gn[NB];
    for (int j = 0; j < NB; j++) {
      gcv[j] = gn[j] = 0.0;
      for (int d = 0; d < 3; d++) {
        gcv[j] += a.grad[(gc * NB + j) * 3 + d] * nrm[d];
        if (n >= 0) gn[j] += a.grad[((int64_t)a.gcls[n] * NB + j) * 3 + d] * nrm[d];
      }
    }
    const double hc = 2.0 * a.dist[2 * k];
    if (n >= 0 && !is_masked(a, n)) {
      const double hn = 2.0 * a.dist[2 * k + 1];
      double dcf = dc, dn = dcoef(a, n);
      // a face touching a bridged void: both sides take the face's harmonic
      // D, in the averages and in kappa (the same from either cell)
      if (a.bridged && (is_void(a, c) || is_void(a, n))) {
        dcf = 2.0 * dc * dn / (dc + dn);
        dn = dcf;
      }
      const double kappa = fmax(0.5 * a.C * (dcf / hc + dn / hn), 0.25);
      for (int i = 0; i < NB; i++) {
        const double mi = scale * Fo[i * NB];
        for (int j = 0; j < NB; j++) {
          own[i][j] += kappa * scale * Fo[i * NB + j] - 0.5 * dcf * gcv[j] * mi - 0.5 * dcf * gcv[i] * scale * Fo[j];
          nbk[i * NB + j] = -kappa * scale * Fu[i * NB + j] - 0.5 * dn * gn[j] * mi + 0.5 * dcf * gcv[i] * scale * Fu[j];
        }
      }
    } else {  // a vacuum face, or one into a masked void
      const double kappa = fmax(a.C * dc / hc, 0.25);
      for (int i = 0; i < NB; i++) {
        const double mi = scale * Fo[i * NB];
        for (int j = 0; j < NB; j++)
          own[i][j] += kappa * scale * Fo[i * NB + j] - 0.5 * dc * gcv[j] * mi - 0.5 * dc * gcv[i] * scale * Fo[j];
      }
    }
  }
  for (int i = 0; i < NB; i++)
    for (int j = 0; j < NB; j++) a.own[(c * NB + i) * NB + j] = own[i][j];
}

// block-Jacobi: invert each cell's own block (Gauss-Jordan, partial pivoting)
template <int NB>
__global__ void jacobi_invert(const double *own, double *jinv, int64_t n_cells, int *n_bad) {
  const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= n_cells) return;
  double m[NB][NB], w[NB][NB];
  for (int i = 0; i < NB; i++)
    for (int j = 0; j < NB; j++) {
      m[i][j] = own[(c * NB + i) * NB + j];
      w[i][j] = i == j ? 1.0 : 0.0;
    }
  bool ok = true;
  for (int k = 0; k < NB && ok; k++) {
    int piv = k;
    for (int i = k + 1; i < NB; i++)
      if (fabs(m[i][k]) > fabs(m[piv][k])) piv = i;
    if (!(fabs(m[piv][k]) > 0.0)) {
      ok = false;
      break;
    }
    if (piv != k)
      for (int j = 0; j < NB; j++) {
        double t = m[k][j];
        m[k][j] = m[piv][j];
        m[piv][j] = t;
        t = w[k][j];
        w[k][j] = w[piv][j];
        w[piv][j] = t;
      }
    const double ip = 1.0 / m[k][k];
    for (int j = 0; j < NB; j++) {
      m[k][j] *= ip;
      w[k][j] *= ip;
    }
    for (int i = 0; i < NB; i++) {
      if (i == k) continue;
      const double f = m[i][k];
      for (int j = 0; j < NB; j++) {
        m[i][j] -= f * m[k][j];
        w[i][j] -= f * w[k][j];
      }
    }
  }
  if (!ok) atomicAdd(n_bad, 1);
  for (int i = 0; i < NB; i++)
    for (int j = 0; j < NB; j++) jinv[(c * NB + i) * NB + j] = ok ? w[i][j] : 0.0;
}

struct MvArgs {
  int nb, ns;
  int64_t n_cells;
  const int *nbr;
  const double *own, *off, *jinv;
};

// y = D x for node (c, i); NB compile-time
template <int NB>
__device__ inline double mv_row(const MvArgs &m, const double *__restrict__ x, int64_t c, int i) {
  double v = 0.0;
#pragma unroll
  for (int j = 0; j < NB; j++) v += m.own[(c * NB + i) * NB + j] * x[c * NB + j];
  for (int s = 0; s < m.ns; s++) {
    const int n = m.nbr[c * m.ns + s];
    if (n < 0) continue;
    const double *o = m.off + ((c * m.ns + s) * NB + i) * NB;
#pragma unroll
    for (int j = 0; j < NB; j++) v += o[j] * x[(int64_t)n * NB + j];
  }
  return v;
}

// (D p)(c, i) with p = z + beta p_old formed on the fly; pc: p at (c, i)
template <int NB>
__device__ inline double mv_row_dir(const MvArgs &m, const double *__restrict__ z, const double *__restrict__ po,
                                    double beta, int64_t c, int i, double &pc) {
  double v = 0.0;
#pragma unroll
  for (int j = 0; j < NB; j++) {
    const double pj = fma(beta, po[c * NB + j], z[c * NB + j]);
    if (j == i) pc = pj;
    v += m.own[(c * NB + i) * NB + j] * pj;
  }
  for (int s = 0; s < m.ns; s++) {
    const int n = m.nbr[c * m.ns + s];
    if (n < 0) continue;
    const double *o = m.off + ((c * m.ns + s) * NB + i) * NB;
#pragma unroll
    for (int j = 0; j < NB; j++) v += o[j] * fma(beta, po[(int64_t)n * NB + j], z[(int64_t)n * NB + j]);
  }
  return v;
}

template <int NB>
__global__ void dsa_matvec(MvArgs m, const double *__restrict__ x, double *__restrict__ y) {
  const int64_t node = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (node >= m.n_cells * NB) return;
  y[node] = mv_row<NB>(m, x, node / NB, (int)(node % NB));
}

// ---------------------------------------------------------------------------
// CG with device scalars. scal: 0 rz, 1 rz0, 2 alpha, 3 beta, 4 done, 5 its,
// 6 rtol^2, 7 the reductions' ticket (an int). Partial sums per block,
// reduced in fixed order by the block that finishes last. Two launches per
// iteration: the new direction p = z + beta p_old is formed inside the
// matvec (every thread forms it for the nodes it reads, with one fma, so all
// agree bit for bit, and writes its own node's into the other p buffer),
// then the update.
// ---------------------------------------------------------------------------
constexpr int kCgThreads = 256;

__device__ inline double block_sum(double v) {
  __shared__ double sh[kCgThreads / 64];
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
  if (l == 0) sh[w] = v;
  __syncthreads();
  double t = 0.0;
  if (threadIdx.x == 0)
    for (int i = 0; i < (int)(blockDim.x >> 6); i++) t += sh[i];
  return t;
}

// z = M r (block Jacobi) for node (c, i)
template <int NB>
__device__ inline double prec_row(const MvArgs &m, const double *__restrict__ r, int64_t c, int i) {
  double v = 0.0;
#pragma unroll
  for (int j = 0; j < NB; j++) v += m.jinv[(c * NB + i) * NB + j] * r[c * NB + j];
  return v;
}

// cg_init and cg_update run a thread per CELL, so the block-Jacobi step reads
// only the r values its own thread just wrote
template <int NB>
__global__ void cg_init(MvArgs m, const double *__restrict__ b, const double *__restrict__ wgt, double *x, double *r,
                        double *z, double *p, double *partial) {
  const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  double v = 0.0;
  if (c < m.n_cells) {
#pragma unroll
    for (int i = 0; i < NB; i++) {
      const int64_t n = c * NB + i;
      x[n] = 0.0;
      r[n] = wgt ? b[n] * wgt[n] : b[n];
    }
#pragma unroll
    for (int i = 0; i < NB; i++) {
      const int64_t n = c * NB + i;
      const double zz = prec_row<NB>(m, r, c, i);
      z[n] = zz;
      p[n] = zz;
      v += r[n] * zz;
    }
  }
  v = block_sum(v);
  if (threadIdx.x == 0) partial[blockIdx.x] = v;
}

// the sum of the block partials by one wavefront (lane-strided, then a
// butterfly: a fixed order, so the result is reproducible); every lane holds it
__device__ inline double wave_total(const double *partial, int nblk) {
  double s = 0.0;
  for (int i = threadIdx.x; i < nblk; i += 64) s += partial[i];
  for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o, 64);
  return s;
}

__global__ void cg_reduce_init(const double *partial, int nblk, double *scal, double rtol2, int *ticket) {
  const double s = wave_total(partial, nblk);
  if (threadIdx.x != 0) return;
  *ticket = 0;
  scal[0] = s;
  scal[1] = s;
  scal[3] = 0.0;  // the first direction is z
  scal[4] = (s == 0.0) ? 1.0 : 0.0;
  scal[5] = 0.0;
  scal[6] = rtol2;
}

// The block that finishes last (an atomic ticket taken after its partial is
// published) sums every block's partial in index order - one wavefront's
// lane-strided sum and butterfly, the fixed order the separate reduction
// kernel used - so the reduction needs no launch of its own; the ticket is
// reset for the next one. True on that block's thread 0 only, with the total.
__device__ inline bool last_block_total(double v, double *partial, int *ticket, double &total) {
  __shared__ int last;
  if (threadIdx.x == 0) {
    __hip_atomic_store(&partial[blockIdx.x], v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __threadfence();
    last = atomicAdd(ticket, 1) == (int)gridDim.x - 1;
  }
  __syncthreads();
  if (!last || threadIdx.x >= 64) return false;
  __threadfence();
  double s = 0.0;
  for (int i = threadIdx.x; i < (int)gridDim.x; i += 64)
    s += __hip_atomic_load(&partial[i], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o, 64);
  if (threadIdx.x != 0) return false;
  __hip_atomic_store(ticket, 0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  total = s;
  return true;
}

// p = z + beta p_old (beta 0 on the first iteration), q = D p and p.q; the
// last block forms alpha
template <int NB>
__global__ void cg_matvec_dot(MvArgs m, const double *__restrict__ z, const double *__restrict__ po,
                              double *__restrict__ p, double *__restrict__ q, double *partial, double *scal,
                              int *ticket) {
  if (scal[4] != 0.0) return;
  const double beta = scal[3];
  const int64_t n = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  double v = 0.0;
  if (n < m.n_cells * NB) {
    double pn = 0.0;
    const double qq = mv_row_dir<NB>(m, z, po, beta, n / NB, (int)(n % NB), pn);
    p[n] = pn;
    q[n] = qq;
    v = pn * qq;
  }
  v = block_sum(v);
  double s;
  if (last_block_total(v, partial, ticket, s)) {
    scal[2] = s != 0.0 ? scal[0] / s : 0.0;
    if (s == 0.0) scal[4] = 1.0;
  }
}

// x += alpha p, r -= alpha q, z = M r and r.z; the last block forms beta and
// the convergence flag
template <int NB>
__global__ void cg_update(MvArgs m, double *x, double *r, double *z, const double *p, const double *q, double *partial,
                          double *scal, int *ticket) {
  if (scal[4] != 0.0) return;
  const double alpha = scal[2];
  const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  double v = 0.0;
  if (c < m.n_cells) {
#pragma unroll
    for (int i = 0; i < NB; i++) {
      const int64_t n = c * NB + i;
      x[n] += alpha * p[n];
      r[n] -= alpha * q[n];
    }
#pragma unroll
    for (int i = 0; i < NB; i++) {
      const int64_t n = c * NB + i;
      const double zz = prec_row<NB>(m, r, c, i);
      z[n] = zz;
      v += r[n] * zz;
    }
  }
  v = block_sum(v);
  double s;
  if (last_block_total(v, partial, ticket, s)) {
    scal[3] = scal[0] != 0.0 ? s / scal[0] : 0.0;
    scal[0] = s;
    scal[5] += 1.0;
    if (s <= scal[6] * scal[1]) scal[4] = 1.0;
  }
}

// the void census of one group: void cells (sigma_t <= the threshold), the first of
// them, absorbing cells (sigma_t > sigma_s(g -> g))
__global__ void census_init(int *flags) {
  flags[0] = 0;
  flags[1] = 0;
  flags[2] = 0x7fffffff;
  flags[3] = 0;
}

__global__ void void_census(const double *__restrict__ st, const double *__restrict__ ss, int64_t n, double void_st,
                            int *flags) {
  const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= n) return;
  if (!(st[c] > void_st)) {
    atomicAdd(&flags[1], 1);
    atomicMin(&flags[2], (int)c);
  }
  if (st[c] - ss[c] > 0.0) atomicAdd(&flags[3], 1);
}

// The voids' volume V and surface S (their faces onto a non-void cell or a
// vacuum face) for the mean chord L = 4 V / S: per block in a fixed order,
// then one block over the partials, so the sums do not depend on scheduling
constexpr int kVoidBlocks = 1024;

__global__ void __launch_bounds__(256) void_vs_partial(AsmArgs a, double *part) {
  __shared__ double sv[256], sf[256];
  double v = 0.0, f = 0.0;
  for (int64_t c = (int64_t)blockIdx.x * 256 + threadIdx.x; c < a.n_cells; c += (int64_t)gridDim.x * 256) {
    if (!is_void(a, c)) continue;
    if (a.kind == DISC_FD) {
      double vol = 1.0;
      for (int d = 0; d < a.dim; d++) vol *= a.h[d];
      v += vol;
      for (int d = 0; d < a.dim; d++)
        for (int side = 0; side < 2; side++) {
          const int n = a.nbr[c * a.ns + 2 * d + side];
          if (n >= 0 ? !is_void(a, n) : a.vac[d][side] != 0) f += vol / a.h[d];
        }
    } else {
      v += a.volume[c];
      for (int k = 0; k < a.nf; k++) {
        const int n = a.face_nb[c * a.nf + k];
        if (n >= 0 ? !is_void(a, n) : n == -1) f += a.area[c * a.nf + k];
      }
    }
  }
  sv[threadIdx.x] = v;
  sf[threadIdx.x] = f;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) {
      sv[threadIdx.x] += sv[threadIdx.x + w];
      sf[threadIdx.x] += sf[threadIdx.x + w];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    part[2 * blockIdx.x] = sv[0];
    part[2 * blockIdx.x + 1] = sf[0];
  }
}

__global__ void __launch_bounds__(256) void_vs_final(double *part, int nblk) {
  __shared__ double sv[256], sf[256];
  double v = 0.0, f = 0.0;
  for (int b = threadIdx.x; b < nblk; b += 256) {
    v += part[2 * b];
    f += part[2 * b + 1];
  }
  sv[threadIdx.x] = v;
  sf[threadIdx.x] = f;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) {
      sv[threadIdx.x] += sv[threadIdx.x + w];
      sf[threadIdx.x] += sf[threadIdx.x + w];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    part[2 * nblk] = sv[0];
    part[2 * nblk + 1] = sf[0];
  }
}

__global__ void clear_int(int *p) { *p = 0; }

// the restriction / prolongation weight with the masked void nodes zeroed
__global__ void mask_weight(const double *__restrict__ weight, const double *__restrict__ st, double void_st, int nb,
                            int64_t n_nodes, double *__restrict__ wm) {
  const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (r < n_nodes) wm[r] = st[r / nb] > void_st ? weight[r] : 0.0;
}

template <class T>
__global__ void fill_zero(T *__restrict__ y, int64_t n) {
  const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (r < n) y[r] = T(0);
}

template <class T>
__global__ void prolong(const double *__restrict__ sol, T *__restrict__ y, int64_t rows, int A, double inv_w) {
  const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (r < rows) y[r] = (T)(sol[r / A] * inv_w);
}

// the adjoint prolongation: y(node, a) = w_a * weight(node) * sol(node) * inv_w
template <class T>
__global__ void prolong_t(const double *__restrict__ sol, const double *__restrict__ weight,
                          const double *__restrict__ w, T *__restrict__ y, int64_t rows, int A, double inv_w) {
  const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (r < rows) {
    const int64_t node = r / A;
    y[r] = (T)(w[r - node * A] * weight[node] * sol[node] * inv_w);
  }
}

double half_range_current(const Quadrature &q, const double *n) {
  double s = 0.0;
  for (int a = 0; a < q.n_angles; a++) {
    double c = 0.0;
    for (int d = 0; d < 3; d++) c += q.omega[3 * a + d] * n[d];
    if (c > 0.0) s += q.w[a] * c;
  }
  return s / q.sum_weights;
}

}  // namespace

// ---------------------------------------------------------------------------
void DSA::create(Operator *o) {
  op = o;
  const Disc &D = *o->disc;
  nb = D.nb;
  n_cells = D.n_cells;
  n_nodes = D.n_nodes();
  std::vector<int> nb_tab;
  std::vector<double> w(n_nodes, 1.0);
  if (D.kind == DISC_FD) {
    ns = 2 * D.dim;
    nb_tab.assign((size_t)n_cells * ns, -1);
    const int64_t stride[3] = {1, D.n[0], (int64_t)D.n[0] * D.n[1]};
    for (int64_t c = 0; c < n_cells; c++) {
      int64_t cc = c;
      int idx[3];
      idx[0] = (int)(cc % D.n[0]);
      cc /= D.n[0];
      idx[1] = (int)(cc % D.n[1]);
      idx[2] = (int)(cc / D.n[1]);
      for (int d = 0; d < D.dim; d++) {
        if (idx[d] > 0) nb_tab[c * ns + 2 * d] = (int)(c - stride[d]);
        if (idx[d] < D.n[d] - 1) nb_tab[c * ns + 2 * d + 1] = (int)(c + stride[d]);
      }
    }
    for (int d = 0; d < D.dim; d++) {
      double e[3] = {0, 0, 0};
      e[d] = 1.0;
      half_range[d] = half_range_current(D.quad, e);
      for (int s = 0; s < 2; s++) {
        vacuum[d][s] = D.face_type[d][s] == UBT_BC_VACUUM;
        any_vacuum = any_vacuum || vacuum[d][s];
      }
    }
  } else {
    ns = D.nf;
    nb_tab.resize(D.face_nb.size());
    std::vector<double> hr(D.face_nb.size(), 0.0);
    for (size_t k = 0; k < D.face_nb.size(); k++) {
      nb_tab[k] = D.face_nb[k] >= 0 ? D.face_nb[k] : -1;
      any_vacuum = any_vacuum || D.face_nb[k] == -1;
      double nrm[3];
      for (int d = 0; d < 3; d++) nrm[d] = D.face_nA[3 * k + d] / D.face_area[k];
      hr[k] = half_range_current(D.quad, nrm);
    }
    slot_hr.upload(hr);
    for (int64_t c = 0; c < n_cells; c++)
      for (int i = 0; i < nb; i++) w[c * nb + i] = D.volume[c];
  }
  nbr.upload(nb_tab);
  weight.upload(w);
  own.alloc(sizeof(double) * n_cells * nb * nb);
  off.alloc(sizeof(double) * n_cells * ns * nb * nb);
  jinv.alloc(sizeof(double) * n_cells * nb * nb);
  const size_t nodes = sizeof(double) * n_nodes;
  rhs.alloc(nodes);
  sol.alloc(nodes);
  r.alloc(nodes);
  z.alloc(nodes);
  p.alloc(nodes);
  p2.alloc(nodes);
  q.alloc(nodes);
  const int64_t nblk = (n_nodes + kCgThreads - 1) / kCgThreads;
  partial.alloc(sizeof(double) * nblk);
  scal.alloc(sizeof(double) * 8);
  vsum.alloc(sizeof(double) * (2 * kVoidBlocks + 2));
  wmask.alloc(sizeof(double) * n_nodes);
  UBT_HIP(hipHostMalloc((void **)&flag_host, sizeof(double), hipHostMallocDefault));
}

DSA::~DSA() {
  if (flag_host) (void)hipHostFree(flag_host);
}

void DSA::set_group(int g, hipStream_t s) {
  Operator &o = *op;
  const Disc &D = *o.disc;
  UBT_CHECK(o.xs_set, UBT_ERR_ARG_WRONGSTATE, "set the cross sections before the DSA group");
  UBT_CHECK(g >= 0 && g < o.n_groups, UBT_ERR_ARG_OUTOFRANGE, "group %d out of range", g);
  group = -1;  // until this group's operator is in place
  const double *st = o.t64.sigt.as<double>() + (int64_t)g * n_cells;
  const double *ssg = o.t64.sigs.as<double>() + ((int64_t)g * o.n_groups + g) * n_cells;
  // flags (read back below, on the caller's stream): 0 singular diagonal
  // blocks, 1 void cells (sigma_t <= the void threshold), 2 the first void
  // cell, 3 absorbing cells (sigma_a > 0)
  int *flags = o.flag.as<int>();
  const unsigned grid = (unsigned)((n_cells + 255) / 256);
  hipLaunchKernelGGL(census_init, dim3(1), dim3(1), 0, s, flags);
  hipLaunchKernelGGL(void_census, dim3(grid), dim3(256), 0, s, st, ssg, n_cells, opt.void_sigma_t, flags);
  AsmArgs a{};
  a.kind = D.kind;
  a.dim = D.dim;
  a.nb = nb;
  a.nf = D.nf;
  a.ns = ns;
  a.n_cells = (int)n_cells;
  a.cons = opt.consistent_d;
  a.p = opt.consistent_power;
  a.C = opt.mip_penalty;
  a.void_st = opt.void_sigma_t;
  a.void_d = opt.void_d;
  a.sigt = st;
  a.sigs = ssg;
  a.nbr = nbr.as<int>();
  for (int d = 0; d < 3; d++) {
    a.h[d] = D.h[d];
    a.hr[d] = half_range[d];
    a.vac[d][0] = vacuum[d][0];
    a.vac[d][1] = vacuum[d][1];
  }
  a.own = own.as<double>();
  a.off = off.as<double>();
  if (D.kind != DISC_FD) {
    a.volume = o.volume_d.as<double>();
    a.area = o.face_area_d.as<double>();
    a.dist = o.face_dist_d.as<double>();
    a.nA = o.face_nA_raw.as<double>();
    a.slot_hr = slot_hr.as<double>();
    a.face_own = o.t64.face_own.as<double>();
    a.face_up = o.t64.face_up.as<double>();
    a.grad = o.t64.grad.as<double>();
    a.face_nb = o.face_nb.as<int>();
    a.gcls = o.gcls.as<int>();
  }
  // the operator and its block-Jacobi inverses (singular blocks counted in flags[0])
  auto assemble = [&]() {
    if (D.kind == DISC_FD)
      hipLaunchKernelGGL(asm_structured, dim3(grid), dim3(256), 0, s, a);
    else if (nb == 1)
      hipLaunchKernelGGL(asm_dg0, dim3(grid), dim3(256), 0, s, a);
    else if (nb == 3)
      hipLaunchKernelGGL(asm_dg1<3>, dim3(grid), dim3(256), 0, s, a);
    else
      hipLaunchKernelGGL(asm_dg1<4>, dim3(grid), dim3(256), 0, s, a);
    UBT_HIP(hipGetLastError());
    if (nb == 1)
      hipLaunchKernelGGL(jacobi_invert<1>, dim3(grid), dim3(256), 0, s, own.as<double>(), jinv.as<double>(), n_cells, flags);
    else if (nb == 3)
      hipLaunchKernelGGL(jacobi_invert<3>, dim3(grid), dim3(256), 0, s, own.as<double>(), jinv.as<double>(), n_cells, flags);
    else
      hipLaunchKernelGGL(jacobi_invert<4>, dim3(grid), dim3(256), 0, s, own.as<double>(), jinv.as<double>(), n_cells, flags);
    UBT_HIP(hipGetLastError());
  };
  // without voids (the common case) one assembly and one read-back
  assemble();
  int h[4] = {0, 0, 0, 0};
  UBT_HIP(hipMemcpyAsync(h, flags, sizeof(h), hipMemcpyDeviceToHost, s));
  UBT_HIP(hipStreamSynchronize(s));
  n_void = h[1];
  inv_chord = 0.0;
  bridged = masked = all_void = false;
  if (n_void == n_cells) {
    // void everywhere: nothing on the far side of a void to couple to, so the
    // whole operator is masked and the correction is zero
    all_void = true;
    group = g;
    return;
  }
  if (n_void > 0) {
    // bridged when asked and something is not void and a vacuum face or some
    // absorption holds the operator, and a face of the voids ends a flight;
    // masked otherwise
    if (opt.void_bridge && (any_vacuum || h[3] > 0)) {
      double *vs = vsum.as<double>();
      const int nblk = (int)std::min<int64_t>(grid, kVoidBlocks);
      hipLaunchKernelGGL(void_vs_partial, dim3(nblk), dim3(256), 0, s, a, vs);
      hipLaunchKernelGGL(void_vs_final, dim3(1), dim3(256), 0, s, vs, nblk);
      UBT_HIP(hipGetLastError());
      double vsh[2] = {0.0, 0.0};
      UBT_HIP(hipMemcpyAsync(vsh, vs + 2 * nblk, sizeof(vsh), hipMemcpyDeviceToHost, s));
      UBT_HIP(hipStreamSynchronize(s));
      if (vsh[1] > 0.0) {
        const double chord = 4.0 * vsh[0] / vsh[1];
        inv_chord = 1.0 / chord;
        bridged = true;
      }
    }
    masked = !bridged;
    a.bridged = bridged;
    a.masked = masked;
    a.inv_chord = inv_chord;
    if (masked) {
      const unsigned gn = (unsigned)((n_nodes + 255) / 256);
      hipLaunchKernelGGL(mask_weight, dim3(gn), dim3(256), 0, s, weight.as<double>(), st, opt.void_sigma_t, nb,
                         n_nodes, wmask.as<double>());
    }
    hipLaunchKernelGGL(clear_int, dim3(1), dim3(1), 0, s, flags);
    assemble();
    UBT_HIP(hipMemcpyAsync(h, flags, sizeof(int), hipMemcpyDeviceToHost, s));
    UBT_HIP(hipStreamSynchronize(s));
  } else {
    UBT_CHECK(any_vacuum || h[3] > 0, UBT_ERR_ARG_WRONGSTATE,
              "the DSA diffusion operator is singular: every face is reflective and sigma_a is zero everywhere");
  }
  UBT_CHECK(h[0] == 0, UBT_ERR_MAT_LU_ZRPVT, "%d diffusion diagonal blocks are singular", h[0]);
  group = g;
}

static MvArgs mv_args(const DSA &d) {
  MvArgs m;
  m.nb = d.nb;
  m.ns = d.ns;
  m.n_cells = d.n_cells;
  m.nbr = d.nbr.as<int>();
  m.own = d.own.as<double>();
  m.off = d.off.as<double>();
  m.jinv = d.jinv.as<double>();
  return m;
}

void DSA::matvec(const double *x, double *y, hipStream_t s) const {
  UBT_CHECK(group >= 0, UBT_ERR_ARG_WRONGSTATE, "set the DSA group first");
  UBT_CHECK(!all_void, UBT_ERR_ARG_WRONGSTATE, "group %d is void everywhere: its diffusion operator is masked out",
            group);
  const MvArgs m = mv_args(*this);
  const unsigned grid = (unsigned)((n_nodes + 255) / 256);
  if (nb == 1)
    hipLaunchKernelGGL(dsa_matvec<1>, dim3(grid), dim3(256), 0, s, m, x, y);
  else if (nb == 3)
    hipLaunchKernelGGL(dsa_matvec<3>, dim3(grid), dim3(256), 0, s, m, x, y);
  else
    hipLaunchKernelGGL(dsa_matvec<4>, dim3(grid), dim3(256), 0, s, m, x, y);
  UBT_HIP(hipGetLastError());
}

// The inner CG's workspace: five node vectors (r, z, the two direction
// buffers, q), the per-block partial sums and the CG scalars (every entry is
// written before it is read)
struct CgWork {
  double *r, *z, *p, *p2, *q, *part, *sc;
  int *ticket;  // the last-block reductions' counter
};

static CgWork own_work(DSA &d) {
  return {d.r.as<double>(), d.z.as<double>(), d.p.as<double>(), d.p2.as<double>(), d.q.as<double>(),
          d.partial.as<double>(), d.scal.as<double>(), reinterpret_cast<int *>(d.scal.as<double>() + 7)};
}

template <int NB>
static void cg_run(DSA &d, const double *b, const double *wgt, double *x, hipStream_t s, const CgWork &w) {
  const MvArgs m = mv_args(d);
  const int64_t N = d.n_nodes;
  const unsigned gnode = (unsigned)((N + kCgThreads - 1) / kCgThreads);
  const unsigned gcell = (unsigned)((d.n_cells + kCgThreads - 1) / kCgThreads);
  double *sc = w.sc, *part = w.part;
  double *r = w.r, *z = w.z, *po = w.p, *p = w.p2, *q = w.q;
  hipLaunchKernelGGL(cg_init<NB>, dim3(gcell), dim3(kCgThreads), 0, s, m, b, wgt, x, r, z, po, part);
  hipLaunchKernelGGL(cg_reduce_init, dim3(1), dim3(64), 0, s, part, (int)gcell, sc, d.opt.rtol * d.opt.rtol,
                     w.ticket);
  // two launches per iteration: the new direction with the matvec and dot
  // (alpha), the update and preconditioner (beta), each reduced by its last
  // block; the two direction buffers swap roles every iteration
  for (int it = 0; it < d.opt.max_it; it++) {
    hipLaunchKernelGGL(cg_matvec_dot<NB>, dim3(gnode), dim3(kCgThreads), 0, s, m, z, po, p, q, part, sc, w.ticket);
    hipLaunchKernelGGL(cg_update<NB>, dim3(gcell), dim3(kCgThreads), 0, s, m, x, r, z, p, q, part, sc, w.ticket);
    std::swap(po, p);
    if (d.opt.check_every > 0 && (it + 1) % d.opt.check_every == 0 && it + 1 < d.opt.max_it) {
      UBT_HIP(hipMemcpyAsync(d.flag_host, sc + 4, sizeof(double), hipMemcpyDeviceToHost, s));
      UBT_HIP(hipStreamSynchronize(s));
      if (*d.flag_host != 0.0) break;
    }
  }
  UBT_HIP(hipGetLastError());
}

static void cg_dispatch(DSA &d, const double *b, const double *wgt, double *x, hipStream_t s, const CgWork &w) {
  if (d.nb == 1)
    cg_run<1>(d, b, wgt, x, s, w);
  else if (d.nb == 3)
    cg_run<3>(d, b, wgt, x, s, w);
  else
    cg_run<4>(d, b, wgt, x, s, w);
}

void DSA::solve(const double *b, double *x, hipStream_t s) {
  UBT_CHECK(group == op->group, UBT_ERR_ARG_WRONGSTATE,
            "the DSA operator was set for group %d but the transport operator is at group %d", group, op->group);
  UBT_CHECK(!all_void, UBT_ERR_ARG_WRONGSTATE, "group %d is void everywhere: its diffusion operator is masked out",
            group);
  cg_dispatch(*this, b, nullptr, x, s, own_work(*this));
}

// ---------------------------------------------------------------------------
// The allocation-free inner solve. Its workspace is the file-scope array
// below: one copy per GPU, which the HIP runtime places when it loads this
// code object on that GPU and which lives until the process ends. In host
// code the array's name is only a stand-in, so its device address is looked
// up once per GPU (hipGetSymbolAddress, with that GPU current) and that
// address is what the kernels get. Every call on one GPU shares the array:
// calls that overlap on one GPU race.
// ---------------------------------------------------------------------------
constexpr int64_t kNoAllocBlocks = (kNoAllocNodes + kCgThreads - 1) / kCgThreads;
constexpr int64_t kNoAllocDoubles = 5 * kNoAllocNodes + kNoAllocBlocks + 8;
static __device__ __align__(256) double scratch[kNoAllocDoubles];

static std::mutex scratch_mu;
static std::vector<double *> scratch_ptr;  // per device: the array's address there

static double *scratch_on(int dev) {
  std::lock_guard<std::mutex> lk(scratch_mu);
  if ((int)scratch_ptr.size() <= dev) scratch_ptr.resize(dev + 1, nullptr);
  if (!scratch_ptr[dev]) {
    int cur = 0;
    UBT_HIP(hipGetDevice(&cur));
    if (cur != dev) UBT_HIP(hipSetDevice(dev));
    void *ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (cur != dev) UBT_HIP(hipSetDevice(cur));
    UBT_HIP(e);
    UBT_CHECK(ptr != nullptr, UBT_ERR_GPU, "no device address for the inner-solve workspace on device %d", dev);
    scratch_ptr[dev] = static_cast<double *>(ptr);
  }
  return scratch_ptr[dev];
}

void DSA::solve_noalloc(const double *b, double *x, hipStream_t s) {
  UBT_CHECK(group == op->group, UBT_ERR_ARG_WRONGSTATE,
            "the DSA operator was set for group %d but the transport operator is at group %d", group, op->group);
  UBT_CHECK(!all_void, UBT_ERR_ARG_WRONGSTATE, "group %d is void everywhere: its diffusion operator is masked out",
            group);
  UBT_CHECK(n_nodes <= kNoAllocNodes, UBT_ERR_ARG_SIZ,
            "the allocation-free inner solve takes at most %lld diffusion nodes, this operator has %lld",
            (long long)kNoAllocNodes, (long long)n_nodes);
  double *base = scratch_on(op->device);
  const int64_t N = n_nodes, nblk = (N + kCgThreads - 1) / kCgThreads;
  double *sc = base + 5 * N + nblk;
  const CgWork w = {base, base + N, base + 2 * N, base + 3 * N, base + 4 * N, base + 5 * N, sc,
                    reinterpret_cast<int *>(sc + 7)};
  cg_dispatch(*this, b, nullptr, x, s, w);
}

void DSA::apply(int prec, const void *x, void *y, hipStream_t s, bool transpose) {
  UBT_CHECK(group == op->group, UBT_ERR_ARG_WRONGSTATE,
            "the DSA operator was set for group %d but the transport operator is at group %d", group, op->group);
  UBT_CHECK(x && y, UBT_ERR_ARG_NULL, "x and y are required");
  UBT_CHECK(x != y, UBT_ERR_ARG_IDN, "the DSA apply needs distinct x and y");
  if (all_void) {  // the masked operator's correction: zero (a fill kernel, so the apply can be captured)
    const int64_t rows = op->disc->n_rows();
    const unsigned grid = (unsigned)((rows + 255) / 256);
    if (prec == UBT_FP64)
      hipLaunchKernelGGL(fill_zero<double>, dim3(grid), dim3(256), 0, s, (double *)y, rows);
    else
      hipLaunchKernelGGL(fill_zero<float>, dim3(grid), dim3(256), 0, s, (float *)y, rows);
    UBT_HIP(hipGetLastError());
    return;
  }
  // R: the 0th moment per node (matrix cores), weighted inside the CG init;
  // the adjoint sums unweighted and moves the node weight to the prolongation
  op->moments_f64(prec, x, rhs.as<double>(), s, transpose);
  double *xs = sol.as<double>();
  // masked void nodes are neither restricted nor corrected: their weight is zero
  const double *wn = masked ? wmask.as<double>() : weight.as<double>();
  const double *w = transpose ? nullptr : wn;
  cg_dispatch(*this, rhs.as<double>(), w, xs, s, own_work(*this));
  const int64_t rows = op->disc->n_rows();
  const unsigned grid = (unsigned)((rows + 255) / 256);
  const double inv_w = 1.0 / op->disc->quad.sum_weights;
  if (transpose) {
    const double *wa = op->t64.w.as<double>();
    if (prec == UBT_FP64)
      hipLaunchKernelGGL(prolong_t<double>, dim3(grid), dim3(256), 0, s, xs, wn, wa, (double *)y, rows,
                         op->disc->n_angles, inv_w);
    else
      hipLaunchKernelGGL(prolong_t<float>, dim3(grid), dim3(256), 0, s, xs, wn, wa, (float *)y, rows,
                         op->disc->n_angles, inv_w);
  } else if (prec == UBT_FP64)
    hipLaunchKernelGGL(prolong<double>, dim3(grid), dim3(256), 0, s, xs, (double *)y, rows, op->disc->n_angles, inv_w);
  else
    hipLaunchKernelGGL(prolong<float>, dim3(grid), dim3(256), 0, s, xs, (float *)y, rows, op->disc->n_angles, inv_w);
  UBT_HIP(hipGetLastError());
}

}  // namespace ubt
