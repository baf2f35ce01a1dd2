// Batched problems: work lists, the batch autotuner, and FGMRES with one
// Hessenberg per problem held on the device.
//
// Vectors of a batch are the problems' group vectors laid end to end (row
// offsets fixed at creation; problems without the current group keep their
// span and are skipped). Every reduction of the Krylov solver is segmented
// by problem: the vector is cut into chunks that never straddle two problems,
// a block reduces one chunk, and a second kernel sums each problem's chunks.
// The Givens rotations, the residual estimates and the convergence flags of
// every problem live in device arrays, so a whole restart cycle - restart
// Arnoldi steps, each a preconditioner apply, an operator apply, two
// classical Gram-Schmidt passes and the rotation - has no host round trip and
// is captured once as a HIP graph and replayed; the host reads the flags once
// per cycle. A problem that converges mid-cycle freezes its rotation state;
// the rest of the cycle still runs over its (inert) span.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>

#include "batch.hpp"

namespace ubt {

namespace {

constexpr int kS = 32;         // per-problem stride of the coefficient arrays (restart + 1 at most)
constexpr int kChunk = 16384;  // rows per reduction chunk
constexpr int kT = 256;

// 16 bytes of T: the unit a thread loads and stores
template <class T> struct alignas(16) Vec16 {
  static constexpr int N = 16 / sizeof(T);
  T v[N];
};

// a chunk of x (x the whole vector) moves in 16-byte units when its start is
// on a 16-byte boundary and its length a whole number of units (always on
// these layouts: 3 x angles rows per cell, vectors a multiple of that apart);
// otherwise row by row
template <class T> __device__ bool units_ok(const T *x, int64_t s0, int len) {
  return ((reinterpret_cast<uintptr_t>(x + s0) & 15) == 0) && len % Vec16<T>::N == 0;
}

// part[ch][j] = V_j . w over chunk ch for the K vectors (K a template
// constant, so every load of a unit - w and the K vectors - is
// unconditional and issued together)
template <class T, int K>
__global__ void __launch_bounds__(kT) k_bdot(int64_t n, const T *__restrict__ V, const T *__restrict__ w,
                                             const int64_t *__restrict__ cstart, const int *__restrict__ clen,
                                             double *__restrict__ part) {
  constexpr int VW = Vec16<T>::N;
  const int ch = blockIdx.x;
  const int64_t s0 = cstart[ch];
  const int len = clen[ch];
  double acc[K];
#pragma unroll
  for (int j = 0; j < K; j++) acc[j] = 0.0;
  bool vec = units_ok(w, s0, len);
#pragma unroll
  for (int j = 0; j < K; j++) vec = vec && units_ok(V + (int64_t)j * n, s0, len);
  if (vec) {
    for (int i = threadIdx.x * VW; i < len; i += kT * VW) {
      const Vec16<T> wu = *reinterpret_cast<const Vec16<T> *>(w + s0 + i);
      Vec16<T> vu[K];
#pragma unroll
      for (int j = 0; j < K; j++) vu[j] = *reinterpret_cast<const Vec16<T> *>(V + (int64_t)j * n + s0 + i);
#pragma unroll
      for (int j = 0; j < K; j++)
#pragma unroll
        for (int e = 0; e < VW; e++) acc[j] += (double)vu[j].v[e] * (double)wu.v[e];
    }
  } else {
    for (int i = threadIdx.x; i < len; i += kT) {
      const double wi = (double)w[s0 + i];
#pragma unroll
      for (int j = 0; j < K; j++) acc[j] += (double)V[(int64_t)j * n + s0 + i] * wi;
    }
  }
  __shared__ double sh[K][kT / 64];
  const int wv = threadIdx.x >> 6, l = threadIdx.x & 63;
#pragma unroll
  for (int j = 0; j < K; j++) {
    double v = acc[j];
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    if (l == 0) sh[j][wv] = v;
  }
  __syncthreads();
  if ((int)threadIdx.x < K) {
    double t = 0.0;
    for (int q = 0; q < kT / 64; q++) t += sh[threadIdx.x][q];
    part[(int64_t)ch * kS + threadIdx.x] = t;
  }
}

// out[p * stride + j] = (sqrt of) the sum of problem p's chunk partials; one
// wavefront per (problem, column): lane-strided sums, then a butterfly
__global__ void k_bdot_final(int k, const int *__restrict__ cfirst, const double *__restrict__ part,
                             double *__restrict__ out, int stride, int take_sqrt) {
  const int p = blockIdx.x, j = blockIdx.y;
  if (j >= k) return;
  double t = 0.0;
  for (int ch = cfirst[p] + (int)threadIdx.x; ch < cfirst[p + 1]; ch += 64) t += part[(int64_t)ch * kS + j];
  for (int o = 32; o > 0; o >>= 1) t += __shfl_xor(t, o, 64);
  if (threadIdx.x == 0) out[(int64_t)p * stride + j] = take_sqrt ? sqrt(t) : t;
}

// w += sign * sum_j coef[p][j] V[j] over each chunk of problem p, j < k (and
// j < kcount[p] when given: basis vectors past a problem's own count may never
// have been written this cycle)
template <class T>
__global__ void __launch_bounds__(kT) k_bmaxpy(int64_t n, int k, const T *__restrict__ V, const double *__restrict__ coef,
                                               int stride, double sign, T *__restrict__ w,
                                               const int64_t *__restrict__ cstart, const int *__restrict__ clen,
                                               const int *__restrict__ cprob, const int *__restrict__ kcount) {
  const int ch = blockIdx.x;
  const int64_t s0 = cstart[ch];
  const int len = clen[ch];
  const double *h = coef + (int64_t)cprob[ch] * stride;
  if (kcount) k = min(k, kcount[cprob[ch]]);
  for (int i = threadIdx.x; i < len; i += kT) {
    double v = (double)w[s0 + i];
    for (int j = 0; j < k; j++) v += sign * h[j] * (double)V[(int64_t)j * n + s0 + i];
    w[s0 + i] = (T)v;
  }
}

// Projections of the Arnoldi step j on the stored basis (V_i = s_i W_i, the
// vector w in the scale a_p = s_j (first pass, w = A M^{-1} W_j) or 1 (second
// pass, w already the true vector)): d_i = W_i . w summed over problem p's
// chunks, h_i = s_i a_p d_i (the Hessenberg entry V_i . (a_p w)), and the
// update coefficient c_i = -s_i h_i of W_i in a_p w - sum_i h_i V_i
__global__ void k_bproj_final(int k, const int *__restrict__ cfirst, const double *__restrict__ part,
                              const double *__restrict__ vs, int j, int first_pass, double *__restrict__ h,
                              double *__restrict__ c) {
  const int p = blockIdx.x, i = blockIdx.y;
  if (i >= k) return;
  double t = 0.0;
  for (int ch = cfirst[p] + (int)threadIdx.x; ch < cfirst[p + 1]; ch += 64) t += part[(int64_t)ch * kS + i];
  for (int o = 32; o > 0; o >>= 1) t += __shfl_xor(t, o, 64);
  if (threadIdx.x == 0) {
    const double si = vs[(int64_t)p * kS + i], ap = first_pass ? vs[(int64_t)p * kS + j] : 1.0;
    const double hi = si * ap * t;
    h[(int64_t)p * kS + i] = hi;
    c[(int64_t)p * kS + i] = -si * hi;
  }
}

// w = a_p w + sum_i c[p][i] V[i], i < K, over each chunk of problem p (a_p =
// vs[p][j] when vs is given, else 1), and with part, each chunk's sum of the
// new w^2; K a template constant, so a unit's loads are issued together
template <class T, int K>
__global__ void __launch_bounds__(kT) k_bgs(int64_t n, const T *__restrict__ V, const double *__restrict__ coef,
                                            const double *__restrict__ vs, int j, T *__restrict__ w,
                                            const int64_t *__restrict__ cstart, const int *__restrict__ clen,
                                            const int *__restrict__ cprob, double *__restrict__ part) {
  constexpr int VW = Vec16<T>::N;
  const int ch = blockIdx.x;
  const int64_t s0 = cstart[ch];
  const int len = clen[ch];
  const int p = cprob[ch];
  double h[K];
#pragma unroll
  for (int q = 0; q < K; q++) h[q] = coef[(int64_t)p * kS + q];
  const double a = vs ? vs[(int64_t)p * kS + j] : 1.0;
  double ss = 0.0;
  bool vec = units_ok((const T *)w, s0, len);
#pragma unroll
  for (int q = 0; q < K; q++) vec = vec && units_ok(V + (int64_t)q * n, s0, len);
  if (vec) {
    for (int i = threadIdx.x * VW; i < len; i += kT * VW) {
      const Vec16<T> wu = *reinterpret_cast<const Vec16<T> *>(w + s0 + i);
      Vec16<T> vu[K];
#pragma unroll
      for (int q = 0; q < K; q++) vu[q] = *reinterpret_cast<const Vec16<T> *>(V + (int64_t)q * n + s0 + i);
      Vec16<T> o;
#pragma unroll
      for (int e = 0; e < VW; e++) {
        double v = a * (double)wu.v[e];
#pragma unroll
        for (int q = 0; q < K; q++) v += h[q] * (double)vu[q].v[e];
        o.v[e] = (T)v;
        ss += (double)o.v[e] * (double)o.v[e];
      }
      *reinterpret_cast<Vec16<T> *>(w + s0 + i) = o;
    }
  } else {
    for (int i = threadIdx.x; i < len; i += kT) {
      double v = a * (double)w[s0 + i];
#pragma unroll
      for (int q = 0; q < K; q++) v += h[q] * (double)V[(int64_t)q * n + s0 + i];
      const T o = (T)v;
      w[s0 + i] = o;
      ss += (double)o * (double)o;
    }
  }
  if (!part) return;
  __shared__ double sh[kT / 64];
  for (int o = 32; o > 0; o >>= 1) ss += __shfl_xor(ss, o, 64);
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = ss;
  __syncthreads();
  if (threadIdx.x == 0) {
    double t = 0.0;
    for (int q = 0; q < kT / 64; q++) t += sh[q];
    part[(int64_t)ch * kS] = t;
  }
}

// launchers: the vector count k (1 .. kS - 1) picks the instance
#define UBT_K_CASES(X) \
  X(1) X(2) X(3) X(4) X(5) X(6) X(7) X(8) X(9) X(10) X(11) X(12) X(13) X(14) X(15) X(16) X(17) X(18) X(19) X(20) \
  X(21) X(22) X(23) X(24) X(25) X(26) X(27) X(28) X(29) X(30) X(31)

template <class T>
void launch_bdot(int k, unsigned nc, hipStream_t s, int64_t n, const T *V, const T *w, const int64_t *cst,
                 const int *cln, double *pt) {
  switch (k) {
#define UBT_BDOT(K) \
  case K: hipLaunchKernelGGL((k_bdot<T, K>), dim3(nc), dim3(kT), 0, s, n, V, w, cst, cln, pt); break;
    UBT_K_CASES(UBT_BDOT)
#undef UBT_BDOT
    default: fail(UBT_ERR_ARG_OUTOFRANGE, "%d Krylov vectors (1 to %d)", k, kS - 1);
  }
}

template <class T>
void launch_bgs(int k, unsigned nc, hipStream_t s, int64_t n, const T *V, const double *coef, const double *vs, int j,
                T *w, const int64_t *cst, const int *cln, const int *cpr, double *pt) {
  switch (k) {
#define UBT_BGS(K) \
  case K: hipLaunchKernelGGL((k_bgs<T, K>), dim3(nc), dim3(kT), 0, s, n, V, coef, vs, j, w, cst, cln, cpr, pt); break;
    UBT_K_CASES(UBT_BGS)
#undef UBT_BGS
    default: fail(UBT_ERR_ARG_OUTOFRANGE, "%d Krylov vectors (1 to %d)", k, kS - 1);
  }
}

// v = w * (d[p] != 0 ? 1 / d[p] : 0) (inv) or v = w * d[p] (!inv), converting precision
template <class Ti, class To>
__global__ void __launch_bounds__(kT) k_bscale(const Ti *__restrict__ w, const double *__restrict__ d, int inv,
                                               To *__restrict__ v, const int64_t *__restrict__ cstart,
                                               const int *__restrict__ clen, const int *__restrict__ cprob) {
  const int ch = blockIdx.x;
  const int64_t s0 = cstart[ch];
  const int len = clen[ch];
  const double dd = d[cprob[ch]];
  const double sc = inv ? (dd != 0.0 ? 1.0 / dd : 0.0) : dd;
  for (int i = threadIdx.x; i < len; i += kT) v[s0 + i] = (To)(sc * (double)w[s0 + i]);
}

// x += d[p] * u over problem p's chunks (the refinement update)
template <class T>
__global__ void __launch_bounds__(kT) k_badd(const T *__restrict__ u, const double *__restrict__ d, double *__restrict__ x,
                                             const int64_t *__restrict__ cstart, const int *__restrict__ clen,
                                             const int *__restrict__ cprob) {
  const int ch = blockIdx.x;
  const int64_t s0 = cstart[ch];
  const int len = clen[ch];
  const double dd = d[cprob[ch]];
  for (int i = threadIdx.x; i < len; i += kT) x[s0 + i] += dd * (double)u[s0 + i];
}

template <class T> __global__ void k_fill(int64_t n, T v, T *__restrict__ y) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) y[i] = v;
}

// y = a x + b y; with b = 0, y is not read (it may hold anything, NaN included)
template <class T> __global__ void k_axpby_b(int64_t n, double a, const T *__restrict__ x, double b, T *__restrict__ y) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    y[i] = (T)(b == 0.0 ? a * (double)x[i] : a * (double)x[i] + b * (double)y[i]);
}

// cycle start: g = beta e_0, done where beta <= tol (or the problem is idle)
__global__ void k_bstart(int np, int m, const double *__restrict__ beta, const double *__restrict__ tol,
                         const int *__restrict__ live, double *__restrict__ g, double *__restrict__ vs,
                         int *__restrict__ done, int *__restrict__ jdone) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= np) return;
  for (int i = 0; i < kS; i++) g[(int64_t)p * kS + i] = 0.0;
  g[(int64_t)p * kS] = beta[p];
  vs[(int64_t)p * kS] = beta[p] > 0.0 ? 1.0 / beta[p] : 0.0;  // W_0 is the residual itself
  const bool d = !live[p] || !(beta[p] > tol[p]);
  done[p] = d ? 1 : 0;
  jdone[p] = d ? 0 : m;
}

// Arnoldi column j of every live problem: assemble, rotate, test
__global__ void k_bgivens(int np, int j, const double *__restrict__ h1, const double *__restrict__ h2,
                          const double *__restrict__ hn, double *__restrict__ H, double *__restrict__ cs,
                          double *__restrict__ sn, double *__restrict__ g, double *__restrict__ res,
                          const double *__restrict__ tol, double *__restrict__ vs, int *__restrict__ done,
                          int *__restrict__ jdone, int *__restrict__ its) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= np) return;
  // W_{j+1} is the orthogonalised vector itself, of norm h_{j+1,j} (a
  // finished problem keeps a finite scale too: its span still runs, inert)
  vs[(int64_t)p * kS + j + 1] = hn[p] > 0.0 ? 1.0 / hn[p] : 0.0;
  if (done[p]) return;
  double *Hp = H + (int64_t)p * kS * kS, *c = cs + (int64_t)p * kS, *s = sn + (int64_t)p * kS,
         *gp = g + (int64_t)p * kS;
  for (int i = 0; i <= j; i++) Hp[i * kS + j] = h1[(int64_t)p * kS + i] + (h2 ? h2[(int64_t)p * kS + i] : 0.0);
  const double hnext = hn[p];
  for (int i = 0; i < j; i++) {
    const double a = Hp[i * kS + j], b = Hp[(i + 1) * kS + j];
    Hp[i * kS + j] = c[i] * a + s[i] * b;
    Hp[(i + 1) * kS + j] = -s[i] * a + c[i] * b;
  }
  const double a = Hp[j * kS + j];
  const double rr = hypot(a, hnext);
  c[j] = rr != 0.0 ? a / rr : 1.0;
  s[j] = rr != 0.0 ? hnext / rr : 0.0;
  Hp[j * kS + j] = rr;
  gp[j + 1] = -s[j] * gp[j];
  gp[j] = c[j] * gp[j];
  res[p] = fabs(gp[j + 1]);
  its[p] += 1;
  if (!(res[p] > tol[p]) || hnext == 0.0) {
    done[p] = 1;
    jdone[p] = j + 1;
  }
}

// y = H^{-1} g on the leading jdone x jdone triangle; zero past it
__global__ void k_btri(int np, const double *__restrict__ H, const double *__restrict__ g, const int *__restrict__ jdone,
                       const double *__restrict__ vs, double *__restrict__ y) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= np) return;
  const double *Hp = H + (int64_t)p * kS * kS, *gp = g + (int64_t)p * kS;
  double *yp = y + (int64_t)p * kS;
  const int jd = jdone[p];
  for (int i = kS - 1; i >= 0; i--) {
    if (i >= jd) {
      yp[i] = 0.0;
      continue;
    }
    double v = gp[i];
    for (int k = i + 1; k < jd; k++) v -= Hp[i * kS + k] * yp[k];
    yp[i] = Hp[i * kS + i] != 0.0 ? v / Hp[i * kS + i] : 0.0;
  }
  // Z_i was formed from W_i = V_i / s_i: its coefficient carries s_i
  for (int i = 0; i < jd; i++) yp[i] *= vs[(int64_t)p * kS + i];
}

unsigned grid_n(int64_t n) { return (unsigned)std::max<int64_t>(1, std::min<int64_t>((n + kT - 1) / kT, 8192)); }

template <class T> int prec_id() { return sizeof(T) == 8 ? UBT_FP64 : UBT_FP32; }

}  // namespace

// ---------------------------------------------------------------------------
// Batch
// ---------------------------------------------------------------------------
void Batch::create(const std::vector<Operator *> &o) {
  UBT_CHECK(!o.empty(), UBT_ERR_ARG_SIZ, "a batch needs at least one problem");
  ops = o;
  const Disc &D0 = *o[0]->disc;
  UBT_CHECK(D0.kind == DISC_DG, UBT_ERR_SUP, "batches hold DG problems");
  dim = D0.dim;
  nb = D0.nb;
  nf = D0.nf;
  off.assign(1, 0);
  for (size_t p = 0; p < o.size(); p++) {
    const Disc &D = *o[p]->disc;
    UBT_CHECK(D.kind == DISC_DG && D.dim == dim && D.nb == nb && D.nf == nf, UBT_ERR_ARG_INCOMP,
              "problem %zu is not a %dD DG problem with %d basis functions and %d faces per cell like problem 0", p,
              dim, nb, nf);
    UBT_CHECK(o[p]->xs_set, UBT_ERR_ARG_WRONGSTATE, "set the cross sections of problem %zu first", p);
    UBT_CHECK(o[p]->device == o[0]->device, UBT_ERR_ARG_INCOMP, "the problems of a batch share one device");
    off.push_back(off.back() + D.n_rows());
  }
  n_rows = off.back();
  sweeps.clear();
  for (auto *op : ops) {
    sweeps.emplace_back(new Sweep());
    sweeps.back()->create(op, true);
  }
  int dev = 0;
  UBT_HIP(hipGetDevice(&dev));
  UBT_HIP(hipDeviceGetAttribute(&n_cu, hipDeviceAttributeMultiprocessorCount, dev));
  set_group(0);
  tune();
}

void Batch::set_group(int g) {
  UBT_CHECK(g >= 0, UBT_ERR_ARG_OUTOFRANGE, "group %d is negative", g);
  bool any = false;
  active.assign(ops.size(), 0);
  for (size_t p = 0; p < ops.size(); p++) {
    if (ops[p]->n_groups > g) {
      active[p] = 1;
      ops[p]->group = g;
      any = true;
    }
  }
  UBT_CHECK(any, UBT_ERR_ARG_OUTOFRANGE, "no problem of the batch has group %d", g);
  group = g;
  build_lists();
  gen++;
}

template <class T>
static void build_items(Batch &b, int body) {
  const int pi = sizeof(T) == 8 ? 0 : 1;
  std::vector<BatchItem<T>> items;
  std::vector<int64_t> tiles(1, 0), edges(1, 0);
  size_t lds = 0;
  bool ok = true;
  for (size_t p = 0; p < b.ops.size() && ok; p++) {
    if (!b.active[p]) continue;
    int cpb = 1, lanes = 1;
    size_t l = 0;
    int64_t nt = 0;
    BatchItem<T> it = {};
    if (!batch_body_layout(*b.ops[p], pi == 0 ? UBT_FP64 : UBT_FP32, body, &cpb, &lanes, &l, &nt, &it.box)) {
      ok = false;
      break;
    }
    it.g = b.ops[p]->args<T>();
    it.g.cpb = cpb;
    it.g.red_lanes = lanes;
    it.off = b.off[p];
    items.push_back(it);
    tiles.push_back(tiles.back() + nt);
    edges.push_back(edges.back() + (body >= 2 ? it.box.edge_tiles : 0));
    lds = std::max(lds, l);
  }
  b.body_ok[pi][body] = ok;
  if (!ok) return;
  b.items[pi][body].upload(items);
  b.tiles[pi][body].upload(tiles);
  if (body >= 2) {
    b.edge_off[pi][body].upload(edges);
    b.n_edge[pi][body] = edges.back();
  }
  b.n_items[pi][body] = (int)items.size();
  b.n_tiles[pi][body] = tiles.back();
  b.lds[pi][body] = lds;
}

void Batch::build_lists() {
  for (int body_i = 0; body_i < kBatchBodies; body_i++) {
    build_items<double>(*this, body_i);
    build_items<float>(*this, body_i);
  }
  UBT_CHECK(body_ok[0][0] && body_ok[1][0], UBT_ERR_SUP, "a problem of the batch has more than 256 angles");
  for (int pi = 0; pi < 2; pi++)
    if (!body_ok[pi][body[pi]]) body[pi] = 0;
  // element blocks: prefix of cells x angles over the active problems
  std::vector<int64_t> blk(1, 0);
  for (size_t p = 0; p < ops.size(); p++)
    if (active[p]) blk.push_back(blk.back() + (int64_t)ops[p]->disc->n_cells * ops[p]->disc->n_angles);
  bblk.upload(blk);
  n_blocks = blk.back();
  // sweep, per precision: one workgroup per direction group of every active
  // problem (kBatchSweepGroup directions of one octant), the problems'
  // schedules concatenated; the workgroups are listed longest first (most
  // levels), so the groups that bound the sweep are dispatched before the
  // short ones
  int nfu = 1;
  for (size_t p = 0; p < ops.size(); p++)
    if (active[p]) nfu = std::max(nfu, sweeps[p]->nfu);
  for (int pi = 0; pi < 2; pi++) {
    BatchSweep &w = sweep[pi];
    struct Grp {
      int q, g, nlev;
      size_t p;
    };
    std::vector<Grp> grps;
    std::vector<SweepLists> lists(ops.size());
    std::vector<int64_t> item0(ops.size(), 0);
    std::vector<int> ce, an;
    std::vector<unsigned char> up;
    int q = 0;
    int64_t widest_level = 1;
    for (size_t p = 0; p < ops.size(); p++) {
      if (!active[p]) continue;
      SweepLists &L = lists[p];
      group_schedule(*sweeps[p], kBatchSweepGroup[pi], L);
      item0[p] = (int64_t)ce.size();
      ce.insert(ce.end(), L.cell.begin(), L.cell.end());
      an.insert(an.end(), L.angle.begin(), L.angle.end());
      up.insert(up.end(), L.up.begin(), L.up.end());
      widest_level = std::max(widest_level, L.widest);
      for (int g = 0; g + 1 < (int)L.base.size(); g++) grps.push_back({q, g, L.base[g + 1] - L.base[g] - 1, p});
      q++;
    }
    std::stable_sort(grps.begin(), grps.end(), [](const Grp &x, const Grp &y) { return x.nlev > y.nlev; });
    std::vector<int> wp, bs(1, 0);
    std::vector<int64_t> lv;
    w.widest = 0;
    for (const Grp &d : grps) {
      const SweepLists &L = lists[d.p];
      wp.push_back(d.q);
      for (int k = L.base[d.g]; k < L.base[d.g + 1]; k++) lv.push_back(item0[d.p] + L.lvl[k]);
      bs.push_back((int)lv.size());
      w.widest = std::max<int64_t>(w.widest, L.lvl[L.base[d.g + 1] - 1] - L.lvl[L.base[d.g]]);
    }
    w.n_wg = (int)wp.size();
    w.n_items = (int64_t)ce.size();
    w.threads = (int)std::min<int64_t>(1024, std::max<int64_t>(64, (widest_level + 63) / 64 * 64));
    w.slots = sweep_slots(nf, nfu);
    w.prob.upload(wp);
    w.angle.upload(an);
    w.cells.upload(ce);
    w.upwind.upload(up);
    w.lvl.upload(lv);
    w.base.upload(bs);
    w.rec_tag = ~0ull;
  }
  build_box_sweep();
  sweep_records(nullptr);
  UBT_HIP(hipStreamSynchronize(nullptr));
}

// The record-free box sweep's work list, per precision: every active problem
// a uniform all-vacuum 2D DG1 box (and the strip body's layout, whose work
// items the sweep reads, in place), at most 256 cells along x (the LDS
// diagonal); one workgroup per problem, quadrant and group of the quadrant's
// angles. The groups of one (problem, quadrant) share lines of the
// quadrant-major copies, so they go to one XCD: the sets, longest march
// first, are dealt to the XCD with the least work so far, and workgroup
// i = 8 j + x is the j-th of XCD x's list (idle slots pad the shorter lists)
void Batch::build_box_sweep() {
  constexpr int kXcd = 8;
  for (int pi = 0; pi < 2; pi++) {
    BoxSweepPlan &w = box_sweep[pi];
    const int AG = pi == 0 ? kBoxSweepAngles<double> : kBoxSweepAngles<float>;
    w.ok = dim == 2 && nb == 3 && body_ok[pi][3] && box_sweep_enabled;
    w.nx_max = 0;
    struct Set {
      int diagonals;
      double work;
      std::vector<BoxSweepGroup> g;
    };
    std::vector<Set> sets;
    std::vector<int> qpos;
    std::vector<int64_t> qoff(1, 0);
    int item = 0;
    for (size_t p = 0; p < ops.size() && w.ok; p++) {
      if (!active[p]) continue;
      const Operator &o = *ops[p];
      const Quadrature &q = o.disc->quad;
      w.ok = o.box_cls >= 0 && o.n_box_edge == 0 && o.box_n[0] <= 256 && q.n_angles % 4 == 0;
      if (!w.ok) break;
      w.nx_max = std::max(w.nx_max, o.box_n[0]);
      const int Aq = q.n_angles / 4;
      const int64_t nc = o.disc->n_cells;
      // a diagonal holds at most min(nx, ny) cells, 256 / AG per pass
      const int need = (std::min(o.box_n[0], o.box_n[1]) + 256 / AG - 1) / (256 / AG);
      const int ipt = need <= 1 ? 1 : need <= 2 ? 2 : need <= 4 ? 4 : 8;
      std::vector<int> pos(q.n_angles, -1);
      for (int sx = 0; sx < 2; sx++)
        for (int sy = 0; sy < 2; sy++) {
          const int quad = 2 * sx + sy;
          std::vector<int> an;
          for (int a = 0; a < q.n_angles; a++)
            if ((q.omega[3 * a] > 0.0) == (sx == 0) && (q.omega[3 * a + 1] > 0.0) == (sy == 0)) an.push_back(a);
          if ((int)an.size() != Aq) {
            w.ok = false;
            break;
          }
          for (int j = 0; j < Aq; j++) pos[an[j]] = quad * Aq + j;
          Set st;
          st.diagonals = o.box_n[0] + o.box_n[1] - 1;
          st.work = (double)nc * Aq;
          for (size_t k0 = 0; k0 < an.size(); k0 += AG) {
            BoxSweepGroup gw;
            gw.item = item;
            gw.q = quad;
            gw.k0 = (int)k0;
            for (int k = 0; k < 8; k++) gw.angle[k] = k < AG && k0 + k < an.size() ? an[k0 + k] : -1;
            gw.ipt = ipt;
            gw.rbase = off[p] + (int64_t)quad * nc * 3 * Aq + (int64_t)k0;
            st.g.push_back(gw);
          }
          sets.push_back(std::move(st));
        }
      qpos.insert(qpos.end(), pos.begin(), pos.end());
      qoff.push_back(qoff.back() + q.n_angles);
      item++;
    }
    if (!w.ok) {
      w.n_wg = 0;
      continue;
    }
    std::stable_sort(sets.begin(), sets.end(), [](const Set &x, const Set &y) { return x.diagonals > y.diagonals; });
    std::vector<std::vector<BoxSweepGroup>> lists(kXcd);
    std::vector<double> load(kXcd, 0.0);
    for (const Set &st : sets) {
      const int x = (int)(std::min_element(load.begin(), load.end()) - load.begin());
      load[x] += st.work;
      lists[x].insert(lists[x].end(), st.g.begin(), st.g.end());
    }
    size_t len = 0;
    for (const auto &l : lists) len = std::max(len, l.size());
    BoxSweepGroup idle{};
    idle.item = -1;
    std::vector<BoxSweepGroup> g(len * kXcd, idle);
    for (int x = 0; x < kXcd; x++)
      for (size_t j = 0; j < lists[x].size(); j++) g[j * kXcd + x] = lists[x][j];
    w.n_wg = (int)g.size();
    w.groups.upload(g);
    w.qpos.upload(qpos);
    w.qoff.upload(qoff);
    const size_t ts = pi == 0 ? sizeof(double) : sizeof(float);
    if (w.rq.bytes < ts * (size_t)n_rows) {
      w.rq.alloc(ts * (size_t)n_rows);
      w.pq.alloc(ts * (size_t)n_rows);
    }
  }
}

uint64_t Batch::xs_tag() const {
  uint64_t t = (uint64_t)group * 0x9E3779B97F4A7C15ull;
  for (size_t p = 0; p < ops.size(); p++) t = t * 31 + ops[p]->xs_gen;
  return t;
}

void Batch::sweep_records(hipStream_t s) const {
  const uint64_t tag = xs_tag();
  for (int pi = 0; pi < 2; pi++) {
    const BatchSweep &w = sweep[pi];
    if (box_sweep[pi].ok) continue;  // the box sweep forms its blocks on the fly
    if (w.rec_tag == tag || w.n_items == 0) continue;
    // 64-item blocks (see the sweep's record layout)
    const size_t blocks64 = (size_t)((w.n_items + 63) / 64) * 64;
    const size_t per = (size_t)nb * nb * (1 + w.slots);
    const size_t ts = pi == 0 ? sizeof(double) : sizeof(float);
    if (w.rec.bytes < ts * per * blocks64) {
      w.rec.alloc(ts * per * blocks64);
      w.rnb.alloc(sizeof(int) * (size_t)w.slots * blocks64);
    }
    batch_records_launch(*this, pi, s);
    w.rec_tag = tag;
  }
}

void Batch::mult(int prec, const void *x, void *y, bool scatter, hipStream_t s) const {
  UBT_CHECK(x && y, UBT_ERR_ARG_NULL, "x and y are required");
  UBT_CHECK(x != y, UBT_ERR_ARG_IDN, "the batched apply needs distinct x and y");
  if (prec == UBT_FP64)
    batch_apply_launch<double>(*this, body[0], (const double *)x, (double *)y, scatter, s);
  else if (prec == UBT_FP32)
    batch_apply_launch<float>(*this, body[1], (const float *)x, (float *)y, scatter, s);
  else
    fail(UBT_ERR_ARG_OUTOFRANGE, "unknown precision %d", prec);
}

void Batch::blocks_setup(hipStream_t s) {
  for (size_t p = 0; p < ops.size(); p++) {
    if (!active[p]) continue;
    BlockInverse bi;
    bi.op = ops[p];
    bi.setup(s);  // UBT_ERR_MAT_LU_ZRPVT on a singular block
  }
}

void Batch::blocks_apply(int prec, const void *x, void *y, hipStream_t s) const {
  if (prec == UBT_FP64)
    batch_blocks_launch<double>(*this, (const double *)x, (double *)y, s);
  else
    batch_blocks_launch<float>(*this, (const float *)x, (float *)y, s);
}

void Batch::sweep_apply(int prec, const void *r, void *psi, hipStream_t s) const {
  const int pi = prec == UBT_FP64 ? 0 : 1;
  if (box_sweep[pi].ok) {
    if (prec == UBT_FP64)
      batch_box_sweep_launch<double>(*this, (const double *)r, (double *)psi, s);
    else
      batch_box_sweep_launch<float>(*this, (const float *)r, (float *)psi, s);
    return;
  }
  sweep_records(s);
  if (prec == UBT_FP64)
    batch_sweep_launch<double>(*this, (const double *)r, (double *)psi, s);
  else
    batch_sweep_launch<float>(*this, (const float *)r, (float *)psi, s);
}

bool Batch::sweep_exact(int prec) const {
  if (box_sweep[prec == UBT_FP64 ? 0 : 1].ok) return true;  // uniform all-vacuum boxes: no cycle, no reflection
  for (size_t p = 0; p < ops.size(); p++)
    if (active[p] && !sweeps[p]->exact) return false;
  return true;
}

void Batch::rhs(int prec, void *b, hipStream_t s) const {
  const size_t ts = prec == UBT_FP64 ? 8 : 4;
  for (size_t p = 0; p < ops.size(); p++) {
    char *bp = (char *)b + off[p] * ts;
    const int64_t n = off[p + 1] - off[p];
    if (active[p]) {
      ops[p]->rhs(group, prec, bp, s);
    } else if (prec == UBT_FP64) {
      hipLaunchKernelGGL(k_fill<double>, dim3(grid_n(n)), dim3(kT), 0, s, n, 0.0, (double *)bp);
    } else {
      hipLaunchKernelGGL(k_fill<float>, dim3(grid_n(n)), dim3(kT), 0, s, n, 0.f, (float *)bp);
    }
  }
  UBT_HIP(hipGetLastError());
}

void Batch::set_scalar_flux(int prec, const void *psi, hipStream_t s) const {
  const size_t ts = prec == UBT_FP64 ? 8 : 4;
  for (size_t p = 0; p < ops.size(); p++)
    if (active[p]) ops[p]->set_scalar_flux(group, prec, (const char *)psi + off[p] * ts, s);
}

void Batch::tune() {
  hipStream_t s = nullptr;
  UBT_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
  hipEvent_t e0, e1;
  UBT_HIP(hipEventCreate(&e0));
  UBT_HIP(hipEventCreate(&e1));
  DevBuf xb, yb;
  xb.alloc(sizeof(double) * n_rows);
  yb.alloc(sizeof(double) * n_rows);
  hipLaunchKernelGGL(k_fill<double>, dim3(grid_n(n_rows)), dim3(kT), 0, s, n_rows, 0.0, xb.as<double>());
  const int persists[3] = {2, 4, 8};
  for (int pi = 0; pi < 2; pi++) {
    float best = 1e30f;
    int best_b = 0, best_p = 8;
    for (int bo = 0; bo < kBatchBodies; bo++) {
      tuned_us[pi][bo] = 0.f;
      if (!body_ok[pi][bo]) continue;
      float bb = 1e30f;
      for (int pp : persists) {
        body[pi] = bo;
        persist[pi] = pp;
        const int prec = pi == 0 ? UBT_FP64 : UBT_FP32;
        for (int it = 0; it < 2; it++) mult(prec, xb.ptr, yb.ptr, true, s);
        UBT_HIP(hipEventRecord(e0, s));
        for (int it = 0; it < 5; it++) mult(prec, xb.ptr, yb.ptr, true, s);
        UBT_HIP(hipEventRecord(e1, s));
        UBT_HIP(hipEventSynchronize(e1));
        float ms = 0.f;
        UBT_HIP(hipEventElapsedTime(&ms, e0, e1));
        const float us = ms * 1000.f / 5.f;
        bb = std::min(bb, us);
        if (us < best) {
          best = us;
          best_b = bo;
          best_p = pp;
        }
      }
      tuned_us[pi][bo] = bb;
    }
    body[pi] = best_b;
    persist[pi] = best_p;
  }
  UBT_HIP(hipStreamSynchronize(s));
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  (void)hipStreamDestroy(s);
}

// ---------------------------------------------------------------------------
// BatchSolver
// ---------------------------------------------------------------------------
void BatchSolver::drop_graphs() {
  for (auto &v : graphs) {
    for (auto g : v) (void)hipGraphExecDestroy(g);
    v.clear();
  }
}

BatchSolver::~BatchSolver() {
  drop_graphs();
  if (cap) (void)hipStreamDestroy(cap);
}

void BatchSolver::create(Batch *b, const SolverOptions &o) {
  batch = b;
  opt = o;
  UBT_CHECK(opt.restart >= 1 && opt.restart + 1 <= kS, UBT_ERR_ARG_OUTOFRANGE, "restart must be 1 to %d, got %d", kS - 1,
            opt.restart);
  UBT_CHECK(opt.rtol > 0.0 && opt.max_it >= 1, UBT_ERR_ARG_OUTOFRANGE, "rtol must be positive and max_it >= 1");
  n = b->n_rows;
  np = (int)b->ops.size();
  const int m = opt.restart;
  if (!opt.mixed) {
    V64.alloc(sizeof(double) * n * (m + 1));
    Z64.alloc(sizeof(double) * n * m);
    w64.alloc(sizeof(double) * n);
    t64.alloc(sizeof(double) * n);
    u64.alloc(sizeof(double) * n);
    b32.alloc(sizeof(double) * n);  // the internal right-hand side (FP64 here)
    d32.alloc(sizeof(double) * n);  // the internal solution (FP64 here)
  } else {
    V32.alloc(sizeof(float) * n * (m + 1));
    Z32.alloc(sizeof(float) * n * m);
    t32.alloc(sizeof(float) * n);
    u32.alloc(sizeof(float) * n);
    d32.alloc(sizeof(float) * n);
    b32.alloc(sizeof(float) * n);
    r64.alloc(sizeof(double) * n);
  }
  const size_t P = (size_t)np;
  H.alloc(sizeof(double) * P * kS * kS);
  cs.alloc(sizeof(double) * P * kS);
  sn.alloc(sizeof(double) * P * kS);
  gv.alloc(sizeof(double) * P * kS);
  hc.alloc(sizeof(double) * P * kS * 4);  // h and c of both Gram-Schmidt passes
  vsc.alloc(sizeof(double) * P * kS);
  yv.alloc(sizeof(double) * P * kS);
  nrm.alloc(sizeof(double) * P);
  tol.alloc(sizeof(double) * P);
  res.alloc(sizeof(double) * P);
  onrm.alloc(sizeof(double) * P);
  bnorm.alloc(sizeof(double) * P);
  done.alloc(sizeof(int) * P * 4);  // done, jdone, its, live
  UBT_HIP(hipStreamCreateWithFlags(&cap, hipStreamNonBlocking));
}

void BatchSolver::refresh(hipStream_t s) {
  batch->blocks_setup(s);
  // chunks of the active problems; the live flags
  std::vector<int64_t> st;
  std::vector<int> ln, pr, first(np + 1, 0), live(np, 0);
  for (int p = 0; p < np; p++) {
    first[p] = (int)st.size();
    if (!batch->active[p]) continue;
    live[p] = 1;
    for (int64_t r = batch->off[p]; r < batch->off[p + 1]; r += kChunk) {
      st.push_back(r);
      ln.push_back((int)std::min<int64_t>(kChunk, batch->off[p + 1] - r));
      pr.push_back(p);
    }
  }
  first[np] = (int)st.size();
  n_chunks = (int)st.size();
  chunk_start.upload(st);
  chunk_len.upload(ln);
  chunk_prob.upload(pr);
  part.alloc(sizeof(double) * (size_t)std::max(1, n_chunks) * kS);
  cfirst.upload(first);
  UBT_HIP(hipMemcpyAsync(done.as<int>() + 3 * np, live.data(), sizeof(int) * np, hipMemcpyHostToDevice, s));
  UBT_HIP(hipStreamSynchronize(s));
  batch->sweep_records(s);
  drop_graphs();
  gen = batch->gen;
  xs_tag = batch->xs_tag();
}

// y = M^{-1} r: element blocks, then the sweep on the L-residual. Where the
// sweep is L's exact inverse S the composite is the sweep alone:
// y + S (r - L y) = S r for any first stage y, so the blocks, the residual
// apply and the two vector updates are left out (the same result up to
// rounding)
template <class T> void BatchSolver::apply_pc(const T *r, T *y, hipStream_t s) {
  T *t = sizeof(T) == 8 ? t64.as<T>() : t32.as<T>();
  T *u = sizeof(T) == 8 ? u64.as<T>() : u32.as<T>();
  const int prec = prec_id<T>();
  const unsigned gn = grid_n(n);
  bool first = true;
  if (opt.pc_blocks && !(opt.pc_sweep && batch->sweep_exact(prec))) {
    batch->blocks_apply(prec, r, y, s);
    first = false;
  }
  if (opt.pc_sweep) {
    if (first) {
      batch->sweep_apply(prec, r, y, s);
      first = false;
    } else {
      batch->mult(prec, y, t, false, s);
      hipLaunchKernelGGL(k_axpby_b<T>, dim3(gn), dim3(kT), 0, s, n, 1.0, r, -1.0, t);
      batch->sweep_apply(prec, t, u, s);
      hipLaunchKernelGGL(k_axpby_b<T>, dim3(gn), dim3(kT), 0, s, n, 1.0, (const T *)u, 1.0, y);
    }
  }
  if (first) hipLaunchKernelGGL(k_axpby_b<T>, dim3(gn), dim3(kT), 0, s, n, 1.0, r, 0.0, y);
  UBT_HIP(hipGetLastError());
}

// A restart cycle in three parts: start (r0, beta, V0), Arnoldi steps
// [j0, j1), end (the least-squares update of x). b, x are internal vectors.
template <class T> void BatchSolver::cycle_start(const T *b, const T *x, hipStream_t s) {
  const int m = opt.restart, prec = prec_id<T>();
  T *V = sizeof(T) == 8 ? V64.as<T>() : V32.as<T>();
  const int64_t *cst = chunk_start.as<int64_t>();
  const int *cln = chunk_len.as<int>(), *cfp = cfirst.as<int>();
  int *dn = done.as<int>(), *jd = dn + np, *live = dn + 3 * np;
  double *hn = nrm.as<double>(), *pt = part.as<double>();
  const unsigned gn = grid_n(n), gp = (unsigned)((np + 63) / 64), nc = (unsigned)n_chunks;
  // W_0 = b - A x in V's first slot, s_0 = 1 / beta
  batch->mult(prec, x, V, true, s);
  hipLaunchKernelGGL(k_axpby_b<T>, dim3(gn), dim3(kT), 0, s, n, 1.0, b, -1.0, V);
  launch_bdot<T>(1, nc, s, n, (const T *)V, (const T *)V, cst, cln, pt);
  hipLaunchKernelGGL(k_bdot_final, dim3(np, 1), dim3(64), 0, s, 1, cfp, (const double *)pt, hn, 1, 1);
  hipLaunchKernelGGL(k_bstart, dim3(gp), dim3(64), 0, s, np, m, (const double *)hn, (const double *)tol.as<double>(),
                     (const int *)live, gv.as<double>(), vsc.as<double>(), dn, jd);
  UBT_HIP(hipGetLastError());
}

template <class T> void BatchSolver::cycle_steps(int j0, int j1, hipStream_t s) {
  const int prec = prec_id<T>();
  T *V = sizeof(T) == 8 ? V64.as<T>() : V32.as<T>();
  T *Z = sizeof(T) == 8 ? Z64.as<T>() : Z32.as<T>();
  const int64_t *cst = chunk_start.as<int64_t>();
  const int *cln = chunk_len.as<int>(), *cpr = chunk_prob.as<int>(), *cfp = cfirst.as<int>();
  int *dn = done.as<int>(), *jd = dn + np, *its = dn + 2 * np;
  const size_t ps = (size_t)np * kS;
  double *h1 = hc.as<double>(), *h2 = h1 + ps, *c1 = h1 + 2 * ps, *c2 = h1 + 3 * ps;
  double *hn = nrm.as<double>(), *pt = part.as<double>(), *vs = vsc.as<double>();
  const unsigned gp = (unsigned)((np + 63) / 64), nc = (unsigned)n_chunks;
  const bool two = opt.gs_passes > 1;
  for (int j = j0; j < j1; j++) {
    // W_{j+1} is built in place in V's slot j + 1: A M^{-1} W_j, then the
    // projections removed and the scale s_j applied in one pass (the last
    // pass also sums its squares for h_{j+1,j})
    T *wj = V + (int64_t)j * n, *zj = Z + (int64_t)j * n, *w = V + (int64_t)(j + 1) * n;
    apply_pc<T>(wj, zj, s);
    batch->mult(prec, zj, w, true, s);
    launch_bdot<T>(j + 1, nc, s, n, (const T *)V, (const T *)w, cst, cln, pt);
    hipLaunchKernelGGL(k_bproj_final, dim3(np, j + 1), dim3(64), 0, s, j + 1, cfp, (const double *)pt,
                       (const double *)vs, j, 1, h1, c1);
    launch_bgs<T>(j + 1, nc, s, n, (const T *)V, (const double *)c1, (const double *)vs, j, w, cst, cln, cpr,
                  two ? nullptr : pt);
    if (two) {  // classical Gram-Schmidt twice
      launch_bdot<T>(j + 1, nc, s, n, (const T *)V, (const T *)w, cst, cln, pt);
      hipLaunchKernelGGL(k_bproj_final, dim3(np, j + 1), dim3(64), 0, s, j + 1, cfp, (const double *)pt,
                         (const double *)vs, j, 0, h2, c2);
      launch_bgs<T>(j + 1, nc, s, n, (const T *)V, (const double *)c2, (const double *)nullptr, 0, w, cst, cln,
                    cpr, pt);
    }
    hipLaunchKernelGGL(k_bdot_final, dim3(np, 1), dim3(64), 0, s, 1, cfp, (const double *)pt, hn, 1, 1);
    hipLaunchKernelGGL(k_bgivens, dim3(gp), dim3(64), 0, s, np, j, (const double *)h1,
                       (const double *)(two ? h2 : nullptr),
                       (const double *)hn, H.as<double>(), cs.as<double>(), sn.as<double>(), gv.as<double>(),
                       res.as<double>(), (const double *)tol.as<double>(), vs, dn, jd, its);
  }
  UBT_HIP(hipGetLastError());
}

template <class T> void BatchSolver::cycle_end(T *x, hipStream_t s) {
  const int m = opt.restart;
  T *Z = sizeof(T) == 8 ? Z64.as<T>() : Z32.as<T>();
  const int64_t *cst = chunk_start.as<int64_t>();
  const int *cln = chunk_len.as<int>(), *cpr = chunk_prob.as<int>();
  const int *jd = done.as<int>() + np;
  const unsigned gp = (unsigned)((np + 63) / 64), nc = (unsigned)n_chunks;
  hipLaunchKernelGGL(k_btri, dim3(gp), dim3(64), 0, s, np, (const double *)H.as<double>(),
                     (const double *)gv.as<double>(), jd, (const double *)vsc.as<double>(), yv.as<double>());
  hipLaunchKernelGGL(k_bmaxpy<T>, dim3(nc), dim3(kT), 0, s, n, m, (const T *)Z, (const double *)yv.as<double>(), kS, 1.0,
                     x, cst, cln, cpr, jd);
  UBT_HIP(hipGetLastError());
}

// Record what fn enqueues as a graph (on the capture stream)
template <class F> static hipGraphExec_t capture(hipStream_t cap, F &&fn) {
  hipGraph_t gr = nullptr;
  hipGraphExec_t ex = nullptr;
  UBT_HIP(hipStreamBeginCapture(cap, hipStreamCaptureModeThreadLocal));
  fn();
  UBT_HIP(hipStreamEndCapture(cap, &gr));
  UBT_HIP(hipGraphInstantiate(&ex, gr, nullptr, nullptr, 0));
  (void)hipGraphDestroy(gr);
  return ex;
}

// Restart cycles until every problem is done: the start, blocks of
// kStepsPerCheck Arnoldi steps, and the end are each a graph; the host reads
// the done flags after every block, so a cycle stops as soon as every
// problem has converged instead of running all restart steps.
template <class T>
void BatchSolver::fgmres(const T *b, T *x, const double *tol_abs, int max_cycles, hipStream_t s) {
  const int pi = sizeof(T) == 8 ? 0 : 1, m = opt.restart;
  const int nblk = (m + kStepsPerCheck - 1) / kStepsPerCheck;
  UBT_HIP(hipMemcpyAsync(tol.as<double>(), tol_abs, sizeof(double) * np, hipMemcpyHostToDevice, s));
  UBT_HIP(hipStreamSynchronize(s));
  auto &gx = graphs[pi];
  if (use_graph && gx.empty()) {
    gx.push_back(capture(cap, [&] { cycle_start<T>(b, x, cap); }));
    for (int k = 0; k < nblk; k++)
      gx.push_back(capture(cap, [&] { cycle_steps<T>(k * kStepsPerCheck, std::min(m, (k + 1) * kStepsPerCheck), cap); }));
    gx.push_back(capture(cap, [&] { cycle_end<T>(x, cap); }));
  }
  std::vector<int> dn(np);
  auto all_done = [&] {
    UBT_HIP(hipMemcpyAsync(dn.data(), done.as<int>(), sizeof(int) * np, hipMemcpyDeviceToHost, cap));
    UBT_HIP(hipStreamSynchronize(cap));
    for (int p = 0; p < np; p++)
      if (!dn[p]) return false;
    return true;
  };
  for (int c = 0; c < max_cycles; c++) {
    if (use_graph)
      UBT_HIP(hipGraphLaunch(gx[0], cap));
    else
      cycle_start<T>(b, x, cap);
    for (int k = 0; k < nblk; k++) {
      if (use_graph)
        UBT_HIP(hipGraphLaunch(gx[1 + k], cap));
      else
        cycle_steps<T>(k * kStepsPerCheck, std::min(m, (k + 1) * kStepsPerCheck), cap);
      if (all_done()) break;
    }
    if (use_graph)
      UBT_HIP(hipGraphLaunch(gx[1 + nblk], cap));
    else
      cycle_end<T>(x, cap);
    cycles++;
    if (all_done()) break;
  }
}

void BatchSolver::solve(const double *b, double *x, hipStream_t s) {
  UBT_CHECK(b && x, UBT_ERR_ARG_NULL, "b and x are required");
  UBT_CHECK(b != x, UBT_ERR_ARG_IDN, "the solve needs distinct b and x");
  UBT_CHECK(n_chunks > 0 && gen == batch->gen, UBT_ERR_ARG_WRONGSTATE,
            "refresh the batch solver after setting the batch group");
  const int m = opt.restart;
  const int max_cycles = std::max(1, (opt.max_it + m - 1) / m);
  const int64_t *cst = chunk_start.as<int64_t>();
  const int *cln = chunk_len.as<int>(), *cpr = chunk_prob.as<int>(), *cfp = cfirst.as<int>();
  const unsigned gn = grid_n(n), nc = (unsigned)n_chunks;
  double *pt = part.as<double>();
  // the caller's stream, then the capture stream
  UBT_HIP(hipStreamSynchronize(s));
  // a member's cross sections changed since the refresh (ubt_op_set_xs): the
  // captured cycles would replay the old sweep records, so check the blocks,
  // rebuild the records outside any capture and capture the cycles again
  const uint64_t tag = batch->xs_tag();
  if (tag != xs_tag) {
    drop_graphs();
    batch->blocks_setup(cap);
    batch->sweep_records(cap);
    xs_tag = tag;
  }
  cycles = outer = 0;
  iterations.assign(np, 0);
  reason.assign(np, 0);
  rel_residual.assign(np, 0.0);
  UBT_HIP(hipMemsetAsync(done.as<int>() + 2 * np, 0, sizeof(int) * np, cap));  // iteration counters
  // ||b|| per problem
  launch_bdot<double>(1, nc, cap, n, b, b, cst, cln, pt);
  hipLaunchKernelGGL(k_bdot_final, dim3(np, 1), dim3(64), 0, cap, 1, cfp, (const double *)pt, bnorm.as<double>(), 1, 1);
  std::vector<double> bn(np), rn(np), tl(np);
  std::vector<int> live(np);
  UBT_HIP(hipMemcpyAsync(bn.data(), bnorm.as<double>(), sizeof(double) * np, hipMemcpyDeviceToHost, cap));
  UBT_HIP(hipMemcpyAsync(live.data(), done.as<int>() + 3 * np, sizeof(int) * np, hipMemcpyDeviceToHost, cap));
  UBT_HIP(hipStreamSynchronize(cap));
  if (!opt.mixed) {
    double *bi = b32.as<double>(), *xi = d32.as<double>();
    hipLaunchKernelGGL(k_axpby_b<double>, dim3(gn), dim3(kT), 0, cap, n, 1.0, b, 0.0, bi);
    hipLaunchKernelGGL(k_fill<double>, dim3(gn), dim3(kT), 0, cap, n, 0.0, xi);
    for (int p = 0; p < np; p++) tl[p] = opt.rtol * bn[p];
    fgmres<double>(bi, xi, tl.data(), max_cycles, cap);
    // true residuals
    double *w = w64.as<double>();
    batch->mult(UBT_FP64, xi, w, true, cap);
    hipLaunchKernelGGL(k_axpby_b<double>, dim3(gn), dim3(kT), 0, cap, n, 1.0, b, -1.0, w);
    launch_bdot<double>(1, nc, cap, n, (const double *)w, (const double *)w, cst, cln, pt);
    hipLaunchKernelGGL(k_bdot_final, dim3(np, 1), dim3(64), 0, cap, 1, cfp, (const double *)pt, nrm.as<double>(), 1, 1);
    hipLaunchKernelGGL(k_axpby_b<double>, dim3(gn), dim3(kT), 0, cap, n, 1.0, (const double *)xi, 0.0, x);
    UBT_HIP(hipMemcpyAsync(rn.data(), nrm.as<double>(), sizeof(double) * np, hipMemcpyDeviceToHost, cap));
    outer = 1;
  } else {
    double *r = r64.as<double>();
    float *bf = b32.as<float>(), *d = d32.as<float>();
    hipLaunchKernelGGL(k_fill<double>, dim3(gn), dim3(kT), 0, cap, n, 0.0, x);
    for (outer = 0; outer <= opt.max_outer; outer++) {
      batch->mult(UBT_FP64, x, r, true, cap);
      hipLaunchKernelGGL(k_axpby_b<double>, dim3(gn), dim3(kT), 0, cap, n, 1.0, b, -1.0, r);
      launch_bdot<double>(1, nc, cap, n, (const double *)r, (const double *)r, cst, cln, pt);
      hipLaunchKernelGGL(k_bdot_final, dim3(np, 1), dim3(64), 0, cap, 1, cfp, (const double *)pt, onrm.as<double>(), 1, 1);
      UBT_HIP(hipMemcpyAsync(rn.data(), onrm.as<double>(), sizeof(double) * np, hipMemcpyDeviceToHost, cap));
      UBT_HIP(hipStreamSynchronize(cap));
      bool all = true;
      for (int p = 0; p < np; p++) {
        const bool conv = !live[p] || rn[p] <= opt.rtol * bn[p];
        all = all && conv;
        tl[p] = conv ? 1e300 : opt.inner_rtol;  // a converged problem sits the inner solve out
      }
      int its_max = 0;
      {
        std::vector<int> it(np);
        UBT_HIP(hipMemcpyAsync(it.data(), done.as<int>() + 2 * np, sizeof(int) * np, hipMemcpyDeviceToHost, cap));
        UBT_HIP(hipStreamSynchronize(cap));
        for (int v : it) its_max = std::max(its_max, v);
      }
      if (all || outer == opt.max_outer || its_max >= opt.max_it) break;
      // correction: A d = r / |r| per problem, FP32 storage, to inner_rtol
      hipLaunchKernelGGL((k_bscale<double, float>), dim3(nc), dim3(kT), 0, cap, (const double *)r,
                         (const double *)onrm.as<double>(), 1, bf, cst, cln, cpr);
      hipLaunchKernelGGL(k_fill<float>, dim3(gn), dim3(kT), 0, cap, n, 0.f, d);
      fgmres<float>(bf, d, tl.data(), max_cycles, cap);
      // the inner cycles reuse nrm; the outer scale has its own buffer
      hipLaunchKernelGGL(k_badd<float>, dim3(nc), dim3(kT), 0, cap, (const float *)d, (const double *)onrm.as<double>(),
                         x, cst, cln, cpr);
      UBT_HIP(hipGetLastError());
    }
  }
  std::vector<int> it(np);
  UBT_HIP(hipMemcpyAsync(it.data(), done.as<int>() + 2 * np, sizeof(int) * np, hipMemcpyDeviceToHost, cap));
  UBT_HIP(hipStreamSynchronize(cap));
  for (int p = 0; p < np; p++) {
    iterations[p] = it[p];
    rel_residual[p] = live[p] ? (bn[p] > 0.0 ? rn[p] / bn[p] : 0.0) : 0.0;
    reason[p] = !live[p] ? 0 : (rel_residual[p] <= opt.rtol ? 1 : -1);
  }
}

template void BatchSolver::apply_pc<double>(const double *, double *, hipStream_t);
template void BatchSolver::apply_pc<float>(const float *, float *, hipStream_t);

}  // namespace ubt
