// Right-preconditioned flexible GMRES on the device over np problems laid
// end to end (the batch's members, or one problem): the segmented vector
// kernels, the device-side Hessenberg and rotations, and the restart cycles
// captured as HIP graphs. Dot products accumulate in FP64 whatever the vector
// storage, in a fixed order (deterministic).
//
// Mixed precision: the FP64 true residual r = b - A x drives an iterative
// refinement whose correction solves A d = r / ||r|| with FP32 vectors and
// FP32 applies and preconditioners, FP64 orthogonalisation arithmetic.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>

#include "fgmres.hpp"

namespace ubt {

namespace {

constexpr int kS = kKrylovStride;

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

// part[j][ch] = V_j . w over chunk ch for the K vectors (K a template
// constant, so every load of a unit - w and the K vectors - is
// unconditional and issued together)
template <class T, int K>
__global__ void __launch_bounds__(kT) k_bdot(int64_t n, const T *__restrict__ V, const T *__restrict__ w,
                                             const int64_t *__restrict__ cstart, const int *__restrict__ clen,
                                             double *__restrict__ part, int nch) {
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
    part[(int64_t)threadIdx.x * nch + ch] = t;
  }
}

// the sum of column col's partials over problem p's chunks, a whole block:
// thread-strided (consecutive chunks, consecutive addresses), then the
// wavefronts' butterflies and their sums in order on thread 0 (the order is
// fixed: deterministic)
__device__ double problem_sum(const double *__restrict__ part, int nch, int col, int first, int last) {
  double t = 0.0;
  const double *pc = part + (int64_t)col * nch;
  for (int ch = first + (int)threadIdx.x; ch < last; ch += kT) t += pc[ch];
  for (int o = 32; o > 0; o >>= 1) t += __shfl_xor(t, o, 64);
  __shared__ double sh[kT / 64];
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = t;
  __syncthreads();
  double r = 0.0;
  if (threadIdx.x == 0)
    for (int q = 0; q < kT / 64; q++) r += sh[q];
  return r;
}

// out[p * stride + j] = (sqrt of) the sum of problem p's chunk partials of
// column j; one block per (problem, column)
__global__ void __launch_bounds__(kT) k_bdot_final(int k, const int *__restrict__ cfirst, const double *__restrict__ part,
                                                   int nch, double *__restrict__ out, int stride, int take_sqrt) {
  const int p = blockIdx.x, j = blockIdx.y;
  if (j >= k) return;
  const double t = problem_sum(part, nch, j, cfirst[p], cfirst[p + 1]);
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
__global__ void __launch_bounds__(kT) k_bproj_final(int k, const int *__restrict__ cfirst,
                                                    const double *__restrict__ part, int nch,
                                                    const double *__restrict__ vs, int j, int first_pass,
                                                    double *__restrict__ h, double *__restrict__ c) {
  const int p = blockIdx.x, i = blockIdx.y;
  if (i >= k) return;
  const double t = problem_sum(part, nch, i, cfirst[p], cfirst[p + 1]);
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
    part[ch] = t;
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
  case K: hipLaunchKernelGGL((k_bdot<T, K>), dim3(nc), dim3(kT), 0, s, n, V, w, cst, cln, pt, (int)nc); break;
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
                         const int *__restrict__ live, const int *__restrict__ its, int max_it,
                         double *__restrict__ g, double *__restrict__ vs, int *__restrict__ done,
                         int *__restrict__ jdone) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= np) return;
  for (int i = 0; i < kS; i++) g[(int64_t)p * kS + i] = 0.0;
  g[(int64_t)p * kS] = beta[p];
  vs[(int64_t)p * kS] = beta[p] > 0.0 ? 1.0 / beta[p] : 0.0;  // W_0 is the residual itself
  const bool d = !live[p] || !(beta[p] > tol[p]) || its[p] >= max_it;
  done[p] = d ? 1 : 0;
  jdone[p] = d ? 0 : m;
}

// Arnoldi column j of every live problem: assemble, rotate, test
__global__ void k_bgivens(int np, int j, const double *__restrict__ h1, const double *__restrict__ h2,
                          const double *__restrict__ hn, double *__restrict__ H, double *__restrict__ cs,
                          double *__restrict__ sn, double *__restrict__ g, double *__restrict__ res,
                          const double *__restrict__ tol, double *__restrict__ vs, int *__restrict__ done,
                          int *__restrict__ jdone, int *__restrict__ its, int max_it, double *__restrict__ hist,
                          int hcap) {
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
  if (hist && its[p] < hcap) hist[(int64_t)p * hcap + its[p]] = res[p];
  its[p] += 1;
  if (!(res[p] > tol[p]) || hnext == 0.0 || its[p] >= max_it) {
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


}  // namespace

template <class T> void vec_axpby(int64_t n, double a, const T *x, double b, T *y, hipStream_t s) {
  hipLaunchKernelGGL(k_axpby_b<T>, dim3(grid_n(n)), dim3(kT), 0, s, n, a, x, b, y);
  UBT_HIP(hipGetLastError());
}
template <class T> void vec_fill(int64_t n, T v, T *y, hipStream_t s) {
  hipLaunchKernelGGL(k_fill<T>, dim3(grid_n(n)), dim3(kT), 0, s, n, v, y);
  UBT_HIP(hipGetLastError());
}
template void vec_axpby<double>(int64_t, double, const double *, double, double *, hipStream_t);
template void vec_axpby<float>(int64_t, double, const float *, double, float *, hipStream_t);
template void vec_fill<double>(int64_t, double, double *, hipStream_t);
template void vec_fill<float>(int64_t, float, float *, hipStream_t);

void Fgmres::drop_graphs() {
  for (auto &v : graphs) {
    for (auto g : v) (void)hipGraphExecDestroy(g);
    v.clear();
  }
}

Fgmres::~Fgmres() {
  drop_graphs();
  if (cap) (void)hipStreamDestroy(cap);
}

void Fgmres::create(KrylovSystem *system, int64_t rows, int problems, const SolverOptions &o, int hcap) {
  sys = system;
  opt = o;
  UBT_CHECK(opt.restart >= 1 && opt.restart + 1 <= kS, UBT_ERR_ARG_OUTOFRANGE, "restart must be 1 to %d, got %d", kS - 1,
            opt.restart);
  UBT_CHECK(opt.rtol > 0.0 && opt.max_it >= 1, UBT_ERR_ARG_OUTOFRANGE, "rtol must be positive and max_it >= 1");
  n = rows;
  np = problems;
  const int m = opt.restart;
  if (!opt.mixed) {
    V64.alloc(sizeof(double) * n * (m + 1));
    Z64.alloc(sizeof(double) * n * m);
    b32.alloc(sizeof(double) * n);  // the internal right-hand side (FP64 here)
    d32.alloc(sizeof(double) * n);  // the internal solution (FP64 here)
  } else {
    V32.alloc(sizeof(float) * n * (m + 1));
    Z32.alloc(sizeof(float) * n * m);
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
  hist_cap = hcap;
  if (hist_cap > 0) hist.alloc(sizeof(double) * P * hist_cap);
  // about 2048 reduction chunks over the vector (enough blocks to fill the
  // GPU on a single problem), 1024 to 16384 rows, a multiple of 256
  chunk = std::min<int64_t>(16384, std::max<int64_t>(1024, (n / 2048 + 255) / 256 * 256));
  if (!cap) UBT_HIP(hipStreamCreateWithFlags(&cap, hipStreamNonBlocking));
}

void Fgmres::set_problems(const std::vector<int64_t> &off, const std::vector<char> &active, hipStream_t s) {
  std::vector<int64_t> st;
  std::vector<int> ln, pr, first(np + 1, 0);
  live.assign(np, 0);
  for (int p = 0; p < np; p++) {
    first[p] = (int)st.size();
    if (!active[p]) continue;
    live[p] = 1;
    for (int64_t r = off[p]; r < off[p + 1]; r += chunk) {
      st.push_back(r);
      ln.push_back((int)std::min<int64_t>(chunk, off[p + 1] - r));
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
  drop_graphs();
}

// A restart cycle in three parts: start (W_0 = b - A x, beta, s_0), Arnoldi
// steps [j0, j1), end (the least-squares update of x). b, x are internal
// vectors.
template <class T> void Fgmres::cycle_start(const T *b, const T *x, hipStream_t s) {
  const int m = opt.restart, prec = prec_id<T>();
  T *V = sizeof(T) == 8 ? V64.as<T>() : V32.as<T>();
  const int64_t *cst = chunk_start.as<int64_t>();
  const int *cln = chunk_len.as<int>(), *cfp = cfirst.as<int>();
  int *dn = done.as<int>(), *jd = dn + np, *its = dn + 2 * np, *lv = dn + 3 * np;
  double *hn = nrm.as<double>(), *pt = part.as<double>();
  const unsigned gn = grid_n(n), gp = (unsigned)((np + 63) / 64), nc = (unsigned)n_chunks;
  // W_0 = b - A x in V's first slot, s_0 = 1 / beta
  sys->mult(prec, x, V, s);
  hipLaunchKernelGGL(k_axpby_b<T>, dim3(gn), dim3(kT), 0, s, n, 1.0, b, -1.0, V);
  launch_bdot<T>(1, nc, s, n, (const T *)V, (const T *)V, cst, cln, pt);
  hipLaunchKernelGGL(k_bdot_final, dim3(np, 1), dim3(kT), 0, s, 1, cfp, (const double *)pt, n_chunks, hn, 1, 1);
  hipLaunchKernelGGL(k_bstart, dim3(gp), dim3(64), 0, s, np, m, (const double *)hn, (const double *)tol.as<double>(),
                     (const int *)lv, (const int *)its, opt.max_it, gv.as<double>(), vsc.as<double>(), dn, jd);
  UBT_HIP(hipGetLastError());
}

template <class T> void Fgmres::cycle_steps(int j0, int j1, hipStream_t s) {
  const int prec = prec_id<T>();
  T *V = sizeof(T) == 8 ? V64.as<T>() : V32.as<T>();
  T *Z = sizeof(T) == 8 ? Z64.as<T>() : Z32.as<T>();
  const int64_t *cst = chunk_start.as<int64_t>();
  const int *cln = chunk_len.as<int>(), *cpr = chunk_prob.as<int>(), *cfp = cfirst.as<int>();
  int *dn = done.as<int>(), *jd = dn + np, *its = dn + 2 * np;
  const size_t ps = (size_t)np * kS;
  double *h1 = hc.as<double>(), *h2 = h1 + ps, *c1 = h1 + 2 * ps, *c2 = h1 + 3 * ps;
  double *hn = nrm.as<double>(), *pt = part.as<double>(), *vs = vsc.as<double>();
  double *hs = hist_cap > 0 ? hist.as<double>() : nullptr;
  const unsigned gp = (unsigned)((np + 63) / 64), nc = (unsigned)n_chunks;
  const bool two = opt.gs_passes > 1;
  for (int j = j0; j < j1; j++) {
    // W_{j+1} is built in place in V's slot j + 1: A M^{-1} W_j, then the
    // projections removed and the scale s_j applied in one pass (the last
    // pass also sums its squares for h_{j+1,j})
    T *wj = V + (int64_t)j * n, *zj = Z + (int64_t)j * n, *w = V + (int64_t)(j + 1) * n;
    sys->pc(prec, wj, zj, s);
    sys->mult(prec, zj, w, s);
    launch_bdot<T>(j + 1, nc, s, n, (const T *)V, (const T *)w, cst, cln, pt);
    hipLaunchKernelGGL(k_bproj_final, dim3(np, j + 1), dim3(kT), 0, s, j + 1, cfp, (const double *)pt, n_chunks,
                       (const double *)vs, j, 1, h1, c1);
    launch_bgs<T>(j + 1, nc, s, n, (const T *)V, (const double *)c1, (const double *)vs, j, w, cst, cln, cpr,
                  two ? nullptr : pt);
    if (two) {  // classical Gram-Schmidt twice
      launch_bdot<T>(j + 1, nc, s, n, (const T *)V, (const T *)w, cst, cln, pt);
      hipLaunchKernelGGL(k_bproj_final, dim3(np, j + 1), dim3(kT), 0, s, j + 1, cfp, (const double *)pt, n_chunks,
                         (const double *)vs, j, 0, h2, c2);
      launch_bgs<T>(j + 1, nc, s, n, (const T *)V, (const double *)c2, (const double *)nullptr, 0, w, cst, cln,
                    cpr, pt);
    }
    hipLaunchKernelGGL(k_bdot_final, dim3(np, 1), dim3(kT), 0, s, 1, cfp, (const double *)pt, n_chunks, hn, 1, 1);
    hipLaunchKernelGGL(k_bgivens, dim3(gp), dim3(64), 0, s, np, j, (const double *)h1,
                       (const double *)(two ? h2 : nullptr), (const double *)hn, H.as<double>(), cs.as<double>(),
                       sn.as<double>(), gv.as<double>(), res.as<double>(), (const double *)tol.as<double>(), vs, dn,
                       jd, its, opt.max_it, hs, hist_cap);
  }
  UBT_HIP(hipGetLastError());
}

template <class T> void Fgmres::cycle_end(T *x, hipStream_t s) {
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
// the done flags after the start and after every block, so a cycle stops as
// soon as every problem has converged instead of running all restart steps.
template <class T>
void Fgmres::fgmres(const T *b, T *x, const double *tol_abs, int max_cycles, bool true_stop, hipStream_t s) {
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
  for (int c = 0;; c++) {
    if (use_graph)
      UBT_HIP(hipGraphLaunch(gx[0], cap));
    else
      cycle_start<T>(b, x, cap);
    if (all_done() || c >= max_cycles) break;
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
    if (!true_stop && all_done()) break;
  }
}

void Fgmres::solve(const double *b, double *x, uint64_t state_tag, hipStream_t s, bool guess, const double *rtol) {
  UBT_CHECK(b && x, UBT_ERR_ARG_NULL, "b and x are required");
  UBT_CHECK(b != x, UBT_ERR_ARG_IDN, "the solve needs distinct b and x");
  UBT_CHECK(n_chunks > 0, UBT_ERR_ARG_WRONGSTATE, "no problem takes part in the solve");
  if (state_tag != tag) {
    drop_graphs();
    tag = state_tag;
  }
  // every cycle but a last one takes at least one iteration of a live problem
  const int max_cycles = opt.max_it + 1;
  const int64_t *cst = chunk_start.as<int64_t>();
  const int *cln = chunk_len.as<int>(), *cpr = chunk_prob.as<int>(), *cfp = cfirst.as<int>();
  const unsigned nc = (unsigned)n_chunks;
  double *pt = part.as<double>();
  // the caller's stream, then the capture stream
  UBT_HIP(hipStreamSynchronize(s));
  cycles = outer = 0;
  iterations.assign(np, 0);
  reason.assign(np, 0);
  rel_residual.assign(np, 0.0);
  history.clear();
  UBT_HIP(hipMemsetAsync(done.as<int>() + 2 * np, 0, sizeof(int) * np, cap));  // iteration counters
  // ||b|| per problem
  launch_bdot<double>(1, nc, cap, n, b, b, cst, cln, pt);
  hipLaunchKernelGGL(k_bdot_final, dim3(np, 1), dim3(kT), 0, cap, 1, cfp, (const double *)pt, n_chunks, bnorm.as<double>(), 1, 1);
  std::vector<double> bn(np), rn(np), tl(np), rt(np, opt.rtol);
  if (rtol)
    for (int p = 0; p < np; p++) {
      UBT_CHECK(rtol[p] > 0.0, UBT_ERR_ARG_OUTOFRANGE, "rtol of problem %d must be positive, got %g", p, rtol[p]);
      rt[p] = rtol[p];
    }
  UBT_HIP(hipMemcpyAsync(bn.data(), bnorm.as<double>(), sizeof(double) * np, hipMemcpyDeviceToHost, cap));
  UBT_HIP(hipStreamSynchronize(cap));
  if (!opt.mixed) {
    double *bi = b32.as<double>(), *xi = d32.as<double>();
    vec_axpby<double>(n, 1.0, b, 0.0, bi, cap);
    if (guess)
      vec_axpby<double>(n, 1.0, (const double *)x, 0.0, xi, cap);
    else
      vec_fill<double>(n, 0.0, xi, cap);
    for (int p = 0; p < np; p++) tl[p] = rt[p] * bn[p];
    fgmres<double>(bi, xi, tl.data(), max_cycles, true, cap);
    // the last start's true residual norms
    UBT_HIP(hipMemcpyAsync(rn.data(), nrm.as<double>(), sizeof(double) * np, hipMemcpyDeviceToHost, cap));
    vec_axpby<double>(n, 1.0, (const double *)xi, 0.0, x, cap);
    outer = 1;
  } else {
    double *r = r64.as<double>();
    float *bf = b32.as<float>(), *d = d32.as<float>();
    if (!guess) vec_fill<double>(n, 0.0, x, cap);
    for (outer = 0; outer <= opt.max_outer; outer++) {
      sys->mult(UBT_FP64, x, r, cap);
      vec_axpby<double>(n, 1.0, b, -1.0, r, cap);
      launch_bdot<double>(1, nc, cap, n, (const double *)r, (const double *)r, cst, cln, pt);
      hipLaunchKernelGGL(k_bdot_final, dim3(np, 1), dim3(kT), 0, cap, 1, cfp, (const double *)pt, n_chunks, onrm.as<double>(), 1, 1);
      std::vector<int> it(np);
      UBT_HIP(hipMemcpyAsync(rn.data(), onrm.as<double>(), sizeof(double) * np, hipMemcpyDeviceToHost, cap));
      UBT_HIP(hipMemcpyAsync(it.data(), done.as<int>() + 2 * np, sizeof(int) * np, hipMemcpyDeviceToHost, cap));
      UBT_HIP(hipStreamSynchronize(cap));
      bool all = true;
      int its_max = 0;
      for (int p = 0; p < np; p++) {
        const bool conv = !live[p] || rn[p] <= rt[p] * bn[p];
        all = all && conv;
        tl[p] = conv ? 1e300 : opt.inner_rtol;  // a converged problem sits the inner solve out
        its_max = std::max(its_max, it[p]);
      }
      history.push_back(bn[0] > 0.0 ? rn[0] / bn[0] : 0.0);
      if (all || outer == opt.max_outer || its_max >= opt.max_it) break;
      // correction: A d = r / |r| per problem, FP32 storage, to inner_rtol
      hipLaunchKernelGGL((k_bscale<double, float>), dim3(nc), dim3(kT), 0, cap, (const double *)r,
                         (const double *)onrm.as<double>(), 1, bf, cst, cln, cpr);
      vec_fill<float>(n, 0.f, d, cap);
      fgmres<float>(bf, d, tl.data(), max_cycles, false, cap);
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
    reason[p] = !live[p] ? 0 : (rel_residual[p] <= rt[p] ? 1 : -1);
  }
  if (!opt.mixed && hist_cap > 0 && np >= 1) {
    const int nh = std::min(it[0], hist_cap);
    history.resize(nh);
    if (nh > 0) UBT_HIP(hipMemcpy(history.data(), hist.as<double>(), sizeof(double) * nh, hipMemcpyDeviceToHost));
    for (auto &h : history) h = bn[0] > 0.0 ? h / bn[0] : 0.0;
  }
}

}  // namespace ubt
