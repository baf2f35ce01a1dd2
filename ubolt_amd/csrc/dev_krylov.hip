// Krylov solver on gfx950: right-preconditioned flexible GMRES(m) with
// classical Gram-Schmidt applied twice (CGS2). The Gram-Schmidt coefficients,
// the norms and the basis scaling stay on the device; one copy of the new
// Hessenberg column per iteration feeds the Givens rotations and the
// convergence test on the host. Dot products accumulate in FP64 whatever the
// vector storage, in a fixed order (deterministic).
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

namespace {

constexpr int kMaxK = 32;  // restart + 1 at most
constexpr int kThreads = 256;
constexpr int kMaxBlocks = 1024;

// y = a x + b y; with b = 0, y is not read (it may hold anything, NaN included)
template <class T> __global__ void k_axpby(int64_t n, double a, const T *__restrict__ x, double b, T *__restrict__ y) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    y[i] = (T)(b == 0.0 ? a * (double)x[i] : a * (double)x[i] + b * (double)y[i]);
}

template <class Ti, class To> __global__ void k_convert(int64_t n, double a, const Ti *__restrict__ x, To *__restrict__ y) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    y[i] = (To)(a * (double)x[i]);
}

template <class Ti> __global__ void k_add_scaled(int64_t n, double a, const Ti *__restrict__ d, double *__restrict__ x) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    x[i] += a * (double)d[i];
}

// partial[b * k + j] = sum over this block's rows of V[j][i] w[i]
template <class T>
__global__ void __launch_bounds__(kThreads) k_mdot_partial(int64_t n, int k, const T *__restrict__ V,
                                                           const T *__restrict__ w, double *partial) {
  double acc[kMaxK];
#pragma unroll
  for (int j = 0; j < kMaxK; j++) acc[j] = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    const double wi = (double)w[i];
#pragma unroll
    for (int j = 0; j < kMaxK; j++)
      if (j < k) acc[j] += (double)V[(int64_t)j * n + i] * wi;
  }
  __shared__ double sh[kMaxK][kThreads / 64];
  const int wv = threadIdx.x >> 6, l = threadIdx.x & 63;
#pragma unroll
  for (int j = 0; j < kMaxK; j++) {
    if (j >= k) break;
    double v = acc[j];
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    if (l == 0) sh[j][wv] = v;
  }
  __syncthreads();
  if ((int)threadIdx.x < k) {
    double t = 0.0;
    for (int q = 0; q < (int)(blockDim.x >> 6); q++) t += sh[threadIdx.x][q];
    partial[(int64_t)blockIdx.x * k + threadIdx.x] = t;
  }
}

// out[j] = (accumulate ? out[j] : 0) + sum_b partial[b * k + j]; sqrt_last:
// out[0] = sqrt(sum) (a norm)
// one wavefront per column j (grid = k): lane-strided sums, then a butterfly
__global__ void k_mdot_final(int nblk, int k, const double *partial, double *out, int accumulate, int take_sqrt) {
  const int j = blockIdx.x;
  if (j >= k) return;
  double t = 0.0;
  for (int b = threadIdx.x; b < nblk; b += 64) t += partial[(int64_t)b * k + j];
  for (int o = 32; o > 0; o >>= 1) t += __shfl_xor(t, o, 64);
  if (threadIdx.x != 0) return;
  if (take_sqrt) t = sqrt(t);
  out[j] = accumulate ? out[j] + t : t;
}

// w += sign * sum_j h[j] V[j]
template <class T>
__global__ void k_maxpy(int64_t n, int k, const T *__restrict__ V, const double *__restrict__ h, double sign,
                        T *__restrict__ w) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    double v = (double)w[i];
    for (int j = 0; j < k; j++) v += sign * h[j] * (double)V[(int64_t)j * n + i];
    w[i] = (T)v;
  }
}

// v = w / *d (0 on a zero divisor: a happy breakdown)
template <class T>
__global__ void k_scale_dev(int64_t n, const T *__restrict__ w, const double *__restrict__ d, T *__restrict__ v) {
  const double s = *d != 0.0 ? 1.0 / *d : 0.0;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    v[i] = (T)(s * (double)w[i]);
}

unsigned grid_for(int64_t n) {
  const int64_t g = (n + kThreads - 1) / kThreads;
  return (unsigned)std::max<int64_t>(1, std::min<int64_t>(g, 8 * kMaxBlocks));
}
unsigned grid_dot(int64_t n) {
  const int64_t g = (n + kThreads - 1) / kThreads;
  return (unsigned)std::max<int64_t>(1, std::min<int64_t>(g, kMaxBlocks));
}

template <class T> int prec_of() { return sizeof(T) == 8 ? UBT_FP64 : UBT_FP32; }

}  // namespace

// out[0..k) = V[0..k)^T w on the device
template <class T> static void mdot(const T *V, int k, const T *w, int64_t n, double *out, bool acc, bool sq,
                                    double *partial, hipStream_t s) {
  const unsigned g = grid_dot(n);
  hipLaunchKernelGGL(k_mdot_partial<T>, dim3(g), dim3(kThreads), 0, s, n, k, V, w, partial);
  hipLaunchKernelGGL(k_mdot_final, dim3(k), dim3(64), 0, s, (int)g, k, partial, out, acc ? 1 : 0, sq ? 1 : 0);
}

// pin: pinned host memory for the result (an asynchronous copy; pageable
// memory would be staged)
template <class T>
static double host_norm(const T *w, int64_t n, double *dev_out, double *partial, hipStream_t s, double *pin) {
  mdot<T>(w, 1, w, n, dev_out, false, true, partial, s);
  UBT_HIP(hipMemcpyAsync(pin, dev_out, sizeof(double), hipMemcpyDeviceToHost, s));
  UBT_HIP(hipStreamSynchronize(s));
  return *pin;
}

void Solver::create(Operator *o, DSA *d, const SolverOptions &opts) {
  op = o;
  dsa = d;
  opt = opts;
  UBT_CHECK(opt.restart >= 1 && opt.restart + 1 <= kMaxK, UBT_ERR_ARG_OUTOFRANGE, "restart must be 1 to %d, got %d",
            kMaxK - 1, opt.restart);
  UBT_CHECK(opt.rtol > 0.0 && opt.max_it >= 1, UBT_ERR_ARG_OUTOFRANGE, "rtol must be positive and max_it >= 1");
  n = o->disc->n_rows();
  if (opt.pc_sweep) sweep.create(o);
  blocks.op = o;
  const int m = opt.restart;
  if (!opt.mixed) {
    V64.alloc(sizeof(double) * n * (m + 1));
    Z64.alloc(sizeof(double) * n * m);
    w64.alloc(sizeof(double) * n);
    t64.alloc(sizeof(double) * n);
    u64.alloc(sizeof(double) * n);
  } else {
    V32.alloc(sizeof(float) * n * (m + 1));
    Z32.alloc(sizeof(float) * n * m);
    w32.alloc(sizeof(float) * n);
    t32.alloc(sizeof(float) * n);
    u32.alloc(sizeof(float) * n);
    d32.alloc(sizeof(float) * n);
    b32.alloc(sizeof(float) * n);
    r64.alloc(sizeof(double) * n);
  }
  hcol.alloc(sizeof(double) * 2 * kMaxK);
  partial.alloc(sizeof(double) * kMaxBlocks * kMaxK);
  if (!h_pin) UBT_HIP(hipHostMalloc((void **)&h_pin, sizeof(double) * (2 * kMaxK + 1), hipHostMallocDefault));
  if (!h_ev) UBT_HIP(hipEventCreateWithFlags(&h_ev, hipEventDisableTiming));
}

Solver::~Solver() {
  if (h_pin) (void)hipHostFree(h_pin);
  if (h_ev) (void)hipEventDestroy(h_ev);
}

void Solver::refresh(hipStream_t s) {
  if (opt.pc_blocks) blocks.setup(s);
}

template <class T> void Solver::apply_op(const T *x, T *y, bool scatter, hipStream_t s) {
  if (transpose)
    op->mult_transpose(prec_of<T>(), x, y, scatter, s);
  else
    op->mult(prec_of<T>(), x, y, scatter, s);
}

template <class T> void Solver::apply_pc(const T *r, T *y, hipStream_t s) {
  T *t = sizeof(T) == 8 ? (T *)t64.ptr : (T *)t32.ptr;
  T *u = sizeof(T) == 8 ? (T *)u64.ptr : (T *)u32.ptr;
  const bool with_dsa = dsa && opt.pc_dsa;
  const bool amat = with_dsa && opt.use_amat;
  const int prec = prec_of<T>();
  bool first = true;
  auto stage = [&](auto &&fn) {
    if (first) {
      fn(r, y);
      first = false;
      return;
    }
    apply_op<T>(y, t, amat, s);
    hipLaunchKernelGGL(k_axpby<T>, dim3(grid_for(n)), dim3(kThreads), 0, s, n, 1.0, r, -1.0, t);
    fn((const T *)t, u);
    hipLaunchKernelGGL(k_axpby<T>, dim3(grid_for(n)), dim3(kThreads), 0, s, n, 1.0, (const T *)u, 1.0, y);
  };
  // where the sweep is L's exact inverse S and the stages correct with L, the
  // first two stages are the sweep alone: y + S (r - L y) = S r for any y
  const bool fold = opt.pc_sweep && sweep.exact && !amat;
  if (opt.pc_blocks && !fold) stage([&](const T *a, T *b) { blocks.apply(prec, a, b, s, transpose); });
  if (opt.pc_sweep) stage([&](const T *a, T *b) { sweep.apply(prec, a, b, s, transpose); });
  if (with_dsa) stage([&](const T *a, T *b) { dsa->apply(prec, a, b, s, transpose); });
  if (first) hipLaunchKernelGGL((k_convert<T, T>), dim3(grid_for(n)), dim3(kThreads), 0, s, n, 1.0, r, y);
  UBT_HIP(hipGetLastError());
}

template <class T> double Solver::fgmres(const T *b, T *x, double tol_abs, int max_it, int *its, hipStream_t s) {
  const int m = opt.restart;
  T *V = sizeof(T) == 8 ? (T *)V64.ptr : (T *)V32.ptr;
  T *Z = sizeof(T) == 8 ? (T *)Z64.ptr : (T *)Z32.ptr;
  T *w = sizeof(T) == 8 ? (T *)w64.ptr : (T *)w32.ptr;
  double *hc = hcol.as<double>(), *part = partial.as<double>();
  const unsigned gv = grid_for(n);
  std::vector<double> H((size_t)(m + 1) * m), cs(m), sn(m), g(m + 1), y(m);
  double resid = 0.0;
  while (true) {
    apply_op<T>(x, w, true, s);
    hipLaunchKernelGGL(k_axpby<T>, dim3(gv), dim3(kThreads), 0, s, n, 1.0, b, -1.0, w);
    const double beta = host_norm<T>(w, n, hc, part, s, h_pin + 2 * kMaxK);
    resid = beta;
    if (!(beta > tol_abs) || *its >= max_it) return beta;
    hipLaunchKernelGGL((k_convert<T, T>), dim3(gv), dim3(kThreads), 0, s, n, 1.0 / beta, (const T *)w, V);
    std::fill(g.begin(), g.end(), 0.0);
    g[0] = beta;
    int j = 0;
    bool breakdown = false;
    for (; j < m && *its < max_it;) {
      T *vj = V + (int64_t)j * n, *zj = Z + (int64_t)j * n, *vn = V + (int64_t)(j + 1) * n;
      apply_pc<T>(vj, zj, s);
      apply_op<T>(zj, w, true, s);
      // classical Gram-Schmidt: h = V^T w, w -= V h, once or twice (the
      // second pass's coefficients added to the first's)
      mdot<T>(V, j + 1, w, n, hc, false, false, part, s);
      hipLaunchKernelGGL(k_maxpy<T>, dim3(gv), dim3(kThreads), 0, s, n, j + 1, (const T *)V, (const double *)hc, -1.0, w);
      std::vector<double> h2(j + 1, 0.0);
      if (opt.gs_passes > 1) {
        mdot<T>(V, j + 1, w, n, hc + kMaxK, false, false, part, s);
        hipLaunchKernelGGL(k_maxpy<T>, dim3(gv), dim3(kThreads), 0, s, n, j + 1, (const T *)V,
                           (const double *)(hc + kMaxK), -1.0, w);
      }
      mdot<T>(w, 1, w, n, hc + j + 1, false, true, part, s);
      hipLaunchKernelGGL(k_scale_dev<T>, dim3(gv), dim3(kThreads), 0, s, n, (const T *)w, (const double *)(hc + j + 1), vn);
      UBT_HIP(hipMemcpyAsync(h_pin, hc, sizeof(double) * (j + 2), hipMemcpyDeviceToHost, s));
      if (opt.gs_passes > 1)
        UBT_HIP(hipMemcpyAsync(h_pin + kMaxK, hc + kMaxK, sizeof(double) * (j + 1), hipMemcpyDeviceToHost, s));
      UBT_HIP(hipEventRecord(h_ev, s));
      UBT_HIP(hipEventSynchronize(h_ev));
      if (opt.gs_passes > 1)
        for (int i = 0; i <= j; i++) h2[i] = h_pin[kMaxK + i];
      for (int i = 0; i <= j; i++) H[(size_t)i * m + j] = h_pin[i] + h2[i];
      H[(size_t)(j + 1) * m + j] = h_pin[j + 1];
      for (int i = 0; i < j; i++) {
        const double a = H[(size_t)i * m + j], bb = H[(size_t)(i + 1) * m + j];
        H[(size_t)i * m + j] = cs[i] * a + sn[i] * bb;
        H[(size_t)(i + 1) * m + j] = -sn[i] * a + cs[i] * bb;
      }
      const double a = H[(size_t)j * m + j], bb = H[(size_t)(j + 1) * m + j];
      const double rr = std::hypot(a, bb);
      cs[j] = rr != 0.0 ? a / rr : 1.0;
      sn[j] = rr != 0.0 ? bb / rr : 0.0;
      H[(size_t)j * m + j] = rr;
      H[(size_t)(j + 1) * m + j] = 0.0;
      g[j + 1] = -sn[j] * g[j];
      g[j] = cs[j] * g[j];
      resid = std::fabs(g[j + 1]);
      (*its)++;
      j++;
      history.push_back(resid);
      if (h_pin[j] == 0.0) breakdown = true;
      if (!(resid > tol_abs) || breakdown) break;
    }
    // x += Z y, H y = g on the leading j x j triangle
    for (int i = j - 1; i >= 0; i--) {
      double v = g[i];
      for (int k = i + 1; k < j; k++) v -= H[(size_t)i * m + k] * y[k];
      y[i] = H[(size_t)i * m + i] != 0.0 ? v / H[(size_t)i * m + i] : 0.0;
    }
    std::copy(y.begin(), y.begin() + j, h_pin);  // pinned: an asynchronous copy
    UBT_HIP(hipMemcpyAsync(hc, h_pin, sizeof(double) * j, hipMemcpyHostToDevice, s));
    hipLaunchKernelGGL(k_maxpy<T>, dim3(gv), dim3(kThreads), 0, s, n, j, (const T *)Z, (const double *)hc, 1.0, x);
    UBT_HIP(hipGetLastError());
    UBT_HIP(hipStreamSynchronize(s));  // before h_pin is reused
    if (*its >= max_it && resid > tol_abs) {
      // one last true residual
      apply_op<T>(x, w, true, s);
      hipLaunchKernelGGL(k_axpby<T>, dim3(gv), dim3(kThreads), 0, s, n, 1.0, b, -1.0, w);
      return host_norm<T>(w, n, hc, part, s, h_pin + 2 * kMaxK);
    }
  }
}

void Solver::solve(const double *b, double *x, hipStream_t s, bool adj) {
  UBT_CHECK(b && x, UBT_ERR_ARG_NULL, "b and x are required");
  UBT_CHECK(b != x, UBT_ERR_ARG_IDN, "the solve needs distinct b and x");
  UBT_CHECK(op->xs_set, UBT_ERR_ARG_WRONGSTATE, "set the cross sections first");
  if (opt.pc_blocks) UBT_CHECK(blocks.group_ready == op->group, UBT_ERR_ARG_WRONGSTATE,
                               "refresh the solver after switching the operator to group %d", op->group);
  if (dsa && opt.pc_dsa) UBT_CHECK(dsa->group == op->group, UBT_ERR_ARG_WRONGSTATE,
                                   "set the DSA group to the operator's group %d", op->group);
  history.clear();
  iterations = outer = 0;
  reason = 0;
  transpose = adj;
  const unsigned gv = grid_for(n);
  double *part = partial.as<double>(), *hc = hcol.as<double>();
  hipLaunchKernelGGL((k_convert<double, double>), dim3(gv), dim3(kThreads), 0, s, n, 0.0, b, x);  // x = 0
  const double bnorm = host_norm<double>(b, n, hc, part, s, h_pin + 2 * kMaxK);
  if (bnorm == 0.0) {
    rel_residual = 0.0;
    reason = 1;
    return;
  }
  const double tol = opt.rtol * bnorm;
  if (!opt.mixed) {
    int its = 0;
    const double r = fgmres<double>(b, x, tol, opt.max_it, &its, s);
    iterations = its;
    outer = 1;
    rel_residual = r / bnorm;
    reason = r <= tol ? 1 : -1;
    for (auto &h : history) h /= bnorm;
    return;
  }
  double *r = r64.as<double>();
  float *d = d32.as<float>(), *bf = b32.as<float>();
  std::vector<double> hist;
  for (outer = 0; outer <= opt.max_outer; outer++) {
    apply_op<double>(x, r, true, s);
    hipLaunchKernelGGL(k_axpby<double>, dim3(gv), dim3(kThreads), 0, s, n, 1.0, b, -1.0, r);
    const double rn = host_norm<double>(r, n, hc, part, s, h_pin + 2 * kMaxK);
    rel_residual = rn / bnorm;
    hist.push_back(rel_residual);
    if (rn <= tol) {
      reason = 1;
      break;
    }
    if (outer == opt.max_outer || iterations >= opt.max_it) {
      reason = -1;
      break;
    }
    // correction: A d = r / |r| in FP32 storage, to inner_rtol
    hipLaunchKernelGGL((k_convert<double, float>), dim3(gv), dim3(kThreads), 0, s, n, 1.0 / rn, (const double *)r, bf);
    hipLaunchKernelGGL((k_convert<float, float>), dim3(gv), dim3(kThreads), 0, s, n, 0.0, (const float *)bf, d);
    history.clear();
    int its = iterations;
    fgmres<float>(bf, d, opt.inner_rtol, opt.max_it, &its, s);
    iterations = its;
    hipLaunchKernelGGL(k_add_scaled<float>, dim3(gv), dim3(kThreads), 0, s, n, rn, (const float *)d, x);
    UBT_HIP(hipGetLastError());
  }
  history = hist;
}

template void Solver::apply_pc<double>(const double *, double *, hipStream_t);
template void Solver::apply_pc<float>(const float *, float *, hipStream_t);

}  // namespace ubt
