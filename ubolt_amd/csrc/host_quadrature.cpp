// Angular quadratures on the host: Gauss-Legendre in 1D, the level-symmetric
// (LQn) sets in 2D (xi > 0 half, weights doubled) and 3D (every octant), with
// the ordinate order and reflection partners of the reference
// (src/sn_quadraturek.kokkos.cxx:13-396).
#include <cmath>
#include <cstdarg>
#include <cstdio>

#include "ubt_internal.hpp"

namespace ubt {

namespace {

struct LqnEntry {
  int sn_order, n_levels, n_points;
  double mu[9];
  int triples[3 * 45];
  double weights[45];
};

// The generated constants (lqn_gen.py), S2 to S18
const LqnEntry kLqn[] = {
#include "lqn_table.inc"
};

const LqnEntry *find_lqn(int order) {
  for (const auto &e : kLqn)
    if (e.sn_order == order) return &e;
  return nullptr;
}

// P_n and P_n' by the three-term recurrence
void legendre(int n, double x, double *p, double *dp) {
  double pk = 1.0, pkm1 = 0.0;
  for (int k = 1; k <= n; k++) {
    double pkm2 = pkm1;
    pkm1 = pk;
    pk = ((2 * k - 1) * x * pkm1 - (k - 1) * pkm2) / k;
  }
  *p = pk;
  *dp = n * (x * pk - pkm1) / (x * x - 1.0);
}

// Fill the reflection partners by exact search: a mirrored ordinate is built
// by negating the same stored cosine, so == is the right comparison, and a set
// that is not reflection-symmetric fails here rather than reflecting wrongly
void build_reflections(Quadrature &q) {
  const int A = q.n_angles;
  for (int d = 0; d < 3; d++) q.reflect[d].assign(A, -1);
  for (int d = 0; d < q.dim; d++) {
    for (int a = 0; a < A; a++) {
      double want[3] = {q.omega[3 * a], q.omega[3 * a + 1], q.omega[3 * a + 2]};
      want[d] = -want[d];
      int found = -1;
      for (int b = 0; b < A && found < 0; b++) {
        bool eq = true;
        for (int e = 0; e < q.dim; e++) eq = eq && q.omega[3 * b + e] == want[e];
        if (eq) found = b;
      }
      UBT_CHECK(found >= 0, UBT_ERR_PLIB, "quadrature set is not symmetric: no ordinate mirrors %d on axis %d",
                a, d);
      UBT_CHECK(q.w[found] == q.w[a], UBT_ERR_PLIB, "quadrature weights are not symmetric under a flip");
      q.reflect[d][a] = found;
    }
  }
}

void check_weights(const Quadrature &q) {
  double s = 0.0;
  for (double v : q.w) s += v;
  UBT_CHECK(std::fabs(s - q.sum_weights) <= 1e-12 * q.sum_weights * q.n_angles, UBT_ERR_PLIB,
            "quadrature weights sum to %.17g, not the angular measure %.17g", s, q.sum_weights);
}

}  // namespace

LqnOctant lqn_octant(int N) {
  const LqnEntry *e = find_lqn(N);
  UBT_CHECK(e, UBT_ERR_SUP, "no level-symmetric set for S%d: even orders 2 to 18 only", N);
  LqnOctant o;
  o.mu.assign(e->mu, e->mu + e->n_levels);
  o.triple.assign(e->triples, e->triples + 3 * e->n_points);
  o.weight.assign(e->weights, e->weights + e->n_points);
  return o;
}

Quadrature make_quadrature(int dim, int N) {
  Quadrature q;
  q.dim = dim;
  q.sn_order = N;
  UBT_CHECK(dim >= 1 && dim <= 3, UBT_ERR_ARG_OUTOFRANGE, "quadrature dimension must be 1, 2 or 3, got %d", dim);
  UBT_CHECK(N > 0 && N % 2 == 0, UBT_ERR_SUP, "SN order must be positive and even, got S%d", N);
  if (dim == 1) {
    // N-point Gauss-Legendre: Newton from the asymptotic seed, ascending, the
    // negative half the exact mirror of the positive one
    const int half = N / 2;
    std::vector<double> mu(half), w(half);
    for (int i = 0; i < half; i++) {
      double x = std::cos(M_PI * (i + 0.75) / (N + 0.5)), p, dp;
      int it = 0;
      for (; it < 100; it++) {
        legendre(N, x, &p, &dp);
        double dx = -p / dp;
        x += dx;
        if (std::fabs(dx) <= 4.0 * 2.220446049250313e-16 * std::fabs(x)) break;
      }
      UBT_CHECK(it < 100, UBT_ERR_PLIB, "Gauss-Legendre Newton did not converge");
      legendre(N, x, &p, &dp);
      mu[half - 1 - i] = x;
      w[half - 1 - i] = 2.0 / ((1.0 - x * x) * dp * dp);
    }
    q.n_angles = N;
    q.omega.assign(3 * N, 0.0);
    q.w.assign(N, 0.0);
    for (int t = 0; t < half; t++) {
      q.omega[3 * (half + t)] = mu[t];
      q.w[half + t] = w[t];
      q.omega[3 * (half - 1 - t)] = -mu[t];
      q.w[half - 1 - t] = w[t];
    }
    q.sum_weights = 2.0;
  } else {
    const LqnEntry *e = find_lqn(N);
    UBT_CHECK(e, UBT_ERR_SUP, "%dD SN quadrature is level-symmetric, even orders S2 to S18; got S%d", dim, N);
    for (int p = 0; p < e->n_points; p++) {
      const double bx = e->mu[e->triples[3 * p] - 1];
      const double by = e->mu[e->triples[3 * p + 1] - 1];
      const double bz = e->mu[e->triples[3 * p + 2] - 1];
      if (dim == 2) {
        // quadrant slowest per base point; the 2 stands in for the xi < 0 half
        const double wp = 2.0 * e->weights[p] * (4.0 * M_PI);
        for (int sy = -1; sy <= 1; sy += 2)
          for (int sx = -1; sx <= 1; sx += 2) {
            q.omega.push_back(sx * bx);
            q.omega.push_back(sy * by);
            q.omega.push_back(0.0);
            q.w.push_back(wp);
          }
      } else {
        const double wp = e->weights[p] * (4.0 * M_PI);
        for (int sz = -1; sz <= 1; sz += 2)
          for (int sy = -1; sy <= 1; sy += 2)
            for (int sx = -1; sx <= 1; sx += 2) {
              q.omega.push_back(sx * bx);
              q.omega.push_back(sy * by);
              q.omega.push_back(sz * bz);
              q.w.push_back(wp);
            }
      }
    }
    q.n_angles = (int)q.w.size();
    q.sum_weights = 4.0 * M_PI;
  }
  check_weights(q);
  build_reflections(q);
  return q;
}

// ---------------------------------------------------------------------------
// errors
// ---------------------------------------------------------------------------
namespace {
thread_local int g_last_code = 0;
thread_local char g_last_msg[1024] = "";
}  // namespace

void set_last_error(int code, const char *msg) {
  g_last_code = code;
  std::snprintf(g_last_msg, sizeof(g_last_msg), "%s", msg ? msg : "");
}

void fail(int code, const char *fmt, ...) {
  char buf[1024];
  va_list ap;
  va_start(ap, fmt);
  std::vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  throw Error(code, buf);
}

int last_error_copy(char *buf, size_t len) {
  if (buf && len > 0) std::snprintf(buf, len, "%s", g_last_msg);
  return g_last_code;
}

bool in_window(const BCFace &f, const double *t, int n_t) {
  if (f.n_window_pairs == 0) return true;
  for (int p = 0; p < n_t; p++)
    if (!(t[p] >= f.window[2 * p] && t[p] <= f.window[2 * p + 1])) return false;
  return true;
}

}  // namespace ubt
