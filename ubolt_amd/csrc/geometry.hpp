// The DG geometry of one cell and of one (cell, face) slot, the same code on
// the host and on the device. It follows the reference's construction
// (src/unstructured_dgk.kokkos.cxx:29-218,706-821): exact volume, centroid
// and central second moment of every cell and face by fanning them into
// simplices, an orthonormal modal P1 basis from the Cholesky factor of the
// normalised second moment, and per (cell, face) the matrices
//   F_own(i,j) = (A phi_i(x_f) phi_j(x_f) + b_i^T C_f b_j) / (V A)
// and F_up with the upwind cell's basis for j.
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>

namespace ubt {
namespace geo {

#define UBT_HD __host__ __device__ inline

// Moments of a region about a reference point, accumulated simplex by simplex:
// m0 = int 1, m1 = int y, m2 = int y y^T for y = x - reference
struct Moments {
  double m0 = 0.0, m1[3] = {0, 0, 0}, m2[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
  // The exact moments of a uniform d-simplex with vertices y[0..d]:
  //   int y = |S| mean(y_k),  int y y^T = |S|/((d+1)(d+2)) (sum y_k y_k^T + s s^T), s = sum y_k
  UBT_HD void add(int d, double measure, const double (*y)[3]) {
    double s[3] = {0, 0, 0};
    for (int k = 0; k <= d; k++)
      for (int a = 0; a < 3; a++) s[a] += y[k][a];
    m0 += measure;
    for (int a = 0; a < 3; a++) m1[a] += measure * s[a] / (d + 1);
    const double c = measure / ((d + 1) * (d + 2));
    for (int a = 0; a < 3; a++)
      for (int b = 0; b < 3; b++) {
        double o = s[a] * s[b];
        for (int k = 0; k <= d; k++) o += y[k][a] * y[k][b];
        m2[3 * a + b] += c * o;
      }
  }
  // centroid and the second moment about it
  UBT_HD void central(const double ref[3], double cen[3], double C[9]) const {
    for (int a = 0; a < 3; a++) cen[a] = ref[a] + m1[a] / m0;
    for (int a = 0; a < 3; a++)
      for (int b = 0; b < 3; b++) C[3 * a + b] = m2[3 * a + b] - m1[a] * m1[b] / m0;
  }
};

UBT_HD void sub3(const double *a, const double *b, double *o) {
  for (int d = 0; d < 3; d++) o[d] = a[d] - b[d];
}
UBT_HD void cross3(const double *a, const double *b, double *o) {
  o[0] = a[1] * b[2] - a[2] * b[1];
  o[1] = a[2] * b[0] - a[0] * b[2];
  o[2] = a[0] * b[1] - a[1] * b[0];
}
UBT_HD double norm3(const double *a) { return sqrt(a[0] * a[0] + a[1] * a[1] + a[2] * a[2]); }

// the axis a face normal is closest to
UBT_HD int dominant_axis(const double *n, int dim) {
  int ax = 0;
  for (int d = 1; d < dim; d++)
    if (fabs(n[d]) > fabs(n[ax])) ax = d;
  return ax;
}

// Cell c's volume, centroid and central second moment, fanned off the
// vertex average (inside any star-shaped cell): in 2D a triangle per face,
// in 3D a tetrahedron per face edge through the face's vertex average.
// cv: the cell's vpc vertices; fnv, fv: its nf faces' vertex counts and
// vertices (4 per face, cyclic). False when the volume is not positive.
UBT_HD bool cell_moments(int dim, int nf, int vpc, const int *cv, const int *fnv, const int *fv, const double *X,
                         double *vol, double *cen, double *C) {
  double cref[3] = {0, 0, 0};
  for (int i = 0; i < vpc; i++)
    for (int d = 0; d < 3; d++) cref[d] += X[3 * cv[i] + d] / vpc;
  Moments mc;
  for (int f = 0; f < nf; f++) {
    const int n = fnv[f];
    const int *v = fv + 4 * f;
    if (dim == 2) {
      double y[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
      sub3(X + 3 * v[0], cref, y[1]);
      sub3(X + 3 * v[1], cref, y[2]);
      mc.add(2, 0.5 * fabs(y[1][0] * y[2][1] - y[1][1] * y[2][0]), y);
    } else {
      double fref[3] = {0, 0, 0};
      for (int i = 0; i < n; i++)
        for (int d = 0; d < 3; d++) fref[d] += X[3 * v[i] + d] / n;
      for (int e = 0; e < n; e++) {
        double y[4][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
        sub3(fref, cref, y[1]);
        sub3(X + 3 * v[e], cref, y[2]);
        sub3(X + 3 * v[(e + 1) % n], cref, y[3]);
        double cr[3];
        cross3(y[2], y[3], cr);
        const double det = y[1][0] * cr[0] + y[1][1] * cr[1] + y[1][2] * cr[2];
        mc.add(3, fabs(det) / 6.0, y);
      }
    }
  }
  *vol = mc.m0;
  if (!(mc.m0 > 0.0)) return false;
  mc.central(cref, cen, C);
  return true;
}

// what a slot's geometry can get wrong
enum FaceError { kFaceOk = 0, kFaceArea = 1, kFaceNormal = 2, kFaceOutside = 3 };

// One (cell, face) slot: area, centroid, central second moment, outward nA
// (area-weighted normal), the normal's dominant axis, and the distance from
// the cell centroid cen to the face plane along the normal.
UBT_HD int face_geometry(int dim, int n, const int *v, const double *X, const double *cen, double *area, double *xf,
                         double *Cf, double *nAout, int *axis, double *dist) {
  double fref[3] = {0, 0, 0}, nA[3] = {0, 0, 0};
  for (int i = 0; i < n; i++)
    for (int d = 0; d < 3; d++) fref[d] += X[3 * v[i] + d] / n;
  Moments mf;
  if (dim == 2) {
    double y[2][3], e[3];
    sub3(X + 3 * v[0], fref, y[0]);
    sub3(X + 3 * v[1], fref, y[1]);
    sub3(X + 3 * v[1], X + 3 * v[0], e);
    const double L = norm3(e);
    mf.add(1, L, y);
    nA[0] = e[1];
    nA[1] = -e[0];
  } else {
    for (int e = 0; e < n; e++) {
      double y[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}}, cr[3];
      sub3(X + 3 * v[e], fref, y[1]);
      sub3(X + 3 * v[(e + 1) % n], fref, y[2]);
      cross3(y[1], y[2], cr);
      mf.add(2, 0.5 * norm3(cr), y);
      for (int d = 0; d < 3; d++) nA[d] += 0.5 * cr[d];
    }
  }
  const double A = mf.m0;
  *area = A;
  if (!(A > 0.0)) return kFaceArea;
  mf.central(fref, xf, Cf);
  const double nrm = norm3(nA);
  if (!(nrm > 0.0)) return kFaceNormal;
  double out = 0.0;
  for (int d = 0; d < dim; d++) out += nA[d] * (xf[d] - cen[d]);
  const double sgn = out < 0.0 ? -1.0 : 1.0;
  for (int d = 0; d < 3; d++) nAout[d] = d < dim ? sgn * A * nA[d] / nrm : 0.0;
  *axis = dominant_axis(nAout, dim);
  double dd = 0.0;
  for (int d = 0; d < dim; d++) dd += nAout[d] / A * (xf[d] - cen[d]);
  *dist = dd;
  return dd > 0.0 ? kFaceOk : kFaceOutside;
}

// the distance from the face plane to the neighbour's centroid along the
// outward normal (positive when the neighbour's centroid lies beyond it)
UBT_HD double neighbour_distance(int dim, const double *nA, double A, const double *cen_nb, const double *xf) {
  double dn = 0.0;
  for (int d = 0; d < dim; d++) dn += nA[d] / A * (cen_nb[d] - xf[d]);
  return dn;
}

// The P1 basis gradients beta (rows: grad phi_1..phi_dim) from the central
// second moment C of a cell of volume V: C / V = L L^T, beta = L^{-1}. False
// when the second moment is degenerate.
UBT_HD bool basis_beta(int dim, const double *C, double V, double *beta) {
  double L[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
  for (int k = 0; k < 9; k++) beta[k] = 0.0;
  for (int a = 0; a < dim; a++)
    for (int b = 0; b <= a; b++) {
      double s = C[3 * a + b] / V;
      for (int k = 0; k < b; k++) s -= L[3 * a + k] * L[3 * b + k];
      if (a == b) {
        if (!(s > 0.0)) return false;
        L[3 * a + a] = sqrt(s);
      } else {
        L[3 * a + b] = s / L[3 * b + b];
      }
    }
  for (int col = 0; col < dim; col++)
    for (int a = col; a < dim; a++) {
      double s = (a == col) ? 1.0 : 0.0;
      for (int k = col; k < a; k++) s -= L[3 * a + k] * beta[3 * k + col];
      beta[3 * a + col] = s / L[3 * a + a];
    }
  return true;
}

// basis function i of a cell (beta, centroid cen) at x: 1, then beta_i . (x - cen)
UBT_HD double basis_value(const double *beta, const double *cen, int i, const double *x) {
  if (i == 0) return 1.0;
  double v = 0.0;
  for (int d = 0; d < 3; d++) v += beta[3 * (i - 1) + d] * (x[d] - cen[d]);
  return v;
}
UBT_HD double basis_grad(const double *beta, int i, int d) { return i == 0 ? 0.0 : beta[3 * (i - 1) + d]; }

// One slot's face matrices (nb x nb each, row i, column j) and basis values
// at the face centroid: own with the cell's basis for j, up with the upwind
// cell's (beta_up null: no upwind cell, up left as is)
UBT_HD void face_matrices(int nb, const double *beta, const double *cen, const double *beta_up, const double *cen_up,
                          const double *xf, const double *Cf, double A, double V, double *own, double *up,
                          double *bval) {
  const double scale = 1.0 / (V * A);
  for (int i = 0; i < nb; i++) {
    const double vi = basis_value(beta, cen, i, xf);
    bval[i] = vi;
    for (int j = 0; j < nb; j++) {
      double o = A * vi * basis_value(beta, cen, j, xf);
      for (int a = 0; a < 3; a++)
        for (int b = 0; b < 3; b++) o += basis_grad(beta, i, a) * Cf[3 * a + b] * basis_grad(beta, j, b);
      own[i * nb + j] = o * scale;
      if (!beta_up) continue;
      double cr = A * vi * basis_value(beta_up, cen_up, j, xf);
      for (int a = 0; a < 3; a++)
        for (int b = 0; b < 3; b++) cr += basis_grad(beta, i, a) * Cf[3 * a + b] * basis_grad(beta_up, j, b);
      up[i * nb + j] = cr * scale;
    }
  }
}

#undef UBT_HD

}  // namespace geo
}  // namespace ubt
