// Discretisations on the host: the structured box (cell-centred upwind FD) and
// the unstructured mesh with its DG0/DG1 geometry. Everything here runs once
// per problem; the operator uploads the resulting tables to HBM.
//
// DG geometry follows the reference's construction
// (src/unstructured_dgk.kokkos.cxx:29-218,706-821): exact volume, centroid and
// central second moment of every cell and face by fanning them into simplices,
// an orthonormal modal P1 basis from the Cholesky factor of the normalised
// second moment, and per (cell, face) the matrices
//   F_own(i,j) = (A phi_i(x_f) phi_j(x_f) + b_i^T C_f b_j) / (V A)
// and F_up with the upwind cell's basis for j.
#include <algorithm>
#include <array>
#include <cmath>

#include "ubt_internal.hpp"

namespace ubt {

namespace {

// Local faces of each supported cell type, vertices in cyclic order
struct CellType {
  int dim, nv, nf;
  int fnv[6];
  int fv[6][4];
};
const CellType kTri = {2, 3, 3, {2, 2, 2}, {{0, 1}, {1, 2}, {2, 0}}};
const CellType kQuad = {2, 4, 4, {2, 2, 2, 2}, {{0, 1}, {1, 2}, {2, 3}, {3, 0}}};
const CellType kTet = {3, 4, 4, {3, 3, 3, 3}, {{0, 1, 2}, {0, 1, 3}, {0, 2, 3}, {1, 2, 3}}};
const CellType kHex = {3, 8, 6, {4, 4, 4, 4, 4, 4},
                       {{0, 1, 2, 3}, {4, 5, 6, 7}, {0, 1, 5, 4}, {1, 2, 6, 5}, {2, 3, 7, 6}, {3, 0, 4, 7}}};

const CellType &cell_type(int dim, int nv) {
  if (dim == 2 && nv == 3) return kTri;
  if (dim == 2 && nv == 4) return kQuad;
  if (dim == 3 && nv == 4) return kTet;
  if (dim == 3 && nv == 8) return kHex;
  fail(UBT_ERR_SUP, "no %dD cell type with %d vertices (triangles, quads, tetrahedra, hexahedra)", dim, nv);
}

using Key = std::array<int, 4>;
Key face_key(const int *v, int n) {
  Key k = {-1, -1, -1, -1};
  for (int i = 0; i < n; i++) k[i] = v[i];
  std::sort(k.begin(), k.begin() + n);
  return k;
}

// Moments of a region about a reference point, accumulated simplex by simplex:
// m0 = int 1, m1 = int y, m2 = int y y^T for y = x - reference
struct Moments {
  double m0 = 0.0, m1[3] = {0, 0, 0}, m2[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
  // The exact moments of a uniform d-simplex with vertices y[0..d]:
  //   int y = |S| mean(y_k),  int y y^T = |S|/((d+1)(d+2)) (sum y_k y_k^T + s s^T), s = sum y_k
  void add(int d, double measure, const double (*y)[3]) {
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
  void central(const double ref[3], double cen[3], double C[9]) const {
    for (int a = 0; a < 3; a++) cen[a] = ref[a] + m1[a] / m0;
    for (int a = 0; a < 3; a++)
      for (int b = 0; b < 3; b++) C[3 * a + b] = m2[3 * a + b] - m1[a] * m1[b] / m0;
  }
};

inline void sub3(const double *a, const double *b, double *o) {
  for (int d = 0; d < 3; d++) o[d] = a[d] - b[d];
}
inline void cross3(const double *a, const double *b, double *o) {
  o[0] = a[1] * b[2] - a[2] * b[1];
  o[1] = a[2] * b[0] - a[0] * b[2];
  o[2] = a[0] * b[1] - a[1] * b[0];
}
inline double norm3(const double *a) { return std::sqrt(a[0] * a[0] + a[1] * a[1] + a[2] * a[2]); }

int dominant_axis(const double *n, int dim) {
  int best = 0;
  for (int d = 1; d < dim; d++)
    if (std::fabs(n[d]) > std::fabs(n[best])) best = d;
  return best;
}

// Tangential coordinates of a point on a face normal to `axis`: the other axes, ascending
int tangential(const double *x, int axis, int dim, double *t) {
  int n = 0;
  for (int d = 0; d < dim; d++)
    if (d != axis) t[n++] = x[d];
  return n;
}

void check_bc_table(const BCTable &bcs, int dim, bool structured) {
  for (auto &e : bcs.faces) {
    if (structured)
      UBT_CHECK(e.first >= 1 && e.first <= 2 * dim, UBT_ERR_ARG_WRONG,
                "a boundary condition was given for face id %d but a %dD box has faces 1 to %d", e.first, dim,
                2 * dim);
    UBT_CHECK(e.second.type == UBT_BC_VACUUM || e.second.type == UBT_BC_REFLECT, UBT_ERR_ARG_WRONG,
              "face %d: unknown boundary type %d", e.first, e.second.type);
    UBT_CHECK(e.second.n_window_pairs == 0 || e.second.n_window_pairs == dim - 1, UBT_ERR_ARG_WRONG,
              "a %dD face takes %d tangential [lo, hi] window pair(s), face %d was given %d", dim, dim - 1,
              e.first, e.second.n_window_pairs);
  }
}

}  // namespace

// ---------------------------------------------------------------------------
// Mesh
// ---------------------------------------------------------------------------
Mesh make_mesh(int dim, int nvtx, const double *coords, int ncell, int vpc, const int *cv) {
  UBT_CHECK(dim == 2 || dim == 3, UBT_ERR_ARG_OUTOFRANGE, "mesh dimension must be 2 or 3, got %d", dim);
  UBT_CHECK(nvtx > 0 && ncell > 0, UBT_ERR_ARG_OUTOFRANGE, "a mesh needs vertices and cells");
  const CellType &ct = cell_type(dim, vpc);
  Mesh m;
  m.dim = dim;
  m.n_vertices = nvtx;
  m.n_cells = ncell;
  m.verts_per_cell = vpc;
  m.faces_per_cell = ct.nf;
  m.coords.assign(3 * (size_t)nvtx, 0.0);
  for (int v = 0; v < nvtx; v++)
    for (int d = 0; d < dim; d++) m.coords[3 * v + d] = coords[dim * v + d];
  m.cell_verts.assign(cv, cv + (size_t)ncell * vpc);
  for (int v : m.cell_verts) UBT_CHECK(v >= 0 && v < nvtx, UBT_ERR_ARG_OUTOFRANGE, "cell vertex %d out of range", v);
  const size_t slots = (size_t)ncell * ct.nf;
  m.face_verts.assign(4 * slots, -1);
  m.face_nverts.assign(slots, 0);
  m.face_neighbour.assign(slots, -1);
  m.face_label.assign(slots, -1);
  std::vector<std::pair<Key, int64_t>> keys;
  keys.reserve(slots);
  for (int c = 0; c < ncell; c++)
    for (int f = 0; f < ct.nf; f++) {
      const int64_t k = (int64_t)c * ct.nf + f;
      int v[4];
      for (int i = 0; i < ct.fnv[f]; i++) v[i] = cv[(size_t)c * vpc + ct.fv[f][i]];
      for (int i = 0; i < ct.fnv[f]; i++) m.face_verts[4 * k + i] = v[i];
      m.face_nverts[k] = ct.fnv[f];
      keys.push_back({face_key(v, ct.fnv[f]), k});
    }
  std::sort(keys.begin(), keys.end());
  for (size_t i = 0; i < keys.size();) {
    size_t j = i + 1;
    while (j < keys.size() && keys[j].first == keys[i].first) j++;
    UBT_CHECK(j - i <= 2, UBT_ERR_ARG_WRONG, "a face is shared by %zu cells", j - i);
    if (j - i == 2) {
      const int64_t a = keys[i].second, b = keys[i + 1].second;
      m.face_neighbour[a] = (int)(b / ct.nf);
      m.face_neighbour[b] = (int)(a / ct.nf);
    }
    i = j;
  }
  return m;
}

void mesh_label_box(Mesh &m) {
  double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300};
  for (int v = 0; v < m.n_vertices; v++)
    for (int d = 0; d < m.dim; d++) {
      lo[d] = std::min(lo[d], m.coords[3 * v + d]);
      hi[d] = std::max(hi[d], m.coords[3 * v + d]);
    }
  // PETSc box "Face Sets": 2D bottom 1, right 2, top 3, left 4; 3D bottom(z-) 1,
  // top 2, front(y-) 3, back 4, right(x+) 5, left 6
  const int id2[2][2] = {{4, 2}, {1, 3}};
  const int id3[3][2] = {{6, 5}, {3, 4}, {1, 2}};
  for (size_t k = 0; k < m.face_neighbour.size(); k++) {
    if (m.face_neighbour[k] >= 0) continue;
    for (int d = 0; d < m.dim; d++) {
      const double tol = 1e-10 * (hi[d] - lo[d]);
      for (int side = 0; side < 2; side++) {
        const double target = side ? hi[d] : lo[d];
        bool on = true;
        for (int i = 0; i < m.face_nverts[k]; i++)
          on = on && std::fabs(m.coords[3 * m.face_verts[4 * k + i] + d] - target) <= tol;
        if (on) m.face_label[k] = (m.dim == 2) ? id2[d][side] : id3[d][side];
      }
    }
  }
}

void mesh_set_boundary_labels(Mesh &m, int n, int vpf, const int *fv, const int *labels) {
  UBT_CHECK(vpf >= 2 && vpf <= 4, UBT_ERR_ARG_OUTOFRANGE, "a face has 2 to 4 vertices, got %d", vpf);
  std::vector<std::pair<Key, int>> want(n);
  for (int i = 0; i < n; i++) want[i] = {face_key(fv + (size_t)i * vpf, vpf), labels[i]};
  std::sort(want.begin(), want.end());
  for (size_t k = 0; k < m.face_neighbour.size(); k++) {
    if (m.face_neighbour[k] >= 0) continue;
    Key key = face_key(&m.face_verts[4 * k], m.face_nverts[k]);
    auto it = std::lower_bound(want.begin(), want.end(), std::make_pair(key, -2147483647));
    if (it != want.end() && it->first == key) m.face_label[k] = it->second;
  }
}

// ---------------------------------------------------------------------------
// Structured box
// ---------------------------------------------------------------------------
std::shared_ptr<Disc> make_structured(const Quadrature &q, int dim, const int n[3], const double len[3],
                                      const BCTable &bcs) {
  UBT_CHECK(dim == 2 || dim == 3, UBT_ERR_SUP, "the structured backend is 2D or 3D here, got %dD", dim);
  UBT_CHECK(q.dim == dim, UBT_ERR_ARG_INCOMP, "a %dD quadrature was handed to a %dD box", q.dim, dim);
  check_bc_table(bcs, dim, true);
  auto D = std::make_shared<Disc>();
  D->kind = DISC_FD;
  D->dim = dim;
  D->order = 0;
  D->nb = 1;
  D->quad = q;
  D->bcs = bcs;
  D->n_angles = q.n_angles;
  int64_t cells = 1;
  for (int d = 0; d < dim; d++) {
    UBT_CHECK(n[d] > 0 && len[d] > 0, UBT_ERR_ARG_OUTOFRANGE, "axis %d needs a positive cell count and length", d);
    D->n[d] = n[d];
    D->len[d] = len[d];
    D->h[d] = len[d] / n[d];
    cells *= n[d];
  }
  UBT_CHECK(cells * q.n_angles < (int64_t)1 << 40, UBT_ERR_ARG_OUTOFRANGE, "phase space too large");
  D->n_cells = (int)cells;
  const int id2[2][2] = {{4, 2}, {1, 3}};
  const int id3[3][2] = {{6, 5}, {3, 4}, {1, 2}};
  for (int d = 0; d < dim; d++)
    for (int s = 0; s < 2; s++) {
      const int id = dim == 2 ? id2[d][s] : id3[d][s];
      D->face_bc[d][s] = bcs.get(id);
      D->face_type[d][s] = D->face_bc[d][s].type;
      if (D->face_type[d][s] == UBT_BC_REFLECT)
        UBT_CHECK(n[d] > 1, UBT_ERR_SUP, "a reflective face on a single-cell-wide axis is not supported");
    }
  return D;
}

// ---------------------------------------------------------------------------
// Unstructured DG0 / DG1
// ---------------------------------------------------------------------------
std::shared_ptr<Disc> make_unstructured(const Quadrature &q, const Mesh &m, int order, const BCTable &bcs) {
  UBT_CHECK(order == 0 || order == 1, UBT_ERR_ARG_OUTOFRANGE, "the unstructured backend is DG0 or DG1, got order %d",
            order);
  UBT_CHECK(q.dim == m.dim, UBT_ERR_ARG_INCOMP, "a %dD quadrature was handed to a %dD mesh", q.dim, m.dim);
  check_bc_table(bcs, m.dim, false);
  auto D = std::make_shared<Disc>();
  Disc &G = *D;
  G.kind = DISC_DG;
  G.dim = m.dim;
  G.order = order;
  G.nb = order == 0 ? 1 : m.dim + 1;
  G.quad = q;
  G.bcs = bcs;
  G.n_angles = q.n_angles;
  G.n_cells = m.n_cells;
  G.nf = m.faces_per_cell;
  const int dim = m.dim, nf = G.nf, nc = m.n_cells, nb = G.nb;
  const size_t slots = (size_t)nc * nf;
  UBT_CHECK((int64_t)nc * nb * q.n_angles < (int64_t)1 << 40, UBT_ERR_ARG_OUTOFRANGE, "phase space too large");

  G.volume.assign(nc, 0.0);
  G.centroid.assign(3 * (size_t)nc, 0.0);
  G.face_nA.assign(3 * slots, 0.0);
  G.face_centroid.assign(3 * slots, 0.0);
  G.face_area.assign(slots, 0.0);
  G.face_distance.assign(2 * slots, 0.0);
  G.face_nb.assign(slots, -1);
  G.face_label = m.face_label;
  G.face_axis.assign(slots, 0);
  G.face_inflow.assign(slots, 0.0);
  std::vector<double> central(9 * (size_t)nc), face_C(9 * slots);

  auto X = [&](int v) { return &m.coords[3 * (size_t)v]; };

  // Per cell: fan moments off the vertex average (inside any star-shaped cell)
  for (int c = 0; c < nc; c++) {
    const int *cv = &m.cell_verts[(size_t)c * m.verts_per_cell];
    double cref[3] = {0, 0, 0};
    for (int i = 0; i < m.verts_per_cell; i++)
      for (int d = 0; d < 3; d++) cref[d] += X(cv[i])[d] / m.verts_per_cell;
    Moments mc;
    for (int f = 0; f < nf; f++) {
      const size_t k = (size_t)c * nf + f;
      const int fnv = m.face_nverts[k];
      const int *fv = &m.face_verts[4 * k];
      if (dim == 2) {
        double y[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
        sub3(X(fv[0]), cref, y[1]);
        sub3(X(fv[1]), cref, y[2]);
        mc.add(2, 0.5 * std::fabs(y[1][0] * y[2][1] - y[1][1] * y[2][0]), y);
      } else {
        double fref[3] = {0, 0, 0};
        for (int i = 0; i < fnv; i++)
          for (int d = 0; d < 3; d++) fref[d] += X(fv[i])[d] / fnv;
        for (int e = 0; e < fnv; e++) {
          double y[4][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
          sub3(fref, cref, y[1]);
          sub3(X(fv[e]), cref, y[2]);
          sub3(X(fv[(e + 1) % fnv]), cref, y[3]);
          double cr[3];
          cross3(y[2], y[3], cr);
          const double det = y[1][0] * cr[0] + y[1][1] * cr[1] + y[1][2] * cr[2];
          mc.add(3, std::fabs(det) / 6.0, y);
        }
      }
    }
    UBT_CHECK(mc.m0 > 0.0, UBT_ERR_ARG_WRONG, "cell %d has non-positive volume %g", c, mc.m0);
    G.volume[c] = mc.m0;
    mc.central(cref, &G.centroid[3 * (size_t)c], &central[9 * (size_t)c]);
  }

  // Per (cell, face): area, centroid, central second moment, outward nA
  for (int c = 0; c < nc; c++) {
    const double *cen = &G.centroid[3 * (size_t)c];
    for (int f = 0; f < nf; f++) {
      const size_t k = (size_t)c * nf + f;
      const int fnv = m.face_nverts[k];
      const int *fv = &m.face_verts[4 * k];
      double fref[3] = {0, 0, 0}, nA[3] = {0, 0, 0};
      for (int i = 0; i < fnv; i++)
        for (int d = 0; d < 3; d++) fref[d] += X(fv[i])[d] / fnv;
      Moments mf;
      if (dim == 2) {
        double y[2][3], e[3];
        sub3(X(fv[0]), fref, y[0]);
        sub3(X(fv[1]), fref, y[1]);
        sub3(X(fv[1]), X(fv[0]), e);
        const double L = norm3(e);
        mf.add(1, L, y);
        nA[0] = e[1];
        nA[1] = -e[0];
      } else {
        for (int e = 0; e < fnv; e++) {
          double y[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}}, cr[3];
          sub3(X(fv[e]), fref, y[1]);
          sub3(X(fv[(e + 1) % fnv]), fref, y[2]);
          cross3(y[1], y[2], cr);
          mf.add(2, 0.5 * norm3(cr), y);
          for (int d = 0; d < 3; d++) nA[d] += 0.5 * cr[d];
        }
      }
      const double area = mf.m0;
      UBT_CHECK(area > 0.0, UBT_ERR_ARG_WRONG, "face %d of cell %d has non-positive area", f, c);
      double *xf = &G.face_centroid[3 * k];
      mf.central(fref, xf, &face_C[9 * k]);
      const double nrm = norm3(nA);
      UBT_CHECK(nrm > 0.0, UBT_ERR_ARG_WRONG, "face %d of cell %d has no normal", f, c);
      double out = 0.0;
      for (int d = 0; d < dim; d++) out += nA[d] * (xf[d] - cen[d]);
      const double sgn = out < 0.0 ? -1.0 : 1.0;
      for (int d = 0; d < dim; d++) G.face_nA[3 * k + d] = sgn * area * nA[d] / nrm;
      G.face_area[k] = area;
      G.face_axis[k] = dominant_axis(&G.face_nA[3 * k], dim);
      double dist = 0.0;
      for (int d = 0; d < dim; d++) dist += G.face_nA[3 * k + d] / area * (xf[d] - cen[d]);
      UBT_CHECK(dist > 0.0, UBT_ERR_ARG_WRONG, "the centroid of cell %d is not inside its face %d", c, f);
      G.face_distance[2 * k] = dist;
      const int nbr = m.face_neighbour[k];
      if (nbr >= 0) {
        G.face_nb[k] = nbr;
        double dn = 0.0;
        for (int d = 0; d < dim; d++) dn += G.face_nA[3 * k + d] / area * (G.centroid[3 * (size_t)nbr + d] - xf[d]);
        UBT_CHECK(dn > 0.0, UBT_ERR_ARG_WRONG, "the centroid of cell %d is not inside its face", nbr);
        G.face_distance[2 * k + 1] = dn;
      } else {
        const BCFace bc = bcs.get(G.face_label[k]);
        if (bc.type == UBT_BC_REFLECT) {
          UBT_CHECK(std::fabs(G.face_nA[3 * k + G.face_axis[k]]) >= (1.0 - 1e-10) * area, UBT_ERR_ARG_WRONG,
                    "face %d of cell %d (label %d) is reflective but not axis-aligned", f, c, G.face_label[k]);
          G.face_nb[k] = -2 - G.face_axis[k];
        } else {
          G.face_nb[k] = -1;
          double t[2] = {0, 0};
          const int nt = tangential(xf, G.face_axis[k], dim, t);
          if (in_window(bc, t, nt)) G.face_inflow[k] = bc.inflow / q.sum_weights;
        }
      }
    }
  }
  // Every label the BC table names must be on some boundary face
  for (auto &e : bcs.faces) {
    bool seen = false;
    for (size_t k = 0; k < slots && !seen; k++) seen = G.face_nb[k] < 0 && G.face_label[k] == e.first;
    UBT_CHECK(seen, UBT_ERR_ARG_WRONG, "a boundary condition was given for label %d but no boundary face carries it",
              e.first);
  }

  if (order == 1) {
    // beta rows: gradients of phi_1..phi_dim, from M = central / V = L L^T, beta = L^{-1}
    G.basis_beta.assign(9 * (size_t)nc, 0.0);
    for (int c = 0; c < nc; c++) {
      double L[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
      const double *C = &central[9 * (size_t)c];
      for (int a = 0; a < dim; a++)
        for (int b = 0; b <= a; b++) {
          double s = C[3 * a + b] / G.volume[c];
          for (int k = 0; k < b; k++) s -= L[3 * a + k] * L[3 * b + k];
          if (a == b) {
            UBT_CHECK(s > 0.0, UBT_ERR_ARG_WRONG, "cell %d has a degenerate second moment", c);
            L[3 * a + a] = std::sqrt(s);
          } else {
            L[3 * a + b] = s / L[3 * b + b];
          }
        }
      double *beta = &G.basis_beta[9 * (size_t)c];
      for (int col = 0; col < dim; col++)
        for (int a = col; a < dim; a++) {
          double s = (a == col) ? 1.0 : 0.0;
          for (int k = col; k < a; k++) s -= L[3 * a + k] * beta[3 * k + col];
          beta[3 * a + col] = s / L[3 * a + a];
        }
    }
    auto value = [&](int cc, int i, const double *x) {
      if (i == 0) return 1.0;
      double v = 0.0;
      for (int d = 0; d < 3; d++) v += G.basis_beta[9 * (size_t)cc + 3 * (i - 1) + d] * (x[d] - G.centroid[3 * (size_t)cc + d]);
      return v;
    };
    auto grad = [&](int cc, int i, int d) { return i == 0 ? 0.0 : G.basis_beta[9 * (size_t)cc + 3 * (i - 1) + d]; };
    G.face_own.assign(slots * nb * nb, 0.0);
    G.face_up.assign(slots * nb * nb, 0.0);
    G.face_basis_value.assign(slots * nb, 0.0);
    G.face_nb_grad.assign(slots * nb * 3, 0.0);
    G.basis_grad.assign((size_t)nc * nb * 3, 0.0);
    for (int c = 0; c < nc; c++) {
      for (int i = 0; i < nb; i++)
        for (int d = 0; d < 3; d++) G.basis_grad[((size_t)c * nb + i) * 3 + d] = grad(c, i, d);
      for (int f = 0; f < nf; f++) {
        const size_t k = (size_t)c * nf + f;
        const double *xf = &G.face_centroid[3 * k], *Cf = &face_C[9 * k];
        const double A = G.face_area[k];
        const int nbr = G.face_nb[k];
        const int up = nbr >= 0 ? nbr : (nbr <= -2 ? c : -1);
        if (nbr >= 0)
          for (int j = 0; j < nb; j++)
            for (int d = 0; d < 3; d++) G.face_nb_grad[(k * nb + j) * 3 + d] = grad(nbr, j, d);
        const double scale = 1.0 / (G.volume[c] * A);
        for (int i = 0; i < nb; i++) {
          const double vi = value(c, i, xf);
          G.face_basis_value[k * nb + i] = vi;
          for (int j = 0; j < nb; j++) {
            double own = A * vi * value(c, j, xf);
            for (int a = 0; a < 3; a++)
              for (int b = 0; b < 3; b++) own += grad(c, i, a) * Cf[3 * a + b] * grad(c, j, b);
            G.face_own[(k * nb + i) * nb + j] = own * scale;
            if (up < 0) continue;
            double cr = A * vi * value(up, j, xf);
            for (int a = 0; a < 3; a++)
              for (int b = 0; b < 3; b++) cr += grad(c, i, a) * Cf[3 * a + b] * grad(up, j, b);
            G.face_up[(k * nb + i) * nb + j] = cr * scale;
          }
        }
      }
    }
  } else {
    G.face_basis_value.assign(slots, 1.0);
  }
  return D;
}

}  // namespace ubt
