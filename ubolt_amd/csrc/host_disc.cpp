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

#include "geometry.hpp"
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
std::shared_ptr<Disc> dg_begin(const Quadrature &q, const Mesh &m, int order, const BCTable &bcs) {
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
  const int nc = m.n_cells;
  const size_t slots = (size_t)nc * G.nf;
  UBT_CHECK((int64_t)nc * G.nb * q.n_angles < (int64_t)1 << 40, UBT_ERR_ARG_OUTOFRANGE, "phase space too large");
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
  if (order == 0) G.face_basis_value.assign(slots, 1.0);
  return D;
}

void dg_geometry_error(int what, int c, int f) {
  switch (what) {
    case 1: fail(UBT_ERR_ARG_WRONG, "cell %d has non-positive volume", c);
    case 2: fail(UBT_ERR_ARG_WRONG, "face %d of cell %d has non-positive area", f, c);
    case 3: fail(UBT_ERR_ARG_WRONG, "face %d of cell %d has no normal", f, c);
    case 4: fail(UBT_ERR_ARG_WRONG, "the centroid of cell %d is not inside its face %d", c, f);
    case 5: fail(UBT_ERR_ARG_WRONG, "the centroid of cell %d is not inside its face", c);
    case 6: fail(UBT_ERR_ARG_WRONG, "cell %d has a degenerate second moment", c);
    default: fail(UBT_ERR_PLIB, "geometry error %d at face %d of cell %d", what, f, c);
  }
}

void dg_geometry_host(Disc &G, const Mesh &m, std::vector<double> &central, std::vector<double> &face_C) {
  const int dim = m.dim, nf = G.nf, nc = m.n_cells;
  const size_t slots = (size_t)nc * nf;
  central.assign(9 * (size_t)nc, 0.0);
  face_C.assign(9 * slots, 0.0);
  for (int c = 0; c < nc; c++) {
    const size_t k0 = (size_t)c * nf;
    if (!geo::cell_moments(dim, nf, m.verts_per_cell, &m.cell_verts[(size_t)c * m.verts_per_cell],
                           &m.face_nverts[k0], &m.face_verts[4 * k0], m.coords.data(), &G.volume[c],
                           &G.centroid[3 * (size_t)c], &central[9 * (size_t)c]))
      dg_geometry_error(1, c, 0);
  }
  for (int c = 0; c < nc; c++)
    for (int f = 0; f < nf; f++) {
      const size_t k = (size_t)c * nf + f;
      const int e = geo::face_geometry(dim, m.face_nverts[k], &m.face_verts[4 * k], m.coords.data(),
                                       &G.centroid[3 * (size_t)c], &G.face_area[k], &G.face_centroid[3 * k],
                                       &face_C[9 * k], &G.face_nA[3 * k], &G.face_axis[k], &G.face_distance[2 * k]);
      if (e != geo::kFaceOk) dg_geometry_error(1 + e, c, f);
      const int nbr = m.face_neighbour[k];
      if (nbr < 0) continue;
      const double dn = geo::neighbour_distance(dim, &G.face_nA[3 * k], G.face_area[k], &G.centroid[3 * (size_t)nbr],
                                                &G.face_centroid[3 * k]);
      if (!(dn > 0.0)) dg_geometry_error(5, nbr, f);
      G.face_distance[2 * k + 1] = dn;
    }
}

void dg_boundaries(Disc &G, const Mesh &m, const BCTable &bcs, const Quadrature &q) {
  const int dim = m.dim;
  const size_t slots = (size_t)m.n_cells * G.nf;
  for (size_t k = 0; k < slots; k++) {
    const int nbr = m.face_neighbour[k];
    const int c = (int)(k / G.nf), f = (int)(k % G.nf);
    if (nbr >= 0) {
      G.face_nb[k] = nbr;
      continue;
    }
    const BCFace bc = bcs.get(G.face_label[k]);
    if (bc.type == UBT_BC_REFLECT) {
      UBT_CHECK(std::fabs(G.face_nA[3 * k + G.face_axis[k]]) >= (1.0 - 1e-10) * G.face_area[k], UBT_ERR_ARG_WRONG,
                "face %d of cell %d (label %d) is reflective but not axis-aligned", f, c, G.face_label[k]);
      G.face_nb[k] = -2 - G.face_axis[k];
    } else {
      G.face_nb[k] = -1;
      double t[2] = {0, 0};
      const int nt = tangential(&G.face_centroid[3 * k], G.face_axis[k], dim, t);
      if (in_window(bc, t, nt)) G.face_inflow[k] = bc.inflow / q.sum_weights;
    }
  }
  // Every label the BC table names must be on some boundary face
  for (auto &e : bcs.faces) {
    bool seen = false;
    for (size_t k = 0; k < slots && !seen; k++) seen = G.face_nb[k] < 0 && G.face_label[k] == e.first;
    UBT_CHECK(seen, UBT_ERR_ARG_WRONG, "a boundary condition was given for label %d but no boundary face carries it",
              e.first);
  }
}

void dg_basis_host(Disc &G, const std::vector<double> &central, const std::vector<double> &face_C) {
  const int dim = G.dim, nf = G.nf, nc = G.n_cells, nb = G.nb;
  const size_t slots = (size_t)nc * nf;
  G.basis_beta.assign(9 * (size_t)nc, 0.0);
  for (int c = 0; c < nc; c++)
    if (!geo::basis_beta(dim, &central[9 * (size_t)c], G.volume[c], &G.basis_beta[9 * (size_t)c]))
      dg_geometry_error(6, c, 0);
  G.face_own.assign(slots * nb * nb, 0.0);
  G.face_up.assign(slots * nb * nb, 0.0);
  G.face_basis_value.assign(slots * nb, 0.0);
  G.face_nb_grad.assign(slots * nb * 3, 0.0);
  G.basis_grad.assign((size_t)nc * nb * 3, 0.0);
  for (int c = 0; c < nc; c++) {
    const double *beta = &G.basis_beta[9 * (size_t)c], *cen = &G.centroid[3 * (size_t)c];
    for (int i = 0; i < nb; i++)
      for (int d = 0; d < 3; d++) G.basis_grad[((size_t)c * nb + i) * 3 + d] = geo::basis_grad(beta, i, d);
    for (int f = 0; f < nf; f++) {
      const size_t k = (size_t)c * nf + f;
      const int nbr = G.face_nb[k];
      const int up = nbr >= 0 ? nbr : (nbr <= -2 ? c : -1);
      if (nbr >= 0)
        for (int j = 0; j < nb; j++)
          for (int d = 0; d < 3; d++)
            G.face_nb_grad[(k * nb + j) * 3 + d] = geo::basis_grad(&G.basis_beta[9 * (size_t)nbr], j, d);
      geo::face_matrices(nb, beta, cen, up >= 0 ? &G.basis_beta[9 * (size_t)up] : nullptr,
                         up >= 0 ? &G.centroid[3 * (size_t)up] : nullptr, &G.face_centroid[3 * k], &face_C[9 * k],
                         G.face_area[k], G.volume[c], &G.face_own[k * nb * nb], &G.face_up[k * nb * nb],
                         &G.face_basis_value[k * nb]);
    }
  }
}

std::shared_ptr<Disc> make_unstructured(const Quadrature &q, const Mesh &m, int order, const BCTable &bcs) {
  auto D = dg_begin(q, m, order, bcs);
  std::vector<double> central, face_C;
  dg_geometry_host(*D, m, central, face_C);
  dg_boundaries(*D, m, bcs, q);
  if (order == 1) dg_basis_host(*D, central, face_C);
  return D;
}

}  // namespace ubt
