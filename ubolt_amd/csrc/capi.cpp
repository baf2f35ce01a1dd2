// The extern "C" boundary: every entry point catches what the C++ below throws
// and turns it into a ubt_status plus a message for ubt_last_error().
#include <cstring>

#include "batch.hpp"
#include "operator.hpp"
#include "solver.hpp"
#include "ubt_internal.hpp"

namespace ubt {
int last_error_copy(char *buf, size_t len);
}

struct ubt_quad_s {
  ubt::Quadrature q;
};
struct ubt_mesh_s {
  ubt::Mesh m;
};
struct ubt_disc_s {
  std::shared_ptr<ubt::Disc> d;
};
struct ubt_op_s {
  std::unique_ptr<ubt::Operator> op;
};
struct ubt_blockinv_s {
  ubt::BlockInverse bi;
};
struct ubt_dsa_s {
  ubt::DSA d;
};
struct ubt_sweep_s {
  ubt::Sweep s;
};
struct ubt_solver_s {
  ubt::Solver s;
};
struct ubt_batch_s {
  ubt::Batch b;
};
struct ubt_batch_solver_s {
  ubt::BatchSolver s;
};

namespace {

template <class F> int guard(F &&f) {
  try {
    f();
    ubt::set_last_error(UBT_OK, "");
    return UBT_OK;
  } catch (const ubt::Error &e) {
    ubt::set_last_error(e.code, e.what());
    return e.code;
  } catch (const std::bad_alloc &) {
    ubt::set_last_error(UBT_ERR_MEM, "host allocation failed");
    return UBT_ERR_MEM;
  } catch (const std::exception &e) {
    ubt::set_last_error(UBT_ERR_PLIB, e.what());
    return UBT_ERR_PLIB;
  }
}

#define NEED(p) UBT_CHECK((p) != nullptr, UBT_ERR_ARG_NULL, "%s is NULL", #p)

ubt::BCTable bc_table(const ubt_bc *bcs, int n) {
  ubt::BCTable t;
  UBT_CHECK(n >= 0 && (n == 0 || bcs), UBT_ERR_ARG_NULL, "boundary conditions: %d entries but no array", n);
  for (int i = 0; i < n; i++) {
    ubt::BCFace f;
    f.type = bcs[i].type;
    f.inflow = bcs[i].inflow;
    f.n_window_pairs = bcs[i].n_window_pairs;
    for (int k = 0; k < 4; k++) f.window[k] = bcs[i].window[k];
    UBT_CHECK(f.n_window_pairs >= 0 && f.n_window_pairs <= 2, UBT_ERR_ARG_WRONG, "a face takes at most 2 window pairs");
    for (auto &e : t.faces)
      UBT_CHECK(e.first != bcs[i].label, UBT_ERR_ARG_WRONG, "label %d is given two boundary conditions", bcs[i].label);
    t.faces.push_back({bcs[i].label, f});
  }
  return t;
}

template <class V> void copy_out(const std::vector<V> &v, void *out, size_t *nbytes) {
  const size_t n = v.size() * sizeof(V);
  if (out) {
    UBT_CHECK(*nbytes >= n, UBT_ERR_ARG_SIZ, "buffer of %zu bytes, the table needs %zu", *nbytes, n);
    if (n) std::memcpy(out, v.data(), n);
  }
  *nbytes = n;
}

ubt::GroupOptions group_options(const ubt_group_options *o) {
  ubt::GroupOptions g;
  if (o) {
    g.outer_rtol = o->outer_rtol;
    g.max_outer = o->max_outer;
    g.inner_gamma = o->inner_gamma;
    g.inner_rtol = o->inner_rtol;
    g.store_moments = o->store_moments;
    g.recon_steps = o->recon_steps;
  }
  return g;
}
void group_result(const ubt::GroupReport &r, ubt_group_result *out, int *its, double *res, int *reason, int stride) {
  if (out) {
    out->g_up = r.g_up;
    out->sweeps = r.sweeps;
    out->status = r.status;
    out->change = r.change;
    out->rho = r.rho;
    out->err_est = r.err_est;
    out->iterations = r.iterations;
    out->recon_steps = r.recon_steps;
  }
  for (size_t g = 0; g < r.group_iterations.size(); g++) {
    if (its) its[g] = r.group_iterations[g];
    if (res) res[g] = r.group_residual[g];
    if (reason) reason[g] = r.group_reason[g];
  }
  for (int g = (int)r.group_iterations.size(); g < stride; g++) {
    if (its) its[g] = 0;
    if (res) res[g] = 0.0;
    if (reason) reason[g] = 0;
  }
}

}  // namespace

extern "C" {

int ubt_last_error(char *buf, size_t len) { return ubt::last_error_copy(buf, len); }

int ubt_device_count(int *count) {
  return guard([&] {
    NEED(count);
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) n = 0;
    *count = n;
  });
}

// quadrature ---------------------------------------------------------------
int ubt_quad_create(int dim, int sn_order, ubt_quad **quad) {
  return guard([&] {
    NEED(quad);
    auto q = std::make_unique<ubt_quad_s>();
    q->q = ubt::make_quadrature(dim, sn_order);
    *quad = q.release();
  });
}
int ubt_quad_size(const ubt_quad *quad, int *n_angles, double *sum_weights) {
  return guard([&] {
    NEED(quad);
    if (n_angles) *n_angles = quad->q.n_angles;
    if (sum_weights) *sum_weights = quad->q.sum_weights;
  });
}
int ubt_quad_get(const ubt_quad *quad, double *omega, double *w, int *reflect) {
  return guard([&] {
    NEED(quad);
    const auto &q = quad->q;
    if (omega) std::memcpy(omega, q.omega.data(), q.omega.size() * sizeof(double));
    if (w) std::memcpy(w, q.w.data(), q.w.size() * sizeof(double));
    if (reflect)
      for (int a = 0; a < q.n_angles; a++)
        for (int d = 0; d < 3; d++) reflect[3 * a + d] = q.reflect[d][a];
  });
}
int ubt_quad_destroy(ubt_quad *quad) {
  delete quad;
  return UBT_OK;
}
int ubt_lqn_octant(int sn_order, int *n_levels, double *mu, int *n_points, int *triples, double *weights) {
  return guard([&] {
    auto o = ubt::lqn_octant(sn_order);
    if (n_levels) *n_levels = (int)o.mu.size();
    if (n_points) *n_points = (int)o.weight.size();
    if (mu) std::memcpy(mu, o.mu.data(), o.mu.size() * sizeof(double));
    if (triples) std::memcpy(triples, o.triple.data(), o.triple.size() * sizeof(int));
    if (weights) std::memcpy(weights, o.weight.data(), o.weight.size() * sizeof(double));
  });
}

// mesh ---------------------------------------------------------------------
int ubt_mesh_create(int dim, int n_vertices, const double *coords, int n_cells, int verts_per_cell,
                    const int *cell_verts, ubt_mesh **mesh) {
  return guard([&] {
    NEED(coords);
    NEED(cell_verts);
    NEED(mesh);
    auto m = std::make_unique<ubt_mesh_s>();
    m->m = ubt::make_mesh(dim, n_vertices, coords, n_cells, verts_per_cell, cell_verts);
    *mesh = m.release();
  });
}
int ubt_mesh_label_box(ubt_mesh *mesh) {
  return guard([&] {
    NEED(mesh);
    ubt::mesh_label_box(mesh->m);
  });
}
int ubt_mesh_set_boundary_labels(ubt_mesh *mesh, int n_faces, int verts_per_face, const int *face_verts,
                                 const int *labels) {
  return guard([&] {
    NEED(mesh);
    UBT_CHECK(n_faces >= 0, UBT_ERR_ARG_OUTOFRANGE, "negative face count");
    if (n_faces == 0) return;
    NEED(face_verts);
    NEED(labels);
    ubt::mesh_set_boundary_labels(mesh->m, n_faces, verts_per_face, face_verts, labels);
  });
}
int ubt_mesh_size(const ubt_mesh *mesh, int *n_cells, int *faces_per_cell, int *n_boundary_faces) {
  return guard([&] {
    NEED(mesh);
    if (n_cells) *n_cells = mesh->m.n_cells;
    if (faces_per_cell) *faces_per_cell = mesh->m.faces_per_cell;
    if (n_boundary_faces) {
      int nb = 0;
      for (int v : mesh->m.face_neighbour) nb += v < 0;
      *n_boundary_faces = nb;
    }
  });
}
int ubt_mesh_destroy(ubt_mesh *mesh) {
  delete mesh;
  return UBT_OK;
}

// discretisation -------------------------------------------------------------
int ubt_disc_create_structured(const ubt_quad *quad, int dim, const int *n_cells, const double *lengths,
                               const ubt_bc *bcs, int n_bcs, ubt_disc **disc) {
  return guard([&] {
    NEED(quad);
    NEED(n_cells);
    NEED(lengths);
    NEED(disc);
    int n[3] = {1, 1, 1};
    double len[3] = {1, 1, 1};
    UBT_CHECK(dim == 2 || dim == 3, UBT_ERR_SUP, "the structured backend is 2D or 3D, got %d", dim);
    for (int d = 0; d < dim; d++) {
      n[d] = n_cells[d];
      len[d] = lengths[d];
    }
    auto h = std::make_unique<ubt_disc_s>();
    h->d = ubt::make_structured(quad->q, dim, n, len, bc_table(bcs, n_bcs));
    *disc = h.release();
  });
}
int ubt_disc_create_dg(const ubt_quad *quad, const ubt_mesh *mesh, int order, const ubt_bc *bcs, int n_bcs,
                       ubt_disc **disc) {
  return guard([&] {
    NEED(quad);
    NEED(mesh);
    NEED(disc);
    auto h = std::make_unique<ubt_disc_s>();
    h->d = ubt::make_unstructured(quad->q, mesh->m, order, bc_table(bcs, n_bcs));
    *disc = h.release();
  });
}
int ubt_disc_create_dg_device(const ubt_quad *quad, const ubt_mesh *mesh, int order, const ubt_bc *bcs, int n_bcs,
                              void *stream, ubt_disc **disc) {
  return guard([&] {
    NEED(quad);
    NEED(mesh);
    NEED(disc);
    auto h = std::make_unique<ubt_disc_s>();
    h->d = ubt::make_unstructured_device(quad->q, mesh->m, order, bc_table(bcs, n_bcs), (hipStream_t)stream);
    *disc = h.release();
  });
}
int ubt_disc_size(const ubt_disc *disc, int *dim, int *n_cells, int *n_basis, int *n_angles, int *faces_per_cell) {
  return guard([&] {
    NEED(disc);
    const auto &D = *disc->d;
    if (dim) *dim = D.dim;
    if (n_cells) *n_cells = D.n_cells;
    if (n_basis) *n_basis = D.nb;
    if (n_angles) *n_angles = D.n_angles;
    if (faces_per_cell) *faces_per_cell = D.kind == ubt::DISC_FD ? 2 * D.dim : D.nf;
  });
}
int ubt_disc_get(const ubt_disc *disc, const char *name, void *out, size_t *nbytes) {
  return guard([&] {
    NEED(disc);
    NEED(name);
    NEED(nbytes);
    const auto &D = *disc->d;
    const std::string n(name);
    if (n == "volume") copy_out(D.volume, out, nbytes);
    else if (n == "centroid") copy_out(D.centroid, out, nbytes);
    else if (n == "face_nb") copy_out(D.face_nb, out, nbytes);
    else if (n == "face_label") copy_out(D.face_label, out, nbytes);
    else if (n == "face_nA") copy_out(D.face_nA, out, nbytes);
    else if (n == "face_area") copy_out(D.face_area, out, nbytes);
    else if (n == "face_centroid") copy_out(D.face_centroid, out, nbytes);
    else if (n == "face_distance") copy_out(D.face_distance, out, nbytes);
    else if (n == "face_inflow") copy_out(D.face_inflow, out, nbytes);
    else if (n == "face_own") copy_out(D.face_own, out, nbytes);
    else if (n == "face_up") copy_out(D.face_up, out, nbytes);
    else if (n == "basis_grad") copy_out(D.basis_grad, out, nbytes);
    else if (n == "face_basis_value") copy_out(D.face_basis_value, out, nbytes);
    else if (n == "face_nb_grad") copy_out(D.face_nb_grad, out, nbytes);
    else if (n == "h") copy_out(std::vector<double>(D.h, D.h + 3), out, nbytes);
    else if (n == "n") copy_out(std::vector<int>(D.n, D.n + 3), out, nbytes);
    else if (n == "bc_face") {
      std::vector<double> v;
      for (int d = 0; d < 3; d++)
        for (int s = 0; s < 2; s++) {
          const auto &f = D.face_bc[d][s];
          v.push_back(D.face_type[d][s]);
          v.push_back(f.inflow / D.quad.sum_weights);
          v.push_back(f.n_window_pairs);
          for (int k = 0; k < 4; k++) v.push_back(f.window[k]);
        }
      copy_out(v, out, nbytes);
    } else {
      ubt::fail(UBT_ERR_ARG_WRONG, "no geometry table called \"%s\"", name);
    }
  });
}
int ubt_disc_destroy(ubt_disc *disc) {
  delete disc;
  return UBT_OK;
}

// operator -------------------------------------------------------------------
int ubt_op_create(const ubt_disc *disc, int n_groups, ubt_op **op) {
  return guard([&] {
    NEED(disc);
    NEED(op);
    auto h = std::make_unique<ubt_op_s>();
    h->op = ubt::make_operator(disc->d, n_groups);
    *op = h.release();
  });
}
int ubt_op_set_xs(ubt_op *op, const double *sigma_t, const double *sigma_s, const double *source) {
  return guard([&] {
    NEED(op);
    ubt::operator_set_xs(*op->op, sigma_t, sigma_s, source);
  });
}
int ubt_op_set_group(ubt_op *op, int g) {
  return guard([&] {
    NEED(op);
    UBT_CHECK(g >= 0 && g < op->op->n_groups, UBT_ERR_ARG_OUTOFRANGE, "group %d out of range, n_groups is %d", g,
              op->op->n_groups);
    op->op->group = g;
  });
}
int ubt_op_size(const ubt_op *op, int64_t *n_rows, int64_t *n_nodes, int *n_groups) {
  return guard([&] {
    NEED(op);
    if (n_rows) *n_rows = op->op->disc->n_rows();
    if (n_nodes) *n_nodes = op->op->disc->n_nodes();
    if (n_groups) *n_groups = op->op->n_groups;
  });
}
int ubt_op_mult(ubt_op *op, int prec, const void *x, void *y, void *stream) {
  return guard([&] {
    NEED(op);
    op->op->mult(prec, x, y, true, (hipStream_t)stream);
  });
}
int ubt_op_mult_unscattered(ubt_op *op, int prec, const void *x, void *y, void *stream) {
  return guard([&] {
    NEED(op);
    op->op->mult(prec, x, y, false, (hipStream_t)stream);
  });
}
int ubt_op_geometry_classes(const ubt_op *op, int *n_classes) {
  return guard([&] {
    NEED(op);
    NEED(n_classes);
    *n_classes = op->op->n_geo_classes;
  });
}
int ubt_op_mult_transpose(ubt_op *op, int prec, const void *x, void *y, void *stream) {
  return guard([&] {
    NEED(op);
    op->op->mult_transpose(prec, x, y, true, (hipStream_t)stream);
  });
}
int ubt_op_mult_unscattered_transpose(ubt_op *op, int prec, const void *x, void *y, void *stream) {
  return guard([&] {
    NEED(op);
    op->op->mult_transpose(prec, x, y, false, (hipStream_t)stream);
  });
}
int ubt_op_diagonal(ubt_op *op, int prec, void *d, void *stream) {
  return guard([&] {
    NEED(op);
    NEED(d);
    op->op->diagonal(prec, d, (hipStream_t)stream);
  });
}
int ubt_op_set_variant(ubt_op *op, int prec, int variant) {
  return guard([&] {
    NEED(op);
    UBT_CHECK(prec == UBT_FP64 || prec == UBT_FP32, UBT_ERR_ARG_OUTOFRANGE, "unknown precision %d", prec);
    UBT_CHECK(variant >= -1 && variant < ubt::kApplyVariants, UBT_ERR_ARG_OUTOFRANGE,
              "apply variant must be -1 (tune) or 0 to %d", ubt::kApplyVariants - 1);
    UBT_CHECK(op->op->xs_set, UBT_ERR_ARG_WRONGSTATE, "set the cross sections first");
    if (variant == -1)
      op->op->tune();
    else
      op->op->variant[prec == UBT_FP64 ? 0 : 1] = variant;
  });
}
int ubt_op_tuning(const ubt_op *op, int *variant, float *us) {
  return guard([&] {
    NEED(op);
    if (variant)
      for (int p = 0; p < 2; p++) variant[p] = op->op->variant[p];
    if (us)
      for (int p = 0; p < 2; p++)
        for (int v = 0; v < ubt::kApplyVariants; v++) us[ubt::kApplyVariants * p + v] = op->op->tuned_us[p][v];
  });
}
int ubt_op_destroy(ubt_op *op) {
  delete op;
  return UBT_OK;
}

// moments and sources ----------------------------------------------------------
int ubt_moments(ubt_op *op, int prec, const void *psi, void *phi, int n_moments, void *stream) {
  return guard([&] {
    NEED(op);
    op->op->moments(prec, psi, phi, n_moments, (hipStream_t)stream);
  });
}
int ubt_set_scalar_flux(ubt_op *op, int g, int prec, const void *psi, void *stream) {
  return guard([&] {
    NEED(op);
    NEED(psi);
    op->op->set_scalar_flux(g, prec, psi, (hipStream_t)stream);
  });
}
int ubt_scatter_source(ubt_op *op, int g_from, int g_to, int prec, void *b, void *stream) {
  return guard([&] {
    NEED(op);
    NEED(b);
    op->op->scatter_source(g_from, g_to, prec, b, (hipStream_t)stream);
  });
}
int ubt_op_scalar_flux(const ubt_op *op, int g, double *phi, void *stream) {
  return guard([&] {
    NEED(op);
    UBT_CHECK(phi, UBT_ERR_ARG_NULL, "phi is required");
    const auto &o = *op->op;
    UBT_CHECK(g >= 0 && g < o.n_groups, UBT_ERR_ARG_OUTOFRANGE, "group %d out of range, n_groups is %d", g, o.n_groups);
    UBT_CHECK(o.phi_set[g], UBT_ERR_ARG_WRONGSTATE, "no scalar flux cached for group %d", g);
    const int64_t nodes = o.disc->n_nodes();
    UBT_HIP(hipMemcpyAsync(phi, o.phi_cache.as<double>() + (int64_t)g * nodes, sizeof(double) * (size_t)nodes,
                           hipMemcpyDeviceToDevice, (hipStream_t)stream));
  });
}
int ubt_rhs(ubt_op *op, int g, int prec, void *b, void *stream) {
  return guard([&] {
    NEED(op);
    NEED(b);
    op->op->rhs(g, prec, b, (hipStream_t)stream);
  });
}

// block inverse ------------------------------------------------------------------
int ubt_blockinv_create(ubt_op *op, ubt_blockinv **bi) {
  return guard([&] {
    NEED(op);
    NEED(bi);
    UBT_CHECK(op->op->disc->nb <= 4, UBT_ERR_SUP, "element blocks of up to 4 basis functions");
    auto h = std::make_unique<ubt_blockinv_s>();
    h->bi.op = op->op.get();
    *bi = h.release();
  });
}
int ubt_blockinv_setup(ubt_blockinv *bi, void *stream) {
  return guard([&] {
    NEED(bi);
    bi->bi.setup((hipStream_t)stream);
  });
}
int ubt_blockinv_apply(ubt_blockinv *bi, int prec, const void *x, void *y, void *stream) {
  return guard([&] {
    NEED(bi);
    bi->bi.apply(prec, x, y, (hipStream_t)stream);
  });
}
int ubt_blockinv_apply_transpose(ubt_blockinv *bi, int prec, const void *x, void *y, void *stream) {
  return guard([&] {
    NEED(bi);
    bi->bi.apply(prec, x, y, (hipStream_t)stream, true);
  });
}
int ubt_blockinv_destroy(ubt_blockinv *bi) {
  delete bi;
  return UBT_OK;
}

// DSA -----------------------------------------------------------------------------
int ubt_dsa_create(ubt_op *op, ubt_dsa **dsa) {
  return guard([&] {
    NEED(op);
    NEED(dsa);
    auto h = std::make_unique<ubt_dsa_s>();
    h->d.create(op->op.get());
    *dsa = h.release();
  });
}
int ubt_dsa_set_options(ubt_dsa *dsa, int consistent_d, double consistent_power, double mip_penalty, int inner_max_it,
                        double inner_rtol, int inner_check_every) {
  return guard([&] {
    NEED(dsa);
    UBT_CHECK(consistent_power >= 1.0, UBT_ERR_ARG_OUTOFRANGE, "the consistent-D power must be at least 1, got %g",
              consistent_power);
    UBT_CHECK(mip_penalty > 0.0, UBT_ERR_ARG_OUTOFRANGE, "the MIP penalty must be positive, got %g", mip_penalty);
    UBT_CHECK(inner_max_it >= 1 && inner_rtol >= 0.0 && inner_check_every >= 0, UBT_ERR_ARG_OUTOFRANGE,
              "inner solve needs max_it >= 1, rtol >= 0, check_every >= 0");
    dsa->d.opt.consistent_d = consistent_d;
    dsa->d.opt.consistent_power = consistent_power;
    dsa->d.opt.mip_penalty = mip_penalty;
    dsa->d.opt.max_it = inner_max_it;
    dsa->d.opt.rtol = inner_rtol;
    dsa->d.opt.check_every = inner_check_every;
  });
}
int ubt_dsa_set_void_policy(ubt_dsa *dsa, int bridge, double void_sigma_t, double void_d) {
  return guard([&] {
    NEED(dsa);
    UBT_CHECK(void_sigma_t >= 0.0 && void_d >= 0.0, UBT_ERR_ARG_OUTOFRANGE,
              "the void threshold and the void D must not be negative, got %g and %g", void_sigma_t, void_d);
    dsa->d.opt.void_bridge = bridge != 0;
    dsa->d.opt.void_sigma_t = void_sigma_t;
    dsa->d.opt.void_d = void_d;
  });
}
int ubt_dsa_set_group(ubt_dsa *dsa, int g, void *stream) {
  return guard([&] {
    NEED(dsa);
    dsa->d.set_group(g, (hipStream_t)stream);
  });
}
int ubt_dsa_apply(ubt_dsa *dsa, int prec, const void *x, void *y, void *stream) {
  return guard([&] {
    NEED(dsa);
    dsa->d.apply(prec, x, y, (hipStream_t)stream);
  });
}
int ubt_dsa_apply_transpose(ubt_dsa *dsa, int prec, const void *x, void *y, void *stream) {
  return guard([&] {
    NEED(dsa);
    dsa->d.apply(prec, x, y, (hipStream_t)stream, true);
  });
}
int ubt_dsa_matvec(ubt_dsa *dsa, const double *x, double *y, void *stream) {
  return guard([&] {
    NEED(dsa);
    NEED(x);
    NEED(y);
    dsa->d.matvec(x, y, (hipStream_t)stream);
  });
}
int ubt_dsa_solve(ubt_dsa *dsa, const double *b, double *x, void *stream) {
  return guard([&] {
    NEED(dsa);
    NEED(b);
    NEED(x);
    dsa->d.solve(b, x, (hipStream_t)stream);
  });
}
int ubt_dsa_solve_noalloc(ubt_dsa *dsa, const double *b, double *x, void *stream) {
  return guard([&] {
    NEED(dsa);
    NEED(b);
    NEED(x);
    dsa->d.solve_noalloc(b, x, (hipStream_t)stream);
  });
}
int ubt_dsa_info(ubt_dsa *dsa, int *inner_iterations, int *n_nodes) {
  return guard([&] {
    NEED(dsa);
    if (inner_iterations) {
      double it = 0.0;
      UBT_HIP(hipMemcpy(&it, dsa->d.scal.as<double>() + 5, sizeof(double), hipMemcpyDeviceToHost));
      *inner_iterations = (int)it;
    }
    if (n_nodes) *n_nodes = (int)dsa->d.n_nodes;
  });
}
int ubt_dsa_destroy(ubt_dsa *dsa) {
  delete dsa;
  return UBT_OK;
}

// sweep ---------------------------------------------------------------------------
int ubt_sweep_create(ubt_op *op, ubt_sweep **sw) {
  return guard([&] {
    NEED(op);
    NEED(sw);
    auto h = std::make_unique<ubt_sweep_s>();
    h->s.create(op->op.get());
    *sw = h.release();
  });
}
int ubt_sweep_create_opts(ubt_op *op, int lds_traces, ubt_sweep **sw) {
  return guard([&] {
    NEED(op);
    NEED(sw);
    auto h = std::make_unique<ubt_sweep_s>();
    h->s.create(op->op.get(), false, lds_traces != 0);
    *sw = h.release();
  });
}
int ubt_sweep_trace_slots(const ubt_sweep *sw, int *slots) {
  return guard([&] {
    NEED(sw);
    NEED(slots);
    *slots = sw->s.trace_slots;
  });
}
int ubt_sweep_info(const ubt_sweep *sw, int *n_levels, int64_t *n_lagged) {
  return guard([&] {
    NEED(sw);
    if (n_levels) *n_levels = sw->s.n_levels;
    if (n_lagged) *n_lagged = sw->s.n_lagged;
  });
}
int ubt_sweep_apply(ubt_sweep *sw, int prec, const void *r, void *psi, void *stream) {
  return guard([&] {
    NEED(sw);
    sw->s.apply(prec, r, psi, (hipStream_t)stream);
  });
}
int ubt_sweep_apply_transpose(ubt_sweep *sw, int prec, const void *r, void *psi, void *stream) {
  return guard([&] {
    NEED(sw);
    sw->s.apply(prec, r, psi, (hipStream_t)stream, true);
  });
}
int ubt_sweep_destroy(ubt_sweep *sw) {
  delete sw;
  return UBT_OK;
}

// solver --------------------------------------------------------------------------
int ubt_solver_create(ubt_op *op, ubt_dsa *dsa, const ubt_solver_options *options, ubt_solver **solver) {
  return guard([&] {
    NEED(op);
    NEED(solver);
    ubt::SolverOptions o;
    if (options) {
      o.rtol = options->rtol;
      o.max_it = options->max_it;
      o.restart = options->restart;
      o.mixed = options->mixed;
      o.inner_rtol = options->inner_rtol;
      o.max_outer = options->max_outer;
      o.pc_blocks = options->pc_blocks;
      o.pc_sweep = options->pc_sweep;
      o.pc_dsa = options->pc_dsa;
      o.use_amat = options->use_amat;
      o.gs_passes = options->gs_passes > 0 ? options->gs_passes : 2;
    }
    UBT_CHECK(o.gs_passes == 1 || o.gs_passes == 2, UBT_ERR_ARG_OUTOFRANGE, "gs_passes must be 1 or 2, got %d",
              o.gs_passes);
    UBT_CHECK(!o.mixed || (o.inner_rtol > 0.0 && o.inner_rtol < 1.0 && o.max_outer >= 1), UBT_ERR_ARG_OUTOFRANGE,
              "mixed mode needs 0 < inner_rtol < 1 and max_outer >= 1");
    auto h = std::make_unique<ubt_solver_s>();
    h->s.create(op->op.get(), dsa ? &dsa->d : nullptr, o);
    *solver = h.release();
  });
}
int ubt_solver_refresh(ubt_solver *solver, void *stream) {
  return guard([&] {
    NEED(solver);
    solver->s.refresh((hipStream_t)stream);
  });
}
int ubt_solver_solve(ubt_solver *solver, const double *b, double *x, void *stream) {
  return guard([&] {
    NEED(solver);
    solver->s.solve(b, x, (hipStream_t)stream);
  });
}
int ubt_solver_solve_transpose(ubt_solver *solver, const double *b, double *x, void *stream) {
  return guard([&] {
    NEED(solver);
    solver->s.solve(b, x, (hipStream_t)stream, true);
  });
}
int ubt_solver_apply_pc(ubt_solver *solver, int prec, const void *r, void *y, void *stream) {
  return guard([&] {
    NEED(solver);
    NEED(r);
    NEED(y);
    UBT_CHECK(r != y, UBT_ERR_ARG_IDN, "the preconditioner apply needs distinct r and y");
    auto &s = solver->s;
    s.transpose = false;
    UBT_CHECK(s.has_precision(prec), UBT_ERR_ARG_INCOMP, "this solver holds %s work vectors",
              s.opt.mixed ? "FP32" : "FP64");
    if (prec == UBT_FP64)
      s.apply_pc<double>((const double *)r, (double *)y, (hipStream_t)stream);
    else
      s.apply_pc<float>((const float *)r, (float *)y, (hipStream_t)stream);
  });
}
int ubt_solver_result(const ubt_solver *solver, int *iterations, int *outer, double *rel_residual, int *reason,
                      double *history, int max_history, int *n_history) {
  return guard([&] {
    NEED(solver);
    const auto &s = solver->s;
    if (iterations) *iterations = s.iterations;
    if (outer) *outer = s.outer;
    if (rel_residual) *rel_residual = s.rel_residual;
    if (reason) *reason = s.reason;
    const int nh = (int)std::min<size_t>(s.history.size(), (size_t)std::max(0, max_history));
    if (history)
      for (int i = 0; i < nh; i++) history[i] = s.history[i];
    if (n_history) *n_history = (int)s.history.size();
  });
}
int ubt_solver_destroy(ubt_solver *solver) {
  delete solver;
  return UBT_OK;
}

// batches -------------------------------------------------------------------------
int ubt_batch_create(ubt_op *const *ops, int n, ubt_batch **batch) {
  return guard([&] {
    NEED(ops);
    NEED(batch);
    UBT_CHECK(n >= 1, UBT_ERR_ARG_SIZ, "a batch needs at least one problem, got %d", n);
    std::vector<ubt::Operator *> o;
    for (int p = 0; p < n; p++) {
      UBT_CHECK(ops[p], UBT_ERR_ARG_NULL, "operator %d is NULL", p);
      o.push_back(ops[p]->op.get());
    }
    auto h = std::make_unique<ubt_batch_s>();
    h->b.create(o);
    *batch = h.release();
  });
}
int ubt_batch_size(const ubt_batch *batch, int64_t *n_rows, int *n_problems) {
  return guard([&] {
    NEED(batch);
    if (n_rows) *n_rows = batch->b.n_rows;
    if (n_problems) *n_problems = (int)batch->b.ops.size();
  });
}
int ubt_batch_offsets(const ubt_batch *batch, int64_t *off) {
  return guard([&] {
    NEED(batch);
    NEED(off);
    for (size_t p = 0; p < batch->b.off.size(); p++) off[p] = batch->b.off[p];
  });
}
int ubt_batch_set_group(ubt_batch *batch, int g) {
  return guard([&] {
    NEED(batch);
    batch->b.set_group(g);
  });
}
int ubt_batch_mult(ubt_batch *batch, int prec, const void *x, void *y, void *stream) {
  return guard([&] {
    NEED(batch);
    batch->b.mult(prec, x, y, true, (hipStream_t)stream);
  });
}
int ubt_batch_tuning(const ubt_batch *batch, int *body, int *persist, float *us) {
  return guard([&] {
    NEED(batch);
    for (int p = 0; p < 2; p++) {
      if (body) body[p] = batch->b.body[p];
      if (persist) persist[p] = batch->b.persist[p];
      if (us)
        for (int v = 0; v < ubt::kBatchBodies; v++) us[ubt::kBatchBodies * p + v] = batch->b.tuned_us[p][v];
    }
  });
}
int ubt_batch_set_body(ubt_batch *batch, int prec, int body) {
  return guard([&] {
    NEED(batch);
    UBT_CHECK(prec == UBT_FP64 || prec == UBT_FP32, UBT_ERR_ARG_OUTOFRANGE, "unknown precision %d", prec);
    UBT_CHECK(body >= 0 && body < ubt::kBatchBodies, UBT_ERR_ARG_OUTOFRANGE, "apply body must be 0 to %d",
              ubt::kBatchBodies - 1);
    const int pi = prec == UBT_FP64 ? 0 : 1;
    UBT_CHECK(batch->b.body_ok[pi][body], UBT_ERR_SUP, "apply body %d does not apply to this batch", body);
    batch->b.body[pi] = body;
  });
}
int ubt_batch_set_sweep(ubt_batch *batch, int box) {
  return guard([&] {
    NEED(batch);
    batch->b.box_sweep_enabled = box != 0;
    batch->b.build_box_sweep();
    batch->b.gen++;  // a solver refreshes (and recaptures) before its next solve
  });
}
int ubt_batch_sweep_info(const ubt_batch *batch, int *box, int *exact) {
  return guard([&] {
    NEED(batch);
    NEED(box);
    box[0] = batch->b.box_sweep[0].ok ? 1 : 0;
    box[1] = batch->b.box_sweep[1].ok ? 1 : 0;
    if (exact) {
      exact[0] = batch->b.sweep_exact(UBT_FP64) ? 1 : 0;
      exact[1] = batch->b.sweep_exact(UBT_FP32) ? 1 : 0;
    }
  });
}
int ubt_batch_rhs(ubt_batch *batch, int prec, void *b, void *stream) {
  return guard([&] {
    NEED(batch);
    NEED(b);
    batch->b.rhs(prec, b, (hipStream_t)stream);
  });
}
int ubt_batch_set_scalar_flux(ubt_batch *batch, int prec, const void *psi, void *stream) {
  return guard([&] {
    NEED(batch);
    NEED(psi);
    batch->b.set_scalar_flux(prec, psi, (hipStream_t)stream);
  });
}
int ubt_batch_blocks_apply(ubt_batch *batch, int prec, const void *x, void *y, void *stream) {
  return guard([&] {
    NEED(batch);
    NEED(x);
    NEED(y);
    UBT_CHECK(x != y, UBT_ERR_ARG_IDN, "the block inverse apply needs distinct x and y");
    batch->b.blocks_apply(prec, x, y, (hipStream_t)stream);
  });
}
int ubt_batch_sweep_apply(ubt_batch *batch, int prec, const void *r, void *psi, void *stream) {
  return guard([&] {
    NEED(batch);
    NEED(r);
    NEED(psi);
    UBT_CHECK(r != psi, UBT_ERR_ARG_IDN, "the sweep needs distinct r and psi");
    batch->b.sweep_apply(prec, r, psi, (hipStream_t)stream);
  });
}
int ubt_batch_destroy(ubt_batch *batch) {
  delete batch;
  return UBT_OK;
}
int ubt_batch_solver_create(ubt_batch *batch, const ubt_solver_options *options, ubt_batch_solver **solver) {
  return guard([&] {
    NEED(batch);
    NEED(solver);
    ubt::SolverOptions o;
    if (options) {
      o.rtol = options->rtol;
      o.max_it = options->max_it;
      o.restart = options->restart;
      o.mixed = options->mixed;
      o.inner_rtol = options->inner_rtol;
      o.max_outer = options->max_outer;
      o.pc_blocks = options->pc_blocks;
      o.pc_sweep = options->pc_sweep;
      o.pc_dsa = 0;
      o.use_amat = options->use_amat;
      o.gs_passes = options->gs_passes > 0 ? options->gs_passes : 2;
    }
    UBT_CHECK(o.gs_passes == 1 || o.gs_passes == 2, UBT_ERR_ARG_OUTOFRANGE, "gs_passes must be 1 or 2, got %d",
              o.gs_passes);
    UBT_CHECK(!o.mixed || (o.inner_rtol > 0.0 && o.inner_rtol < 1.0 && o.max_outer >= 1), UBT_ERR_ARG_OUTOFRANGE,
              "mixed mode needs 0 < inner_rtol < 1 and max_outer >= 1");
    auto h = std::make_unique<ubt_batch_solver_s>();
    h->s.create(&batch->b, o);
    *solver = h.release();
  });
}
int ubt_batch_solver_refresh(ubt_batch_solver *solver, void *stream) {
  return guard([&] {
    NEED(solver);
    solver->s.refresh((hipStream_t)stream);
  });
}
int ubt_batch_solver_solve(ubt_batch_solver *solver, const double *b, double *x, void *stream) {
  return guard([&] {
    NEED(solver);
    solver->s.solve(b, x, (hipStream_t)stream);
  });
}
int ubt_batch_solver_result(const ubt_batch_solver *solver, int *iterations, double *rel_residual, int *reason,
                            int *outer, int *cycles) {
  return guard([&] {
    NEED(solver);
    const auto &s = solver->s.eng;
    for (size_t p = 0; p < s.iterations.size(); p++) {
      if (iterations) iterations[p] = s.iterations[p];
      if (rel_residual) rel_residual[p] = s.rel_residual[p];
      if (reason) reason[p] = s.reason[p];
    }
    if (outer) *outer = s.outer;
    if (cycles) *cycles = s.cycles;
  });
}

int ubt_solver_solve_groups(ubt_solver *solver, const ubt_group_options *options, double *psi, void *stream) {
  return guard([&] {
    NEED(solver);
    ubt::solve_groups(solver->s, group_options(options), psi, (hipStream_t)stream, solver->s.groups);
  });
}
int ubt_solver_groups_result(const ubt_solver *solver, ubt_group_result *result, int *iterations,
                             double *rel_residual, int *reason) {
  return guard([&] {
    NEED(solver);
    const auto &r = solver->s.groups;
    group_result(r, result, iterations, rel_residual, reason, (int)r.group_iterations.size());
  });
}
int ubt_batch_solver_solve_groups(ubt_batch_solver *solver, const ubt_group_options *options, double *psi,
                                  void *stream) {
  return guard([&] {
    NEED(solver);
    ubt::solve_groups(solver->s, group_options(options), psi, (hipStream_t)stream, solver->s.groups);
  });
}
int ubt_batch_solver_groups_result(const ubt_batch_solver *solver, ubt_group_result *result, int *iterations,
                                   double *rel_residual, int *reason) {
  return guard([&] {
    NEED(solver);
    const auto &rs = solver->s.groups;
    int G = 0;
    for (const auto *op : solver->s.batch->ops) G = std::max(G, op->n_groups);
    for (size_t p = 0; p < rs.size(); p++)
      group_result(rs[p], result ? result + p : nullptr, iterations ? iterations + p * G : nullptr,
                   rel_residual ? rel_residual + p * G : nullptr, reason ? reason + p * G : nullptr, G);
  });
}

int ubt_batch_solver_set_graph(ubt_batch_solver *solver, int on) {
  return guard([&] {
    NEED(solver);
    solver->s.eng.use_graph = on != 0;
    solver->s.eng.drop_graphs();
  });
}
int ubt_batch_solver_destroy(ubt_batch_solver *solver) {
  delete solver;
  return UBT_OK;
}

}  // extern "C"
