// Internal declarations shared by the host translation units and the HIP
// kernels of libubolt_amd. Nothing here crosses the C ABI: include/ubt.h is the
// public surface, and every entry point there converts a C++ exception thrown
// from below into a status code plus a message for ubt_last_error().
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstring>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/ubt.h"

namespace ubt {

// ---------------------------------------------------------------------------
// Errors: a status class (the ubt_status values, which mirror PETSc's error
// classes so a PETSc adapter can hand them straight back) and a message
// ---------------------------------------------------------------------------
struct Error : std::runtime_error {
  int code;
  Error(int c, const std::string &m) : std::runtime_error(m), code(c) {}
};

[[noreturn]] void fail(int code, const char *fmt, ...) __attribute__((format(printf, 2, 3)));

#define UBT_CHECK(cond, code, ...)                                                                \
  do {                                                                                            \
    if (!(cond)) ::ubt::fail((code), __VA_ARGS__);                                                \
  } while (0)

#define UBT_HIP(call)                                                                             \
  do {                                                                                            \
    hipError_t e_ = (call);                                                                       \
    if (e_ != hipSuccess)                                                                         \
      ::ubt::fail(UBT_ERR_GPU, "%s failed: %s (%s:%d)", #call, hipGetErrorString(e_), __FILE__,  \
                  __LINE__);                                                                      \
  } while (0)

void set_last_error(int code, const char *msg);

// ---------------------------------------------------------------------------
// Device memory, owned by the handles that allocate it at create time. Apply
// paths never allocate.
// ---------------------------------------------------------------------------
struct DevBuf {
  void *ptr = nullptr;
  size_t bytes = 0;
  DevBuf() = default;
  DevBuf(const DevBuf &) = delete;
  DevBuf &operator=(const DevBuf &) = delete;
  DevBuf(DevBuf &&o) noexcept : ptr(o.ptr), bytes(o.bytes) { o.ptr = nullptr; o.bytes = 0; }
  DevBuf &operator=(DevBuf &&o) noexcept {
    if (this != &o) { release(); ptr = o.ptr; bytes = o.bytes; o.ptr = nullptr; o.bytes = 0; }
    return *this;
  }
  ~DevBuf() { release(); }
  void alloc(size_t n) {
    release();
    if (n == 0) return;
    UBT_HIP(hipMalloc(&ptr, n));
    bytes = n;
  }
  void release() {
    if (ptr) (void)hipFree(ptr);
    ptr = nullptr;
    bytes = 0;
  }
  template <class T> T *as() const { return static_cast<T *>(ptr); }
  // A table of the size already held is copied into the existing allocation,
  // so device pointers handed out earlier (a batch's work items hold an
  // operator's cross-section pointers) stay valid across re-uploads.
  template <class T> void upload(const std::vector<T> &h) {
    if (bytes != h.size() * sizeof(T)) alloc(h.size() * sizeof(T));
    if (!h.empty()) UBT_HIP(hipMemcpy(ptr, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice));
  }
  // Upload a double table converted to T (the FP32 copies of the geometry)
  template <class T> void upload_as(const std::vector<double> &h) {
    std::vector<T> c(h.begin(), h.end());
    upload(c);
  }
};

// ---------------------------------------------------------------------------
// Angular quadrature (host). Ordinates carry 3 cosines each (zero past dim),
// weights sum to sum_weights (2 in 1D, 4 pi in 2D/3D), reflect[d][a] is the
// ordinate with cosine d negated.
// ---------------------------------------------------------------------------
struct Quadrature {
  int dim = 0, sn_order = 0, n_angles = 0;
  double sum_weights = 0.0;
  std::vector<double> omega;  // 3 * n_angles
  std::vector<double> w;      // n_angles
  std::vector<int> reflect[3];
};

Quadrature make_quadrature(int dim, int sn_order);
// The level-symmetric octant for an even order (2 <= N <= 12): level cosines
// and, per octant point, its level triple and its weight as a fraction of the
// whole sphere. Exposed for the tests through ubt_lqn_octant.
struct LqnOctant {
  std::vector<double> mu;             // n levels
  std::vector<int> triple;            // 3 per point, 1-based
  std::vector<double> weight;         // per point
};
LqnOctant lqn_octant(int sn_order);

// ---------------------------------------------------------------------------
// Boundary conditions, keyed by face label (PETSc box "Face Sets" ids on the
// generated boxes). Unlisted labels are cold vacuum.
// ---------------------------------------------------------------------------
struct BCFace {
  int type = UBT_BC_VACUUM;
  double inflow = 0.0;
  int n_window_pairs = 0;
  double window[4] = {0, 0, 0, 0};
};

struct BCTable {
  std::vector<std::pair<int, BCFace>> faces;
  BCFace get(int label) const {
    for (auto &f : faces)
      if (f.first == label) return f.second;
    return BCFace();
  }
};

bool in_window(const BCFace &f, const double *t, int n_t);

// ---------------------------------------------------------------------------
// Unstructured mesh (host): vertices and one uniform cell type, faces matched
// between cells, boundary faces labelled.
// ---------------------------------------------------------------------------
struct Mesh {
  int dim = 0, n_vertices = 0, n_cells = 0, verts_per_cell = 0, faces_per_cell = 0;
  std::vector<double> coords;         // 3 per vertex (z = 0 in 2D)
  std::vector<int> cell_verts;        // verts_per_cell per cell
  // Per (cell, local face): the face's vertices in cyclic order (2 in 2D, 3 or
  // 4 in 3D, -1 padded to 4), the neighbour cell (-1 on the boundary) and the
  // boundary label (-1 interior or unlabelled)
  std::vector<int> face_verts;        // 4 per (cell, face)
  std::vector<int> face_nverts;
  std::vector<int> face_neighbour;
  std::vector<int> face_label;
};

Mesh make_mesh(int dim, int n_vertices, const double *coords, int n_cells, int verts_per_cell,
               const int *cell_verts);
void mesh_label_box(Mesh &m);
void mesh_set_boundary_labels(Mesh &m, int n_faces, int verts_per_face, const int *face_verts,
                              const int *labels);

// ---------------------------------------------------------------------------
// The discretisation: structured FD (cell-centred upwind on a box, any dim 2/3)
// or unstructured DG0/DG1 on a Mesh. Host tables in FP64; the operator uploads
// what its kernels read.
// ---------------------------------------------------------------------------
enum DiscKind { DISC_FD = 0, DISC_DG = 1 };

struct Disc {
  int kind = DISC_FD;
  int dim = 0, order = 0, nb = 1, n_cells = 0, n_angles = 0;
  Quadrature quad;
  BCTable bcs;

  // structured
  int n[3] = {1, 1, 1};
  double h[3] = {1, 1, 1}, len[3] = {1, 1, 1};
  int face_type[3][2] = {{0, 0}, {0, 0}, {0, 0}};      // UBT_BC_* per axis side
  BCFace face_bc[3][2];

  // unstructured: nf faces per cell, slot k = cell * nf + f
  int nf = 0;
  std::vector<double> volume, centroid;                  // cells, 3*cells
  std::vector<int> face_nb;                              // >=0 neighbour, -1 vacuum, -2-axis reflect
  std::vector<int> face_label, face_axis;
  std::vector<double> face_nA, face_centroid, face_area; // 3*slots, 3*slots, slots
  std::vector<double> face_distance;                     // 2*slots (own, neighbour)
  std::vector<double> face_inflow;                       // slots: inflow / W on a vacuum face in window, else 0
  // DG1 tables (reference layouts)
  std::vector<double> face_own, face_up;                 // slots * nb * nb, / (V A_f)
  std::vector<double> basis_grad;                        // cells * nb * 3
  std::vector<double> face_basis_value;                  // slots * nb
  std::vector<double> face_nb_grad;                      // slots * nb * 3
  std::vector<double> basis_beta;                        // cells * 9 (rows = grad of phi_1..dim)

  int64_t n_nodes() const { return (int64_t)n_cells * nb; }
  int64_t n_rows() const { return n_nodes() * n_angles; }
};

std::shared_ptr<Disc> make_structured(const Quadrature &q, int dim, const int n[3],
                                      const double len[3], const BCTable &bcs);
std::shared_ptr<Disc> make_unstructured(const Quadrature &q, const Mesh &m, int order,
                                        const BCTable &bcs);
// The same discretisation with the per-cell and per-face geometry (moments,
// normals, distances; at order 1 the basis and the face matrices) computed by
// HIP kernels on the current device and copied into the host tables
std::shared_ptr<Disc> make_unstructured_device(const Quadrature &q, const Mesh &m, int order, const BCTable &bcs,
                                               hipStream_t s);
// the construction's shared stages: the discretisation before any geometry,
// a geometry error by code (1 volume, 2 area, 3 normal, 4 a centroid outside
// its face, 5 a neighbour's centroid outside it, 6 a degenerate second
// moment) of cell c / face f, the boundary classification (neighbours,
// vacuum with its inflow window, reflective axes) from the face geometry
std::shared_ptr<Disc> dg_begin(const Quadrature &q, const Mesh &m, int order, const BCTable &bcs);
[[noreturn]] void dg_geometry_error(int what, int c, int f);
void dg_boundaries(Disc &G, const Mesh &m, const BCTable &bcs, const Quadrature &q);

}  // namespace ubt
