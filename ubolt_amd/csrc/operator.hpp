// The device-resident transport operator: the tables its kernels read, per
// precision, and the host methods the C ABI calls.
#pragma once

#include "ubt_internal.hpp"

namespace ubt {

// Kernel arguments, by value. One struct serves every discretisation; the
// kernels read only their own fields.
template <class T> struct OpArgs {
  int dim, nb, nf, A, n_cells, cpb, red_lanes;
  T inv_sum_w;
  const T *omega;      // 3A
  const T *w;          // A
  const int *reflect;  // 3A (-1 past dim)
  const int *neg;      // A: the ordinate with every cosine negated
  const T *sigt;       // current group, per cell
  const T *sigs;       // current group's within-group scatter, per cell
  // structured
  int n[3];
  T invh[3];
  int face_type[3][2];
  // unstructured. The geometry tables are indexed by geometry class: cells
  // whose geometry agrees to 1e-13 of each quantity's scale share one record
  // (every interior cell of a uniform box is one class), so the applies read
  // a few KB of geometry instead of a record per cell. Neighbours stay per cell.
  const int *face_nb;  // cells * nf (per cell)
  const int *gcls;     // cells: geometry class
  int gcls_identity;   // every cell is its own class, numbered as the cell: no lookup
  const T *face_s;     // classes * nf * 3: nA (DG1) or nA / V (DG0)
  const T *face_own;   // classes * nf * nb * nb
  const T *face_up;
  const T *grad;       // classes * nb * 3
  const T *vol;        // classes: volume (the adjoint's neighbour ratio V_c / V_n)
  // a list of cells for the kernels that run on one (the box-tile kernel's
  // boundary cells), n_list of them
  const int *cells;
  int n_list;
};

// Per precision copies of the geometry and cross sections
struct PrecTables {
  DevBuf omega, w, face_s, face_own, face_up, grad, vol;
  DevBuf wunit;       // [Apad][16]: column 0 = 1, the unweighted angle sum (adjoint restriction)
  DevBuf wext;        // [Apad][16]: w_a {1, Omega_x, Omega_y, Omega_z}, the moments' B operand
  DevBuf sigt, sigs;  // G*cells, G*G*cells
};

// apply kernels the tuner chooses from: 0 tile, 1 wavefront per cell,
// 2 and 3 the multi-cell tile with 2 and 4 cells per thread, 4-6 the pencil
// kernel of a structured box marching 8, 16 or 32 cells per workgroup, 7 the
// four-angle tile, 8 and 9 the box tile with 2 and 4 rows per thread, 10 the
// strip march on a uniform DG1 box, 11 and 12 the z-march on a 3D FD box
// (4 or 8 cells per lane column)
constexpr int kApplyVariants = 13;
constexpr int kPencilMaxClasses = 64;

// The box-tile kernels' view of a uniform DG1 box: a tile is W columns along
// x (thread = column, angle) by K rows along the last axis held in
// registers, at one y index in 3D. Tiles are numbered in bands of kBoxBand
// row blocks: within a band the blocks of one x (and y) position are
// consecutive, then x, then y.
// The cells on the box's boundary (and any of another class) are left to a
// second phase of the same launch that runs the tile kernel's body on their
// list.
constexpr int kBoxThreads = 256;  // strip-march workgroup: W columns x A angles fill at most this
// box-tile workgroup: 1024 threads for 2D FP32 (four times the columns per
// tile of 256, a quarter of the halo columns per cell; on the 64-problem
// ensemble 256 -> 512 -> 1024 ran 511 -> 501 -> 481 us), 256 for FP64 (512
// ran 3% slower there) and for 3D (whose tile holds more registers than
// 1024-thread workgroups allow)
template <class T, int DIM> constexpr int kBoxTileThreads = sizeof(T) == 4 && DIM == 2 ? 1024 : 256;
inline int box_tile_threads(size_t ts, int dim) { return ts == 4 && dim == 2 ? 1024 : 256; }
constexpr int kBoxBand = 2;  // row blocks per band of the box-tile walk (1, 2, 4 measured: 2 and 4 equal, 2-3% ahead of 1)

struct BoxArgs {
  int n[3];         // cells per axis
  int fmap[3][2];   // local face of (axis, side)
  int cls;          // the geometry class of the interior cells
  int W;            // columns per tile
  int tiles[3];     // tiles along x, y (1 in 2D) and the register axis
  int lanes;        // lanes per angular moment in the tile's segmented sum
  int edge_tiles;   // tiles of the boundary-cell list (OpArgs cpb cells each)
  // the strip march: a unit is W columns (at one y index in 3D) by a band of
  // R cells along the last axis; units[] counts strips along x, y indices
  // (1 in 2D) and bands
  int R;
  int units[3];
  // per cell 1 where the box formula holds (interior class, or a boundary
  // cell whose box-boundary faces are all vacuum); null when every cell does
  const unsigned char *ok;
};

// the strip-march unit's band length along the last axis
constexpr int kStripRows = 64;
// blocks of the scalar-flux change reduction (Operator::set_scalar_flux)
constexpr int kMeasureBlocks = 512;

struct Operator {
  std::shared_ptr<Disc> disc;
  int n_groups = 1, group = 0, device = 0;
  // apply kernel variant per precision [fp64, fp32]: 0 tile (LDS-staged, cpb
  // cells per workgroup), 1 wavefront per cell; chosen by tune()
  int variant[2] = {0, 0};
  float tuned_us[2][kApplyVariants] = {};
  bool xs_set = false;
  uint64_t xs_gen = 0;        // bumped whenever the cross sections change
  PrecTables t64, t32;
  DevBuf reflect, face_nb, neg;
  DevBuf gcls;                // int per cell: geometry class (DG)
  int n_geo_classes = 0;
  bool gcls_identity = false;  // class c is cell c for every cell (a mesh with no repeated geometry)
  std::vector<int> geo_rep;   // representative cell of each class
  // a DG mesh that is a lexicographic box (the pencil kernel's precondition)
  bool box = false;
  int box_n[3] = {1, 1, 1};
  int box_fmap[3][2] = {{-1, -1}, {-1, -1}, {-1, -1}};
  // a DG1 box whose interior cells share one geometry class, whose face
  // normals are exactly axis-aligned and whose ordinates have no zero cosine
  // (the box-tile kernel's precondition): that class
  int box_cls = -1;
  DevBuf box_edge;            // int: the cells the box-tile kernel leaves to its list phase
  int n_box_edge = 0;
  DevBuf box_ok;              // uint8 per cell: the box formula holds (see BoxArgs::ok)
  DevBuf face_nA_raw;         // double, 3 per slot (rhs inflow, DSA normals)
  DevBuf volume_d, face_area_d, face_dist_d;  // double: cells, slots, 2*slots (DSA assembly)
  DevBuf inv_volume;          // double, cells (rhs inflow)
  DevBuf face_inflow;         // double, slots
  DevBuf face_basis_value;    // double, slots * nb
  DevBuf source;              // double, G*cells
  DevBuf sigs_d;              // double, G*G*cells (transfer sources)
  DevBuf phi_cache;           // double, G*nodes
  std::vector<char> phi_set;
  // the first group anything upscatters into: the smallest g_to of a nonzero
  // sigma_s[g_from][g_to], g_from > g_to, in any cell (n_groups without upscatter)
  int g_up = 1;
  DevBuf phi_tmp;             // double, nodes: a measured scalar flux before it overwrites the cache
  DevBuf measure_part;        // double, 2 per reduction block
  // the outer iteration's moments store: per group the moments {phi, J} of
  // its last angular flux (G * nodes * (1 + dim)), and the P1 expansion's
  // T^{-1} Omega_a per angle (3 per angle), T = sum_a w_a Omega_a Omega_a^T
  DevBuf mom_store, p1;
  DevBuf norm_sums;           // double, 2: norm2's result
  DevBuf flag;                // int scratch for setup checks

  template <class T> OpArgs<T> args() const;
  void mult(int prec, const void *x, void *y, bool scatter, hipStream_t s) const;
  // y = A^T x (scatter) or L^T x: the adjoint, reversed-direction kernel
  void mult_transpose(int prec, const void *x, void *y, bool scatter, hipStream_t s) const;
  void diagonal(int prec, void *d, hipStream_t s) const;
  void moments(int prec, const void *psi, void *phi, int n_moments, hipStream_t s) const;
  // phi (FP64, one moment) from psi of either precision
  // unit: the unweighted sum over angles (the transposed prolongation)
  void moments_f64(int prec, const void *psi, double *phi, hipStream_t s, bool unit = false) const;
  // sums (device, 2 doubles) or null: adds ||phi_new - phi_old||^2 and
  // ||phi_new||^2 over the group's nodes (the outer iteration's change)
  void set_scalar_flux(int g, int prec, const void *psi, hipStream_t s, double *sums = nullptr);
  // a zero scalar flux cached for group g
  void cache_zero_flux(int g, hipStream_t s);
  // the moments store: {phi, J} of group g from its angular flux (FP64), or
  // zero; psi = phi_g / W + sum_d (T^{-1} Omega_a)_d J_g,d, the P1 expansion
  // of the cached scalar flux and the stored current (reference
  // GroupTransfer::enable_current / expand_p1, src/multigroupk.kokkos.cxx:373-498)
  void store_moments(int g, const double *psi, hipStream_t s);
  void zero_moments(int g, hipStream_t s);
  void expand_p1(int g, double *psi, hipStream_t s);
  // ||x||_2 over the rows (FP64; synchronises s)
  double norm2(const double *x, hipStream_t s);
  void scatter_source(int g_from, int g_to, int prec, void *b, hipStream_t s) const;
  // all_groups: transfers from every group but g (a thermal group under
  // upscatter), else from the groups below g
  void rhs(int g, int prec, void *b, hipStream_t s, bool all_groups = false) const;
  // time every applicable apply variant on scratch vectors and keep the fastest
  void tune();
};

std::unique_ptr<Operator> make_operator(std::shared_ptr<Disc> disc, int n_groups);
void operator_set_xs(Operator &op, const double *sigma_t, const double *sigma_s, const double *source);

// counts o's singular element blocks into *cnt (zeroed by the caller) on s
void blockinv_check_launch(const Operator &o, int *cnt, hipStream_t s);

struct BlockInverse {
  Operator *op = nullptr;
  int group_ready = -1;
  void setup(hipStream_t s);
  // y = D^{-1} x, or the blocks of A^T (transpose) - the forward blocks at -Omega
  void apply(int prec, const void *x, void *y, hipStream_t s, bool transpose = false) const;
};

}  // namespace ubt
