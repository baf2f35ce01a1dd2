/*
 * ubt.h - the C ABI of libubolt_amd, an MI355X-native implementation of
 * UBOLT's matrix-free linear-Boltzmann transport operator and its
 * preconditioner applies.
 *
 * Every entry point is extern "C", takes plain pointers and sizes, and
 * returns an int status: UBT_OK (0) or one of the ubt_status classes below,
 * which mirror the PETSc error classes the reference raises, so a PETSc
 * adapter (MatShell / PCShell callbacks) can return them unchanged. The
 * message of the last failure on the calling thread is read back with
 * ubt_last_error(). No C++ exception crosses this boundary.
 *
 * Vectors are caller-owned DEVICE pointers in the reference row order of one
 * energy group, row = (cell * n_basis + basis) * n_angles + angle (angle
 * fastest; reference include/ubolt/phase_space.hpp:20-50). That order is also
 * the native HBM layout: a (cell, basis) node owns a contiguous run of
 * n_angles values, which is what lets the upwind gathers of the unstructured
 * kernels read whole contiguous runs, so no pack/unpack step exists. A
 * vector's element type is double for UBT_FP64 and float for UBT_FP32.
 *
 * Every device call is stream-ordered on the caller's stream (a hipStream_t
 * passed as void*, NULL for the default stream), allocates nothing and
 * synchronises nothing, so it can be captured into a hipGraph. A handle is not
 * thread-safe: calls on one handle from several host threads must be ordered
 * by the caller.
 *
 * Reference: stevendargaville/UBOLT. Each declaration cites the interface it
 * replaces (paths relative to the reference root).
 */
#ifndef UBT_H
#define UBT_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Status classes - the numeric values of the corresponding PETSC_ERR_* codes */
typedef enum {
  UBT_OK = 0,
  UBT_ERR_MEM = 55,             /* PETSC_ERR_MEM */
  UBT_ERR_SUP = 56,             /* PETSC_ERR_SUP: unsupported order / shape / option */
  UBT_ERR_ARG_SIZ = 60,         /* PETSC_ERR_ARG_SIZ */
  UBT_ERR_ARG_IDN = 61,         /* PETSC_ERR_ARG_IDN: two arguments must not alias */
  UBT_ERR_ARG_WRONG = 62,       /* PETSC_ERR_ARG_WRONG */
  UBT_ERR_ARG_OUTOFRANGE = 63,  /* PETSC_ERR_ARG_OUTOFRANGE */
  UBT_ERR_MAT_LU_ZRPVT = 71,    /* PETSC_ERR_MAT_LU_ZRPVT: a singular element block */
  UBT_ERR_ARG_WRONGSTATE = 73,  /* PETSC_ERR_ARG_WRONGSTATE */
  UBT_ERR_ARG_INCOMP = 75,      /* PETSC_ERR_ARG_INCOMP */
  UBT_ERR_PLIB = 77,            /* PETSC_ERR_PLIB: an internal invariant failed */
  UBT_ERR_ARG_NULL = 85,        /* PETSC_ERR_ARG_NULL */
  UBT_ERR_NOT_CONVERGED = 91,   /* PETSC_ERR_NOT_CONVERGED */
  UBT_ERR_GPU = 97              /* PETSC_ERR_GPU: a HIP runtime call failed */
} ubt_status;

typedef enum { UBT_FP64 = 0, UBT_FP32 = 1 } ubt_prec;

/* Boundary families (reference include/ubolt/bc_spec.hpp:22). Only the
 * ghost-flux vacuum treatment (the reference default) exists here. */
typedef enum { UBT_BC_VACUUM = 0, UBT_BC_REFLECT = 1 } ubt_bc_type;

/* One labelled face's boundary condition (reference BCFace,
 * include/ubolt/bc_spec.hpp:68-119). inflow is the isotropic angle-integrated
 * strength, shared over the ordinates as inflow / sum_weights; window holds
 * n_window_pairs tangential [lo, hi] pairs in ascending axis order. Labels are
 * PETSc's box "Face Sets" ids: 2D 1 bottom, 2 right, 3 top, 4 left; 3D 1
 * bottom (z-min), 2 top, 3 front (y-min), 4 back, 5 right (x-max), 6 left. */
typedef struct {
  int label;
  int type;
  double inflow;
  int n_window_pairs;
  double window[4];
} ubt_bc;

typedef struct ubt_quad_s ubt_quad;
typedef struct ubt_mesh_s ubt_mesh;
typedef struct ubt_disc_s ubt_disc;
typedef struct ubt_op_s ubt_op;
typedef struct ubt_blockinv_s ubt_blockinv;
typedef struct ubt_dsa_s ubt_dsa;
typedef struct ubt_sweep_s ubt_sweep;
typedef struct ubt_solver_s ubt_solver;
typedef struct ubt_batch_s ubt_batch;
typedef struct ubt_batch_solver_s ubt_batch_solver;

/* Solver options (ubt_solver_create). The preconditioner is the reference's
 * multiplicative composite: stage 0 the element block inverse (pc_blocks),
 * stage 1 the streaming stage (pc_sweep: the transport sweep), stage 2 DSA
 * (pc_dsa, when a DSA handle is given); later stages correct the residual with
 * the full operator when DSA is on and use_amat is set, else with L. Where the
 * stages correct with L and the sweep is L's exact inverse S (a DG sweep with
 * no lagged edge and no reflective face), stages 0 and 1 together are the
 * sweep alone - y + S (r - L y) = S r for any stage-0 output y - and stage 0
 * is not run. */
typedef struct {
  double rtol;        /* relative true residual ||b - A x|| / ||b||, e.g. 1e-8 */
  int max_it;         /* Krylov iterations in total */
  int restart;        /* GMRES restart length, 1..31 */
  int mixed;          /* 0: FP64 FGMRES; 1: FP32 inner FGMRES inside FP64 refinement */
  double inner_rtol;  /* inner tolerance in mixed mode, e.g. 1e-4 */
  int max_outer;      /* refinement steps in mixed mode */
  int pc_blocks, pc_sweep, pc_dsa, use_amat;
  int gs_passes;      /* Gram-Schmidt passes per Arnoldi step: 1 classical (PETSc's
                       * default, KSPGMRESClassicalGramSchmidtOrthogonalization with
                       * refinement never), 2 classical twice; 0 means 2 */
} ubt_solver_options;

/* ------------------------------------------------------------------ errors */

/* Copies the last failure's message (NUL-terminated, truncated to len) and
 * returns its status; UBT_OK and an empty string when nothing failed.
 * Replaces PetscCheck/PetscError reporting (AGENTS.md:38). */
int ubt_last_error(char *buf, size_t len);

/* Number of HIP devices visible (0 on a host without a GPU). */
int ubt_device_count(int *count);

/* -------------------------------------------------------------- quadrature */

/* The SN set of a dimension and order: 1D Gauss-Legendre (any even order), 2D
 * level-symmetric (xi > 0 half, weights doubled) and 3D level-symmetric (all
 * octants), S2-S12. Replaces SNQuadrature{,2D,3D}::create
 * (src/sn_quadraturek.kokkos.cxx:181,286,344). */
int ubt_quad_create(int dim, int sn_order, ubt_quad **quad);
int ubt_quad_size(const ubt_quad *quad, int *n_angles, double *sum_weights);
/* omega: 3 per ordinate (zero past dim); w: one per ordinate; reflect: 3 per
 * ordinate, reflect[3*a+d] the ordinate with cosine d negated (-1 past dim).
 * Replaces AngularQuadrature::{omega_host, w_d, reflect_host}
 * (include/ubolt/sn_quadrature.hpp:21-68). Any output may be NULL. */
int ubt_quad_get(const ubt_quad *quad, double *omega, double *w, int *reflect);
int ubt_quad_destroy(ubt_quad *quad);

/* The principal octant of the level-symmetric set of an even order: its level
 * cosines mu (n_levels of them), and per point its 1-based level triple and
 * its weight as a fraction of the whole sphere, points in lexicographic triple
 * order. Buffers hold at least 6 levels and 21 points. Replaces the generated
 * table of src/sn_lqn_table.hpp (the data, regenerated from its defining
 * moment conditions). */
int ubt_lqn_octant(int sn_order, int *n_levels, double *mu, int *n_points, int *triples,
                   double *weights);

/* -------------------------------------------------------------------- mesh */

/* An unstructured mesh of one cell type: 2D triangles (3 vertices) or quads
 * (4, cyclic), 3D tetrahedra (4) or hexahedra (8, VTK order: 0-3 one face
 * cyclic, 4-7 the opposite face). coords holds dim values per vertex. Faces
 * are matched between cells. Replaces the DMPlex mesh the reference builds in
 * UnstructuredDG::create_mesh (src/unstructured_dgk.kokkos.cxx:12). */
int ubt_mesh_create(int dim, int n_vertices, const double *coords, int n_cells,
                    int verts_per_cell, const int *cell_verts, ubt_mesh **mesh);
/* Label the boundary faces lying on the bounding box's sides with PETSc's box
 * "Face Sets" ids (see ubt_bc). Faces elsewhere keep their label. */
int ubt_mesh_label_box(ubt_mesh *mesh);
/* Label boundary faces explicitly (a Gmsh file's physical tags): each face is
 * verts_per_face vertex indices, any order. */
int ubt_mesh_set_boundary_labels(ubt_mesh *mesh, int n_faces, int verts_per_face,
                                 const int *face_verts, const int *labels);
int ubt_mesh_size(const ubt_mesh *mesh, int *n_cells, int *faces_per_cell, int *n_boundary_faces);
int ubt_mesh_destroy(ubt_mesh *mesh);

/* ---------------------------------------------------------- discretisation */

/* Cell-centred upwind finite differences on the box [0, lengths]^dim with
 * n_cells[d] cells per axis (dim 2 or 3), cells x fastest. bcs: one entry per
 * labelled face, unlisted faces cold vacuum. Replaces StructuredFD2D/3D::create
 * and the row classification of src/structured_fd_commonk.hpp:213-397. */
int ubt_disc_create_structured(const ubt_quad *quad, int dim, const int *n_cells,
                               const double *lengths, const ubt_bc *bcs, int n_bcs,
                               ubt_disc **disc);
/* Discontinuous Galerkin of order 0 or 1 on a mesh (modal orthonormal P1
 * basis, n_basis = dim + 1 at order 1). Replaces UnstructuredDG::create and
 * build_dg1_geometry (src/unstructured_dgk.kokkos.cxx:224-821). */
int ubt_disc_create_dg(const ubt_quad *quad, const ubt_mesh *mesh, int order, const ubt_bc *bcs,
                       int n_bcs, ubt_disc **disc);
/* As ubt_disc_create_dg, with the per-cell and per-face geometry - fan
 * moments, normals, distances and, at order 1, the orthonormal basis and the
 * face matrices - computed by kernels on the current device (on stream), as
 * the reference builds it in device kernels (build_dg1_geometry,
 * src/unstructured_dgk.kokkos.cxx:29-218,706-821); the topology and the
 * boundary classification stay on the host. Same tables, same errors. */
int ubt_disc_create_dg_device(const ubt_quad *quad, const ubt_mesh *mesh, int order, const ubt_bc *bcs,
                              int n_bcs, void *stream, ubt_disc **disc);
/* Sizes of the phase space (reference PhaseSpace, include/ubolt/phase_space.hpp:34-87). */
int ubt_disc_size(const ubt_disc *disc, int *dim, int *n_cells, int *n_basis, int *n_angles,
                  int *faces_per_cell);
/* Copy one host geometry table out, by name, in the reference's layouts:
 * "volume", "centroid" (3/cell), "face_nb" (int: >=0 neighbour cell, -1 vacuum,
 * -2-axis reflective), "face_label" (int), "face_nA" (3/slot), "face_area",
 * "face_centroid" (3/slot), "face_distance" (2/slot), "face_inflow" (per slot:
 * inflow / sum_weights if vacuum and in window), "face_own", "face_up"
 * (nb*nb/slot), "basis_grad" (nb*3/cell), "face_basis_value" (nb/slot),
 * "face_nb_grad" (nb*3/slot), "h" (3, structured), "n" (int 3, structured),
 * "bc_face" (structured: per axis side type, inflow/W, window pairs, 4 window
 * values: 7 doubles x 6). A slot is cell * faces_per_cell + local face.
 * *nbytes is set to the table's size; out may be NULL to query it. */
int ubt_disc_get(const ubt_disc *disc, const char *name, void *out, size_t *nbytes);
int ubt_disc_destroy(ubt_disc *disc);

/* ---------------------------------------------------------------- operator */

/* The transport operator on the current GPU: uploads the discretisation's
 * device tables for n_groups energy groups. Replaces TransportOperator::create
 * + add_term (src/transport_operatork.kokkos.cxx:84-120) with the streaming,
 * removal and within-group scattering terms fused into one apply. */
int ubt_op_create(const ubt_disc *disc, int n_groups, ubt_op **op);
/* Host tables, copied to the device once: sigma_t[g][cell],
 * sigma_s[g_from][g_to][cell], source[g][cell] (isotropic, angle-integrated;
 * may be NULL). Replaces GroupXSections::set_from_materials and
 * MaterialSourceTable (src/multigroupk.kokkos.cxx:46-169,
 * src/material_speck.kokkos.cxx:164-213). */
int ubt_op_set_xs(ubt_op *op, const double *sigma_t, const double *sigma_s, const double *source);
/* Point the operator at group g: no allocation, no transfer. Replaces
 * TransportOperator::set_group + assemble (src/transport_operatork.kokkos.cxx:165-206). */
int ubt_op_set_group(ubt_op *op, int g);
int ubt_op_size(const ubt_op *op, int64_t *n_rows, int64_t *n_nodes, int *n_groups);
/* Number of distinct DG geometry records the applies read (cells whose face
 * normals, face matrices, basis gradients and volume agree to 1e-13 of each
 * quantity's scale share one; a uniform box has a handful). 0 for grids. No
 * reference counterpart: the reference stores a record per cell. */
int ubt_op_geometry_classes(const ubt_op *op, int *n_classes);
/* y = A x, A = streaming + removal - within-group scatter of the current
 * group. x and y must not alias. Replaces ShellMatMultApply
 * (src/transport_operatork.kokkos.cxx:49-61), the MATOP_MULT of the MatShell. */
int ubt_op_mult(ubt_op *op, int prec, const void *x, void *y, void *stream);
/* y = L x, the operator without the scatter. Replaces UnscatteredMatMultApply
 * (src/transport_operatork.kokkos.cxx:66-78). */
int ubt_op_mult_unscattered(ubt_op *op, int prec, const void *x, void *y, void *stream);
/* y = A^T x, the adjoint operator: one fused kernel that streams in the
 * reversed direction (A^T = V P A P V^{-1} for the upwind part, P reversing
 * every angle, V the cell volumes) and transposes the scatter (unweighted
 * angle sum, w_a on the output). No reference counterpart: the reference has
 * no adjoint; this is the drop-in MATOP_MULT_TRANSPOSE a PETSc MatShell would
 * register next to ShellMatMultApply (src/transport_operatork.kokkos.cxx:49-61).
 * Exact for vacuum faces and for reflective faces normal to an axis. */
int ubt_op_mult_transpose(ubt_op *op, int prec, const void *x, void *y, void *stream);
/* y = L^T x (no scatter), the transpose of UnscatteredMatMultApply
 * (src/transport_operatork.kokkos.cxx:66-78). */
int ubt_op_mult_unscattered_transpose(ubt_op *op, int prec, const void *x, void *y, void *stream);
/* d = diag(A) composed from the terms (streaming outflow + sigma_t). Replaces
 * TransportOperator::diagonal (src/transport_operatork.kokkos.cxx:240-256). */
int ubt_op_diagonal(ubt_op *op, int prec, void *d, void *stream);
/* The apply kernel variant per precision: 0 the LDS-staged tile kernel (whole
 * cells, every angle, per workgroup), 1 the wavefront-per-cell kernel
 * (per-cell data through the scalar cache, no barrier), 2 and 3 the
 * multi-cell tile kernel (DG: 2 or 4 cells per thread, all loads before the
 * first barrier; structured FD: a run of 4 or 8 cells along x per lane, the x
 * upwind trace from registers), 4, 5 and 6 the pencil kernel (DG on a structured box:
 * workgroups march 8, 16 or 32 cells along the last axis with the upwind
 * traces in registers and LDS), 7 the four-angle tile kernel (a thread owns
 * four consecutive angles of a cell: 16-byte loads, per-cell work shared;
 * DG with A % 4 == 0), 8 and 9 the box-tile kernel (a uniform DG1 box: the
 * interior class folded per angle into three small matrices, 2 or 4 rows
 * per thread in registers), 10 the strip march (a uniform DG1 box: a
 * workgroup walks a strip of cells along the last axis with the neighbouring
 * rows in registers), 11 and 12 the z-march (3D FD: a workgroup walks
 * chunks of 32 planes along z with the z upwind plane in registers and the
 * x neighbours through LDS, 4 or 8 cells per lane column); -1 re-runs the
 * autotuner, which times every applicable
 * kernel on scratch vectors when the cross sections are first set and keeps
 * the fastest. A variant that does not apply falls back to 0. No reference
 * counterpart (the reference has one Kokkos SpMV); exposed so the tests can
 * check every kernel. */
int ubt_op_set_variant(ubt_op *op, int prec, int variant);
/* variant[2] (FP64, FP32) in use and us[26] the tuned times (FP64 variants
 * 0-12, then FP32 variants 0-12; 0 where not applicable). */
int ubt_op_tuning(const ubt_op *op, int *variant, float *us);
int ubt_op_destroy(ubt_op *op);

/* ------------------------------------------------------ moments and sources */

/* phi[node * n_moments + m] = sum_a w_a m_a psi[node, a], m_a = 1 (m = 0) and
 * Omega_a[d] (m = 1 + d), n_moments 1 or 1 + dim, on the matrix cores.
 * Replaces UboltAngularIntegral (src/sn_quadraturek.kokkos.cxx:400-427). */
int ubt_moments(ubt_op *op, int prec, const void *psi, void *phi, int n_moments, void *stream);
/* Cache group g's scalar flux from its angular flux psi (FP64 cache). Replaces
 * GroupTransfer::set_scalar_flux (src/multigroupk.kokkos.cxx:266-299). */
int ubt_set_scalar_flux(ubt_op *op, int g, int prec, const void *psi, void *stream);
/* b += sigma_s(g_from -> g_to) phi_{g_from} / sum_weights on every row.
 * Replaces GroupTransfer::add_transfer (src/multigroupk.kokkos.cxx:306-364). */
int ubt_scatter_source(ubt_op *op, int g_from, int g_to, int prec, void *b, void *stream);
/* phi (FP64 device memory, n_nodes) = the scalar flux cached for group g (the
 * cache ubt_set_scalar_flux fills; the group sweep leaves every group's
 * there). Reads the reference's GroupTransfer::phi_ cache
 * (src/multigroupk.kokkos.cxx:266-299), as its flux output does. */
int ubt_op_scalar_flux(const ubt_op *op, int g, double *phi, void *stream);
/* b = ghost inflow + isotropic source (basis 0) + the transfers from every
 * group g' < g, in one launch: the reference test driver's rhs for group g
 * (tests/transportk.kokkos.cxx:1058-1065; UboltFillInflow
 * src/discretisationk.kokkos.cxx:187-206; add_isotropic
 * src/material_speck.kokkos.cxx:164-187). */
int ubt_rhs(ubt_op *op, int g, int prec, void *b, void *stream);

/* ---------------------------------------------------------- block inverse */

/* The element blocks of the current group's operator, (cell, angle) n_basis x
 * n_basis blocks (streaming own-cell couplings + removal; reflective
 * couplings excluded). Replaces ElementBlockInverse
 * (include/ubolt/block_inverse.hpp:35-81). */
int ubt_blockinv_create(ubt_op *op, ubt_blockinv **bi);
/* Invert the current group's blocks by Gauss-Jordan with partial pivoting;
 * UBT_ERR_MAT_LU_ZRPVT if any block is singular. Replaces
 * ElementBlockInverse::setup (src/block_inversek.kokkos.cxx:83-150,249-277). */
int ubt_blockinv_setup(ubt_blockinv *bi, void *stream);
/* y = D^{-1} x block by block; x and y must not alias. Replaces
 * ElementBlockInverse::apply (src/block_inversek.kokkos.cxx:206-222,281-298),
 * the RemovalPCApply body (src/transport_solverk.kokkos.cxx:84-94). */
int ubt_blockinv_apply(ubt_blockinv *bi, int prec, const void *x, void *y, void *stream);
/* y = D^{-T} x: the element blocks of A^T (the forward blocks at -Omega), the
 * PCApplyTranspose of the RemovalPCApply shell (src/transport_solverk.kokkos.cxx:84-94). */
int ubt_blockinv_apply_transpose(ubt_blockinv *bi, int prec, const void *x, void *y, void *stream);
int ubt_blockinv_destroy(ubt_blockinv *bi);

/* --------------------------------------------------------------------- DSA */

/* The DSA preconditioner stage on the operator's discretisation: structured
 * star, DG0 two-point flux or DG1 MIP diffusion operator, assembled on the
 * device per group, with an FP64 block-Jacobi CG inner solve. Replaces
 * DSAPrecon::create (src/dsak.kokkos.cxx:83-149, include/ubolt/dsa.hpp:63-149).
 * The handle is the caller's, like the reference's DSAPrecon. */
int ubt_dsa_create(ubt_op *op, ubt_dsa **dsa);
/* -dsa_consistent_d[_power], -dsa_mip_penalty and the inner solve's
 * iteration cap and relative tolerance (the reference's inner KSP, one GAMG
 * application, src/dsak.kokkos.cxx:197-212). inner_check_every > 0 reads the
 * inner solve's converged flag back every that many iterations and stops
 * early (a host synchronisation); 0 never synchronises (capturable). */
int ubt_dsa_set_options(ubt_dsa *dsa, int consistent_d, double consistent_power, double mip_penalty,
                        int inner_max_it, double inner_rtol, int inner_check_every);
/* The void policy, the reference's -dsa_void_bridge, -dsa_void_sigma_t and
 * -dsa_void_d (src/dsak.kokkos.cxx:181-183; defaults 1, 0, 0): a cell is void
 * where sigma_t <= void_sigma_t; bridge != 0 bridges voids where the operator
 * allows it, 0 masks them; void_d > 0 fixes a bridged void's D. Takes effect
 * at the next ubt_dsa_set_group. */
int ubt_dsa_set_void_policy(ubt_dsa *dsa, int bridge, double void_sigma_t, double void_d);
/* Assemble group g's diffusion operator from the operator's cross sections:
 * D = 1/(3 sigma_t), sigma_a = sigma_t - sigma_s(g, g). Replaces
 * DSAPrecon::set_group / assemble (src/dsak.kokkos.cxx:216-345) with its void
 * policy: void cells are bridged - D = void_d or 1/(3 (max(sigma_t, 0) +
 * 1/L)), L = 4 V / S the voids' mean chord (V their volume, S their faces onto
 * a non-void cell or a vacuum face), a DG1 face touching a void with the
 * harmonic face D on both sides - when bridging is on, something is not void,
 * a vacuum face or some absorption holds the operator and S > 0; otherwise
 * masked - identity rows (V times identity blocks on the plex) coupled to
 * nothing, a face into one a vacuum face, nothing restricted or corrected
 * there. A group void everywhere is masked whole (ubt_dsa_apply then gives a
 * zero correction, ubt_dsa_matvec and the solves UBT_ERR_ARG_WRONGSTATE). */
int ubt_dsa_set_group(ubt_dsa *dsa, int g, void *stream);
/* y = P D^{-1} R x: the 0th moment (times V on the plex), the inner solve, and
 * the isotropic correction / sum_weights on every row. Replaces
 * DSAPrecon::apply (src/dsak.kokkos.cxx:350-391), the DSAPCApply body. */
int ubt_dsa_apply(ubt_dsa *dsa, int prec, const void *x, void *y, void *stream);
/* y = R^T W D^{-1} P^T x, the transpose of ubt_dsa_apply (the diffusion
 * operators are symmetric): the PCApplyTranspose of the DSAPCApply shell
 * (src/transport_solverk.kokkos.cxx:102-112). */
int ubt_dsa_apply_transpose(ubt_dsa *dsa, int prec, const void *x, void *y, void *stream);
/* y = D_diff x on node vectors, FP64 (the diffusion matvec). */
int ubt_dsa_matvec(ubt_dsa *dsa, const double *x, double *y, void *stream);
/* x = D_diff^{-1} b by the inner solve (x starts at zero). */
int ubt_dsa_solve(ubt_dsa *dsa, const double *b, double *x, void *stream);
/* x = D_diff^{-1} b exactly as ubt_dsa_solve, allocating nothing: the inner
 * CG's workspace (five node vectors, block partial sums, scalars) is a
 * file-scope device array of 2^21 nodes (80 MiB per GPU, held until the
 * process ends), so the call makes no hipMalloc/hipFree or other allocating
 * call. No counterpart in the reference: it is for callers that cannot
 * allocate (code under HIP graph capture - set the inner check interval to 0
 * so no host read is made -, tight serving loops, harnesses that forbid
 * allocation). An operator of more than 2^21 diffusion nodes is refused with
 * UBT_ERR_ARG_SIZ. Every call on one GPU shares that array: calls that
 * overlap on one GPU (two streams, two threads) race. */
int ubt_dsa_solve_noalloc(ubt_dsa *dsa, const double *b, double *x, void *stream);
/* Inner iterations the last solve took. */
int ubt_dsa_info(ubt_dsa *dsa, int *inner_iterations, int *n_nodes);
int ubt_dsa_destroy(ubt_dsa *dsa);

/* ------------------------------------------------------------------- sweep */

/* The transport sweep psi = L^{-1} r (streaming + removal, reflective and
 * cycle-closing couplings lagged): the GPU streaming stage standing in for
 * the reference's PCAIR stage (src/transport_solverk.kokkos.cxx:238), which
 * approximates the same inverse. Level schedule built once. */
int ubt_sweep_create(ubt_op *op, ubt_sweep **sw);
/* As ubt_sweep_create; lds_traces = 0 keeps every upwind trace read in HBM
 * (a diagnostic switch, no counterpart in the reference; by default a DG
 * sweep reads upwind traces from LDS slots wherever they fit 64 KB). */
int ubt_sweep_create_opts(ubt_op *op, int lds_traces, ubt_sweep **sw);
int ubt_sweep_info(const ubt_sweep *sw, int *n_levels, int64_t *n_lagged);
/* LDS trace slots per direction workgroup (0: the traces are read from HBM). */
int ubt_sweep_trace_slots(const ubt_sweep *sw, int *slots);
int ubt_sweep_apply(ubt_sweep *sw, int prec, const void *r, void *psi, void *stream);
/* psi = L^{-T} r: the adjoint sweep, the same level schedule run on the
 * reversed angles' data with the volume scaling of L^T = V P L P V^{-1}. */
int ubt_sweep_apply_transpose(ubt_sweep *sw, int prec, const void *r, void *psi, void *stream);
int ubt_sweep_destroy(ubt_sweep *sw);

/* ------------------------------------------------------------------ solver */

/* Right-preconditioned flexible GMRES on the operator with the composite
 * preconditioner; dsa may be NULL. Replaces TransportSolver::create
 * (src/transport_solverk.kokkos.cxx:196-318). options NULL = defaults. */
int ubt_solver_create(ubt_op *op, ubt_dsa *dsa, const ubt_solver_options *options, ubt_solver **solver);
/* Re-check the element blocks after the group changed. Replaces
 * TransportSolver::refresh (src/transport_solverk.kokkos.cxx:324-331). */
int ubt_solver_refresh(ubt_solver *solver, void *stream);
/* x = A^{-1} b (FP64 device vectors, zero initial guess). Replaces
 * TransportSolver::solve (src/transport_solverk.kokkos.cxx:335-343).
 * Returns UBT_OK also when not converged: read ubt_solver_result. */
int ubt_solver_solve(ubt_solver *solver, const double *b, double *x, void *stream);
/* x = A^{-T} b, the adjoint solve: the same FGMRES / refinement with
 * ubt_op_mult_transpose and every preconditioner stage transposed (KSPSolveTranspose
 * of the reference's KSP, src/transport_solverk.kokkos.cxx:335-343). */
int ubt_solver_solve_transpose(ubt_solver *solver, const double *b, double *x, void *stream);
/* y = M^{-1} r, the composite preconditioner alone (prec of r and y). */
int ubt_solver_apply_pc(ubt_solver *solver, int prec, const void *r, void *y, void *stream);
/* iterations: Krylov iterations; outer: refinement steps (1 in FP64 mode);
 * rel_residual: the final FP64 true relative residual; reason: 1 converged,
 * -1 iteration limit (the sign convention of KSPConvergedReason). history
 * (may be NULL) receives up to max_history per-iteration relative residuals. */
int ubt_solver_result(const ubt_solver *solver, int *iterations, int *outer, double *rel_residual, int *reason,
                      double *history, int max_history, int *n_history);
int ubt_solver_destroy(ubt_solver *solver);

/* ------------------------------------------------------------- group sweep
 * The multigroup driver of the reference's test driver
 * (tests/transportk.kokkos.cxx:925-1208): the Gauss-Seidel sweep over the
 * groups - per group the rhs (inflow, source, the transfers out of the
 * cached scalar flux of the groups above), the solve, the scalar flux cached
 * for the groups below - and, with upscatter, the outer iteration over the
 * thermal block g_up .. G-1 (g_up: the first group anything upscatters into,
 * from the cross sections) until the estimated error change * rho /
 * (1 - rho) of its scalar flux is below outer_rtol; a thermal solve after the
 * first sweep starts from its last psi at max(inner rtol, inner_gamma x the
 * last change), and the iteration stops only on a sweep at the inner rtol.
 * Options replace -outer_rtol, -outer_max_it, -outer_inner_gamma, under
 * upscatter -ksp_rtol, and -outer_store / -outer_recon
 * (tests/transportk.kokkos.cxx:606-680). */
typedef struct {
  double outer_rtol;   /* e.g. 1e-8 */
  int max_outer;       /* sweeps, the first included, e.g. 1000 */
  double inner_gamma;  /* [0, 1), e.g. 0.01; 0: every solve at the inner rtol */
  double inner_rtol;   /* every solve's rtol under upscatter; 0: 1e-10. Without
                        * upscatter the solver's own rtol */
  int store_moments;   /* 0: psi holds every group's angular flux (the angular
                        * store); 1: the moments store - psi is ONE vector every
                        * group's solve reuses, a thermal group keeps its scalar
                        * flux and current, and after the first sweep its solve
                        * starts from their P1 expansion (single problems only) */
  int recon_steps;     /* moments store: at most this many safeguarded
                        * Richardson steps on L_g refine the expansion, to the
                        * solve's own rtol (0: the expansion alone; e.g. 10) */
} ubt_group_options;

typedef struct {
  int g_up;            /* first thermal group (n_groups: no upscatter, one sweep) */
  int sweeps;          /* the first included */
  int status;          /* 1 converged; -1 a solve did not converge (the sweep
                        * stopped at that group); -2 max_outer sweeps ran out */
  double change;       /* the last sweep's ||phi_new - phi_old|| / ||phi_new|| over the thermal block */
  double rho;          /* its ratio to the sweep before's (0 before the third sweep) */
  double err_est;      /* change * rho / (1 - rho); infinite with no rho below 1 */
  int64_t iterations;  /* Krylov iterations over every solve */
  int64_t recon_steps; /* the moments store's Richardson steps */
} ubt_group_result;

/* Every group of the solver's operator: psi (FP64 device memory) receives
 * n_groups angular fluxes of n_rows each, group g at g * n_rows (the moments
 * store: one vector of n_rows, the last group solved). Leaves the
 * operator (and the DSA handle) at the last group solved, every group's
 * scalar flux cached (ubt_set_scalar_flux). options NULL = defaults. Returns
 * UBT_OK also when not converged: read ubt_solver_groups_result. Replaces
 * the group loop of tests/transportk.kokkos.cxx:1019-1208. */
int ubt_solver_solve_groups(ubt_solver *solver, const ubt_group_options *options, double *psi, void *stream);
/* the last group sweep; per group (arrays of n_groups, each may be NULL):
 * Krylov iterations summed over the sweeps, the last solve's relative
 * residual and reason */
int ubt_solver_groups_result(const ubt_solver *solver, ubt_group_result *result, int *iterations,
                             double *rel_residual, int *reason);

/* ---------------------------------------------------------------- batches
 * Many independent problems driven as one (no reference counterpart: the
 * reference solves one problem per process; this is how a caller runs a
 * parameter sweep or an ensemble). A batch references n operators (which
 * must outlive it) of one DG type (dimension, basis, faces per cell); they
 * may differ in mesh, quadrature and group count. Batch vectors are the
 * problems' group vectors end to end (ubt_batch_offsets). */
int ubt_batch_create(ubt_op *const *ops, int n, ubt_batch **batch);
int ubt_batch_size(const ubt_batch *batch, int64_t *n_rows, int *n_problems);
/* off[n + 1]: first row of every problem in a batch vector */
int ubt_batch_offsets(const ubt_batch *batch, int64_t *off);
/* Switch every problem to group g; problems with fewer groups sit out (their
 * span of a batch vector is left alone). Rebuilds the work lists. */
int ubt_batch_set_group(ubt_batch *batch, int g);
/* y = A x for every active problem in one persistent launch. */
int ubt_batch_mult(ubt_batch *batch, int prec, const void *x, void *y, void *stream);
/* body[2] (FP64, FP32: 0 tile, 1 four-angle tile, 2 box tile, 3 strip march,
 * the last two when every problem is a uniform DG1 box) and persist[2]
 * (workgroups per CU) chosen by the batch autotuner, us[8] the best time per
 * body (FP64 bodies 0-3, then FP32 bodies 0-3; 0 if it does not apply). */
int ubt_batch_tuning(const ubt_batch *batch, int *body, int *persist, float *us);
/* Force the apply body of one precision (0 to 3 as above; an error where it
 * does not apply to every active problem). No reference counterpart; exposed
 * so the tests can check every body. */
int ubt_batch_set_body(ubt_batch *batch, int prec, int body);
/* The batched sweep's form: box != 0 (the default) sweeps a batch whose
 * active problems are all uniform all-vacuum 2D DG1 boxes (at most 256 cells
 * along x) without per-item records - the blocks formed on the fly from
 * sigma_t and per-angle matrices, a workgroup per problem, quadrant and angle
 * group marching anti-diagonals -; 0 uses the record sweep on every batch.
 * No reference counterpart; a solver refreshes before its next solve. */
int ubt_batch_set_sweep(ubt_batch *batch, int box);
/* box[2] (FP64, FP32): 1 where the record-free box sweep runs; exact[2]
 * (optional): 1 where the sweep inverts every active problem's L exactly (no
 * lagged edge, no reflective face). There the batch solver's element-block
 * stage is left out: blocks then the sweep on the L-residual equal the sweep
 * alone, y + S (r - L y) = S r. */
int ubt_batch_sweep_info(const ubt_batch *batch, int *box, int *exact);
/* the right-hand side of every active problem (ubt_rhs), zero elsewhere */
int ubt_batch_rhs(ubt_batch *batch, int prec, void *b, void *stream);
int ubt_batch_set_scalar_flux(ubt_batch *batch, int prec, const void *psi, void *stream);
/* the element block inverse / the transport sweep of every active problem */
int ubt_batch_blocks_apply(ubt_batch *batch, int prec, const void *x, void *y, void *stream);
int ubt_batch_sweep_apply(ubt_batch *batch, int prec, const void *r, void *psi, void *stream);
int ubt_batch_destroy(ubt_batch *batch);
/* FGMRES over every active problem at once with one Hessenberg per problem
 * on the device; a restart cycle is captured as a HIP graph and replayed.
 * Preconditioner: element blocks then the transport sweep (pc_dsa is
 * ignored). mixed: FP32 inner solves inside FP64 refinement, per problem. */
int ubt_batch_solver_create(ubt_batch *batch, const ubt_solver_options *options, ubt_batch_solver **solver);
/* after ubt_batch_set_group: block check and per-problem reductions */
int ubt_batch_solver_refresh(ubt_batch_solver *solver, void *stream);
/* x = A^{-1} b per problem (FP64 batch vectors, zero initial guess);
 * synchronises the stream */
int ubt_batch_solver_solve(ubt_batch_solver *solver, const double *b, double *x, void *stream);
/* per problem (arrays of n): Krylov iterations, final FP64 true relative
 * residual, reason (1 converged, -1 limit, 0 inactive); outer refinement
 * steps and restart cycles in total */
int ubt_batch_solver_result(const ubt_batch_solver *solver, int *iterations, double *rel_residual, int *reason,
                            int *outer, int *cycles);
/* 1 (default): replay the restart cycles as captured HIP graphs (captured on
 * the first solve after a refresh); 0: launch every kernel from the host */
int ubt_batch_solver_set_graph(ubt_batch_solver *solver, int on);
/* Every group of every problem (ubt_solver_solve_groups per problem, the
 * problems' solves of one group batched, each problem with its own outer
 * test; a problem whose iteration has ended sits out the later sweeps): psi
 * receives G_max batch vectors (G_max the most groups of any problem), group
 * g at g * n_rows; a problem with fewer groups leaves its span of the later
 * vectors zero. Leaves the batch at the last group swept. No reference
 * counterpart (the reference sweeps one problem per process). */
int ubt_batch_solver_solve_groups(ubt_batch_solver *solver, const ubt_group_options *options, double *psi,
                                  void *stream);
/* per problem: result[n]; iterations, rel_residual, reason [n * G_max],
 * problem p's group g at p * G_max + g (each may be NULL) */
int ubt_batch_solver_groups_result(const ubt_batch_solver *solver, ubt_group_result *result, int *iterations,
                                   double *rel_residual, int *reason);
int ubt_batch_solver_destroy(ubt_batch_solver *solver);

#ifdef __cplusplus
}
#endif

#endif /* UBT_H */
