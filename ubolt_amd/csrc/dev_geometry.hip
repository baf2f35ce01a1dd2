// The DG geometry on the device, as the reference builds it in device
// kernels (src/unstructured_dgk.kokkos.cxx:29-218,706-821): one thread per
// cell for the fan moments and the orthonormal basis, one per (cell, face)
// slot for the face geometry and the face matrices, with the same arithmetic
// as the host construction (geometry.hpp). The mesh topology, the boundary
// classification (neighbours, vacuum inflow windows, reflective axes) and the
// error reports stay on the host; the tables come back into the host
// discretisation, so everything built from it is unchanged.
#include <hip/hip_runtime.h>

#include "geometry.hpp"
#include "ubt_internal.hpp"

namespace ubt {

namespace {

constexpr int kGeoThreads = 128;

__global__ void __launch_bounds__(kGeoThreads)
    k_cell_moments(int dim, int nc, int nf, int vpc, const int *__restrict__ cv, const int *__restrict__ fnv,
                   const int *__restrict__ fv, const double *__restrict__ X, double *__restrict__ vol,
                   double *__restrict__ cen, double *__restrict__ C, int *__restrict__ err) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= nc) return;
  const int64_t k0 = (int64_t)c * nf;
  err[c] = geo::cell_moments(dim, nf, vpc, cv + (int64_t)c * vpc, fnv + k0, fv + 4 * k0, X, vol + c,
                             cen + 3 * (int64_t)c, C + 9 * (int64_t)c)
               ? 0
               : 1;
}

// per slot: the face geometry, then the neighbour's distance (error codes as
// dg_geometry_error numbers them)
__global__ void __launch_bounds__(kGeoThreads)
    k_face_geometry(int dim, int64_t slots, int nf, const int *__restrict__ fnv, const int *__restrict__ fv,
                    const int *__restrict__ fnb, const double *__restrict__ X, const double *__restrict__ cen,
                    double *__restrict__ area, double *__restrict__ xf, double *__restrict__ Cf,
                    double *__restrict__ nA, int *__restrict__ axis, double *__restrict__ dist, int *__restrict__ err) {
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= slots) return;
  const int64_t c = k / nf;
  const int e = geo::face_geometry(dim, fnv[k], fv + 4 * k, X, cen + 3 * c, area + k, xf + 3 * k, Cf + 9 * k,
                                   nA + 3 * k, axis + k, dist + 2 * k);
  int code = e == geo::kFaceOk ? 0 : 1 + e;
  const int nbr = fnb[k];
  dist[2 * k + 1] = 0.0;
  if (code == 0 && nbr >= 0) {
    const double dn = geo::neighbour_distance(dim, nA + 3 * k, area[k], cen + 3 * (int64_t)nbr, xf + 3 * k);
    dist[2 * k + 1] = dn;
    if (!(dn > 0.0)) code = 5;
  }
  err[k] = code;
}

__global__ void __launch_bounds__(kGeoThreads)
    k_basis(int dim, int nc, int nb, const double *__restrict__ C, const double *__restrict__ vol,
            double *__restrict__ beta, double *__restrict__ bgrad, int *__restrict__ err) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= nc) return;
  double *b = beta + 9 * (int64_t)c;
  err[c] = geo::basis_beta(dim, C + 9 * (int64_t)c, vol[c], b) ? 0 : 6;
  for (int i = 0; i < nb; i++)
    for (int d = 0; d < 3; d++) bgrad[((int64_t)c * nb + i) * 3 + d] = geo::basis_grad(b, i, d);
}

// per slot: the face matrices with the cell's basis and the upwind cell's
// (the neighbour, the cell itself across a reflective face, none on vacuum),
// the basis values at the face centroid, the neighbour's basis gradients
__global__ void __launch_bounds__(kGeoThreads)
    k_face_matrices(int64_t slots, int nf, int nb, const int *__restrict__ fnb, const double *__restrict__ beta,
                    const double *__restrict__ cen, const double *__restrict__ xf, const double *__restrict__ Cf,
                    const double *__restrict__ area, const double *__restrict__ vol, double *__restrict__ own,
                    double *__restrict__ up, double *__restrict__ bval, double *__restrict__ nbgrad) {
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= slots) return;
  const int64_t c = k / nf;
  const int nbr = fnb[k];
  const int64_t u = nbr >= 0 ? nbr : (nbr <= -2 ? c : -1);
  for (int i = 0; i < nb * nb; i++) up[k * nb * nb + i] = 0.0;
  for (int j = 0; j < nb; j++)
    for (int d = 0; d < 3; d++)
      nbgrad[(k * nb + j) * 3 + d] = nbr >= 0 ? geo::basis_grad(beta + 9 * (int64_t)nbr, j, d) : 0.0;
  geo::face_matrices(nb, beta + 9 * c, cen + 3 * c, u >= 0 ? beta + 9 * u : nullptr, u >= 0 ? cen + 3 * u : nullptr,
                     xf + 3 * k, Cf + 9 * k, area[k], vol[c], own + k * nb * nb, up + k * nb * nb, bval + k * nb);
}

template <class T> void download(std::vector<T> &h, const DevBuf &d, size_t n, hipStream_t s) {
  h.resize(n);
  if (n) UBT_HIP(hipMemcpyAsync(h.data(), d.ptr, sizeof(T) * n, hipMemcpyDeviceToHost, s));
}

unsigned blocks(int64_t n) { return (unsigned)((n + kGeoThreads - 1) / kGeoThreads); }

}  // namespace

std::shared_ptr<Disc> make_unstructured_device(const Quadrature &q, const Mesh &m, int order, const BCTable &bcs,
                                               hipStream_t s) {
  auto D = dg_begin(q, m, order, bcs);
  Disc &G = *D;
  const int dim = m.dim, nf = G.nf, nc = m.n_cells, nb = G.nb;
  const int64_t slots = (int64_t)nc * nf;
  DevBuf X, cv, fnv, fv, fnb;
  X.upload(m.coords);
  cv.upload(m.cell_verts);
  fnv.upload(m.face_nverts);
  fv.upload(m.face_verts);
  fnb.upload(m.face_neighbour);
  DevBuf vol, cen, C, area, xf, Cf, nA, axis, dist, cerr, ferr;
  vol.alloc(sizeof(double) * nc);
  cen.alloc(sizeof(double) * 3 * nc);
  C.alloc(sizeof(double) * 9 * nc);
  area.alloc(sizeof(double) * slots);
  xf.alloc(sizeof(double) * 3 * slots);
  Cf.alloc(sizeof(double) * 9 * slots);
  nA.alloc(sizeof(double) * 3 * slots);
  axis.alloc(sizeof(int) * slots);
  dist.alloc(sizeof(double) * 2 * slots);
  cerr.alloc(sizeof(int) * nc);
  ferr.alloc(sizeof(int) * slots);
  hipLaunchKernelGGL(k_cell_moments, dim3(blocks(nc)), dim3(kGeoThreads), 0, s, dim, nc, nf, m.verts_per_cell,
                     cv.as<int>(), fnv.as<int>(), fv.as<int>(), X.as<double>(), vol.as<double>(), cen.as<double>(),
                     C.as<double>(), cerr.as<int>());
  hipLaunchKernelGGL(k_face_geometry, dim3(blocks(slots)), dim3(kGeoThreads), 0, s, dim, slots, nf, fnv.as<int>(),
                     fv.as<int>(), fnb.as<int>(), X.as<double>(), cen.as<double>(), area.as<double>(),
                     xf.as<double>(), Cf.as<double>(), nA.as<double>(), axis.as<int>(), dist.as<double>(),
                     ferr.as<int>());
  UBT_HIP(hipGetLastError());
  std::vector<int> ce, fe;
  download(ce, cerr, nc, s);
  download(fe, ferr, slots, s);
  download(G.volume, vol, nc, s);
  download(G.centroid, cen, 3 * (size_t)nc, s);
  download(G.face_area, area, slots, s);
  download(G.face_centroid, xf, 3 * slots, s);
  download(G.face_nA, nA, 3 * slots, s);
  download(G.face_axis, axis, slots, s);
  download(G.face_distance, dist, 2 * slots, s);
  UBT_HIP(hipStreamSynchronize(s));
  // the host construction's error order: cells first, then slots in order
  for (int c = 0; c < nc; c++)
    if (ce[c]) dg_geometry_error(ce[c], c, 0);
  for (int64_t k = 0; k < slots; k++)
    if (fe[k]) dg_geometry_error(fe[k], fe[k] == 5 ? m.face_neighbour[k] : (int)(k / nf), (int)(k % nf));
  dg_boundaries(G, m, bcs, q);
  if (order == 1) {
    fnb.upload(G.face_nb);  // the classified neighbours: reflective faces are their own cell's upwind
    DevBuf beta, bgrad, own, up, bval, nbgrad;
    beta.alloc(sizeof(double) * 9 * nc);
    bgrad.alloc(sizeof(double) * nb * 3 * nc);
    own.alloc(sizeof(double) * nb * nb * slots);
    up.alloc(sizeof(double) * nb * nb * slots);
    bval.alloc(sizeof(double) * nb * slots);
    nbgrad.alloc(sizeof(double) * nb * 3 * slots);
    hipLaunchKernelGGL(k_basis, dim3(blocks(nc)), dim3(kGeoThreads), 0, s, dim, nc, nb, (const double *)C.as<double>(),
                       (const double *)vol.as<double>(), beta.as<double>(), bgrad.as<double>(), cerr.as<int>());
    UBT_HIP(hipGetLastError());
    download(ce, cerr, nc, s);
    UBT_HIP(hipStreamSynchronize(s));
    for (int c = 0; c < nc; c++)
      if (ce[c]) dg_geometry_error(ce[c], c, 0);
    hipLaunchKernelGGL(k_face_matrices, dim3(blocks(slots)), dim3(kGeoThreads), 0, s, slots, nf, nb, fnb.as<int>(),
                       (const double *)beta.as<double>(), (const double *)cen.as<double>(),
                       (const double *)xf.as<double>(), (const double *)Cf.as<double>(),
                       (const double *)area.as<double>(), (const double *)vol.as<double>(), own.as<double>(),
                       up.as<double>(), bval.as<double>(), nbgrad.as<double>());
    UBT_HIP(hipGetLastError());
    download(G.basis_beta, beta, 9 * (size_t)nc, s);
    download(G.basis_grad, bgrad, (size_t)nb * 3 * nc, s);
    download(G.face_own, own, (size_t)nb * nb * slots, s);
    download(G.face_up, up, (size_t)nb * nb * slots, s);
    download(G.face_basis_value, bval, (size_t)nb * slots, s);
    download(G.face_nb_grad, nbgrad, (size_t)nb * 3 * slots, s);
    UBT_HIP(hipStreamSynchronize(s));
  }
  return D;
}

}  // namespace ubt
