// The transport sweep: psi = L^{-1} r for L = streaming + removal (no
// scatter), the streaming stage of the preconditioner. The reference hands
// this stage to PFLARE's PCAIR, a reduction-based AMG that approximates
// exactly this inverse; on the GPU the inverse itself is affordable, in
// wavefronts of cells whose upwind neighbours are all done.
//
// * Unstructured DG0/DG1: per angle, the cells are levelled by a topological
//   sort of the upwind graph (host, once per discretisation). A level step
//   processes the cells of that level for one direction: gather the upwind
//   traces, build the cell's n_basis x n_basis block (the same block the
//   element block inverse solves) and solve it. An edge left
//   on a cycle (possible on a general mesh, not on a Delaunay or box mesh) is
//   lagged: its inflow is taken as zero, so the stage stays a preconditioner.
//   Which faces read an upwind trace is the schedule's own per-item mask, not
//   a sign test in the kernel: a face tangent to the direction has flux 0 on
//   the host but +-1 ulp on the device (FMA, FP32 rounding), and a device-side
//   test would then read a neighbour the schedule has not ordered first.
// * Structured FD: the levels are the diagonal hyperplanes i' + j' + k' = l of
//   each octant, no table needed.
// Reflective couplings are lagged the same way. Directions never read each
// other's traces, so one launch sweeps all of them: a workgroup per direction
// walks its levels with a workgroup barrier between them.
//
// The adjoint sweep solves L^T y = r with L^T = V P L P V^{-1} (P reverses
// every angle, V the cell volumes): the forward schedule of angle b solves for
// u = L^{-1} P V^{-1} r, whose angle b lives in y's angle -b, so the ADJ
// kernels read and write the reversed angle's data and scale by the volumes.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "batch.hpp"
#include "sweep.hpp"

namespace ubt {

namespace {

// one (cell, direction) item of a level: gather, build the block, solve
template <class T, int DIM, int NB, int NF, bool ADJ>
__device__ __forceinline__ void sweep_item(const T *__restrict__ r, T *__restrict__ psi, const OpArgs<T> &g,
                                           int64_t c, int b, unsigned char up) {
  const int a = ADJ ? g.neg[b] : b;        // b: the direction swept; a: the angle whose data it carries
  const int A = g.A;
  T om[3] = {g.omega[3 * b], g.omega[3 * b + 1], DIM == 3 ? g.omega[3 * b + 2] : T(0)};
  T B[NB][NB], v[NB];
  const int64_t base = c * NB * A + a;
  const int64_t gc = g.gcls[c];
  const T inv_vc = ADJ ? T(1) / g.vol[gc] : T(1);
#pragma unroll
  for (int i = 0; i < NB; i++) {
    v[i] = inv_vc * r[base + (int64_t)i * A];
#pragma unroll
    for (int j = 0; j < NB; j++) B[i][j] = T(0);
  }
#pragma unroll
  for (int f = 0; f < NF; f++) {
    const int64_t k = c * NF + f, kg = gc * NF + f;
    T s = T(0);
#pragma unroll
    for (int d = 0; d < DIM; d++) s += om[d] * g.face_s[3 * kg + d];
    if ((up >> f) & 1) {
      const int n = g.face_nb[k];
      const int64_t ub = (int64_t)n * NB * A + a;
      const T inv_vn = ADJ ? T(1) / g.vol[g.gcls[n]] : T(1);
      if (NB == 1) {
        v[0] -= s * inv_vn * psi[ub];
      } else {
        T pu[NB];
#pragma unroll
        for (int j = 0; j < NB; j++) pu[j] = inv_vn * psi[ub + (int64_t)j * A];
        const T *F = g.face_up + kg * NB * NB;
#pragma unroll
        for (int i = 0; i < NB; i++) {
          T t = T(0);
#pragma unroll
          for (int j = 0; j < NB; j++) t += F[i * NB + j] * pu[j];
          v[i] -= s * t;
        }
      }
    } else if (s > T(0)) {  // outflow; an inflow face without the bit is vacuum / reflective / lagged
      if (NB == 1) {
        B[0][0] += s;
      } else {
        const T *F = g.face_own + kg * NB * NB;
#pragma unroll
        for (int i = 0; i < NB; i++)
#pragma unroll
          for (int j = 0; j < NB; j++) B[i][j] += s * F[i * NB + j];
      }
    }
  }
  if (NB > 1) {
#pragma unroll
    for (int i = 1; i < NB; i++) {
      const T *gr = g.grad + (gc * NB + i) * 3;
      T od = T(0);
#pragma unroll
      for (int d = 0; d < DIM; d++) od += om[d] * gr[d];
      B[i][0] -= od;
    }
  }
  const T sig = g.sigt[c];
#pragma unroll
  for (int i = 0; i < NB; i++) B[i][i] += sig;
  // Gauss-Jordan with partial pivoting (the blocks are non-singular by setup)
#pragma unroll
  for (int k = 0; k < NB; k++) {
    int piv = k;
#pragma unroll
    for (int i = k + 1; i < NB; i++)
      if (fabs(B[i][k]) > fabs(B[piv][k])) piv = i;
    if (piv != k) {
#pragma unroll
      for (int j = 0; j < NB; j++) {
        T t = B[k][j];
        B[k][j] = B[piv][j];
        B[piv][j] = t;
      }
      T t = v[k];
      v[k] = v[piv];
      v[piv] = t;
    }
    const T inv = T(1) / B[k][k];
#pragma unroll
    for (int j = 0; j < NB; j++) B[k][j] *= inv;
    v[k] *= inv;
#pragma unroll
    for (int i = 0; i < NB; i++) {
      if (i == k) continue;
      const T f = B[i][k];
#pragma unroll
      for (int j = 0; j < NB; j++) B[i][j] -= f * B[k][j];
      v[i] -= f * v[k];
    }
  }
  const T vc = ADJ ? g.vol[gc] : T(1);
#pragma unroll
  for (int i = 0; i < NB; i++) psi[base + (int64_t)i * A] = vc * v[i];
}

// Persistent per-direction sweep: the sweep of one direction never reads
// another direction's traces, so a workgroup owns one direction and walks all
// of its levels in one launch, a workgroup barrier between levels (the items
// of level l read traces the same workgroup wrote at earlier levels, through
// the CU's own L1/L2, so the barrier is the only synchronisation needed).
// Every direction runs concurrently on its own CU: the launch count of a
// sweep drops from one per level to one.
template <class T, int DIM, int NB, int NF, bool ADJ>
__global__ void __launch_bounds__(1024) sweep_dg_dir(const T *__restrict__ r, T *__restrict__ psi, OpArgs<T> g,
                                                     const int *__restrict__ cell, const unsigned char *__restrict__ up,
                                                     const int64_t *__restrict__ lvl, const int *__restrict__ base) {
  const int b = blockIdx.x;
  const int l0 = base[b], l1 = base[b + 1] - 1;  // levels of direction b: lvl[l0 .. l1]
  for (int l = l0; l < l1; l++) {
    for (int64_t i = lvl[l] + threadIdx.x; i < lvl[l + 1]; i += blockDim.x)
      sweep_item<T, DIM, NB, NF, ADJ>(r, psi, g, cell[i], b, up[i]);
    __syncthreads();
  }
}

// the same over the directions of many problems: workgroup w owns direction
// wdir[w] of the problem at work-list position wprob[w]
template <class T, int DIM, int NB, int NF>
__global__ void __launch_bounds__(256) sweep_dg_dir_batch(const T *__restrict__ r, T *__restrict__ psi,
                                                          const BatchItem<T> *__restrict__ items,
                                                          const int *__restrict__ wprob, const int *__restrict__ wdir,
                                                          const int *__restrict__ cell, const unsigned char *__restrict__ up,
                                                          const int64_t *__restrict__ lvl, const int *__restrict__ base) {
  const int w = blockIdx.x;
  const BatchItem<T> &bi = items[wprob[w]];
  const int b = wdir[w];
  const int l0 = base[w], l1 = base[w + 1] - 1;
  for (int l = l0; l < l1; l++) {
    for (int64_t i = lvl[l] + threadIdx.x; i < lvl[l + 1]; i += blockDim.x)
      sweep_item<T, DIM, NB, NF, false>(r + bi.off, psi + bi.off, bi.g, cell[i], b, up[i]);
    __syncthreads();
  }
}

// Structured, persistent per direction: workgroup = direction, levels in one
// launch with a workgroup barrier between them (see sweep_dg_dir)
template <class T, int DIM, bool ADJ>
__global__ void __launch_bounds__(1024) sweep_fd_dir(const T *__restrict__ r, T *__restrict__ psi, OpArgs<T> g,
                                                     int n_levels) {
  const int b = blockIdx.x;
  const int a = ADJ ? g.neg[b] : b;
  const int nx = g.n[0], ny = g.n[1], nz = DIM == 3 ? g.n[2] : 1;
  const int64_t per = (int64_t)ny * nz;
  const int A = g.A;
  const T cs[3] = {g.omega[3 * b], g.omega[3 * b + 1], DIM == 3 ? g.omega[3 * b + 2] : T(0)};
  const int64_t stride[3] = {1, nx, (int64_t)nx * ny};
  for (int level = 0; level < n_levels; level++) {
    for (int64_t t = threadIdx.x; t < per; t += blockDim.x) {
      const int jp = (int)(t % ny), kp = (int)(t / ny);
      const int ip = level - jp - kp;
      if (ip < 0 || ip >= nx) continue;
      int idx[3];
      idx[0] = cs[0] >= T(0) ? ip : nx - 1 - ip;
      idx[1] = cs[1] >= T(0) ? jp : ny - 1 - jp;
      idx[2] = DIM == 3 ? (cs[2] >= T(0) ? kp : nz - 1 - kp) : 0;
      const int64_t c = idx[0] + (int64_t)nx * (idx[1] + (int64_t)ny * idx[2]);
      T num = r[c * A + a], den = g.sigt[c];
#pragma unroll
      for (int d = 0; d < DIM; d++) {
        const T cd = fabs(cs[d]) * g.invh[d];
        den += cd;
        if (cs[d] == T(0)) continue;
        const int upi = cs[d] > T(0) ? idx[d] - 1 : idx[d] + 1;
        const int nd = d == 0 ? nx : (d == 1 ? ny : nz);
        if (upi >= 0 && upi < nd) num += cd * psi[(c + (cs[d] > T(0) ? -stride[d] : stride[d])) * A + a];
      }
      psi[c * A + a] = num / den;
    }
    __syncthreads();
  }
}

}  // namespace

// ---------------------------------------------------------------------------
// Host: the DG level schedule
// ---------------------------------------------------------------------------
void Sweep::create(Operator *o, bool keep_host) {
  op = o;
  const Disc &D = *o->disc;
  if (D.kind == DISC_FD) {
    n_levels = 0;
    for (int d = 0; d < D.dim; d++) n_levels += D.n[d] - 1;
    n_levels += 1;
    return;
  }
  const int nc = D.n_cells, nf = D.nf, A = D.n_angles;
  std::vector<int> level((size_t)nc), indeg((size_t)nc);
  std::vector<std::vector<int>> by_level;
  // per angle: indegree = interior inflow faces; Kahn by levels; leftovers
  // on a cycle get their remaining inflow edges lagged
  std::vector<std::vector<int64_t>> buckets;  // level -> item ids in the per-angle order
  std::vector<int> item_cell, item_angle;
  std::vector<unsigned char> item_up;  // per item: the faces that read an upwind trace
  std::vector<int64_t> item_level;
  item_cell.reserve((size_t)nc * A);
  item_angle.reserve((size_t)nc * A);
  item_up.reserve((size_t)nc * A);
  item_level.reserve((size_t)nc * A);
  std::vector<int> frontier, next;
  std::vector<char> done((size_t)nc);
  // rev[k]: the neighbour's slot for the same face, so a dependency is always
  // judged by the downstream cell's own inflow test (the two sides' normals
  // are negatives of each other only up to rounding)
  std::vector<int64_t> rev((size_t)nc * nf, -1);
  for (int64_t k = 0; k < (int64_t)nc * nf; k++) {
    const int n = D.face_nb[k];
    if (n < 0) continue;
    const int c = (int)(k / nf);
    for (int f = 0; f < nf; f++)
      if (D.face_nb[(int64_t)n * nf + f] == c) rev[k] = (int64_t)n * nf + f;
    UBT_CHECK(rev[k] >= 0, UBT_ERR_PLIB, "face %lld of cell %d has no matching face in its neighbour", (long long)k, c);
  }
  std::vector<unsigned char> lagmask((size_t)nc);
  n_lagged = 0;
  int max_level = 0;
  for (int a = 0; a < A; a++) {
    const double *om = &D.quad.omega[3 * a];
    auto flux = [&](int64_t k) { return om[0] * D.face_nA[3 * k] + om[1] * D.face_nA[3 * k + 1] + om[2] * D.face_nA[3 * k + 2]; };
    frontier.clear();
    for (int c = 0; c < nc; c++) {
      int in = 0;
      for (int f = 0; f < nf; f++) {
        const int64_t k = (int64_t)c * nf + f;
        if (D.face_nb[k] >= 0 && flux(k) < 0.0) in++;
      }
      indeg[c] = in;
      done[c] = 0;
      lagmask[c] = 0;
      if (in == 0) frontier.push_back(c);
    }
    int lvl = 0, n_done = 0;
    while (n_done < nc) {
      if (frontier.empty()) {
        // a cycle: lag every remaining inflow edge of the not-yet-levelled
        // cell with the fewest, and release it
        int best = -1;
        for (int c = 0; c < nc; c++)
          if (!done[c] && (best < 0 || indeg[c] < indeg[best])) best = c;
        for (int f = 0; f < nf; f++) {
          const int64_t k = (int64_t)best * nf + f;
          const int n = D.face_nb[k];
          if (n >= 0 && flux(k) < 0.0 && !done[n]) {
            lagmask[best] |= (unsigned char)(1u << f);
            n_lagged++;
          }
        }
        indeg[best] = 0;
        frontier.push_back(best);
      }
      next.clear();
      for (int c : frontier) {
        done[c] = 1;
        level[c] = lvl;
        n_done++;
        unsigned char upw = 0;
        for (int f = 0; f < nf; f++) {
          const int64_t k = (int64_t)c * nf + f;
          if (D.face_nb[k] >= 0 && flux(k) < 0.0 && !((lagmask[c] >> f) & 1)) upw |= (unsigned char)(1u << f);
        }
        item_cell.push_back(c);
        item_angle.push_back(a);
        item_up.push_back(upw);
        item_level.push_back(lvl);
      }
      for (int c : frontier)
        for (int f = 0; f < nf; f++) {
          const int64_t k = (int64_t)c * nf + f;
          const int n = D.face_nb[k];
          // n counted c as upwind iff its own test on the face says inflow
          if (n >= 0 && !done[n] && flux(rev[k]) < 0.0 && !((lagmask[n] >> (rev[k] - (int64_t)n * nf)) & 1)) {
            if (--indeg[n] == 0) next.push_back(n);
          }
        }
      std::sort(next.begin(), next.end());
      frontier.swap(next);
      lvl++;
    }
    max_level = std::max(max_level, lvl);
  }
  // per direction: the items in generation order (direction, then level),
  // and each direction's level boundaries
  const int64_t n_items = (int64_t)item_cell.size();
  std::vector<int64_t> alvl;
  std::vector<int> abase(A + 1, 0);
  int64_t widest = 1;
  for (int64_t i = 0, a = 0; a < A; a++) {
    abase[a] = (int)alvl.size();
    int64_t start = i;
    for (; i < n_items && item_angle[i] == a; i++)
      if (i == start || item_level[i] != item_level[i - 1]) {
        if (!alvl.empty() && (int)alvl.size() > abase[a]) widest = std::max(widest, i - alvl.back());
        alvl.push_back(i);
      }
    if ((int)alvl.size() > abase[a]) widest = std::max(widest, i - alvl.back());
    alvl.push_back(i);
  }
  abase[A] = (int)alvl.size();
  n_levels = max_level;
  threads = (int)std::min<int64_t>(1024, std::max<int64_t>(64, (widest + 63) / 64 * 64));
  cells.upload(item_cell);
  upwind.upload(item_up);
  lvl.upload(alvl);
  base.upload(abase);
  if (keep_host) {
    h_cells.swap(item_cell);
    h_up.swap(item_up);
    h_lvl.swap(alvl);
    h_base.swap(abase);
  }
}

void Sweep::apply(int prec, const void *r, void *psi, hipStream_t s, bool adj) const {
  const Operator &o = *op;
  UBT_CHECK(o.xs_set, UBT_ERR_ARG_WRONGSTATE, "set the cross sections before sweeping");
  UBT_CHECK(r && psi, UBT_ERR_ARG_NULL, "r and psi are required");
  UBT_CHECK(r != psi, UBT_ERR_ARG_IDN, "the sweep needs distinct r and psi");
  const Disc &D = *o.disc;
  if (D.kind == DISC_FD) {
    const int64_t per = (int64_t)D.n[1] * (D.dim == 3 ? D.n[2] : 1);
    const int th = (int)std::min<int64_t>(1024, std::max<int64_t>(64, (per + 63) / 64 * 64));
    const unsigned grid = (unsigned)D.n_angles;
#define SWF(T, DIM)                                                                                              \
  if (adj)                                                                                                       \
    hipLaunchKernelGGL((sweep_fd_dir<T, DIM, true>), dim3(grid), dim3(th), 0, s, (const T *)r, (T *)psi,          \
                       o.args<T>(), n_levels);                                                                   \
  else                                                                                                           \
    hipLaunchKernelGGL((sweep_fd_dir<T, DIM, false>), dim3(grid), dim3(th), 0, s, (const T *)r, (T *)psi,         \
                       o.args<T>(), n_levels);
    if (prec == UBT_FP64) {
      if (D.dim == 2) {
        SWF(double, 2)
      } else {
        SWF(double, 3)
      }
    } else {
      if (D.dim == 2) {
        SWF(float, 2)
      } else {
        SWF(float, 3)
      }
    }
#undef SWF
    UBT_HIP(hipGetLastError());
    return;
  }
  const int *ci = cells.as<int>(), *bs = base.as<int>();
  const unsigned char *up = upwind.as<unsigned char>();
  const int64_t *lv = lvl.as<int64_t>();
  const unsigned grid = (unsigned)D.n_angles;
  bool launched = false;
#define SW(T, DIM, NB, NF)                                                                                         \
  if (!launched && D.dim == DIM && D.nb == NB && D.nf == NF) {                                                     \
    if (adj)                                                                                                       \
      hipLaunchKernelGGL((sweep_dg_dir<T, DIM, NB, NF, true>), dim3(grid), dim3(threads), 0, s, (const T *)r,      \
                         (T *)psi, o.args<T>(), ci, up, lv, bs);                                                   \
    else                                                                                                           \
      hipLaunchKernelGGL((sweep_dg_dir<T, DIM, NB, NF, false>), dim3(grid), dim3(threads), 0, s, (const T *)r,     \
                         (T *)psi, o.args<T>(), ci, up, lv, bs);                                                   \
    launched = true;                                                                                               \
  }
#define SW_ALL(T) SW(T, 2, 1, 3) SW(T, 2, 1, 4) SW(T, 3, 1, 4) SW(T, 3, 1, 6) SW(T, 2, 3, 3) SW(T, 2, 3, 4) SW(T, 3, 4, 4) SW(T, 3, 4, 6)
  if (prec == UBT_FP64) {
    SW_ALL(double)
  } else {
    SW_ALL(float)
  }
#undef SW_ALL
#undef SW
  UBT_CHECK(launched, UBT_ERR_SUP, "no sweep kernel for dim %d, %d basis, %d faces", D.dim, D.nb, D.nf);
  UBT_HIP(hipGetLastError());
}

template <class T> void batch_sweep_launch(const Batch &b, const T *r, T *psi, hipStream_t s) {
  const int pi = sizeof(T) == 8 ? 0 : 1;
  const BatchSweep &w = b.sweep;
  if (w.n_wg == 0) return;
  const BatchItem<T> *items = b.items[pi][0].as<BatchItem<T>>();
  bool launched = false;
#define SWB(DIM, NB, NF)                                                                                            \
  if (!launched && b.dim == DIM && b.nb == NB && b.nf == NF) {                                                      \
    hipLaunchKernelGGL((sweep_dg_dir_batch<T, DIM, NB, NF>), dim3(w.n_wg), dim3(256), 0, s, r, psi, items,          \
                       w.prob.as<int>(), w.dir.as<int>(), w.cells.as<int>(), w.upwind.as<unsigned char>(),          \
                       w.lvl.as<int64_t>(), w.base.as<int>());                                                      \
    launched = true;                                                                                                \
  }
  SWB(2, 1, 3) SWB(2, 1, 4) SWB(3, 1, 4) SWB(3, 1, 6) SWB(2, 3, 3) SWB(2, 3, 4) SWB(3, 4, 4) SWB(3, 4, 6)
#undef SWB
  UBT_CHECK(launched, UBT_ERR_SUP, "no batched sweep for dim %d, %d basis, %d faces", b.dim, b.nb, b.nf);
  UBT_HIP(hipGetLastError());
}

template void batch_sweep_launch<double>(const Batch &, const double *, double *, hipStream_t);
template void batch_sweep_launch<float>(const Batch &, const float *, float *, hipStream_t);

}  // namespace ubt
