// The transport sweep: psi = L^{-1} r for L = streaming + removal (no
// scatter), the streaming stage of the preconditioner. The reference hands
// this stage to PFLARE's PCAIR, a reduction-based AMG that approximates
// exactly this inverse; on the GPU the inverse itself is affordable, in
// wavefronts of cells whose upwind neighbours are all done.
//
// * Unstructured DG0/DG1: per angle, the cells are levelled by a topological
//   sort of the upwind graph (host, once per discretisation). Each (cell,
//   direction) item's n_basis x n_basis block (the same block the element
//   block inverse solves) is inverted once per group into its sweep record,
//   with the upwind couplings; a level step then gathers the upwind traces
//   and applies the record. An edge left
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
// u = L^{-1} P V^{-1} r, whose angle b lives in y's angle -b, so the adjoint
// sweep runs the forward records, reading and writing the reversed angle's
// data and scaling by the volumes.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdlib>
#include <cstring>

#include "batch.hpp"
#include "sweep.hpp"

namespace ubt {

namespace {

// ---------------------------------------------------------------------------
// Sweep records. The block of an item (cell c, direction b) and its upwind
// coupling do not change between sweeps, only r does, so they are folded once
// (per group and cross sections) into
//   psi_c = B^{-1} r_c + sum_k H_k psi_{n_k},   H_k = -B^{-1} s_k F_up,k,
// B inverted by Gauss-Jordan with partial pivoting in FP64, the records kept
// in both precisions. A level of the sweep then costs one round trip for the
// upwind traces plus a few multiply-adds, instead of gathering the class
// tables, building the block and solving it on the critical path.
// ---------------------------------------------------------------------------
// Record layout: blocks of 64 items, each block field-major (field e of item
// i at ((i / 64) F + e) 64 + i % 64, F fields per item), so a wavefront's
// load of one field is 64 consecutive values and a level's records span a
// few contiguous blocks - not F streams a whole array apart, which across
// thousands of concurrent directions exhausts the translation caches.
__host__ __device__ inline int64_t rix(int e, int64_t i, int F) { return ((i >> 6) * F + e) * 64 + (i & 63); }
__host__ __device__ inline int64_t rec_len(int64_t n_items, int F) { return ((n_items + 63) >> 6) * 64 * F; }
// The records of a single problem's pipelined sweep group the fields by 16
// bytes instead (W = 4 floats, 2 doubles): in a block of 64 items, 16-byte
// group q of item i at ((i / 64) ceil(F / W) + q) 64 + i % 64 (in 16-byte
// units). A thread reads its record in 16-byte loads, a few instructions
// rather than one per field, so more levels of records fit the 63 vector
// memory instructions a wavefront keeps in flight, and a wavefront's load is
// still 1 KB contiguous. (The plain form over wide levels keeps one field per
// load: with many wavefronts per level its loads are issue-bound, not
// latency-bound, and there the grouped loads measured slower.)
template <class T> __host__ __device__ constexpr int rec_w() { return 16 / (int)sizeof(T); }
template <class T> __host__ __device__ constexpr int rec_pad(int F) { return (F + rec_w<T>() - 1) / rec_w<T>() * rec_w<T>(); }
template <class T> __host__ __device__ inline int64_t vix(int e, int64_t i, int F) {
  constexpr int W = rec_w<T>();
  return (((i >> 6) * (rec_pad<T>(F) / W) + e / W) * 64 + (i & 63)) * W + e % W;
}
template <class T> inline int64_t vrec_len(int64_t n_items, int F) { return ((n_items + 63) >> 6) * 64 * rec_pad<T>(F); }
// field e of item i (F fields) in either layout
template <class T> __host__ __device__ inline int64_t rat(bool grouped, int e, int64_t i, int F) {
  return grouped ? vix<T>(e, i, F) : rix(e, i, F);
}
// 16 bytes from p (16-byte aligned) into out[0 .. 16 / sizeof(T))
__device__ __forceinline__ void load16(const float *p, float *out) {
  const float4 v = *reinterpret_cast<const float4 *>(p);
  out[0] = v.x;
  out[1] = v.y;
  out[2] = v.z;
  out[3] = v.w;
}
__device__ __forceinline__ void load16(const double *p, double *out) {
  const double2 v = *reinterpret_cast<const double2 *>(p);
  out[0] = v.x;
  out[1] = v.y;
}
__device__ __forceinline__ void load16(const int *p, int *out) {
  const int4 v = *reinterpret_cast<const int4 *>(p);
  out[0] = v.x;
  out[1] = v.y;
  out[2] = v.z;
  out[3] = v.w;
}

// g64 / g32: r64 / r32 in the 16-byte grouped layout (else one field per
// load); rnb always one field per load
template <int DIM, int NB, int NF, int NFU>
__device__ __forceinline__ void sweep_record(const OpArgs<double> &g, int64_t i, int64_t n_items, int64_t c, int b,
                                             unsigned char up, double *__restrict__ r64, float *__restrict__ r32,
                                             int *__restrict__ rnb, bool g64, bool g32) {
  constexpr int NB2 = NB * NB;
  const double om[3] = {g.omega[3 * b], g.omega[3 * b + 1], DIM == 3 ? g.omega[3 * b + 2] : 0.0};
  const int64_t gc = g.gcls[c];
  double B[NB][NB], H[NFU][NB][NB];
  int nbk[NFU];
#pragma unroll
  for (int p = 0; p < NB; p++)
#pragma unroll
    for (int q = 0; q < NB; q++) B[p][q] = 0.0;
#pragma unroll
  for (int k = 0; k < NFU; k++) {
    nbk[k] = -1;
#pragma unroll
    for (int p = 0; p < NB; p++)
#pragma unroll
      for (int q = 0; q < NB; q++) H[k][p][q] = 0.0;
  }
  int k = 0;
#pragma unroll
  for (int f = 0; f < NF; f++) {
    const int64_t kg = gc * NF + f;
    double sf = 0.0;
#pragma unroll
    for (int d = 0; d < DIM; d++) sf += om[d] * g.face_s[3 * kg + d];
    if ((up >> f) & 1) {
      // an inflow slot: s F_up for now, scaled by -B^{-1} below
#pragma unroll
      for (int kk = 0; kk < NFU; kk++)
        if (kk == k) {
          nbk[kk] = g.face_nb[c * NF + f];
#pragma unroll
          for (int p = 0; p < NB; p++)
#pragma unroll
            for (int q = 0; q < NB; q++) H[kk][p][q] = NB == 1 ? sf : sf * g.face_up[kg * NB2 + p * NB + q];
        }
      k++;
    } else if (sf > 0.0) {
#pragma unroll
      for (int p = 0; p < NB; p++)
#pragma unroll
        for (int q = 0; q < NB; q++) B[p][q] += NB == 1 ? sf : sf * g.face_own[kg * NB2 + p * NB + q];
    }
  }
  if (NB > 1) {
#pragma unroll
    for (int p = 1; p < NB; p++) {
      double od = 0.0;
#pragma unroll
      for (int d = 0; d < DIM; d++) od += om[d] * g.grad[(gc * NB + p) * 3 + d];
      B[p][0] -= od;
    }
  }
  const double sig = g.sigt[c];
#pragma unroll
  for (int p = 0; p < NB; p++) B[p][p] += sig;
  // Gauss-Jordan on [B | I]
  double Bi[NB][NB];
#pragma unroll
  for (int p = 0; p < NB; p++)
#pragma unroll
    for (int q = 0; q < NB; q++) Bi[p][q] = p == q ? 1.0 : 0.0;
#pragma unroll
  for (int col = 0; col < NB; col++) {
    int piv = col;
#pragma unroll
    for (int p = col + 1; p < NB; p++)
      if (fabs(B[p][col]) > fabs(B[piv][col])) piv = p;
#pragma unroll
    for (int p = 0; p < NB; p++)
      if (p == piv && p != col) {
#pragma unroll
        for (int q = 0; q < NB; q++) {
          double t = B[col][q];
          B[col][q] = B[p][q];
          B[p][q] = t;
          t = Bi[col][q];
          Bi[col][q] = Bi[p][q];
          Bi[p][q] = t;
        }
      }
    const double inv = 1.0 / B[col][col];
#pragma unroll
    for (int q = 0; q < NB; q++) {
      B[col][q] *= inv;
      Bi[col][q] *= inv;
    }
#pragma unroll
    for (int p = 0; p < NB; p++) {
      if (p == col) continue;
      const double fct = B[p][col];
#pragma unroll
      for (int q = 0; q < NB; q++) {
        B[p][q] -= fct * B[col][q];
        Bi[p][q] -= fct * Bi[col][q];
      }
    }
  }
#pragma unroll
  for (int p = 0; p < NB; p++)
#pragma unroll
    for (int q = 0; q < NB; q++) {
      if (r64) r64[rat<double>(g64, p * NB + q, i, NB2 * (1 + NFU))] = Bi[p][q];
      if (r32) r32[rat<float>(g32, p * NB + q, i, NB2 * (1 + NFU))] = (float)Bi[p][q];
    }
#pragma unroll
  for (int kk = 0; kk < NFU; kk++) {
    rnb[rix(kk, i, NFU)] = nbk[kk];
#pragma unroll
    for (int p = 0; p < NB; p++)
#pragma unroll
      for (int q = 0; q < NB; q++) {
        double h = 0.0;
#pragma unroll
        for (int m = 0; m < NB; m++) h -= Bi[p][m] * H[kk][m][q];
        const int e = NB2 + kk * NB2 + p * NB + q;
        if (r64) r64[rat<double>(g64, e, i, NB2 * (1 + NFU))] = h;
        if (r32) r32[rat<float>(g32, e, i, NB2 * (1 + NFU))] = (float)h;
      }
  }
  // the padding of the grouped layout, read and discarded by its loads
  constexpr int F = NB2 * (1 + NFU);
  for (int e = F; e < rec_pad<double>(F); e++)
    if (r64 && g64) r64[vix<double>(e, i, F)] = 0.0;
  for (int e = F; e < rec_pad<float>(F); e++)
    if (r32 && g32) r32[vix<float>(e, i, F)] = 0.0f;
}

// The records of every item: blockIdx.y = schedule workgroup w (a direction,
// of problem wprob[w] in a batch), its items lvl[base[w]] .. lvl[base[w+1]-1].
template <int DIM, int NB, int NF, int NFU, bool BATCH>
__global__ void __launch_bounds__(256) sweep_records(OpArgs<double> g1, const BatchItem<double> *__restrict__ items,
                                                     const int *__restrict__ wprob, const int *__restrict__ ang,
                                                     const int *__restrict__ cell, const unsigned char *__restrict__ up,
                                                     const int64_t *__restrict__ lvl, const int *__restrict__ base,
                                                     int64_t n_items, double *__restrict__ r64, float *__restrict__ r32,
                                                     int *__restrict__ rnb, bool g64, bool g32) {
  const int w = blockIdx.y;
  const OpArgs<double> &g = BATCH ? items[wprob[w]].g : g1;
  const int64_t i0 = lvl[base[w]], i1 = lvl[base[w + 1] - 1];
  for (int64_t i = i0 + (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < i1; i += (int64_t)gridDim.x * blockDim.x)
    sweep_record<DIM, NB, NF, NFU>(g, i, n_items, cell[i], ang[i], up[i], r64, r32, rnb, g64, g32);
}

// One group's levels l0 .. l1-1 from the records (item i: cell cell[i],
// angle ang[i]). With PIPE (narrow levels), the records of a thread's items
// two or three levels ahead are in flight while a level is solved (see the
// pipelined loop below); otherwise (wide levels) a thread's first item of
// the next level has its cell, angle and upwind slots loaded one level
// ahead. Vector memory loads complete in issue order, which shapes both.
// ADJ: the adjoint sweep L^T = V P L P V^{-1} from the same records: item
// (c, b) reads r and writes psi at the reversed angle neg[b], r scaled by
// 1/V_c, the upwind traces by 1/V_n, the result by V_c (see the ADJ kernels'
// note at the top of this file). For ADJ, ang holds each item's reversed
// angle and vol each cell's volume, so no load of the fetch depends on
// another (each such dependence is a wait for every load issued before).
// LDSR: upwind traces read from the LDS slots (own_slot / up_slot) rather
// than from psi in HBM; every item whose trace is read later also writes it
// to its slot
template <class T, int NB, int NFU, bool PIPE, bool ADJ, bool GROUPED, bool LDSR = false>
__device__ __forceinline__ void sweep_rec_run(const T *__restrict__ r, T *__restrict__ psi, int A,
                                              const int *__restrict__ cell, const int *__restrict__ ang,
                                              const T *__restrict__ rec, const int *__restrict__ rnb, int64_t n_items,
                                              const int64_t *__restrict__ lvl, int l0, int l1,
                                              const int *__restrict__ neg = nullptr,
                                              const int *__restrict__ gcls = nullptr,
                                              const T *__restrict__ vol = nullptr,
                                              const int *__restrict__ oslot = nullptr,
                                              const int *__restrict__ uslot = nullptr) {
  constexpr int NB2 = NB * NB;
  const int tid = threadIdx.x, nt = blockDim.x;
  extern __shared__ __align__(16) unsigned char sweep_lds[];
  T *ring = reinterpret_cast<T *>(sweep_lds);
  struct Item {
    int64_t c;
    int a;
    int nbk[NFU];
    int os, us[NFU];  // LDSR: own and upwind trace slots
    T bi[NB2], h[NFU][NB2], rc[NB];
    T vc, ivn[NFU];  // ADJ: V_c, 1 / V_n per slot
  };
  // item i = (cell c, angle b) of the schedule
  // the rest of item i's loads once its cell, angle and upwind cells are
  // in it
  auto fetch_rest = [&](int64_t i, Item &it) {
    const int64_t c = it.c;
    const int a = it.a;
    if (ADJ) {
      it.vc = vol[c];
      if (!LDSR) {
#pragma unroll
        for (int k = 0; k < NFU; k++) it.ivn[k] = it.nbk[k] >= 0 ? T(1) / vol[it.nbk[k]] : T(0);
      }
    }
    if (GROUPED) {
      constexpr int F = NB2 * (1 + NFU), W = rec_w<T>();
      T v[rec_pad<T>(F)];
#pragma unroll
      for (int q = 0; q < rec_pad<T>(F); q += W) load16(rec + vix<T>(q, i, F), v + q);
#pragma unroll
      for (int e = 0; e < NB2; e++) it.bi[e] = v[e];
#pragma unroll
      for (int k = 0; k < NFU; k++)
#pragma unroll
        for (int e = 0; e < NB2; e++) it.h[k][e] = v[NB2 + k * NB2 + e];
    } else {
#pragma unroll
      for (int e = 0; e < NB2; e++) it.bi[e] = rec[rix(e, i, NB2 * (1 + NFU))];
#pragma unroll
      for (int k = 0; k < NFU; k++)
#pragma unroll
        for (int e = 0; e < NB2; e++) it.h[k][e] = rec[rix(NB2 + k * NB2 + e, i, NB2 * (1 + NFU))];
    }
#pragma unroll
    for (int j = 0; j < NB; j++) it.rc[j] = r[(c * NB + j) * A + a];
  };
  auto fetch = [&](int64_t i, int64_t c, int b, Item &it) {
    it.c = c;
    it.a = b;
    if (LDSR) {
      it.os = oslot[i];
#pragma unroll
      for (int k = 0; k < NFU; k++) it.us[k] = uslot[rix(k, i, NFU)];
    }
    if (!LDSR) {
#pragma unroll
      for (int k = 0; k < NFU; k++) it.nbk[k] = rnb[rix(k, i, NFU)];
    }
    fetch_rest(i, it);
  };
  // the upwind traces of an item (issued first: the level's critical loads);
  // an unused slot loads cell 0's row and discards it, so the number of loads
  // in flight is the same on every path and waits can count past them
  auto traces = [&](const Item &it, T (&pu)[NFU][NB]) {
#pragma unroll
    for (int k = 0; k < NFU; k++) {
      if (LDSR) {
        const int sl = it.us[k] >= 0 ? it.us[k] : 0;
#pragma unroll
        for (int j = 0; j < NB; j++) {
          const T v = ring[sl * NB + j];
          pu[k][j] = it.us[k] >= 0 ? v : T(0);
        }
      } else {
        const int64_t n = it.nbk[k] >= 0 ? it.nbk[k] : 0;
#pragma unroll
        for (int j = 0; j < NB; j++) {
          const T v = psi[(n * NB + j) * A + it.a];
          pu[k][j] = it.nbk[k] >= 0 ? v : T(0);
        }
      }
    }
  };
  auto finish = [&](const Item &it, const T (&pu)[NFU][NB]) {
#pragma unroll
    for (int p = 0; p < NB; p++) {
      T v = T(0);
#pragma unroll
      for (int q = 0; q < NB; q++) v += it.bi[p * NB + q] * (ADJ ? it.rc[q] * (T(1) / it.vc) : it.rc[q]);
#pragma unroll
      for (int k = 0; k < NFU; k++)
#pragma unroll
        for (int q = 0; q < NB; q++)
          v += it.h[k][p * NB + q] * (ADJ && !LDSR ? pu[k][q] * it.ivn[k] : pu[k][q]);
      psi[(it.c * NB + p) * A + it.a] = ADJ ? it.vc * v : v;
      // the adjoint's slot holds the trace before the volume scaling: the
      // reader's 1 / V_n is then already applied
      if (LDSR && it.os >= 0) ring[it.os * NB + p] = v;
    }
  };
  auto solve = [&](const Item &it) {
    T pu[NFU][NB];
    traces(it, pu);
    finish(it, pu);
  };
  if (l0 >= l1) return;
  if (!PIPE) {
    // plain form (wide levels): the cell, angle and upwind cells of a
    // thread's first item of level l + 1 are loaded during level l, after
    // level l's loads (loads complete in issue order), so that item issues
    // its trace loads at once, ahead of its records. A thread without such
    // an item loads the group's first one and discards it.
    const int64_t i_any = lvl[l0];
    struct Head {
      int64_t i;
      int c, a, nbk[NFU], os, us[NFU];
      bool h;
    } hd;
    auto head_of = [&](int l) {
      int64_t i = i_any;
      hd.h = false;
      if (l < l1) {
        i = lvl[l] + tid;
        hd.h = i < lvl[l + 1];
        i = hd.h ? i : i_any;
      }
      hd.i = i;
      hd.c = cell[i];
      hd.a = ang[i];
      if (LDSR) {
        hd.os = oslot[i];
#pragma unroll
        for (int k = 0; k < NFU; k++) hd.us[k] = uslot[rix(k, i, NFU)];
      }
      if (!LDSR) {
#pragma unroll
        for (int k = 0; k < NFU; k++) hd.nbk[k] = rnb[rix(k, i, NFU)];
      }
    };
    head_of(l0);
    for (int l = l0; l < l1; l++) {
      Item it;
      it.c = hd.c;
      it.a = hd.a;
      it.os = hd.os;
#pragma unroll
      for (int k = 0; k < NFU; k++) {
        it.nbk[k] = hd.nbk[k];
        it.us[k] = hd.us[k];
      }
      const bool h = hd.h;
      const int64_t i0 = hd.i;
      T pu[NFU][NB];
      traces(it, pu);
      fetch_rest(i0, it);
      head_of(l + 1);
      if (h) finish(it, pu);
      for (int64_t i = lvl[l] + tid + nt; i < lvl[l + 1]; i += nt) {
        Item jt;
        fetch(i, cell[i], ang[i], jt);
        solve(jt);
      }
      __syncthreads();
    }
    return;
  }
  // pipelined DEPTH levels deep: level l's item is solved while the records
  // of levels l + 1 .. l + DEPTH - 2 are in flight and level l + DEPTH - 1's
  // are requested. The items rotate through DEPTH register sets (a copy would
  // wait for its loads). Loads complete in issue order, so the cell and
  // angle that index a fetch are loaded before the records of an earlier
  // level: waiting for them then leaves the younger record loads in flight
  // (and without LDS slots, each level issues its trace loads first). Every
  // load is unconditional: a thread without an item at some level loads the
  // group's first item and discards it, so the count of loads in flight is
  // the same on every path. DEPTH is 4 where three items' loads fit the 63 a
  // wavefront can have in flight (FP32 records of at most 36 values: 9
  // grouped loads), else 3. With DEPTH 4 the cell and angle are loaded two
  // levels before their fetch (alternating between two sets), so every
  // record load has three levels to arrive; with DEPTH 3, one level before,
  // and two levels to arrive.
  constexpr int DEPTH = sizeof(T) == 4 && NB2 * (1 + NFU) <= 36 ? 4 : 3;
  const int64_t i_any = lvl[l0];
  // this thread's item at level l (or item i_any) and its cell and angle
  struct Slot {
    int64_t i;
    int c, b;
    bool h;
  };
  auto slot_of = [&](int l, Slot &sl) {
    int64_t i = i_any;
    sl.h = false;
    if (l >= l0 && l < l1) {
      i = lvl[l] + tid;
      sl.h = i < lvl[l + 1];
      i = sl.h ? i : i_any;
    }
    sl.i = i;
    sl.c = cell[i];
    sl.b = ang[i];
  };
  // The pipeline fills in DEPTH - 1 levels before l0 that solve nothing
  // (items with no upwind slot), inside the loop: fetches issued before it
  // would reach the loop's first level with a different count of loads in
  // flight than its back edge brings, and the compiler would then wait for
  // all of them there on every trip.
  Item x, y, z, u;
  auto none = [](Item &it) {
    it.c = 0;
    it.a = 0;
    it.os = -1;
#pragma unroll
    for (int k = 0; k < NFU; k++) it.nbk[k] = it.us[k] = -1;
  };
  none(x);
  none(y);
  none(z);
  none(u);
  bool hx = false, hy = false, hz = false, hu = false;
  Slot sa, sb;
  slot_of(l0, sa);
  if (DEPTH == 4) slot_of(l0 + 1, sb);
  __builtin_amdgcn_s_waitcnt(0);
  // level l from cur; level l + DEPTH - 1's item, whose slot is sl, loaded
  // into nxt; sl then reloaded for the fetch of level l + 2 (DEPTH 4, sl
  // alternates) or l + 1 (DEPTH 3)
  auto level = [&](int l, const Item &cur, bool hc, Item &nxt, bool &hn, Slot &sl) {
    T pu[NFU][NB];
    traces(cur, pu);
    int cn = sl.c, bn = sl.b;
    const int64_t in = sl.i;
    hn = sl.h;
    slot_of(l + (DEPTH == 4 ? DEPTH + 1 : DEPTH), sl);
    // cn and bn are first used here: without the empty asm the compiler
    // extends them right after their load, and that wait would cover the
    // record loads issued after them
    asm volatile("" : "+v"(cn), "+v"(bn));
    fetch(in, cn, bn, nxt);
    if (hc) finish(cur, pu);
    // items beyond one per thread (a level wider than the workgroup)
    if (l >= l0)
      for (int64_t i = lvl[l] + tid + nt; i < lvl[l + 1]; i += nt) {
        Item it;
        fetch(i, cell[i], ang[i], it);
        solve(it);
      }
    __syncthreads();
  };
  if (DEPTH == 4) {
    for (int l = l0 - 3; l < l1; l += 4) {
      level(l, x, hx, u, hu, sa);
      if (l + 1 < l1) level(l + 1, y, hy, x, hx, sb);
      if (l + 2 < l1) level(l + 2, z, hz, y, hy, sa);
      if (l + 3 < l1) level(l + 3, u, hu, z, hz, sb);
    }
  } else {
    for (int l = l0 - 2; l < l1; l += 3) {
      level(l, x, hx, z, hz, sa);
      if (l + 1 < l1) level(l + 1, y, hy, x, hx, sa);
      if (l + 2 < l1) level(l + 2, z, hz, y, hy, sa);
    }
  }
}

// PIPE holds three or four items' records in registers: at most 256 threads; the plain
// form serves wide levels with up to sweep_max_threads(NB) (one item's
// records of an NB = 4 hex or tet are 64-96 values: 512 threads keep them in
// registers, 1024 would spill)
__host__ __device__ constexpr int sweep_max_threads(int nb) { return nb >= 4 ? 512 : 1024; }

template <class T, int NB, int NFU, bool PIPE, bool ADJ, bool LDSR>
__global__ void __launch_bounds__(PIPE ? 256 : sweep_max_threads(NB)) sweep_dg_rec(const T *__restrict__ r, T *__restrict__ psi, int A,
                                                                  const int *__restrict__ cell, const int *__restrict__ ang,
                                                                  const T *__restrict__ rec, const int *__restrict__ rnb,
                                                                  int64_t n_items, const int64_t *__restrict__ lvl,
                                                                  const int *__restrict__ base, const int *__restrict__ neg,
                                                                  const int *__restrict__ gcls, const T *__restrict__ vol,
                                                                  const int *__restrict__ oslot,
                                                                  const int *__restrict__ uslot) {
  const int w = blockIdx.x;
  sweep_rec_run<T, NB, NFU, PIPE, ADJ, PIPE, LDSR>(r, psi, A, cell, ang, rec, rnb, n_items, lvl, base[w],
                                                   base[w + 1] - 1, neg, gcls, vol, oslot, uslot);
}

// (traces from HBM: LDS trace slots measured slower here, the LDS per
// workgroup costing the occupancy this traffic-bound sweep needs)
template <class T, int NB, int NFU>
__global__ void __launch_bounds__(sweep_max_threads(NB)) sweep_dg_rec_batch(const T *__restrict__ r, T *__restrict__ psi,
                                                           const BatchItem<T> *__restrict__ items,
                                                           const int *__restrict__ wprob, const int *__restrict__ cell,
                                                           const int *__restrict__ ang, const T *__restrict__ rec,
                                                           const int *__restrict__ rnb, int64_t n_items,
                                                           const int64_t *__restrict__ lvl, const int *__restrict__ base) {
  const int w = blockIdx.x;
  const BatchItem<T> &bi = items[wprob[w]];
  sweep_rec_run<T, NB, NFU, false, false, false>(r + bi.off, psi + bi.off, bi.g.A, cell, ang, rec, rnb, n_items, lvl,
                                                 base[w], base[w + 1] - 1);
}

// out[c] = vol[gcls[c]]: the class volumes per cell
template <class T>
__global__ void k_cell_volumes(int n, const int *__restrict__ gcls, const T *__restrict__ vol, T *__restrict__ out) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c < n) out[c] = vol[gcls[c]];
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
// inflow slots of the record layout: the smallest instantiated count that
// holds the most inflow faces of any item
int sweep_slots(int nf, int nfu) {
  if (nf == 3) return 2;
  if (nf == 4) return nfu <= 2 ? 2 : 3;
  return nfu <= 3 ? 3 : 5;
}

// LDS trace slots of the workgroups of one problem's schedule (items cell /
// angle / up in level order, levels lvl[base[g] .. base[g + 1]] of workgroup
// g, item indices local to the list; the list's first item is i0 in own /
// ups). Within a workgroup an item whose trace a later item reads holds a
// slot from its level up to the last level that reads it (register
// allocation over those live ranges: a slot freed by a trace last read at
// level L - 1 is reused from level L, after the barrier that ends L - 1);
// ups[rix(k, i0 + i, slots)] = the slot of item i's k-th upwind item. A
// direction depends only on itself, so every upwind item is in the reader's
// workgroup. Returns the most slots any workgroup holds at once.
int assign_trace_slots(const Disc &D, int slots, const int *cell, const int *angle, const unsigned char *up,
                       const int64_t *lvl, const int *base, int n_wg, int64_t i0, int *own, int *ups) {
  const int nf = D.nf, A = D.n_angles;
  std::vector<int> pos((size_t)D.n_cells * A, -1), last, free_slots;
  std::vector<std::vector<int>> release;
  auto upwind_of = [&](int64_t i, auto &&fn) {
    const int c = cell[i];
    for (int f = 0, k = 0; f < nf; f++)
      if ((up[i] >> f) & 1) fn(k++, pos[(size_t)D.face_nb[(int64_t)c * nf + f] * A + angle[i]]);
  };
  int most = 0;
  for (int g = 0; g < n_wg; g++) {
    const int n_lv = base[g + 1] - 1 - base[g];
    const int64_t a0 = lvl[base[g]], a1 = lvl[base[g + 1] - 1];
    for (int64_t i = a0; i < a1; i++) pos[(size_t)cell[i] * A + angle[i]] = (int)i;
    last.assign((size_t)(a1 - a0), -1);
    for (int l = 0; l < n_lv; l++)
      for (int64_t i = lvl[base[g] + l]; i < lvl[base[g] + l + 1]; i++)
        upwind_of(i, [&](int, int j) { last[j - a0] = std::max(last[j - a0], l); });
    release.assign((size_t)n_lv + 2, {});
    free_slots.clear();
    int used = 0;
    for (int l = 0; l < n_lv; l++) {
      for (int sl : release[l]) free_slots.push_back(sl);
      for (int64_t i = lvl[base[g] + l]; i < lvl[base[g] + l + 1]; i++) {
        const int lu = last[i - a0];
        if (lu < 0) continue;
        int sl;
        if (free_slots.empty()) {
          sl = used++;
        } else {
          sl = free_slots.back();
          free_slots.pop_back();
        }
        own[i0 + i] = sl;
        release[lu + 1].push_back(sl);
      }
    }
    for (int64_t i = a0; i < a1; i++)
      upwind_of(i, [&](int k, int j) { ups[rix(k, i0 + i, slots)] = own[i0 + j]; });
    most = std::max(most, used);
  }
  return most;
}

void Sweep::create(Operator *o, bool keep_host, bool lds_traces) {
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
  exact = n_lagged == 0 && std::none_of(D.face_nb.begin(), D.face_nb.end(), [](int v) { return v < -1; });
  // per direction (one workgroup each): the items in generation order
  // (direction, then level), and each direction's level boundaries
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
  n_groups = A;
  n_levels = max_level;
  nfu = 1;
  for (unsigned char m : item_up) nfu = std::max(nfu, __builtin_popcount(m));
  threads = (int)std::min<int64_t>(1024, std::max<int64_t>(64, (widest + 63) / 64 * 64));
  // the pipelined form holds three items' records in registers (the
  // adjoint's, the larger): levels of at most 256 items and records that
  // fit; otherwise the plain form with up to sweep_max_threads(nb) threads
  {
    const int slots = sweep_slots(nf, nfu), nb = D.nb;
    auto fits = [&](int words) { return threads <= 256 && 3 * ((nb * nb * (1 + slots) + nb + 1 + slots) * words + slots + 3) <= 300; };
    pipe64 = fits(2);
    pipe32 = fits(1);
  }
  // LDS trace slots (see assign_trace_slots): the sweep then reads every
  // upwind trace from LDS, and no trace read waits behind the record loads
  // in flight (vector memory loads complete in issue order). A precision
  // whose slots would not fit kLdsTraceBytes reads the traces from HBM, as
  // does every sweep created with lds_traces = false.
  trace_slots = 0;
  if (!keep_host && lds_traces) {
    const int slots = sweep_slots(nf, nfu);
    std::vector<int> own((size_t)n_items, -1), ups((size_t)rec_len(n_items, slots), -1);
    const int most = assign_trace_slots(D, slots, item_cell.data(), item_angle.data(), item_up.data(), alvl.data(),
                                        abase.data(), A, 0, own.data(), ups.data());
    if ((int64_t)most * D.nb * (int64_t)sizeof(float) <= kLdsTraceBytes) {
      trace_slots = std::max(most, 1);
      own_slot.upload(own);
      up_slot.upload(ups);
    }
  }
  rec_group = -1;
  if (!keep_host) alloc_records();
  {
    // the adjoint sweep's per-item reversed angle and per-cell volumes
    std::vector<int> neg(A), adj((size_t)n_items);
    for (int a = 0; a < A; a++) {
      int b = a;
      for (int d = 0; d < D.dim; d++) b = D.quad.reflect[d][b];
      neg[a] = b;
    }
    for (int64_t i = 0; i < n_items; i++) adj[i] = neg[item_angle[i]];
    angle_adj.upload(adj);
    vcell64.alloc(sizeof(double) * nc);
    vcell32.alloc(sizeof(float) * nc);
    const unsigned g = (unsigned)((nc + 255) / 256);
    hipLaunchKernelGGL(k_cell_volumes<double>, dim3(g), dim3(256), 0, nullptr, nc, o->gcls.as<int>(),
                       o->t64.vol.as<double>(), vcell64.as<double>());
    hipLaunchKernelGGL(k_cell_volumes<float>, dim3(g), dim3(256), 0, nullptr, nc, o->gcls.as<int>(),
                       o->t32.vol.as<float>(), vcell32.as<float>());
    UBT_HIP(hipGetLastError());
    UBT_HIP(hipStreamSynchronize(nullptr));
  }
  cells.upload(item_cell);
  angle.upload(item_angle);
  upwind.upload(item_up);
  lvl.upload(alvl);
  base.upload(abase);
  if (keep_host) {
    h_cells.swap(item_cell);
    h_angle.swap(item_angle);
    h_up.swap(item_up);
    h_lvl.swap(alvl);
    h_base.swap(abase);
  }
}

// Groups of up to G directions of one octant of a box, one workgroup each: a
// group's level l is the union of its directions' levels l (a direction
// depends only on its own earlier levels, so any grouping is a valid order).
// On a box every direction of an octant has the same level structure, so a
// level's items, sorted by (cell, angle), put the angles of a cell in adjacent
// lanes and the rows they read and write are contiguous; fewer, wider
// workgroups trade concurrency for that. Elsewhere (or G = 1) a group is one
// direction. From the per-direction host schedule of s.
void group_schedule(const Sweep &s, int G, SweepLists &out) {
  const Disc &D = *s.op->disc;
  const int A = D.n_angles;
  std::vector<std::vector<int>> groups;
  if (s.op->box && G > 1) {
    std::vector<std::vector<int>> by_oct(8);
    for (int a = 0; a < A; a++) {
      int oc = 0;
      for (int d = 0; d < D.dim; d++) oc |= (D.quad.omega[3 * (size_t)a + d] > 0.0 ? 1 : 0) << d;
      by_oct[oc].push_back(a);
    }
    for (auto &g : by_oct)
      for (size_t k = 0; k < g.size(); k += G)
        groups.push_back(std::vector<int>(g.begin() + k, g.begin() + std::min(g.size(), k + (size_t)G)));
  } else {
    for (int a = 0; a < A; a++) groups.push_back({a});
  }
  out = SweepLists{};
  out.base.push_back(0);
  std::vector<std::pair<int64_t, int64_t>> chunk;  // (cell * A + angle, source item)
  for (const auto &g : groups) {
    for (int l = 0;; l++) {
      chunk.clear();
      for (int a : g) {
        const int k = s.h_base[a] + l;
        if (k + 1 >= s.h_base[a + 1]) continue;  // direction a has fewer levels
        for (int64_t i = s.h_lvl[k]; i < s.h_lvl[k + 1]; i++) chunk.push_back({(int64_t)s.h_cells[i] * A + a, i});
      }
      if (chunk.empty()) break;
      std::sort(chunk.begin(), chunk.end());
      out.lvl.push_back((int64_t)out.cell.size());
      for (auto &e : chunk) {
        out.cell.push_back(s.h_cells[e.second]);
        out.angle.push_back(s.h_angle[e.second]);
        out.up.push_back(s.h_up[e.second]);
      }
      out.widest = std::max<int64_t>(out.widest, (int64_t)chunk.size());
    }
    out.lvl.push_back((int64_t)out.cell.size());
    out.base.push_back((int)out.lvl.size());
  }
}


// the records of every item of `g`'s (or the batch's) schedule
static void launch_records(int dim, int nb, int nf, int slots, const OpArgs<double> &g1,
                           const BatchItem<double> *items, const int *wprob, const int *ang, const int *cell,
                           const unsigned char *up, const int64_t *lvl, const int *base, int n_wg, int64_t widest_dir,
                           int64_t n_items, double *r64, float *r32, int *rnb, bool g64, bool g32, bool batch,
                           hipStream_t s) {
  const unsigned gx = (unsigned)std::max<int64_t>(1, std::min<int64_t>(64, (widest_dir + 255) / 256));
  const dim3 grid(gx, (unsigned)n_wg);
  bool launched = false;
#define RC(DIM, NB, NF, NFU)                                                                                         \
  if (!launched && dim == DIM && nb == NB && nf == NF && slots == NFU) {                                             \
    if (batch)                                                                                                       \
      hipLaunchKernelGGL((sweep_records<DIM, NB, NF, NFU, true>), grid, dim3(256), 0, s, g1, items, wprob, ang,       \
                         cell, up, lvl, base, n_items, r64, r32, rnb, g64, g32);                                     \
    else                                                                                                             \
      hipLaunchKernelGGL((sweep_records<DIM, NB, NF, NFU, false>), grid, dim3(256), 0, s, g1, items, wprob, ang,      \
                         cell, up, lvl, base, n_items, r64, r32, rnb, g64, g32);                                     \
    launched = true;                                                                                                 \
  }
  RC(2, 1, 3, 2) RC(2, 1, 4, 2) RC(2, 1, 4, 3) RC(3, 1, 4, 2) RC(3, 1, 4, 3) RC(3, 1, 6, 3) RC(3, 1, 6, 5)
  RC(2, 3, 3, 2) RC(2, 3, 4, 2) RC(2, 3, 4, 3) RC(3, 4, 4, 2) RC(3, 4, 4, 3) RC(3, 4, 6, 3) RC(3, 4, 6, 5)
#undef RC
  UBT_CHECK(launched, UBT_ERR_SUP, "no sweep records for dim %d, %d basis, %d faces, %d slots", dim, nb, nf, slots);
  UBT_HIP(hipGetLastError());
}

void Sweep::alloc_records() const {
  const Disc &D = *op->disc;
  const int64_t n_items = (int64_t)D.n_cells * D.n_angles;
  const int slots = sweep_slots(D.nf, nfu), per = D.nb * D.nb * (1 + slots);
  rec64.alloc(sizeof(double) * (pipe64 ? vrec_len<double>(n_items, per) : rec_len(n_items, per)));
  rec32.alloc(sizeof(float) * (pipe32 ? vrec_len<float>(n_items, per) : rec_len(n_items, per)));
  rnb.alloc(sizeof(int) * rec_len(n_items, slots));
}

void Sweep::build_records(hipStream_t s) const {
  const Operator &o = *op;
  const Disc &D = *o.disc;
  const int slots = sweep_slots(D.nf, nfu);
  const int64_t n_items = (int64_t)D.n_cells * D.n_angles;
  if (!rec64.ptr) alloc_records();
  launch_records(D.dim, D.nb, D.nf, slots, o.args<double>(), nullptr, nullptr, angle.as<int>(), cells.as<int>(),
                 upwind.as<unsigned char>(), lvl.as<int64_t>(), base.as<int>(), n_groups,
                 (n_items + n_groups - 1) / n_groups, n_items, rec64.as<double>(), rec32.as<float>(), rnb.as<int>(),
                 pipe64, pipe32, false, s);
  rec_group = o.group;
  rec_gen = o.xs_gen;
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
  const int *ci = cells.as<int>(), *bs = base.as<int>(), *an = (adj ? angle_adj : angle).as<int>();
  const int64_t *lv = lvl.as<int64_t>();
  const unsigned grid = (unsigned)n_groups;
  bool launched = false;
  // forward and adjoint sweeps both run from the forward records (the
  // adjoint reverses the angle and scales by the volumes in the kernel)
  if (rec_group != o.group || rec_gen != o.xs_gen) build_records(s);
  const int slots = sweep_slots(D.nf, nfu);
  const int64_t n_items = (int64_t)D.n_cells * D.n_angles;
  const bool pipe = prec == UBT_FP64 ? pipe64 : pipe32;
  const bool lds =
      trace_slots > 0 && (int64_t)trace_slots * D.nb * (prec == UBT_FP64 ? 8 : 4) <= kLdsTraceBytes;
  const int th = std::min(threads, sweep_max_threads(D.nb));
#define SR_K(T, NB, NFU, PIPE, ADJ, TH)                                                                            \
  if (lds)                                                                                                          \
    hipLaunchKernelGGL((sweep_dg_rec<T, NB, NFU, PIPE, ADJ, true>), dim3(grid), dim3(TH),                            \
                       (size_t)trace_slots * NB * sizeof(T), s, (const T *)r, (T *)psi, D.n_angles, ci, an, rc,       \
                       rnb.as<int>(), n_items, lv, bs, o.neg.as<int>(), o.gcls.as<int>(), vl, own_slot.as<int>(),     \
                       up_slot.as<int>());                                                                           \
  else                                                                                                              \
    hipLaunchKernelGGL((sweep_dg_rec<T, NB, NFU, PIPE, ADJ, false>), dim3(grid), dim3(TH), 0, s, (const T *)r,      \
                       (T *)psi, D.n_angles, ci, an, rc, rnb.as<int>(), n_items, lv, bs, o.neg.as<int>(),            \
                       o.gcls.as<int>(), vl, nullptr, nullptr);
#define SR(T, NB, NFU)                                                                                              \
  if (!launched && D.nb == NB && slots == NFU) {                                                                    \
    const T *rc = prec == UBT_FP64 ? rec64.as<T>() : rec32.as<T>();                                                 \
    const T *vl = (prec == UBT_FP64 ? vcell64 : vcell32).as<T>();                                                   \
    if (pipe && adj) {                                                                                              \
      SR_K(T, NB, NFU, true, true, threads)                                                                         \
    } else if (pipe) {                                                                                              \
      SR_K(T, NB, NFU, true, false, threads)                                                                        \
    } else if (adj) {                                                                                               \
      SR_K(T, NB, NFU, false, true, th)                                                                             \
    } else {                                                                                                        \
      SR_K(T, NB, NFU, false, false, th)                                                                            \
    }                                                                                                               \
    launched = true;                                                                                                \
  }
  if (prec == UBT_FP64) {
    SR(double, 1, 2) SR(double, 1, 3) SR(double, 1, 5) SR(double, 3, 2) SR(double, 3, 3) SR(double, 4, 3)
    SR(double, 4, 5)
  } else {
    SR(float, 1, 2) SR(float, 1, 3) SR(float, 1, 5) SR(float, 3, 2) SR(float, 3, 3) SR(float, 4, 3) SR(float, 4, 5)
  }
#undef SR
#undef SR_K
  UBT_CHECK(launched, UBT_ERR_SUP, "no sweep kernel for %d basis functions and %d inflow slots", D.nb, slots);
  UBT_HIP(hipGetLastError());
}

void batch_records_launch(const Batch &b, int pi, hipStream_t s) {
  const BatchSweep &w = b.sweep[pi];
  launch_records(b.dim, b.nb, b.nf, w.slots, OpArgs<double>{}, b.items[0][0].as<BatchItem<double>>(),
                 w.prob.as<int>(), w.angle.as<int>(), w.cells.as<int>(), w.upwind.as<unsigned char>(),
                 w.lvl.as<int64_t>(), w.base.as<int>(), w.n_wg, w.widest, w.n_items,
                 pi == 0 ? w.rec.as<double>() : nullptr, pi == 1 ? w.rec.as<float>() : nullptr, w.rnb.as<int>(), false,
                 false, true, s);
}

template <class T> void batch_sweep_launch(const Batch &b, const T *r, T *psi, hipStream_t s) {
  const int pi = sizeof(T) == 8 ? 0 : 1;
  const BatchSweep &w = b.sweep[pi];
  if (w.n_wg == 0) return;
  const BatchItem<T> *items = b.items[pi][0].as<BatchItem<T>>();
  bool launched = false;
  const T *rec = w.rec.as<T>();
#define SRB(NB, NFU)                                                                                                \
  if (!launched && b.nb == NB && w.slots == NFU) {                                                                  \
    hipLaunchKernelGGL((sweep_dg_rec_batch<T, NB, NFU>), dim3(w.n_wg),                                                \
                       dim3(std::min(w.threads, sweep_max_threads(NB))), 0, s, r, psi, items,                       \
                       w.prob.as<int>(), w.cells.as<int>(), w.angle.as<int>(), rec, w.rnb.as<int>(), w.n_items,      \
                       w.lvl.as<int64_t>(), w.base.as<int>());                                                      \
    launched = true;                                                                                                \
  }
  SRB(1, 2) SRB(1, 3) SRB(1, 5) SRB(3, 2) SRB(3, 3) SRB(4, 3) SRB(4, 5)
#undef SRB
  UBT_CHECK(launched, UBT_ERR_SUP, "no batched sweep for %d basis functions and %d inflow slots", b.nb, w.slots);
  UBT_HIP(hipGetLastError());
}

template void batch_sweep_launch<double>(const Batch &, const double *, double *, hipStream_t);
template void batch_sweep_launch<float>(const Batch &, const float *, float *, hipStream_t);

}  // namespace ubt
