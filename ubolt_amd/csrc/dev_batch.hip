// Batched problems: work lists, the batch autotuner, and the batch's solver
// (the device FGMRES engine over the batch's problems with the batch's
// persistent apply and composite preconditioner).
//
// Vectors of a batch are the problems' group vectors laid end to end (row
// offsets fixed at creation; problems without the current group keep their
// span and are skipped).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>

#include "batch.hpp"

namespace ubt {


// ---------------------------------------------------------------------------
// Batch
// ---------------------------------------------------------------------------
void Batch::create(const std::vector<Operator *> &o) {
  UBT_CHECK(!o.empty(), UBT_ERR_ARG_SIZ, "a batch needs at least one problem");
  ops = o;
  const Disc &D0 = *o[0]->disc;
  UBT_CHECK(D0.kind == DISC_DG, UBT_ERR_SUP, "batches hold DG problems");
  dim = D0.dim;
  nb = D0.nb;
  nf = D0.nf;
  off.assign(1, 0);
  for (size_t p = 0; p < o.size(); p++) {
    const Disc &D = *o[p]->disc;
    UBT_CHECK(D.kind == DISC_DG && D.dim == dim && D.nb == nb && D.nf == nf, UBT_ERR_ARG_INCOMP,
              "problem %zu is not a %dD DG problem with %d basis functions and %d faces per cell like problem 0", p,
              dim, nb, nf);
    UBT_CHECK(o[p]->xs_set, UBT_ERR_ARG_WRONGSTATE, "set the cross sections of problem %zu first", p);
    UBT_CHECK(o[p]->device == o[0]->device, UBT_ERR_ARG_INCOMP, "the problems of a batch share one device");
    off.push_back(off.back() + D.n_rows());
  }
  n_rows = off.back();
  sweeps.clear();
  for (auto *op : ops) {
    sweeps.emplace_back(new Sweep());
    sweeps.back()->create(op, true);
  }
  int dev = 0;
  UBT_HIP(hipGetDevice(&dev));
  UBT_HIP(hipDeviceGetAttribute(&n_cu, hipDeviceAttributeMultiprocessorCount, dev));
  set_group(0);
  tune();
}

void Batch::set_group(int g, const std::vector<char> &want) {
  UBT_CHECK(g >= 0, UBT_ERR_ARG_OUTOFRANGE, "group %d is negative", g);
  bool any = false;
  active.assign(ops.size(), 0);
  for (size_t p = 0; p < ops.size(); p++) {
    if (ops[p]->n_groups > g && (want.empty() || want[p])) {
      active[p] = 1;
      ops[p]->group = g;
      any = true;
    }
  }
  UBT_CHECK(any, UBT_ERR_ARG_OUTOFRANGE, "no problem of the batch has group %d", g);
  group = g;
  build_lists();
  gen++;
}

template <class T>
static void build_items(Batch &b, int body) {
  const int pi = sizeof(T) == 8 ? 0 : 1;
  std::vector<BatchItem<T>> items;
  std::vector<int64_t> tiles(1, 0), edges(1, 0);
  size_t lds = 0;
  bool ok = true;
  for (size_t p = 0; p < b.ops.size() && ok; p++) {
    if (!b.active[p]) continue;
    int cpb = 1, lanes = 1;
    size_t l = 0;
    int64_t nt = 0;
    BatchItem<T> it = {};
    if (!batch_body_layout(*b.ops[p], pi == 0 ? UBT_FP64 : UBT_FP32, body, &cpb, &lanes, &l, &nt, &it.box)) {
      ok = false;
      break;
    }
    it.g = b.ops[p]->args<T>();
    it.g.cpb = cpb;
    it.g.red_lanes = lanes;
    it.off = b.off[p];
    items.push_back(it);
    tiles.push_back(tiles.back() + nt);
    edges.push_back(edges.back() + (body >= 2 ? it.box.edge_tiles : 0));
    lds = std::max(lds, l);
  }
  b.body_ok[pi][body] = ok;
  if (!ok) return;
  b.items[pi][body].upload(items);
  b.tiles[pi][body].upload(tiles);
  if (body >= 2) {
    b.edge_off[pi][body].upload(edges);
    b.n_edge[pi][body] = edges.back();
  }
  b.n_items[pi][body] = (int)items.size();
  b.n_tiles[pi][body] = tiles.back();
  b.lds[pi][body] = lds;
}

void Batch::build_lists() {
  for (int body_i = 0; body_i < kBatchBodies; body_i++) {
    build_items<double>(*this, body_i);
    build_items<float>(*this, body_i);
  }
  UBT_CHECK(body_ok[0][0] && body_ok[1][0], UBT_ERR_SUP, "a problem of the batch has more than 256 angles");
  for (int pi = 0; pi < 2; pi++)
    if (!body_ok[pi][body[pi]]) body[pi] = 0;
  // what follows depends on the active set alone (not on the group's cross
  // sections): kept while the set is unchanged, so a group change costs the
  // apply work lists and the sweep's per-group records only
  if (active == lists_for) {
    sweep_records(nullptr);
    UBT_HIP(hipStreamSynchronize(nullptr));
    return;
  }
  lists_for = active;
  // element blocks: prefix of cells x angles over the active problems
  std::vector<int64_t> blk(1, 0);
  for (size_t p = 0; p < ops.size(); p++)
    if (active[p]) blk.push_back(blk.back() + (int64_t)ops[p]->disc->n_cells * ops[p]->disc->n_angles);
  bblk.upload(blk);
  n_blocks = blk.back();
  build_box_sweep();
  // the record sweep's lists are built when it first runs on this set
  for (int pi = 0; pi < 2; pi++) sweep[pi].lists_for.clear();
  sweep_records(nullptr);
  UBT_HIP(hipStreamSynchronize(nullptr));
}

// The record sweep's work list in precision pi: one workgroup per direction
// group of every active problem (kBatchSweepGroup directions of one octant),
// the problems' schedules concatenated; the workgroups are listed longest
// first (most levels), so the groups that bound the sweep are dispatched
// before the short ones. Each problem's regrouped schedule is computed once
// and kept; the list is rebuilt when the active set changes.
void Batch::record_lists(int pi) const {
  BatchSweep &w = sweep[pi];
  if (w.lists_for == active) return;
  int nfu = 1;
  for (size_t p = 0; p < ops.size(); p++)
    if (active[p]) nfu = std::max(nfu, sweeps[p]->nfu);
  std::vector<SweepLists> &lists = sched[pi];
  if (lists.size() != ops.size()) lists.assign(ops.size(), SweepLists{});
  {
    struct Grp {
      int q, g, nlev;
      size_t p;
    };
    std::vector<Grp> grps;
    std::vector<int64_t> item0(ops.size(), 0);
    std::vector<int> ce, an;
    std::vector<unsigned char> up;
    int q = 0;
    int64_t widest_level = 1;
    for (size_t p = 0; p < ops.size(); p++) {
      if (!active[p]) continue;
      SweepLists &L = lists[p];
      if (L.base.empty()) group_schedule(*sweeps[p], kBatchSweepGroup[pi], L);
      item0[p] = (int64_t)ce.size();
      ce.insert(ce.end(), L.cell.begin(), L.cell.end());
      an.insert(an.end(), L.angle.begin(), L.angle.end());
      up.insert(up.end(), L.up.begin(), L.up.end());
      widest_level = std::max(widest_level, L.widest);
      for (int g = 0; g + 1 < (int)L.base.size(); g++) grps.push_back({q, g, L.base[g + 1] - L.base[g] - 1, p});
      q++;
    }
    std::stable_sort(grps.begin(), grps.end(), [](const Grp &x, const Grp &y) { return x.nlev > y.nlev; });
    std::vector<int> wp, bs(1, 0);
    std::vector<int64_t> lv;
    w.widest = 0;
    for (const Grp &d : grps) {
      const SweepLists &L = lists[d.p];
      wp.push_back(d.q);
      for (int k = L.base[d.g]; k < L.base[d.g + 1]; k++) lv.push_back(item0[d.p] + L.lvl[k]);
      bs.push_back((int)lv.size());
      w.widest = std::max<int64_t>(w.widest, L.lvl[L.base[d.g + 1] - 1] - L.lvl[L.base[d.g]]);
    }
    w.n_wg = (int)wp.size();
    w.n_items = (int64_t)ce.size();
    w.threads = (int)std::min<int64_t>(1024, std::max<int64_t>(64, (widest_level + 63) / 64 * 64));
    w.slots = sweep_slots(nf, nfu);
    w.prob.upload(wp);
    w.angle.upload(an);
    w.cells.upload(ce);
    w.upwind.upload(up);
    w.lvl.upload(lv);
    w.base.upload(bs);
    w.rec_tag = ~0ull;
  }
  w.lists_for = active;
}

// The record-free box sweep's work list, per precision: every active problem
// a uniform all-vacuum 2D DG1 box (and the strip body's layout, whose work
// items the sweep reads, in place), at most 256 cells along x (the LDS
// diagonal); one workgroup per problem, quadrant and group of the quadrant's
// angles. The groups of one (problem, quadrant) share lines of the
// quadrant-major copies, so they go to one XCD: the sets, longest march
// first, are dealt to the XCD with the least work so far, and workgroup
// i = 8 j + x is the j-th of XCD x's list (idle slots pad the shorter lists)
void Batch::build_box_sweep() {
  constexpr int kXcd = 8;
  for (int pi = 0; pi < 2; pi++) {
    BoxSweepPlan &w = box_sweep[pi];
    const int AG = pi == 0 ? kBoxSweepAngles<double> : kBoxSweepAngles<float>;
    w.ok = dim == 2 && nb == 3 && body_ok[pi][3] && box_sweep_enabled;
    w.nx_max = 0;
    struct Set {
      int diagonals;
      double work;
      std::vector<BoxSweepGroup> g;
    };
    std::vector<Set> sets;
    std::vector<int> qpos;
    std::vector<int64_t> qoff(1, 0);
    int item = 0;
    for (size_t p = 0; p < ops.size() && w.ok; p++) {
      if (!active[p]) continue;
      const Operator &o = *ops[p];
      const Quadrature &q = o.disc->quad;
      w.ok = o.box_cls >= 0 && o.n_box_edge == 0 && o.box_n[0] <= 256 && q.n_angles % 4 == 0;
      if (!w.ok) break;
      w.nx_max = std::max(w.nx_max, o.box_n[0]);
      const int Aq = q.n_angles / 4;
      const int64_t nc = o.disc->n_cells;
      // a diagonal holds at most min(nx, ny) cells, 256 / AG per pass
      const int need = (std::min(o.box_n[0], o.box_n[1]) + 256 / AG - 1) / (256 / AG);
      const int ipt = need <= 1 ? 1 : need <= 2 ? 2 : need <= 4 ? 4 : 8;
      std::vector<int> pos(q.n_angles, -1);
      for (int sx = 0; sx < 2; sx++)
        for (int sy = 0; sy < 2; sy++) {
          const int quad = 2 * sx + sy;
          std::vector<int> an;
          for (int a = 0; a < q.n_angles; a++)
            if ((q.omega[3 * a] > 0.0) == (sx == 0) && (q.omega[3 * a + 1] > 0.0) == (sy == 0)) an.push_back(a);
          if ((int)an.size() != Aq) {
            w.ok = false;
            break;
          }
          for (int j = 0; j < Aq; j++) pos[an[j]] = quad * Aq + j;
          Set st;
          st.diagonals = o.box_n[0] + o.box_n[1] - 1;
          st.work = (double)nc * Aq;
          for (size_t k0 = 0; k0 < an.size(); k0 += AG) {
            BoxSweepGroup gw;
            gw.item = item;
            gw.q = quad;
            gw.k0 = (int)k0;
            for (int k = 0; k < 8; k++) gw.angle[k] = k < AG && k0 + k < an.size() ? an[k0 + k] : -1;
            gw.ipt = ipt;
            gw.rbase = off[p] + (int64_t)quad * nc * 3 * Aq + (int64_t)k0;
            st.g.push_back(gw);
          }
          sets.push_back(std::move(st));
        }
      qpos.insert(qpos.end(), pos.begin(), pos.end());
      qoff.push_back(qoff.back() + q.n_angles);
      item++;
    }
    if (!w.ok) {
      w.n_wg = 0;
      continue;
    }
    std::stable_sort(sets.begin(), sets.end(), [](const Set &x, const Set &y) { return x.diagonals > y.diagonals; });
    std::vector<std::vector<BoxSweepGroup>> lists(kXcd);
    std::vector<double> load(kXcd, 0.0);
    for (const Set &st : sets) {
      const int x = (int)(std::min_element(load.begin(), load.end()) - load.begin());
      load[x] += st.work;
      lists[x].insert(lists[x].end(), st.g.begin(), st.g.end());
    }
    size_t len = 0;
    for (const auto &l : lists) len = std::max(len, l.size());
    BoxSweepGroup idle{};
    idle.item = -1;
    std::vector<BoxSweepGroup> g(len * kXcd, idle);
    for (int x = 0; x < kXcd; x++)
      for (size_t j = 0; j < lists[x].size(); j++) g[j * kXcd + x] = lists[x][j];
    w.n_wg = (int)g.size();
    w.groups.upload(g);
    w.qpos.upload(qpos);
    w.qoff.upload(qoff);
    const size_t ts = pi == 0 ? sizeof(double) : sizeof(float);
    if (w.rq.bytes < ts * (size_t)n_rows) {
      w.rq.alloc(ts * (size_t)n_rows);
      w.pq.alloc(ts * (size_t)n_rows);
    }
  }
}

uint64_t Batch::xs_tag() const {
  uint64_t t = (uint64_t)group * 0x9E3779B97F4A7C15ull;
  for (size_t p = 0; p < ops.size(); p++) t = t * 31 + ops[p]->xs_gen;
  return t;
}

void Batch::sweep_records(hipStream_t s) const {
  const uint64_t tag = xs_tag();
  for (int pi = 0; pi < 2; pi++) {
    const BatchSweep &w = sweep[pi];
    if (box_sweep[pi].ok) continue;  // the box sweep forms its blocks on the fly
    record_lists(pi);
    if (w.rec_tag == tag || w.n_items == 0) continue;
    // 64-item blocks (see the sweep's record layout)
    const size_t blocks64 = (size_t)((w.n_items + 63) / 64) * 64;
    const size_t per = (size_t)nb * nb * (1 + w.slots);
    const size_t ts = pi == 0 ? sizeof(double) : sizeof(float);
    if (w.rec.bytes < ts * per * blocks64) {
      w.rec.alloc(ts * per * blocks64);
      w.rnb.alloc(sizeof(int) * (size_t)w.slots * blocks64);
    }
    batch_records_launch(*this, pi, s);
    w.rec_tag = tag;
  }
}

void Batch::mult(int prec, const void *x, void *y, bool scatter, hipStream_t s, bool quad) const {
  UBT_CHECK(x && y, UBT_ERR_ARG_NULL, "x and y are required");
  UBT_CHECK(x != y, UBT_ERR_ARG_IDN, "the batched apply needs distinct x and y");
  // the quadrant-major order has the box body only
  if (prec == UBT_FP64)
    batch_apply_launch<double>(*this, quad ? 2 : body[0], (const double *)x, (double *)y, scatter, s, quad);
  else if (prec == UBT_FP32)
    batch_apply_launch<float>(*this, quad ? 2 : body[1], (const float *)x, (float *)y, scatter, s, quad);
  else
    fail(UBT_ERR_ARG_OUTOFRANGE, "unknown precision %d", prec);
}

void Batch::quad_order(int prec, const void *src, void *dst, bool to_q, hipStream_t s) const {
  UBT_CHECK(quad_ready(prec), UBT_ERR_ARG_WRONGSTATE, "the batch has no quadrant-major order");
  if (prec == UBT_FP64)
    batch_quad_order<double>(*this, (const double *)src, (double *)dst, to_q, s);
  else
    batch_quad_order<float>(*this, (const float *)src, (float *)dst, to_q, s);
}

// every active problem's singular-block check, one counter each, read back
// together (UBT_ERR_MAT_LU_ZRPVT naming the first problem with a singular
// block)
void Batch::blocks_setup(hipStream_t s) {
  const size_t np = ops.size();
  if (block_cnt.bytes < sizeof(int) * np) block_cnt.alloc(sizeof(int) * np);
  int *cnt = block_cnt.as<int>();
  UBT_HIP(hipMemsetAsync(cnt, 0, sizeof(int) * np, s));
  for (size_t p = 0; p < np; p++)
    if (active[p]) blockinv_check_launch(*ops[p], cnt + p, s);
  std::vector<int> h(np);
  UBT_HIP(hipMemcpyAsync(h.data(), cnt, sizeof(int) * np, hipMemcpyDeviceToHost, s));
  UBT_HIP(hipStreamSynchronize(s));
  for (size_t p = 0; p < np; p++)
    UBT_CHECK(!active[p] || h[p] == 0, UBT_ERR_MAT_LU_ZRPVT, "%d of %lld element blocks of problem %zu are singular",
              h[p], (long long)ops[p]->disc->n_cells * ops[p]->disc->n_angles, p);
}

void Batch::blocks_apply(int prec, const void *x, void *y, hipStream_t s) const {
  if (prec == UBT_FP64)
    batch_blocks_launch<double>(*this, (const double *)x, (double *)y, s);
  else
    batch_blocks_launch<float>(*this, (const float *)x, (float *)y, s);
}

void Batch::sweep_apply(int prec, const void *r, void *psi, hipStream_t s, bool staged) const {
  const int pi = prec == UBT_FP64 ? 0 : 1;
  UBT_CHECK(!staged || box_sweep[pi].ok, UBT_ERR_ARG_WRONGSTATE, "a staged sweep needs the box sweep");
  if (box_sweep[pi].ok) {
    if (prec == UBT_FP64)
      batch_box_sweep_launch<double>(*this, (const double *)r, (double *)psi, s, staged);
    else
      batch_box_sweep_launch<float>(*this, (const float *)r, (float *)psi, s, staged);
    return;
  }
  sweep_records(s);
  if (prec == UBT_FP64)
    batch_sweep_launch<double>(*this, (const double *)r, (double *)psi, s);
  else
    batch_sweep_launch<float>(*this, (const float *)r, (float *)psi, s);
}

bool Batch::sweep_exact(int prec) const {
  if (box_sweep[prec == UBT_FP64 ? 0 : 1].ok) return true;  // uniform all-vacuum boxes: no cycle, no reflection
  for (size_t p = 0; p < ops.size(); p++)
    if (active[p] && !sweeps[p]->exact) return false;
  return true;
}

void Batch::rhs(int prec, void *b, hipStream_t s) const {
  const size_t ts = prec == UBT_FP64 ? 8 : 4;
  for (size_t p = 0; p < ops.size(); p++) {
    char *bp = (char *)b + off[p] * ts;
    const int64_t n = off[p + 1] - off[p];
    if (active[p]) {
      ops[p]->rhs(group, prec, bp, s);
    } else if (prec == UBT_FP64) {
      vec_fill<double>(n, 0.0, (double *)bp, s);
    } else {
      vec_fill<float>(n, 0.f, (float *)bp, s);
    }
  }
  UBT_HIP(hipGetLastError());
}

void Batch::set_scalar_flux(int prec, const void *psi, hipStream_t s) const {
  const size_t ts = prec == UBT_FP64 ? 8 : 4;
  for (size_t p = 0; p < ops.size(); p++)
    if (active[p]) ops[p]->set_scalar_flux(group, prec, (const char *)psi + off[p] * ts, s);
}

void Batch::tune() {
  hipStream_t s = nullptr;
  UBT_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
  hipEvent_t e0, e1;
  UBT_HIP(hipEventCreate(&e0));
  UBT_HIP(hipEventCreate(&e1));
  DevBuf xb, yb;
  xb.alloc(sizeof(double) * n_rows);
  yb.alloc(sizeof(double) * n_rows);
  vec_fill<double>(n_rows, 1.0, xb.as<double>(), s);  // tuned on a non-zero input
  const int persists[3] = {2, 4, 8};
  for (int pi = 0; pi < 2; pi++) {
    float best = 1e30f;
    int best_b = 0, best_p = 8;
    for (int bo = 0; bo < kBatchBodies; bo++) {
      tuned_us[pi][bo] = 0.f;
      if (!body_ok[pi][bo]) continue;
      float bb = 1e30f;
      for (int pp : persists) {
        body[pi] = bo;
        persist[pi] = pp;
        const int prec = pi == 0 ? UBT_FP64 : UBT_FP32;
        for (int it = 0; it < 2; it++) mult(prec, xb.ptr, yb.ptr, true, s);
        UBT_HIP(hipEventRecord(e0, s));
        for (int it = 0; it < 5; it++) mult(prec, xb.ptr, yb.ptr, true, s);
        UBT_HIP(hipEventRecord(e1, s));
        UBT_HIP(hipEventSynchronize(e1));
        float ms = 0.f;
        UBT_HIP(hipEventElapsedTime(&ms, e0, e1));
        const float us = ms * 1000.f / 5.f;
        bb = std::min(bb, us);
        if (us < best) {
          best = us;
          best_b = bo;
          best_p = pp;
        }
      }
      tuned_us[pi][bo] = bb;
    }
    body[pi] = best_b;
    persist[pi] = best_p;
  }
  UBT_HIP(hipStreamSynchronize(s));
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  (void)hipStreamDestroy(s);
}

// ---------------------------------------------------------------------------
// BatchSolver
// ---------------------------------------------------------------------------
void BatchSolver::create(Batch *b, const SolverOptions &o) {
  batch = b;
  opt = o;
  const int64_t n = b->n_rows;
  eng.create(this, n, (int)b->ops.size(), o, 0);
  if (!opt.mixed) {
    t64.alloc(sizeof(double) * n);
    u64.alloc(sizeof(double) * n);
  } else {
    t32.alloc(sizeof(float) * n);
    u32.alloc(sizeof(float) * n);
  }
}

void BatchSolver::refresh(hipStream_t s) {
  batch->blocks_setup(s);
  eng.set_problems(batch->off, batch->active, s);
  batch->sweep_records(s);
  gen = batch->gen;
  xs_tag = batch->xs_tag();
  // the quadrant-major order: the preconditioner is then the box sweep alone
  // (it inverts every problem's L exactly), so no block stage is needed
  quad = opt.pc_sweep && batch->quad_ready(UBT_FP64) && (!opt.mixed || batch->quad_ready(UBT_FP32));
  if (quad && bq.bytes < sizeof(double) * (size_t)batch->n_rows) {
    bq.alloc(sizeof(double) * (size_t)batch->n_rows);
    xq.alloc(sizeof(double) * (size_t)batch->n_rows);
  }
}

void BatchSolver::mult(int prec, const void *x, void *y, hipStream_t s) { batch->mult(prec, x, y, true, s, quad); }

void BatchSolver::pc(int prec, const void *r, void *y, hipStream_t s) {
  if (prec == UBT_FP64)
    apply_pc<double>((const double *)r, (double *)y, s);
  else
    apply_pc<float>((const float *)r, (float *)y, s);
}

// y = M^{-1} r: element blocks, then the sweep on the L-residual. Where the
// sweep is L's exact inverse S the composite is the sweep alone:
// y + S (r - L y) = S r for any first stage y, so the blocks, the residual
// apply and the two vector updates are left out (the same result up to
// rounding)
template <class T> void BatchSolver::apply_pc(const T *r, T *y, hipStream_t s) {
  T *t = sizeof(T) == 8 ? t64.as<T>() : t32.as<T>();
  T *u = sizeof(T) == 8 ? u64.as<T>() : u32.as<T>();
  const int prec = prec_id<T>();
  const int64_t n = batch->n_rows;
  if (quad) {  // the box sweep on vectors already in its order
    batch->sweep_apply(prec, r, y, s, true);
    return;
  }
  bool first = true;
  if (opt.pc_blocks && !(opt.pc_sweep && batch->sweep_exact(prec))) {
    batch->blocks_apply(prec, r, y, s);
    first = false;
  }
  if (opt.pc_sweep) {
    if (first) {
      batch->sweep_apply(prec, r, y, s);
      first = false;
    } else {
      batch->mult(prec, y, t, false, s);
      vec_axpby<T>(n, 1.0, r, -1.0, t, s);
      batch->sweep_apply(prec, t, u, s);
      vec_axpby<T>(n, 1.0, (const T *)u, 1.0, y, s);
    }
  }
  if (first) vec_axpby<T>(n, 1.0, r, 0.0, y, s);
}

void BatchSolver::solve(const double *b, double *x, hipStream_t s, bool guess, const double *rtol) {
  UBT_CHECK(eng.n_chunks > 0 && gen == batch->gen, UBT_ERR_ARG_WRONGSTATE,
            "refresh the batch solver after setting the batch group");
  // a member's cross sections changed since the refresh (ubt_op_set_xs):
  // check the blocks and rebuild the sweep records outside any capture; the
  // engine recaptures its cycles under the new tag
  const uint64_t tag = batch->xs_tag();
  if (tag != xs_tag) {
    batch->blocks_setup(s);
    batch->sweep_records(s);
    xs_tag = tag;
  }
  const uint64_t state = (tag * 0x100000001B3ull + (uint64_t)gen) * 2 + (quad ? 1 : 0);
  if (!quad) {
    eng.solve(b, x, state, s, guess, rtol);
    return;
  }
  UBT_CHECK(b && x && b != x, UBT_ERR_ARG_IDN, "the solve needs distinct b and x");
  batch->quad_order(UBT_FP64, b, bq.ptr, true, s);
  if (guess) batch->quad_order(UBT_FP64, x, xq.ptr, true, s);
  eng.solve(bq.as<double>(), xq.as<double>(), state, s, guess, rtol);
  batch->quad_order(UBT_FP64, xq.ptr, x, false, s);
}

template void BatchSolver::apply_pc<double>(const double *, double *, hipStream_t);
template void BatchSolver::apply_pc<float>(const float *, float *, hipStream_t);

}  // namespace ubt
