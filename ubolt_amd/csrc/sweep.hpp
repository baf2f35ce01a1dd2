// The transport sweep (psi = L^{-1} r, streaming + removal) used as the
// preconditioner's streaming stage.
#pragma once

#include "operator.hpp"

namespace ubt {

struct Sweep;

// a sweep schedule as host lists: items (cell, angle, upwind mask) in level
// order, level boundaries lvl[base[g] .. base[g + 1]) of workgroup g
struct SweepLists {
  std::vector<int> cell, angle, base;
  std::vector<unsigned char> up;
  std::vector<int64_t> lvl;
  int64_t widest = 1;  // the widest level
};
// the schedule of s (kept on the host) regrouped into workgroups of up to G
// directions of one octant (box meshes; otherwise one direction each)
void group_schedule(const Sweep &s, int G, SweepLists &out);
// LDS trace slots of a schedule's workgroups (one problem's disc D; items
// and level boundaries local to the list, written at i0 in own / ups); the
// most slots any workgroup holds at once
int assign_trace_slots(const Disc &D, int slots, const int *cell, const int *angle, const unsigned char *up,
                       const int64_t *lvl, const int *base, int n_wg, int64_t i0, int *own, int *ups);

struct Sweep {
  Operator *op = nullptr;
  int n_levels = 0;                 // the most levels of any direction
  int64_t n_lagged = 0;             // upwind edges lagged to break cycles
  // psi = L^{-1} r exactly: a DG schedule with no lagged edge and no
  // reflective face (whose coupling the sweep leaves out)
  bool exact = false;
  int threads = 64;                 // DG: workgroup size (the widest level, at most 1024)
  // DG, per direction (one workgroup each): items (cell, angle, upwind face
  // mask) in level order, level boundaries lvl[base[g] .. base[g + 1]) of
  // direction g
  int n_groups = 0;
  DevBuf cells, angle, upwind, lvl, base;
  std::vector<int> h_cells, h_angle, h_base;  // host copies (kept on request, for batching)
  std::vector<unsigned char> h_up;
  std::vector<int64_t> h_lvl;
  // DG forward sweep records, per item in schedule order (structure of
  // arrays): the block inverse B^{-1} (nb x nb) and, per inflow slot, the
  // coupling -B^{-1} s_f F_up,f and the upwind cell (-1: unused slot), so the
  // sweep of an item is psi = B^{-1} r + sum_k H_k psi_up,k. Built in FP64 for
  // the operator's current group and cross sections, stored in both
  // precisions.
  int nfu = 0;                      // inflow slots per item: the most inflow faces of any item
  // per precision: the pipelined sweep form (narrow levels, records that fit
  // in registers three deep), whose records use the 16-byte grouped layout
  bool pipe64 = false, pipe32 = false;
  // LDS trace slots: per item its own slot (-1: no later item reads it)
  // and, per inflow slot, the upwind item's slot (in the one-field-per-load
  // record layout); trace_slots = the most any direction holds at once (0:
  // traces read from HBM; a precision uses them where they fit
  // kLdsTraceBytes)
  static constexpr int64_t kLdsTraceBytes = 64 * 1024;
  int trace_slots = 0;
  DevBuf own_slot, up_slot;
  // the adjoint sweep's per-item reversed angle and per-cell volumes
  DevBuf angle_adj, vcell64, vcell32;
  mutable DevBuf rec64, rec32, rnb;
  mutable int rec_group = -1;
  mutable uint64_t rec_gen = 0;
  void build_records(hipStream_t s) const;
  void alloc_records() const;
  // keep_host: the schedule stays on the host for a batch, which builds its
  // own records (this sweep's are then allocated by its first forward sweep);
  // otherwise the record storage is allocated here
  // lds_traces = false keeps every upwind trace read in HBM (a diagnostic
  // switch: the LDS path is the default wherever the slots fit)
  void create(Operator *o, bool keep_host = false, bool lds_traces = true);
  // psi = L^{-1} r, or (adj) psi = L^{-T} r
  void apply(int prec, const void *r, void *psi, hipStream_t s, bool adj = false) const;
};

}  // namespace ubt
