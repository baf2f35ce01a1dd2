// The transport sweep (psi = L^{-1} r, streaming + removal) used as the
// preconditioner's streaming stage.
#pragma once

#include "operator.hpp"

namespace ubt {

struct Sweep {
  Operator *op = nullptr;
  int n_levels = 0;                 // the most levels of any direction
  int64_t n_lagged = 0;             // upwind edges lagged to break cycles
  int threads = 64;                 // DG: workgroup size (the widest level, at most 1024)
  // DG, per direction: items (cell, upwind face mask) in level order, level
  // boundaries lvl[base[a] .. base[a + 1]) of direction a
  DevBuf cells, upwind, lvl, base;
  std::vector<int> h_cells, h_base;  // host copies (kept on request, for batching)
  std::vector<unsigned char> h_up;
  std::vector<int64_t> h_lvl;
  void create(Operator *o, bool keep_host = false);
  // psi = L^{-1} r, or (adj) psi = L^{-T} r
  void apply(int prec, const void *r, void *psi, hipStream_t s, bool adj = false) const;
};

}  // namespace ubt
