// The transport sweep (psi = L^{-1} r, streaming + removal) used as the
// preconditioner's streaming stage.
#pragma once

#include "operator.hpp"

namespace ubt {

struct Sweep {
  Operator *op = nullptr;
  int n_levels = 0;
  int64_t n_lagged = 0;             // upwind edges lagged to break cycles
  std::vector<int64_t> level_off;   // DG: items of level l are [level_off[l], level_off[l+1])
  DevBuf cells, angles, upwind;     // DG: per item, level-ordered; upwind = face bit mask
  std::vector<int> h_cells, h_angles;  // host copies (kept on request, for batching)
  std::vector<unsigned char> h_up;
  void create(Operator *o, bool keep_host = false);
  // psi = L^{-1} r, or (adj) psi = L^{-T} r
  void apply(int prec, const void *r, void *psi, hipStream_t s, bool adj = false) const;
};

}  // namespace ubt
