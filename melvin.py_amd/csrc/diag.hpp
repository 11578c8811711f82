// Layout of the on-device diagnostics shared by the fused batches (fully
// spectral and FD-z): per-block partials written by the pass that holds the
// physical velocities, then a header and a ring of tracker records.
#pragma once
#include <stddef.h>

namespace mlv {

// partials: max ux, max uz, then the Neumaier (sum, compensation) pairs of
// sum(ux^2 + uz^2), sum(w^2), sum(f1 * uz): the three sums, then the three
// compensations
constexpr int kDiagPart = 8;
// records: loop, t, dt of the step, KE, max ux, max uz, next dt, enstrophy
// 0.5 mean(w^2), flux mean(f1 * uz) of the first advected scalar (NaN if none)
constexpr int kDiagRec = 9;
constexpr int kDiagCap = 4096;  // records kept per job (the oldest are overwritten)
// header after the partials: record count, CFL breach flag, two spare
constexpr int kDiagHdr = 4;
inline constexpr size_t diag_doubles(int nblocks) {
  return (size_t)kDiagPart * nblocks + kDiagHdr + (size_t)kDiagRec * kDiagCap;
}

}  // namespace mlv
