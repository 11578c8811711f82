// Block -> (job, tile) mapping shared by the fused batch kernels.
#pragma once
#include <hip/hip_runtime.h>

#include "fft.hpp"  // wave_sync

namespace mlv {

// Job of logical block b: pref is the exclusive prefix of blocks per job.
__device__ __forceinline__ int find_job(const int* __restrict__ pref, int njobs, int b) {
  int lo = 0, hi = njobs;  // pref[lo] <= b < pref[hi]
  while (hi - lo > 1) {
    const int mid = (lo + hi) >> 1;
    if (pref[mid] <= b)
      lo = mid;
    else
      hi = mid;
  }
  return lo;
}

__device__ __forceinline__ int ring(int c0, int step, int q) { return (c0 + step) % q; }

// Workgroups of a launch are dealt to the 8 XCDs round robin (block b runs
// on XCD b % 8).  Give each XCD a contiguous range of logical blocks, so
// neighbouring tiles -- which share 128-byte lines whenever a tile is
// narrower than a line -- are fetched through the same L2.  The grid is
// padded to a multiple of 8; returns -1 for padding blocks.
constexpr int kXcds = 8;
__device__ __forceinline__ int xcd_logical(int nlogical) {
  const int per = (nlogical + kXcds - 1) / kXcds;
  const int b = blockIdx.x;
  const int l = (b % kXcds) * per + b / kXcds;
  return l < nlogical ? l : -1;
}
static inline unsigned xcd_grid(int n) { return (unsigned)((n + kXcds - 1) / kXcds * kXcds); }

}  // namespace mlv

#define MLV_SET_LDS_ONCE(kern, bytes)                                                     \
  do {                                                                                    \
    static size_t done_ = 0;                                                              \
    if (done_ < (bytes)) {                                                                \
      (void)hipFuncSetAttribute(reinterpret_cast<const void*>(kern),                      \
                                hipFuncAttributeMaxDynamicSharedMemorySize, (int)(bytes)); \
      done_ = (bytes);                                                                    \
    }                                                                                     \
  } while (0)
