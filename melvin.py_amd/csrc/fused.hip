// Fused fully-spectral step over a job table: the stage sequence and the
// one-time launch set-up (the kernels live in fused_a.hip, fused_b.hip and
// fused_c.hip; see fused.hpp).
#include "fused.hpp"

namespace mlv {

template <typename T>
hipError_t fused_stage(const JobDev<T>* jobs, const BatchLaunch& bl, int stage, int flags,
                       int step, const cx<T>* tw, hipStream_t s, hipEvent_t* ev) {
  if (ev) (void)hipEventRecord(ev[0], s);
  launch_xinv<T>(jobs, bl, stage, step, tw, s);
  if (ev) (void)hipEventRecord(ev[1], s);
  launch_zpass<T>(jobs, bl, stage, flags, tw, s);
  if (ev) (void)hipEventRecord(ev[2], s);
  launch_update<T>(jobs, bl, stage, step, tw, s);
  if (ev) (void)hipEventRecord(ev[3], s);
  return hipGetLastError();
}

// Raise the dynamic-LDS limit of every fused kernel once, outside any
// stream capture.
template <typename T>
hipError_t fused_prepare(const BatchLaunch& bl) {
  prepare_xinv<T>(bl);
  prepare_zpass<T>(bl);
  prepare_update<T>(bl);
  return hipGetLastError();
}

template hipError_t fused_prepare<float>(const BatchLaunch&);
template hipError_t fused_prepare<double>(const BatchLaunch&);
template hipError_t fused_stage<float>(const JobDev<float>*, const BatchLaunch&, int, int, int,
                                       const cx<float>*, hipStream_t, hipEvent_t*);
template hipError_t fused_stage<double>(const JobDev<double>*, const BatchLaunch&, int, int, int,
                                        const cx<double>*, hipStream_t, hipEvent_t*);

}  // namespace mlv
