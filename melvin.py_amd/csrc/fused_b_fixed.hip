// Fused step, z-pass kernel B compiled for one z size at a time (2^9 ..
// 2^11 points): a batch whose jobs all share that size launches these, whose
// register allocation is their own instead of the largest over every size
// the generic kernel (fused_b.hip) handles.  The body is in fused_zpass.hpp.
#include "fused_zpass.hpp"

namespace mlv {

template <typename T, int V, int SPLIT, int PLAN, int LOGN>
__global__ void __launch_bounds__(kNT)
    __attribute__((amdgpu_waves_per_eu(sizeof(T) == 4 ? 4 : 1)))
    k_stepB_fixed(const JobDev<T>* __restrict__ jobs, const int* __restrict__ pref, int njobs,
                  int stage, int flags, const cx<T>* __restrict__ tw) {
  extern __shared__ char smem_raw[];
  cx<T>* lds = reinterpret_cast<cx<T>*>(smem_raw);
  const int lb = xcd_logical(pref[njobs]);
  if (lb < 0) return;
  roleB_fixed<T, V, SPLIT, PLAN, LOGN>(jobs, pref, njobs, lb, stage, flags, tw, lds);
}

template <typename T, int V, int SPLIT, int PLAN, int LOGN>
static void launch_one(const JobDev<T>* jobs, const BatchLaunch& bl, int stage, int flags,
                       const cx<T>* tw, hipStream_t s) {
  MLV_SET_LDS_ONCE((k_stepB_fixed<T, V, SPLIT, PLAN, LOGN>), bl.ldsB[V]);
  hipLaunchKernelGGL((k_stepB_fixed<T, V, SPLIT, PLAN, LOGN>), dim3(xcd_grid(bl.nB[stage][V])),
                     dim3(kNT), bl.ldsB[V], s, jobs, bl.prefB[stage][V], bl.njobs, stage, flags,
                     tw);
}

template <typename T, int V, int SPLIT, int PLAN>
static void launch_logn(const JobDev<T>* jobs, const BatchLaunch& bl, int stage, int flags,
                        const cx<T>* tw, hipStream_t s) {
  switch (bl.lognz_all) {
    case 9:
      launch_one<T, V, SPLIT, PLAN, 9>(jobs, bl, stage, flags, tw, s);
      break;
    case 10:
      launch_one<T, V, SPLIT, PLAN, 10>(jobs, bl, stage, flags, tw, s);
      break;
    case 11:
      launch_one<T, V, SPLIT, PLAN, 11>(jobs, bl, stage, flags, tw, s);
      break;
    default:
      break;
  }
}

template <typename T, int V>
static void launch_v(const JobDev<T>* jobs, const BatchLaunch& bl, int stage, int flags,
                     const cx<T>* tw, hipStream_t s) {
  // single precision with mean profiles: SPLIT = 1 (as in fused_b.hip)
  const bool split = sizeof(T) == 4 && bl.ldsM;
  if (bl.planB == kPlanTwoSweep) {
    if (split)
      launch_logn<T, V, sizeof(T) == 4 ? 1 : 0, kPlanTwoSweep>(jobs, bl, stage, flags, tw, s);
    else
      launch_logn<T, V, 0, kPlanTwoSweep>(jobs, bl, stage, flags, tw, s);
  } else {
    if (split)
      launch_logn<T, V, sizeof(T) == 4 ? 1 : 0, kPlanSplit>(jobs, bl, stage, flags, tw, s);
    else
      launch_logn<T, V, 0, kPlanSplit>(jobs, bl, stage, flags, tw, s);
  }
}

template <typename T>
void launch_zpass_fixed(int v, const JobDev<T>* jobs, const BatchLaunch& bl, int stage, int flags,
                        const cx<T>* tw, hipStream_t s) {
  switch (v) {
    case BV_VORT1:
      launch_v<T, BV_VORT1>(jobs, bl, stage, flags, tw, s);
      break;
    case BV_VORT2:
      launch_v<T, BV_VORT2>(jobs, bl, stage, flags, tw, s);
      break;
    case BV_VORT3:
      launch_v<T, BV_VORT3>(jobs, bl, stage, flags, tw, s);
      break;
    case BV_MHD1:
      launch_v<T, BV_MHD1>(jobs, bl, stage, flags, tw, s);
      break;
    case BV_MHD2:
      launch_v<T, BV_MHD2>(jobs, bl, stage, flags, tw, s);
      break;
    default:
      break;
  }
}

template <typename T, int V, int L>
static void prepare_vl(const BatchLaunch& bl) {
  MLV_SET_LDS_ONCE((k_stepB_fixed<T, V, 0, kPlanSplit, L>), bl.ldsB[V]);
  MLV_SET_LDS_ONCE((k_stepB_fixed<T, V, 0, kPlanTwoSweep, L>), bl.ldsB[V]);
  if constexpr (sizeof(T) == 4) {
    MLV_SET_LDS_ONCE((k_stepB_fixed<T, V, 1, kPlanSplit, L>), bl.ldsB[V]);
    MLV_SET_LDS_ONCE((k_stepB_fixed<T, V, 1, kPlanTwoSweep, L>), bl.ldsB[V]);
  }
}

template <typename T, int V>
static void prepare_v(const BatchLaunch& bl) {
  if (!bl.ldsB[V]) return;
  if (bl.lognz_all == 9) prepare_vl<T, V, 9>(bl);
  if (bl.lognz_all == 10) prepare_vl<T, V, 10>(bl);
  if (bl.lognz_all == 11) prepare_vl<T, V, 11>(bl);
}

template <typename T>
void prepare_zpass_fixed(const BatchLaunch& bl) {
  prepare_v<T, BV_VORT1>(bl);
  prepare_v<T, BV_VORT2>(bl);
  prepare_v<T, BV_VORT3>(bl);
  prepare_v<T, BV_MHD1>(bl);
  prepare_v<T, BV_MHD2>(bl);
}

template void launch_zpass_fixed<float>(int, const JobDev<float>*, const BatchLaunch&, int, int,
                                        const cx<float>*, hipStream_t);
template void launch_zpass_fixed<double>(int, const JobDev<double>*, const BatchLaunch&, int, int,
                                         const cx<double>*, hipStream_t);
template void prepare_zpass_fixed<float>(const BatchLaunch&);
template void prepare_zpass_fixed<double>(const BatchLaunch&);

}  // namespace mlv
