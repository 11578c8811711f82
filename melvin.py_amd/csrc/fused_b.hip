// Fused step, z-pass kernel B; the body is in fused_zpass.hpp.
#include "fused_zpass.hpp"

namespace mlv {

// Single precision is held to 128 VGPRs: four wavefronts per SIMD, as many
// as the four blocks per CU its LDS allows at 1024 points.
template <typename T, int V, int SPLIT, int PLAN>
__global__ void __launch_bounds__(kNT)
    __attribute__((amdgpu_waves_per_eu(sizeof(T) == 4 ? 4 : 1)))
    k_stepB(const JobDev<T>* __restrict__ jobs, const int* __restrict__ pref, int njobs,
            int stage, int flags, const cx<T>* __restrict__ tw) {
  extern __shared__ char smem_raw[];
  cx<T>* lds = reinterpret_cast<cx<T>*>(smem_raw);
  const int lb = xcd_logical(pref[njobs]);
  if (lb < 0) return;
  roleB<T, V, SPLIT, PLAN>(jobs, pref, njobs, lb, stage, flags, tw, lds);
}

// --------------------------------------------------------------------------
// Launchers
// --------------------------------------------------------------------------
template <typename T, int V, int PLAN>
static void launchB_plan(const JobDev<T>* jobs, const BatchLaunch& bl, int stage, int flags,
                         const cx<T>* tw, hipStream_t s) {
  const dim3 grid(xcd_grid(bl.nB[stage][V]));
  const int* pref = bl.prefB[stage][V];
  if constexpr (sizeof(T) == 4) {
    if (bl.ldsM) {  // some job carries mean profiles
      MLV_SET_LDS_ONCE((k_stepB<T, V, 1, PLAN>), bl.ldsB[V]);
      hipLaunchKernelGGL((k_stepB<T, V, 1, PLAN>), grid, dim3(kNT), bl.ldsB[V], s, jobs, pref,
                         bl.njobs, stage, flags, tw);
      return;
    }
  }
  MLV_SET_LDS_ONCE((k_stepB<T, V, 0, PLAN>), bl.ldsB[V]);
  hipLaunchKernelGGL((k_stepB<T, V, 0, PLAN>), grid, dim3(kNT), bl.ldsB[V], s, jobs, pref,
                     bl.njobs, stage, flags, tw);
}

template <typename T, int V>
static void launchB(const JobDev<T>* jobs, const BatchLaunch& bl, int stage, int flags,
                    const cx<T>* tw, hipStream_t s) {
  if (bl.nB[stage][V] == 0) return;
  if (fixed_logn(bl.lognz_all)) {
    launch_zpass_fixed<T>(V, jobs, bl, stage, flags, tw, s);
    return;
  }
  if (bl.planB == kPlanTwoSweep)
    launchB_plan<T, V, kPlanTwoSweep>(jobs, bl, stage, flags, tw, s);
  else
    launchB_plan<T, V, kPlanSplit>(jobs, bl, stage, flags, tw, s);
}

template <typename T>
void launch_zpass(const JobDev<T>* jobs, const BatchLaunch& bl, int stage, int flags,
                  const cx<T>* tw, hipStream_t s) {
  launchB<T, BV_VORT1>(jobs, bl, stage, flags, tw, s);
  launchB<T, BV_VORT2>(jobs, bl, stage, flags, tw, s);
  launchB<T, BV_VORT3>(jobs, bl, stage, flags, tw, s);
  launchB<T, BV_MHD1>(jobs, bl, stage, flags, tw, s);
  launchB<T, BV_MHD2>(jobs, bl, stage, flags, tw, s);
}

template <typename T>
void prepare_zpass(const BatchLaunch& bl) {
#define MLV_B_LDS(V)                                                                     \
  if (bl.ldsB[V]) {                                                                      \
    MLV_SET_LDS_ONCE((k_stepB<T, V, 0, kPlanSplit>), bl.ldsB[V]);                        \
    MLV_SET_LDS_ONCE((k_stepB<T, V, 0, kPlanTwoSweep>), bl.ldsB[V]);                     \
    if constexpr (sizeof(T) == 4) {                                                      \
      MLV_SET_LDS_ONCE((k_stepB<T, V, 1, kPlanSplit>), bl.ldsB[V]);                      \
      MLV_SET_LDS_ONCE((k_stepB<T, V, 1, kPlanTwoSweep>), bl.ldsB[V]);                   \
    }                                                                                    \
  }
  MLV_B_LDS(BV_VORT1)
  MLV_B_LDS(BV_VORT2)
  MLV_B_LDS(BV_VORT3)
  MLV_B_LDS(BV_MHD1)
  MLV_B_LDS(BV_MHD2)
#undef MLV_B_LDS
  if (fixed_logn(bl.lognz_all)) prepare_zpass_fixed<T>(bl);
}

template void launch_zpass<float>(const JobDev<float>*, const BatchLaunch&, int, int,
                                  const cx<float>*, hipStream_t);
template void launch_zpass<double>(const JobDev<double>*, const BatchLaunch&, int, int,
                                   const cx<double>*, hipStream_t);
template void prepare_zpass<float>(const BatchLaunch&);
template void prepare_zpass<double>(const BatchLaunch&);

}  // namespace mlv
