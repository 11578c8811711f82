// Fused step, x-inverse kernels: A (spectral prologue + inverse x pass)
// and M (single-precision x-mean profiles); the bodies are in fused_xinv.hpp.
#include "fused_xinv.hpp"

namespace mlv {

template <typename T, int SHAPE>
__global__ void __launch_bounds__((xinv_threads<T, SHAPE>()))
    k_stepA(const JobDev<T>* __restrict__ jobs, const int* __restrict__ pref, int njobs,
            int stage, int step, const cx<T>* __restrict__ tw) {
  extern __shared__ char smem_raw[];
  cx<T>* lds = reinterpret_cast<cx<T>*>(smem_raw);
  const int lb = xcd_logical(pref[njobs]);
  if (lb < 0) return;
  roleA<T, SHAPE>(jobs, pref, njobs, lb, stage, step, tw, lds);
}

// A batch whose jobs all have 2^LOGN points along x (fixed_logn).
template <typename T, int SHAPE, int LOGN>
__global__ void __launch_bounds__((xinv_threads<T, SHAPE>()))
    k_stepA_fixed(const JobDev<T>* __restrict__ jobs, const int* __restrict__ pref, int njobs,
                  int stage, int step, const cx<T>* __restrict__ tw) {
  extern __shared__ char smem_raw[];
  cx<T>* lds = reinterpret_cast<cx<T>*>(smem_raw);
  const int lb = xcd_logical(pref[njobs]);
  if (lb < 0) return;
  roleA_fixed<T, SHAPE, LOGN>(jobs, pref, njobs, lb, stage, step, tw, lds);
}

// Grid: jobs x (kMeanCh / 2) op pairs.
template <typename T>
__global__ void __launch_bounds__(kNT) k_stepM(const JobDev<T>* __restrict__ jobs, int stage,
                                               int step, const cx<T>* __restrict__ tw) {
  extern __shared__ char smem_raw[];
  cx<T>* lds = reinterpret_cast<cx<T>*>(smem_raw);
  roleM<T>(jobs[blockIdx.x / (kMeanCh / 2)], blockIdx.x % (kMeanCh / 2), stage, step, tw, lds);
}

// --------------------------------------------------------------------------
// Launchers
// --------------------------------------------------------------------------
template <typename T>
void launch_xinv(const JobDev<T>* jobs, const BatchLaunch& bl, int stage, int step,
                 const cx<T>* tw, hipStream_t s) {
  if (bl.nA[stage] > 0) {
    const dim3 grid(xcd_grid(bl.nA[stage]));
    const int* pref = bl.prefA[stage];
    if (fixed_logn(bl.lognx_all)) {
#define MLV_A_FIXED(SH, L)                                                                  \
  if (bl.shapeA == SH && bl.lognx_all == L) {                                               \
    MLV_SET_LDS_ONCE((k_stepA_fixed<T, SH, L>), bl.ldsA);                                   \
    hipLaunchKernelGGL((k_stepA_fixed<T, SH, L>), grid, dim3(xinv_threads<T, SH>()), bl.ldsA, \
                       s, jobs, pref, bl.njobs, stage, step, tw);                           \
  }
      MLV_A_FIXED(0, 9) MLV_A_FIXED(0, 10) MLV_A_FIXED(0, 11)
      MLV_A_FIXED(1, 9) MLV_A_FIXED(1, 10) MLV_A_FIXED(1, 11)
      if constexpr (sizeof(T) == 4) {
        MLV_A_FIXED(2, 9) MLV_A_FIXED(2, 10)
      }
#undef MLV_A_FIXED
    } else if (bl.shapeA == 1) {
      MLV_SET_LDS_ONCE((k_stepA<T, 1>), bl.ldsA);
      hipLaunchKernelGGL((k_stepA<T, 1>), grid, dim3(xinv_threads<T, 1>()), bl.ldsA, s, jobs,
                         pref, bl.njobs, stage, step, tw);
    } else {
      MLV_SET_LDS_ONCE((k_stepA<T, 0>), bl.ldsA);
      hipLaunchKernelGGL((k_stepA<T, 0>), grid, dim3(xinv_threads<T, 0>()), bl.ldsA, s, jobs,
                         pref, bl.njobs, stage, step, tw);
    }
  }
  if constexpr (sizeof(T) == 4) {
    if (bl.ldsM) {
      MLV_SET_LDS_ONCE(k_stepM<T>, bl.ldsM);
      hipLaunchKernelGGL(k_stepM<T>, dim3(bl.njobs * (kMeanCh / 2)), dim3(kNT), bl.ldsM, s, jobs,
                         stage, step, tw);
    }
  }
}

template <typename T>
void prepare_xinv(const BatchLaunch& bl) {
  if (bl.ldsAp[0]) MLV_SET_LDS_ONCE((k_stepA<T, 0>), bl.ldsAp[0]);
  if (bl.okA[1] && bl.ldsAp[1]) MLV_SET_LDS_ONCE((k_stepA<T, 1>), bl.ldsAp[1]);
#define MLV_A_FIXED(SH, L)                                                   \
  if (bl.lognx_all == L && bl.ldsAp[SH] && bl.okA[SH])                        \
    MLV_SET_LDS_ONCE((k_stepA_fixed<T, SH, L>), bl.ldsAp[SH]);
  MLV_A_FIXED(0, 9) MLV_A_FIXED(0, 10) MLV_A_FIXED(0, 11)
  MLV_A_FIXED(1, 9) MLV_A_FIXED(1, 10) MLV_A_FIXED(1, 11)
  if constexpr (sizeof(T) == 4) {
    MLV_A_FIXED(2, 9) MLV_A_FIXED(2, 10)
  }
#undef MLV_A_FIXED
  if constexpr (sizeof(T) == 4) {
    if (bl.ldsM) MLV_SET_LDS_ONCE(k_stepM<T>, bl.ldsM);
  }
}

template void launch_xinv<float>(const JobDev<float>*, const BatchLaunch&, int, int,
                                 const cx<float>*, hipStream_t);
template void launch_xinv<double>(const JobDev<double>*, const BatchLaunch&, int, int,
                                  const cx<double>*, hipStream_t);
template void prepare_xinv<float>(const BatchLaunch&);
template void prepare_xinv<double>(const BatchLaunch&);

}  // namespace mlv
