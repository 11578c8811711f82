// Fused step, update kernel C, the dt setter and the device CFL /
// diagnostics kernel; the update body is in fused_update.hpp.
#include "fused_update.hpp"

namespace mlv {

// Narrow single-precision blocks are held to 128 VGPRs (four wavefronts per
// SIMD, matching the four blocks per CU their LDS allows).
template <typename T, int WIDE>
__global__ void __launch_bounds__(kNT)
    __attribute__((amdgpu_waves_per_eu((WIDE || sizeof(T) == 8) ? 1 : 4)))
    k_stepC(const JobDev<T>* __restrict__ jobs, JobDev<T>* clock, const int* __restrict__ pref,
            int njobs, int stage, int step, const cx<T>* __restrict__ tw) {
  extern __shared__ char smem_raw[];
  cx<T>* lds = reinterpret_cast<cx<T>*>(smem_raw);
  const int lb = xcd_logical(pref[njobs]);
  if (lb < 0) return;
  roleC<T, WIDE>(jobs, clock, pref, njobs, lb, stage, step, tw, lds);
}

// A batch whose jobs all have 2^LOGN points along x (fixed_logn), tile
// shape SHAPE (cshape_*).
template <typename T, int SHAPE, int LOGN>
__global__ void __launch_bounds__(cshape_threads(SHAPE))
    __attribute__((amdgpu_waves_per_eu((SHAPE == 1 || sizeof(T) == 8) ? 1 : 4)))
    k_stepC_fixed(const JobDev<T>* __restrict__ jobs, JobDev<T>* clock,
                  const int* __restrict__ pref, int njobs, int stage, int step,
                  const cx<T>* __restrict__ tw) {
  extern __shared__ char smem_raw[];
  cx<T>* lds = reinterpret_cast<cx<T>*>(smem_raw);
  const int lb = xcd_logical(pref[njobs]);
  if (lb < 0) return;
  roleC_fixed<T, SHAPE, LOGN>(jobs, clock, pref, njobs, lb, stage, step, tw, lds);
}

template <typename T>
__global__ void k_set_dt(JobDev<T>* jobs, int j, double dt) {
  if (threadIdx.x == 0 && blockIdx.x == 0) jobs[j].dt = dt;
}

// One block per job, after a step whose z pass wrote diagnostics partials:
// tracker record and CFL rule (diag_apply).  The first advected scalar
// exists in vorticity and DDC jobs of two or more fields.
template <typename T>
__global__ void __launch_bounds__(kNT) k_diag(JobDev<T>* jobs, double cutoff, int cadence) {
  JobDev<T>& J = jobs[blockIdx.x];
  // loops has been advanced by this step's update kernel: the step was this
  // job's check loop when the count before it was a multiple of cadence
  if ((J.loops - 1) % cadence != 0) return;
  double dt = J.dt;
  const bool flux = J.problem != MLV_PROB_MHD && J.nfield >= 2;
  if (diag_apply<kNT>(J.diag, J.nbB, J.dx, J.dz, (double)J.nx * (double)J.nz, flux, cutoff, J.t,
                      J.loops, dt))
    J.dt = dt;
}


// --------------------------------------------------------------------------
// Launchers
// --------------------------------------------------------------------------
template <typename T>
void launch_update(const JobDev<T>* jobs, const BatchLaunch& bl, int stage, int step,
                   const cx<T>* tw, hipStream_t s) {
  if (bl.nC[stage] == 0) return;
  if (fixed_logn(bl.lognx_all)) {
#define MLV_C_FIXED(W, L)                                                                     \
  if (bl.wideC == W && bl.lognx_all == L) {                                                   \
    MLV_SET_LDS_ONCE((k_stepC_fixed<T, W, L>), bl.ldsC);                                      \
    hipLaunchKernelGGL((k_stepC_fixed<T, W, L>), dim3(xcd_grid(bl.nC[stage])),                \
                       dim3(cshape_threads(W)),                                               \
                       bl.ldsC, s, jobs, const_cast<JobDev<T>*>(jobs),                        \
                       (const int*)bl.prefC[stage], bl.njobs, stage, step, tw);               \
  }
    MLV_C_FIXED(0, 9) MLV_C_FIXED(0, 10) MLV_C_FIXED(0, 11)
    MLV_C_FIXED(1, 9) MLV_C_FIXED(1, 10) MLV_C_FIXED(1, 11)
    if constexpr (sizeof(T) == 4) {
      MLV_C_FIXED(2, 9) MLV_C_FIXED(2, 10) MLV_C_FIXED(2, 11)
    }
#undef MLV_C_FIXED
    return;
  }
  if (bl.wideC) {
    MLV_SET_LDS_ONCE((k_stepC<T, 1>), bl.ldsC);
    hipLaunchKernelGGL((k_stepC<T, 1>), dim3(xcd_grid(bl.nC[stage])), dim3(kNT), bl.ldsC, s,
                       jobs, const_cast<JobDev<T>*>(jobs), (const int*)bl.prefC[stage], bl.njobs,
                       stage, step, tw);
  } else {
    MLV_SET_LDS_ONCE((k_stepC<T, 0>), bl.ldsC);
    hipLaunchKernelGGL((k_stepC<T, 0>), dim3(xcd_grid(bl.nC[stage])), dim3(kNT), bl.ldsC, s,
                       jobs, const_cast<JobDev<T>*>(jobs), (const int*)bl.prefC[stage], bl.njobs,
                       stage, step, tw);
  }
}

template <typename T>
void prepare_update(const BatchLaunch& bl) {
  if (bl.ldsCp[1]) MLV_SET_LDS_ONCE((k_stepC<T, 1>), bl.ldsCp[1]);
  if (bl.ldsCp[0]) MLV_SET_LDS_ONCE((k_stepC<T, 0>), bl.ldsCp[0]);
#define MLV_C_FIXED(W, L) \
  if (bl.lognx_all == L && bl.ldsCp[W] && bl.okC[W])                                       \
    MLV_SET_LDS_ONCE((k_stepC_fixed<T, W, L>), bl.ldsCp[W]);
  MLV_C_FIXED(0, 9) MLV_C_FIXED(0, 10) MLV_C_FIXED(0, 11)
  MLV_C_FIXED(1, 9) MLV_C_FIXED(1, 10) MLV_C_FIXED(1, 11)
  if constexpr (sizeof(T) == 4) {
    MLV_C_FIXED(2, 9) MLV_C_FIXED(2, 10) MLV_C_FIXED(2, 11)
  }
#undef MLV_C_FIXED
}

template <typename T>
hipError_t fused_set_dt(JobDev<T>* jobs, int job, double dt, hipStream_t s) {
  hipLaunchKernelGGL(k_set_dt<T>, dim3(1), dim3(64), 0, s, jobs, job, dt);
  return hipGetLastError();
}

template <typename T>
hipError_t fused_diag(JobDev<T>* jobs, int njobs, double cutoff, int cadence, hipStream_t s) {
  hipLaunchKernelGGL(k_diag<T>, dim3(njobs), dim3(kNT), 0, s, jobs, cutoff, cadence);
  return hipGetLastError();
}

template void launch_update<float>(const JobDev<float>*, const BatchLaunch&, int, int,
                                   const cx<float>*, hipStream_t);
template void launch_update<double>(const JobDev<double>*, const BatchLaunch&, int, int,
                                    const cx<double>*, hipStream_t);
template void prepare_update<float>(const BatchLaunch&);
template void prepare_update<double>(const BatchLaunch&);
template hipError_t fused_diag<float>(JobDev<float>*, int, double, int, hipStream_t);
template hipError_t fused_diag<double>(JobDev<double>*, int, double, int, hipStream_t);
template hipError_t fused_set_dt<float>(JobDev<float>*, int, double, hipStream_t);
template hipError_t fused_set_dt<double>(JobDev<double>*, int, double, hipStream_t);

}  // namespace mlv
