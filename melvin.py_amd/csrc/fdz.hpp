// Fused Fourier-x / finite-difference-z step (the Rayleigh-Benard body of
// rayleigh_benard_convection.py:95-145) over a table of jobs.
//
// Spectral state (nn, nz) complex per field (omega, T): row = x mode
// kx in [0, nn), column = z grid point.  One evaluation of the explicit
// terms ("phase") is two kernels:
//
//   cols  per block of z columns: Hermitian-packed inverse x transforms of
//         (ux_s + i uz_s) and (omega + i T), the four products ux*f, uz*f
//         in physical space, two packed forward x transforms, keep
//         kx < nn, scale 1/nx  ->  Qx_f, Qz_f.
//   rows  one wavefront per (x mode, field): the flux divergence
//         D = i sx(kx) Qx + Dz(Qz) (the periodic x difference is its exact
//         modified wavenumber; the z difference acts along the row, and both
//         commute with the x transform), buoyancy -Pr Ra i kx T for omega,
//         the Adams-Bashforth / Adams-Moulton ring, the update (the
//         reference's explicit one, Integrator.py:53-56, or theta-implicit
//         diffusion solved per mode), the spectral wall writes
//         (rayleigh_benard_convection.py:113-142), and for omega the FP64
//         Poisson solve psi = solve(-omega) of the new state
//         (LaplacianSolver.py:70-78) with ux_s = -Dz(psi), uz_s = i kx psi
//         (utility.py:62-79): the velocity of the next evaluation.
//
// Schemes: EXPLICIT and SEMI_IMPLICIT run one phase per step; PC runs two
// (AB predictor into a scratch state, AM corrector from the old state).
// The tridiagonal solves run in FP64 whatever the state precision: each
// lane of the wave owns a contiguous chunk of the row, Thomas's two
// first-order recurrences are composed into affine maps per chunk and
// scanned across the 64 lanes with shuffles.
#pragma once
#include <hip/hip_runtime.h>
#include "../../include/melvin_hip.h"
#include "diag.hpp"
#include "transforms.hpp"

namespace mlv {

enum FdzStage : int { FDZ_MAIN = 0, FDZ_PRED = 1, FDZ_CORR = 2, FDZ_PROLOGUE = 3 };
constexpr int kFdzWaves = 4;   // rows (mode, field) per row block of one wavefront per row
// Row kernel variants: KT row elements per thread and WPR wavefronts per
// row.  A row of nz <= 256 points takes one wavefront (K = ceil(nz / 64)
// elements per lane, blocks of kFdzWaves rows); a longer row takes two
// (K = ceil(nz / 128), a block per row), which halves each wavefront's
// serial chain of instructions -- a single job has fewer rows than the
// device has SIMDs, so that chain sets the kernel's time.  The choice
// depends on the job alone, so a job computes the same bits in any batch.
// A job runs in the smallest variant of its WPR with KT >= its K.
constexpr int kFdzVariants = 5;
constexpr int kFdzKts[kFdzVariants] = {1, 2, 4, 4, 6};
constexpr int kFdzWprs[kFdzVariants] = {1, 1, 1, 2, 2};
constexpr int kFdzMaxNz = 768;
__host__ __device__ constexpr int fdz_wpr(int nz) { return nz > 256 ? 2 : 1; }
// rows per block and threads per block of a variant with wpr wavefronts per row
__host__ __device__ constexpr int fdz_rows_per_block(int wpr) { return wpr == 1 ? kFdzWaves : 1; }
__host__ __device__ constexpr int fdz_row_threads(int wpr) {
  return 64 * wpr * fdz_rows_per_block(wpr);
}

template <typename T>
struct FdzJob {
  int nx, nz, lognx, nn, K, kt, wpr, fd_order, scheme, q, curr0, par0;
  double fx, fx2, dx, dz, dt, alpha, Pr, Ra;
  double lo[3];            // interior off-diagonal of each tridiagonal operator
  cx<T>* state[2][2];      // [buffer][omega, T]
  cx<T>* star[2];          // predicted state (PC)
  cx<T>* hist[2];          // ring per field: q x (nn, nz)
  cx<T>* comp[2];          // single precision: residual of the state (value = state - comp)
  cx<T>* psi[2];           // per buffer: walled psi of that buffer's omega, or NULL
  cx<T>* vel[3][2];        // [buffer 0, 1, predicted][ux_s, uz_s]
  cx<T>* qw[4];            // Qx_omega, Qz_omega, Qx_T, Qz_T
  T* phys[4];              // optional physical ux, uz, omega, T
  double2* fac[3];         // (cp, 1/den) per operator: Poisson, theta(omega), theta(T)
  // device-side time keeping and diagnostics (diag.hpp): the row kernel of
  // a step's last phase advances t and loops; diag holds the column kernel's
  // per-block partials (nbC blocks of phase 0), the header and the tracker
  // records; refactor is set by the CFL update when it changed dt (the
  // theta tables are rebuilt on the device right after)
  double t;
  long long loops;
  int nbC, refactor;
  double* diag;
};

// flags of phase 0's column kernel: physical outputs, diagnostics partials
// (with the CFL cadence from bit kFdzDiagShift up: a job writes partials on
// the loops whose count before the loop is a multiple of it)
constexpr int kFdzWritePhys = 1, kFdzDiag = 2, kFdzDiagShift = 8;

struct FdzLaunch {
  int njobs;
  int ncol[2];                       // column blocks of phases 0 and 1
  int nrow[3][kFdzVariants];         // row blocks of phases 0, 1 and the prologue
  size_t ldsrow[kFdzVariants], ldscol;  // LDS of each row variant's blocks
  int zstride;                       // LDS elements per row of the row kernel
  int *prefcol[2];
  int *prefrow[3][kFdzVariants];
};

// Points per thread of the column kernel: 16 below N = 2048 (two columns
// per block at 1024, so a single job's z columns spread over twice as many
// CUs as at 32 points), 32 at 2048 (a block must hold a column pair).
template <typename T>
__host__ __device__ constexpr int fdz_points(int lognx) {
  return lognx >= 11 ? 32 : 16;
}
// Columns per block of the column kernel (two LDS slots each, in pairs of
// columns); 0 when the LDS slots do not hold two columns.
template <typename T>
inline int fdz_cols_per_block(int lognx) {
  const int s = ((kNT * fdz_points<T>(lognx)) >> lognx) / 2;
  return s < 2 ? 0 : (s > kNT ? kNT : s);
}
template <typename T>
inline size_t fdz_cols_lds_bytes(int lognx) {
  return (size_t)((kNT * fdz_points<T>(lognx)) >> lognx) * fft_ldsn(1 << lognx) * sizeof(cx<T>);
}

template <typename T>
hipError_t fdz_phase(const FdzJob<T>* jobs, const FdzLaunch& fl, int phase, int flags, int step,
                     const cx<T>* tw, hipStream_t s, hipEvent_t* ev = nullptr);
// After a step whose phase-0 column kernel ran with kFdzDiag: for each job
// whose check loop it was, the tracker record and CFL rule (dt, theta
// bands), then the theta tables of the jobs whose dt changed (max_nn =
// largest nn of the batch).
template <typename T>
hipError_t fdz_diag(FdzJob<T>* jobs, int njobs, int max_nn, double cutoff, int cadence,
                    hipStream_t s);
template <typename T>
hipError_t fdz_prologue(const FdzJob<T>* jobs, const FdzLaunch& fl, int step, hipStream_t s);
// (Re)build the factor tables of one job: which = 0 Poisson, 1 theta(omega),
// 2 theta(T); theta tables depend on dt.
hipError_t fdz_factors(double2* fac, int nn, int nz, int K, int wpr, double fxabs, double dz,
                       int which, double a, hipStream_t s);
template <typename T>
hipError_t fdz_set_dt(FdzJob<T>* jobs, int job, double dt, double lo_w, double lo_t,
                      hipStream_t s);
template <typename T>
hipError_t fdz_prepare(const FdzLaunch& fl);

}  // namespace mlv
