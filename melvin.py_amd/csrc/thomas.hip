// Per-mode tridiagonal solve of the FD-z Poisson problem
// (LaplacianSolver.py:21-56, 70-78): for every x mode n a Thomas sweep
// along z with the host-precomputed elimination factors of the constant
// matrix (diagonal -((2*pi*n/lx)^2 + 2/dz^2), off-diagonals 1/dz^2,
// identity wall rows).  FP64 throughout; the state type only changes the
// load and store.  The arithmetic matches a NumPy Thomas sweep:
//   d_j = (r_j - l_j d_{j-1}) * (1/den_j),   x_j = d_j - c_j x_{j+1}.
#include "ops.hpp"

namespace mlv {

template <typename T>
__global__ void __launch_bounds__(64) k_thomas(const cx<T>* __restrict__ rhs,
                                               cx<T>* __restrict__ out,
                                               double2* __restrict__ dwork,
                                               const double* __restrict__ cp,
                                               const double* __restrict__ rden, double lower,
                                               double sign, int nn, int nz) {
  const int n = blockIdx.x * 64 + threadIdx.x;
  if (n >= nn) return;
  const size_t row = (size_t)n * nz;
  double2 d;
  {
    const cx<T> r = rhs[row];
    const double s = rden[row];
    d = make_double2((sign * (double)r.x) * s, (sign * (double)r.y) * s);
    dwork[row] = d;
  }
  for (int j = 1; j < nz; ++j) {
    const cx<T> r = rhs[row + j];
    const double l = (j == nz - 1) ? 0.0 : lower;
    const double s = rden[row + j];
    d = make_double2((sign * (double)r.x - l * d.x) * s, (sign * (double)r.y - l * d.y) * s);
    dwork[row + j] = d;
  }
  double2 x = d;
  out[row + nz - 1] = mk<T>((T)x.x, (T)x.y);
  for (int j = nz - 2; j >= 0; --j) {
    const double c = cp[row + j];
    const double2 dj = dwork[row + j];
    x = make_double2(dj.x - c * x.x, dj.y - c * x.y);
    out[row + j] = mk<T>((T)x.x, (T)x.y);
  }
}

template <typename T>
hipError_t launch_thomas(const void* rhs, void* out, void* dwork, const double* cp,
                         const double* rden, double lower, double sign, int nn, int nz,
                         hipStream_t s) {
  hipLaunchKernelGGL(k_thomas<T>, dim3((nn + 63) / 64), dim3(64), 0, s, (const cx<T>*)rhs,
                     (cx<T>*)out, (double2*)dwork, cp, rden, lower, sign, nn, nz);
  return hipGetLastError();
}

template hipError_t launch_thomas<float>(const void*, void*, void*, const double*, const double*,
                                         double, double, int, int, hipStream_t);
template hipError_t launch_thomas<double>(const void*, void*, void*, const double*,
                                          const double*, double, double, int, int, hipStream_t);

}  // namespace mlv
