// Host-side launchers for the spectral transform kernels.
#pragma once
#include <hip/hip_runtime.h>
#include "fft.hpp"

namespace mlv {

// Threads per block of every FFT-carrying kernel.
constexpr int kNT = 256;

// LDS bytes of one block transforming slots_for(logn) sequences of 2^logn.
template <typename T>
inline size_t fft_lds_bytes(int logn) {
  const int slots = (kNT * fft_points(logn)) >> logn;
  return (size_t)slots * fft_ldsn(1 << logn) * sizeof(cx<T>);
}
template <typename T>
inline int slots_for(int logn) {
  return (kNT * fft_points(logn)) >> logn;
}

template <typename T>
hipError_t launch_xinv(int logn, const void* in, void* out, int nn, int nm, const void* tw,
                       hipStream_t s);
template <typename T>
hipError_t launch_xfwd(int logn, const void* in, void* out, int nn, int nm, double scale,
                       const void* tw, hipStream_t s);
template <typename T>
hipError_t launch_zc2r(int logn, const void* in, void* out, int nx, int nm, const void* tw,
                       hipStream_t s);
template <typename T>
hipError_t launch_zr2c(int logn, const void* in, void* out, int nx, int nm, const void* tw,
                       hipStream_t s);
template <typename T>
hipError_t launch_xc2r_cols(int logn, const void* in, void* out, int nn, int nz,
                            const void* tw, hipStream_t s);
template <typename T>
hipError_t launch_xr2c_cols(int logn, const void* in, void* out, int nn, int nz,
                            double scale, const void* tw, hipStream_t s);

}  // namespace mlv
