// Host-side launchers for the spectral transform kernels.
#pragma once
#include <hip/hip_runtime.h>
#include "fft.hpp"

namespace mlv {

// Threads per block of every FFT-carrying kernel.
constexpr int kNT = 256;

// Sequences of 2^logn one block of a column (x) pass transforms, and the
// LDS they take; the _row versions are for the row (z) passes.
template <typename T>
inline int slots_for(int logn) {
  return (kNT * fft_points_col<T>(logn)) >> logn;
}
template <typename T>
inline size_t fft_lds_bytes(int logn) {
  return (size_t)slots_for<T>(logn) * fft_ldsn(1 << logn) * sizeof(cx<T>);
}
template <typename T>
inline int slots_row(int logn) {
  return (kNT * fft_points_row<T>(logn)) >> logn;
}
template <typename T>
inline size_t fft_lds_bytes_row(int logn) {
  return (size_t)slots_row<T>(logn) * fft_ldsn(1 << logn) * sizeof(cx<T>);
}

// nrow: rows of the x-pass spectrum (default 2 nn + 1, the compact layout;
// nx for a full spectrum, rows in FFT order); scale multiplies the input
// of launch_xinv and the outputs of the c2r passes.
template <typename T>
hipError_t launch_xinv(int logn, const void* in, void* out, int nn, int nm, const void* tw,
                       hipStream_t s, int nrow = -1, double scale = 1.0);
template <typename T>
hipError_t launch_xfwd(int logn, const void* in, void* out, int nn, int nm, double scale,
                       const void* tw, hipStream_t s, int nrow = -1);
template <typename T>
hipError_t launch_zc2r(int logn, const void* in, void* out, int nx, int nm, const void* tw,
                       hipStream_t s, double scale = 1.0);
template <typename T>
hipError_t launch_zr2c(int logn, const void* in, void* out, int nx, int nm, const void* tw,
                       hipStream_t s);
template <typename T>
hipError_t launch_xc2r_cols(int logn, const void* in, void* out, int nn, int nz,
                            const void* tw, hipStream_t s, double scale = 1.0);
template <typename T>
hipError_t launch_xr2c_cols(int logn, const void* in, void* out, int nn, int nz,
                            double scale, const void* tw, hipStream_t s);

}  // namespace mlv
