// Operator-level spectral transforms (the SpectralTransformer mirror).
//
// Fully spectral 2-D, spectral (2nn+1, nm) <-> physical (nx, nz):
//   to_physical = k_xinv   (x pass: zero-pad rows to nx, inverse c2c along x)
//                 k_zc2r   (z pass: Hermitian-extend nm columns, inverse along z)
//   to_spectral = k_zr2c   (z pass: forward along z, keep kz < nm)
//                 k_xfwd   (x pass: forward along x, keep |kx| <= nn, scale)
// The intermediate between the passes is (nx, nm) complex, row-major, and is
// the only workspace.  Zero padding and truncation never touch memory.
//
// Fourier-x / FD-z, spectral (nn, nz) <-> physical (nx, nz):
//   k_xc2r_cols / k_xr2c_cols: one pass over z-column pairs along x.
//
// Real transforms pack two real sequences into one complex transform
// (rows x, x+1 for the z pass; columns z, z+1 for the FD-z x pass).  As in
// numpy's c2r, the imaginary parts of each sequence's DC bin and, when the
// input holds it, its Nyquist bin are ignored.
//
// The same kernels give numpy.fft's rfft / irfft / rfft2 / irfft2 (full
// half-spectra, unnormalised forward, 1/n inverse): nm = nz/2 + 1 or
// nn = nx/2 + 1 bins, and the x passes over all nx rows (nrow = nx).
#include "transforms.hpp"

namespace mlv {

// --------------------------------------------------------------------------
// x pass, spectral -> intermediate (inverse, zero padded)
// --------------------------------------------------------------------------
template <typename T, int LOGN>
__global__ void __launch_bounds__(kNT) k_xinv(const cx<T>* __restrict__ in,
                                              cx<T>* __restrict__ out, int nn, int nrow, int nm,
                                              T scale, const cx<T>* __restrict__ tw) {
  constexpr int N = 1 << LOGN;
  constexpr int P = fft_points_col<T>(LOGN);
  constexpr int TC = fft_slots<LOGN, kNT, P>();
  constexpr int LDSN = fft_ldsn(N);
  extern __shared__ char smem_raw[];
  cx<T>* lds = reinterpret_cast<cx<T>*>(smem_raw);
  const int c0 = blockIdx.x * TC;
  const int ncol = min(TC, nm - c0);
  for (int e = threadIdx.x; e < N * TC; e += kNT) {
    const int c = e % TC, x = e / TC;
    cx<T> v = czero<T>();
    if (c < ncol) {
      const int r = x <= nn ? x : (x >= N - nn ? x - N + nrow : -1);
      if (r >= 0) v = cscale<T>(in[(size_t)r * nm + c0 + c], scale);
    }
    lds[c * LDSN + fft_pad(x)] = v;
  }
  __syncthreads();
  fft_lds<T, LOGN, +1, kNT, P>(lds, ncol, tw);
  for (int e = threadIdx.x; e < N * TC; e += kNT) {
    const int c = e % TC, x = e / TC;
    if (c < ncol) out[(size_t)x * nm + c0 + c] = lds[c * LDSN + fft_pad(x)];
  }
}

// --------------------------------------------------------------------------
// x pass, intermediate -> spectral (forward, truncated, scaled)
// --------------------------------------------------------------------------
template <typename T, int LOGN>
__global__ void __launch_bounds__(kNT) k_xfwd(const cx<T>* __restrict__ in,
                                              cx<T>* __restrict__ out, int nn, int nrow, int nm,
                                              T scale, const cx<T>* __restrict__ tw) {
  constexpr int N = 1 << LOGN;
  constexpr int P = fft_points_col<T>(LOGN);
  constexpr int TC = fft_slots<LOGN, kNT, P>();
  constexpr int LDSN = fft_ldsn(N);
  extern __shared__ char smem_raw[];
  cx<T>* lds = reinterpret_cast<cx<T>*>(smem_raw);
  const int c0 = blockIdx.x * TC;
  const int ncol = min(TC, nm - c0);
  for (int e = threadIdx.x; e < N * TC; e += kNT) {
    const int c = e % TC, x = e / TC;
    cx<T> v = czero<T>();
    if (c < ncol) v = in[(size_t)x * nm + c0 + c];
    lds[c * LDSN + fft_pad(x)] = v;
  }
  __syncthreads();
  fft_lds<T, LOGN, -1, kNT, P>(lds, ncol, tw);
  for (int e = threadIdx.x; e < nrow * TC; e += kNT) {
    const int c = e % TC, r = e / TC;
    if (c < ncol) {
      const int x = r <= nn ? r : r + N - nrow;
      out[(size_t)r * nm + c0 + c] = cscale<T>(lds[c * LDSN + fft_pad(x)], scale);
    }
  }
}

// --------------------------------------------------------------------------
// z pass, intermediate -> physical (c2r of row pairs)
// --------------------------------------------------------------------------
template <typename T, int LOGN>
__global__ void __launch_bounds__(kNT) k_zc2r(const cx<T>* __restrict__ in,
                                              T* __restrict__ out, int nx, int nm, T scale,
                                              const cx<T>* __restrict__ tw) {
  constexpr int N = 1 << LOGN;  // nz
  constexpr int P = fft_points_row<T>(LOGN);
  constexpr int S = fft_slots<LOGN, kNT, P>();
  constexpr int LDSN = fft_ldsn(N);
  extern __shared__ char smem_raw[];
  cx<T>* lds = reinterpret_cast<cx<T>*>(smem_raw);
  const int x0 = blockIdx.x * 2 * S;
  const int active = min(S, (nx - x0 + 1) / 2);
  // zero the band the compact spectrum does not cover: k in [nm, N - nm]
  const int nzero = N - 2 * nm + 1;
  for (int e = threadIdx.x; e < active * nzero; e += kNT) {
    const int s = e / nzero, k = nm + e % nzero;
    lds[s * LDSN + fft_pad(k)] = czero<T>();
  }
  for (int e = threadIdx.x; e < active * nm; e += kNT) {
    const int s = e / nm, k = e % nm;
    const int xa = x0 + 2 * s, xb = xa + 1;
    const cx<T> a = in[(size_t)xa * nm + k];
    const cx<T> b = xb < nx ? in[(size_t)xb * nm + k] : czero<T>();
    cx<T>* z = lds + s * LDSN;
    if (k == 0 || 2 * k == N) {
      z[fft_pad(k)] = mk<T>(a.x, b.x);
    } else {
      z[fft_pad(k)] = mk<T>(a.x - b.y, a.y + b.x);
      z[fft_pad(N - k)] = mk<T>(a.x + b.y, b.x - a.y);
    }
  }
  __syncthreads();
  fft_lds<T, LOGN, +1, kNT, P>(lds, active, tw);
  for (int e = threadIdx.x; e < active * N; e += kNT) {
    const int s = e / N, n = e % N;
    const int xa = x0 + 2 * s, xb = xa + 1;
    const cx<T> v = lds[s * LDSN + fft_pad(n)];
    out[(size_t)xa * N + n] = v.x * scale;
    if (xb < nx) out[(size_t)xb * N + n] = v.y * scale;
  }
}

// --------------------------------------------------------------------------
// z pass, physical -> intermediate (r2c of row pairs, keep kz < nm)
// --------------------------------------------------------------------------
template <typename T, int LOGN>
__global__ void __launch_bounds__(kNT) k_zr2c(const T* __restrict__ in,
                                              cx<T>* __restrict__ out, int nx, int nm,
                                              const cx<T>* __restrict__ tw) {
  constexpr int N = 1 << LOGN;
  constexpr int P = fft_points_row<T>(LOGN);
  constexpr int S = fft_slots<LOGN, kNT, P>();
  constexpr int LDSN = fft_ldsn(N);
  extern __shared__ char smem_raw[];
  cx<T>* lds = reinterpret_cast<cx<T>*>(smem_raw);
  const int x0 = blockIdx.x * 2 * S;
  const int active = min(S, (nx - x0 + 1) / 2);
  for (int e = threadIdx.x; e < active * N; e += kNT) {
    const int s = e / N, n = e % N;
    const int xa = x0 + 2 * s, xb = xa + 1;
    const T a = in[(size_t)xa * N + n];
    const T b = xb < nx ? in[(size_t)xb * N + n] : T(0);
    lds[s * LDSN + fft_pad(n)] = mk<T>(a, b);
  }
  __syncthreads();
  fft_lds<T, LOGN, -1, kNT, P>(lds, active, tw);
  for (int e = threadIdx.x; e < active * nm; e += kNT) {
    const int s = e / nm, k = e % nm;
    const int xa = x0 + 2 * s, xb = xa + 1;
    const cx<T>* z = lds + s * LDSN;
    const cx<T> y = z[fft_pad(k)];
    const cx<T> ym = z[fft_pad((N - k) & (N - 1))];
    out[(size_t)xa * nm + k] = mk<T>(T(0.5) * (y.x + ym.x), T(0.5) * (y.y - ym.y));
    if (xb < nx) out[(size_t)xb * nm + k] = mk<T>(T(0.5) * (y.y + ym.y), T(0.5) * (ym.x - y.x));
  }
}

// --------------------------------------------------------------------------
// FD-z x transforms: spectral (nn, nz) <-> physical (nx, nz), column pairs
// --------------------------------------------------------------------------
template <typename T, int LOGN>
__global__ void __launch_bounds__(kNT) k_xc2r_cols(const cx<T>* __restrict__ in,
                                                   T* __restrict__ out, int nn, int nz, T scale,
                                                   const cx<T>* __restrict__ tw) {
  constexpr int N = 1 << LOGN;  // nx
  constexpr int P = fft_points_col<T>(LOGN);
  constexpr int S = fft_slots<LOGN, kNT, P>();
  constexpr int LDSN = fft_ldsn(N);
  extern __shared__ char smem_raw[];
  cx<T>* lds = reinterpret_cast<cx<T>*>(smem_raw);
  const int z0 = blockIdx.x * 2 * S;
  const int active = min(S, (nz - z0 + 1) / 2);
  const int ncols = 2 * active;
  const int nzero = N - 2 * nn + 1;  // k in [nn, N - nn]
  for (int e = threadIdx.x; e < active * nzero; e += kNT) {
    const int s = e % active, k = nn + e / active;
    lds[s * LDSN + fft_pad(k)] = czero<T>();
  }
  for (int e = threadIdx.x; e < active * nn; e += kNT) {
    const int s = e % active, k = e / active;
    const int za = z0 + 2 * s, zb = za + 1;
    const cx<T> a = in[(size_t)k * nz + za];
    const cx<T> b = zb < nz ? in[(size_t)k * nz + zb] : czero<T>();
    cx<T>* q = lds + s * LDSN;
    if (k == 0 || 2 * k == N) {
      q[fft_pad(k)] = mk<T>(a.x, b.x);
    } else {
      q[fft_pad(k)] = mk<T>(a.x - b.y, a.y + b.x);
      q[fft_pad(N - k)] = mk<T>(a.x + b.y, b.x - a.y);
    }
  }
  __syncthreads();
  fft_lds<T, LOGN, +1, kNT, P>(lds, active, tw);
  for (int e = threadIdx.x; e < ncols * N; e += kNT) {
    const int zc = e % ncols, x = e / ncols;
    const int z = z0 + zc;
    if (z < nz) {
      const cx<T> v = lds[(zc >> 1) * LDSN + fft_pad(x)];
      out[(size_t)x * nz + z] = ((zc & 1) ? v.y : v.x) * scale;
    }
  }
}

template <typename T, int LOGN>
__global__ void __launch_bounds__(kNT) k_xr2c_cols(const T* __restrict__ in,
                                                   cx<T>* __restrict__ out, int nn, int nz,
                                                   T scale, const cx<T>* __restrict__ tw) {
  constexpr int N = 1 << LOGN;
  constexpr int P = fft_points_col<T>(LOGN);
  constexpr int S = fft_slots<LOGN, kNT, P>();
  constexpr int LDSN = fft_ldsn(N);
  extern __shared__ char smem_raw[];
  cx<T>* lds = reinterpret_cast<cx<T>*>(smem_raw);
  const int z0 = blockIdx.x * 2 * S;
  const int active = min(S, (nz - z0 + 1) / 2);
  const int ncols = 2 * active;
  for (int e = threadIdx.x; e < ncols * N; e += kNT) {
    const int zc = e % ncols, x = e / ncols;
    const int z = z0 + zc;
    const T v = z < nz ? in[(size_t)x * nz + z] : T(0);
    T* q = reinterpret_cast<T*>(lds + (zc >> 1) * LDSN + fft_pad(x));
    q[zc & 1] = v;
  }
  __syncthreads();
  fft_lds<T, LOGN, -1, kNT, P>(lds, active, tw);
  for (int e = threadIdx.x; e < ncols * nn; e += kNT) {
    const int zc = e % ncols, k = e / ncols;
    const int z = z0 + zc;
    if (z < nz) {
      const cx<T>* q = lds + (zc >> 1) * LDSN;
      const cx<T> y = q[fft_pad(k)];
      const cx<T> ym = q[fft_pad((N - k) & (N - 1))];
      cx<T> r = (zc & 1) ? mk<T>(T(0.5) * (y.y + ym.y), T(0.5) * (ym.x - y.x))
                         : mk<T>(T(0.5) * (y.x + ym.x), T(0.5) * (y.y - ym.y));
      out[(size_t)k * nz + z] = cscale<T>(r, scale);
    }
  }
}

// --------------------------------------------------------------------------
// Host launchers
// --------------------------------------------------------------------------
#define MLV_LOGN_CASES(X) \
  X(2) X(3) X(4) X(5) X(6) X(7) X(8) X(9) X(10) X(11) X(12) X(13)

#define MLV_SET_LDS(kern, bytes)                                                          \
  do {                                                                                    \
    static bool done_ = false;                                                            \
    if (!done_) {                                                                         \
      (void)hipFuncSetAttribute(reinterpret_cast<const void*>(kern),                         \
                          hipFuncAttributeMaxDynamicSharedMemorySize, (int)(bytes));      \
      done_ = true;                                                                       \
    }                                                                                     \
  } while (0)

template <typename T>
hipError_t launch_xinv(int logn, const void* in, void* out, int nn, int nm, const void* tw,
                       hipStream_t s, int nrow, double scale) {
  if (nrow < 0) nrow = 2 * nn + 1;
  const int tc = slots_for<T>(logn);
  const dim3 grid((nm + tc - 1) / tc);
  const size_t lds = fft_lds_bytes<T>(logn);
  switch (logn) {
#define X(L)                                                                               \
  case L:                                                                                  \
    MLV_SET_LDS((k_xinv<T, L>), lds);                                                      \
    hipLaunchKernelGGL((k_xinv<T, L>), grid, dim3(kNT), lds, s, (const cx<T>*)in,          \
                       (cx<T>*)out, nn, nrow, nm, (T)scale, (const cx<T>*)tw);             \
    break;
    MLV_LOGN_CASES(X)
#undef X
    default:
      return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

template <typename T>
hipError_t launch_xfwd(int logn, const void* in, void* out, int nn, int nm, double scale,
                       const void* tw, hipStream_t s, int nrow) {
  if (nrow < 0) nrow = 2 * nn + 1;
  const int tc = slots_for<T>(logn);
  const dim3 grid((nm + tc - 1) / tc);
  const size_t lds = fft_lds_bytes<T>(logn);
  switch (logn) {
#define X(L)                                                                               \
  case L:                                                                                  \
    MLV_SET_LDS((k_xfwd<T, L>), lds);                                                      \
    hipLaunchKernelGGL((k_xfwd<T, L>), grid, dim3(kNT), lds, s, (const cx<T>*)in,          \
                       (cx<T>*)out, nn, nrow, nm, (T)scale, (const cx<T>*)tw);             \
    break;
    MLV_LOGN_CASES(X)
#undef X
    default:
      return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

template <typename T>
hipError_t launch_zc2r(int logn, const void* in, void* out, int nx, int nm, const void* tw,
                       hipStream_t s, double scale) {
  const int rows = 2 * slots_row<T>(logn);
  const dim3 grid((nx + rows - 1) / rows);
  const size_t lds = fft_lds_bytes_row<T>(logn);
  switch (logn) {
#define X(L)                                                                               \
  case L:                                                                                  \
    MLV_SET_LDS((k_zc2r<T, L>), lds);                                                      \
    hipLaunchKernelGGL((k_zc2r<T, L>), grid, dim3(kNT), lds, s, (const cx<T>*)in, (T*)out, \
                       nx, nm, (T)scale, (const cx<T>*)tw);                                \
    break;
    MLV_LOGN_CASES(X)
#undef X
    default:
      return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

template <typename T>
hipError_t launch_zr2c(int logn, const void* in, void* out, int nx, int nm, const void* tw,
                       hipStream_t s) {
  const int rows = 2 * slots_row<T>(logn);
  const dim3 grid((nx + rows - 1) / rows);
  const size_t lds = fft_lds_bytes_row<T>(logn);
  switch (logn) {
#define X(L)                                                                               \
  case L:                                                                                  \
    MLV_SET_LDS((k_zr2c<T, L>), lds);                                                      \
    hipLaunchKernelGGL((k_zr2c<T, L>), grid, dim3(kNT), lds, s, (const T*)in,              \
                       (cx<T>*)out, nx, nm, (const cx<T>*)tw);                             \
    break;
    MLV_LOGN_CASES(X)
#undef X
    default:
      return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

template <typename T>
hipError_t launch_xc2r_cols(int logn, const void* in, void* out, int nn, int nz,
                            const void* tw, hipStream_t s, double scale) {
  const int cols = 2 * slots_for<T>(logn);
  const dim3 grid((nz + cols - 1) / cols);
  const size_t lds = fft_lds_bytes<T>(logn);
  switch (logn) {
#define X(L)                                                                               \
  case L:                                                                                  \
    MLV_SET_LDS((k_xc2r_cols<T, L>), lds);                                                 \
    hipLaunchKernelGGL((k_xc2r_cols<T, L>), grid, dim3(kNT), lds, s, (const cx<T>*)in,     \
                       (T*)out, nn, nz, (T)scale, (const cx<T>*)tw);                       \
    break;
    MLV_LOGN_CASES(X)
#undef X
    default:
      return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

template <typename T>
hipError_t launch_xr2c_cols(int logn, const void* in, void* out, int nn, int nz,
                            double scale, const void* tw, hipStream_t s) {
  const int cols = 2 * slots_for<T>(logn);
  const dim3 grid((nz + cols - 1) / cols);
  const size_t lds = fft_lds_bytes<T>(logn);
  switch (logn) {
#define X(L)                                                                               \
  case L:                                                                                  \
    MLV_SET_LDS((k_xr2c_cols<T, L>), lds);                                                 \
    hipLaunchKernelGGL((k_xr2c_cols<T, L>), grid, dim3(kNT), lds, s, (const T*)in,         \
                       (cx<T>*)out, nn, nz, (T)scale, (const cx<T>*)tw);                   \
    break;
    MLV_LOGN_CASES(X)
#undef X
    default:
      return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

#define MLV_INST(T)                                                                         \
  template hipError_t launch_xinv<T>(int, const void*, void*, int, int, const void*,        \
                                     hipStream_t, int, double);                             \
  template hipError_t launch_xfwd<T>(int, const void*, void*, int, int, double,             \
                                     const void*, hipStream_t, int);                        \
  template hipError_t launch_zc2r<T>(int, const void*, void*, int, int, const void*,        \
                                     hipStream_t, double);                                  \
  template hipError_t launch_zr2c<T>(int, const void*, void*, int, int, const void*,        \
                                     hipStream_t);                                          \
  template hipError_t launch_xc2r_cols<T>(int, const void*, void*, int, int, const void*,   \
                                          hipStream_t, double);                             \
  template hipError_t launch_xr2c_cols<T>(int, const void*, void*, int, int, double,        \
                                          const void*, hipStream_t);
MLV_INST(float)
MLV_INST(double)

}  // namespace mlv
