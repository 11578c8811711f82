// Workgroup-level batched power-of-two FFT held in LDS.
//
// Layout: `slots` transforms of length N = 2^LOGN, slot f starts at
// lds + f * ldsn(N); element i of a slot lives at pad(i) so the strided
// Stockham writes do not pile onto one LDS bank.
//
// Algorithm: Stockham autosort, radix-16 passes followed by one radix
// 2/4/8 pass for the remainder.  Every thread owns P = 16 or 32 points per
// pass (P/R butterflies of radix R), so a block of NT threads transforms
// NT*P/N slots at once.  A pass reads its R inputs into registers,
// barriers, and writes the R outputs back in place, so one LDS buffer
// suffices.  Twiddles come from an FP64-tabulated table (rounded once to
// float for single precision); the in-register radix-R kernels use exact
// constants for the trivial angles.
#pragma once
#include "common.hpp"

namespace mlv {

__host__ __device__ constexpr int fft_pad(int i) { return i + (i >> 4); }
// Slot stride: the padded length plus one element, so that the same index
// of consecutive slots falls on different LDS banks.
__host__ __device__ constexpr int fft_ldsn(int n) { return n + (n >> 4) + 1; }

// cos(2*pi*j/16), j = 0..15.
__device__ constexpr double kCos16[16] = {
    1.0,
    0.92387953251128675613,
    0.70710678118654752440,
    0.38268343236508977173,
    0.0,
    -0.38268343236508977173,
    -0.70710678118654752440,
    -0.92387953251128675613,
    -1.0,
    -0.92387953251128675613,
    -0.70710678118654752440,
    -0.38268343236508977173,
    0.0,
    0.38268343236508977173,
    0.70710678118654752440,
    0.92387953251128675613};

// x * exp(DIR*2*pi*i*k/R) for compile-time R <= 16 (k known after unrolling).
template <int R, int DIR, typename T>
__device__ __forceinline__ cx<T> rot(int k, cx<T> x) {
  if (k == 0) return x;
  if (4 * k == R) return DIR > 0 ? mk<T>(-x.y, x.x) : mk<T>(x.y, -x.x);
  const int j = k * (16 / R);
  const T c = T(kCos16[j]);
  const T s = T(DIR) * T(kCos16[(j + 12) & 15]);
  return mk<T>(x.x * c - x.y * s, x.x * s + x.y * c);
}

// In-register radix-R DFT, R a power of two <= 16 (recursive radix-2 DIT).
template <int R, int DIR, typename T>
struct Dft {
  __device__ __forceinline__ static void run(cx<T>* v) {
    cx<T> e[R / 2], o[R / 2];
#pragma unroll
    for (int i = 0; i < R / 2; ++i) {
      e[i] = v[2 * i];
      o[i] = v[2 * i + 1];
    }
    Dft<R / 2, DIR, T>::run(e);
    Dft<R / 2, DIR, T>::run(o);
#pragma unroll
    for (int k = 0; k < R / 2; ++k) {
      const cx<T> t = rot<R, DIR, T>(k, o[k]);
      v[k] = cadd<T>(e[k], t);
      v[k + R / 2] = csub<T>(e[k], t);
    }
  }
};
template <int DIR, typename T>
struct Dft<1, DIR, T> {
  __device__ __forceinline__ static void run(cx<T>*) {}
};

// Per-pass twiddle tables.  After the NMAX-entry table exp(-2 pi i k/NMAX)
// the twiddle buffer holds, for every length 2^logn (logn = 3..LOG_NMAX) and
// every pass with Ns > 1, the (R-1) x Ns table  T[r-1][jm] = W_{Ns R}^{jm r},
// r = 1..R-1, jm = 0..Ns-1.  A pass's butterflies on consecutive lanes have
// consecutive jm, so each twiddle load of a wave reads consecutive entries.
__host__ __device__ constexpr int fft_pass_logr(int logn, int logns) {
  return logn - logns >= 4 ? 4 : logn - logns;
}
__host__ __device__ constexpr int fft_tables_size(int logn) {
  int size = 0, ns = 0;
  while (ns < logn) {
    const int lr = fft_pass_logr(logn, ns);
    if (ns > 0) size += ((1 << lr) - 1) << ns;
    ns += lr;
  }
  return size;
}
// Offset (in complex entries from the start of the twiddle buffer) of the
// table of the pass of length 2^logn whose sub-transform length is 2^logns.
__host__ __device__ constexpr int fft_pass_table(int logn, int logns) {
  int off = NMAX;
  for (int l = 3; l < logn; ++l) off += fft_tables_size(l);
  int ns = 0;
  while (ns < logns) {
    const int lr = fft_pass_logr(logn, ns);
    if (ns > 0) off += ((1 << lr) - 1) << ns;
    ns += lr;
  }
  return off;
}
__host__ __device__ constexpr int fft_twiddle_entries() {
  int n = NMAX;
  for (int l = 3; l <= LOG_NMAX; ++l) n += fft_tables_size(l);
  return n;
}

// One Stockham pass: radix 2^LOGR, sub-transform length Ns = 2^LOGNS.
// Element i of a slot is at fft_pad(i) = i + i/16.  When the stride between
// a butterfly's R inputs (or outputs) is a multiple of 16 the padded
// positions are a fixed stride apart, so every access is one base address
// plus a compile-time offset.
template <typename T, int LOGN, int LOGR, int LOGNS, int DIR, int NT, int P>
__device__ __forceinline__ void fft_pass(cx<T>* __restrict__ lds, int active,
                                         const cx<T>* __restrict__ tw) {
  constexpr int N = 1 << LOGN;
  constexpr int R = 1 << LOGR;
  constexpr int NS = 1 << LOGNS;
  constexpr int BF = N / R;          // butterflies per transform
  constexpr int ITEMS = P / R;       // butterflies per thread
  constexpr int LDSN = fft_ldsn(N);
  constexpr bool RSEP = (BF % 16) == 0;
  constexpr int RSTR = BF + BF / 16;
  constexpr bool WSEP = (NS % 16) == 0 || (NS == 1 && R == 16);
  constexpr int WSTR = NS == 1 ? 1 : NS + NS / 16;
  constexpr int TOFF = LOGNS > 0 ? fft_pass_table(LOGN, LOGNS) : 0;
  const int tid = threadIdx.x;
  cx<T> v[ITEMS][R];
#pragma unroll
  for (int it = 0; it < ITEMS; ++it) {
    const int w = tid + it * NT;
    const int f = w / BF;
    const int j = w % BF;
    if (f < active) {
      const cx<T>* src = lds + f * LDSN;
      if constexpr (RSEP) {
        const cx<T>* p = src + fft_pad(j);
#pragma unroll
        for (int r = 0; r < R; ++r) v[it][r] = p[r * RSTR];
      } else {
#pragma unroll
        for (int r = 0; r < R; ++r) v[it][r] = src[fft_pad(j + r * BF)];
      }
      if constexpr (NS > 1) {
        const int jm = j & (NS - 1);
        const cx<T>* t = tw + TOFF + jm;
#pragma unroll
        for (int r = 1; r < R; ++r) {
          cx<T> wr = t[(r - 1) * NS];
          if (DIR > 0) wr.y = -wr.y;
          v[it][r] = cmul<T>(v[it][r], wr);
        }
      }
      Dft<R, DIR, T>::run(v[it]);
    }
  }
  __syncthreads();
#pragma unroll
  for (int it = 0; it < ITEMS; ++it) {
    const int w = tid + it * NT;
    const int f = w / BF;
    const int j = w % BF;
    if (f < active) {
      cx<T>* dst = lds + f * LDSN;
      const int jm = j & (NS - 1);
      const int base = (j - jm) * R + jm;
      if constexpr (WSEP) {
        cx<T>* p = dst + fft_pad(base);
#pragma unroll
        for (int r = 0; r < R; ++r) p[r * WSTR] = v[it][r];
      } else {
#pragma unroll
        for (int r = 0; r < R; ++r) dst[fft_pad(base + r * NS)] = v[it][r];
      }
    }
  }
  __syncthreads();
}

template <typename T, int LOGN, int LOGNS, int DIR, int NT, int P>
struct FftPasses {
  static constexpr int REM = LOGN - LOGNS;
  static constexpr int LOGR = REM >= 4 ? 4 : REM;
  __device__ __forceinline__ static void run(cx<T>* lds, int active, const cx<T>* tw) {
    if constexpr (REM > 0) {
      fft_pass<T, LOGN, LOGR, LOGNS, DIR, NT, P>(lds, active, tw);
      FftPasses<T, LOGN, LOGNS + LOGR, DIR, NT, P>::run(lds, active, tw);
    }
  }
};

// Unnormalised DFT of `active` slots in place.  DIR = -1 forward
// (exp(-i...)), +1 inverse.  Requires NT*P/N >= active and a barrier
// between the writes that filled the slots and this call.
template <typename T, int LOGN, int DIR, int NT, int P>
__device__ __forceinline__ void fft_lds(cx<T>* lds, int active, const cx<T>* tw) {
  static_assert(P == 16 || P == 32, "points per thread");
  static_assert((NT * P) % (1 << LOGN) == 0 || (1 << LOGN) > NT * P, "slots");
  FftPasses<T, LOGN, 0, DIR, NT, P>::run(lds, active, tw);
}

// Slots available to a block of NT threads at P points per thread.
template <int LOGN, int NT, int P>
__host__ __device__ constexpr int fft_slots() {
  return (NT * P) >> LOGN;
}

// Points per thread.  Column passes (along x, strided) want enough slots per
// block that every row is read in runs of at least 64 bytes (8 float or 4
// double complex values), so single precision uses 32 points from N = 1024.
// Row passes read contiguous rows and keep 16 points up to N = 1024, which
// leaves more blocks and a smaller register file.
template <typename T>
__host__ __device__ constexpr int fft_points_col(int logn) {
  return sizeof(T) == 4 ? (logn >= 10 ? 32 : 16) : (logn >= 11 ? 32 : 16);
}
template <typename T>
__host__ __device__ constexpr int fft_points_row(int logn) {
  return logn >= 11 ? 32 : 16;
}

}  // namespace mlv
