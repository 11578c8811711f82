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
// synchronises, and writes the R outputs back in place, so one LDS buffer
// suffices.  The threads of a slot stay with it through every pass; when
// a slot fits one wavefront (N/P <= 64) the synchronisation is the
// wavefront's own, not a workgroup barrier.  Twiddles come from an FP64-tabulated table (rounded once to
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
// End of the standard plan's tables.
__host__ __device__ constexpr int fft_std_entries() {
  int n = NMAX;
  for (int l = 3; l <= LOG_NMAX; ++l) n += fft_tables_size(l);
  return n;
}
// Tables for passes outside the standard plan (other orders of the radices),
// keyed by (logns, logr) alone -- a pass's twiddles depend only on Ns and R:
// T[r-1][jm] = W_{Ns R}^{jm r}, for 1 <= logr <= 4, 1 <= logns, logns + logr
// <= LOG_NMAX, placed after the standard tables.  With no match the
// function returns the end of the keyed tables.
__host__ __device__ constexpr int fft_keyed_table(int logns, int logr) {
  int off = fft_std_entries();
  for (int lr = 1; lr <= 4; ++lr)
    for (int ln = 1; ln + lr <= LOG_NMAX; ++ln) {
      if (lr == logr && ln == logns) return off;
      off += ((1 << lr) - 1) << ln;
    }
  return off;
}
__host__ __device__ constexpr int fft_twiddle_entries() { return fft_keyed_table(0, 0); }

// Threads that share one transform of N points at P points per thread (0
// when a thread holds several whole transforms, N < P).  Thread t works on
// slot t / TPS in every pass, so a slot is touched by TPS consecutive
// threads only.
template <int N, int P>
__host__ __device__ constexpr int fft_tps() {
  return N >= P ? N / P : 0;
}
// Every slot lies inside one wavefront: the passes need no workgroup
// barrier, only the wavefront's own ordering of its LDS accesses.
template <int N, int P>
__host__ __device__ constexpr bool fft_wave_local() {
  return fft_tps<N, P>() > 0 && fft_tps<N, P>() <= 64;
}

// Orders the LDS accesses of one wavefront: the compiler may not move
// memory operations across it, and the LDS executes a wavefront's
// instructions in issue order, so a read after it sees the wavefront's
// earlier writes and a write after it cannot overtake earlier reads.
__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

template <int N, int P>
__device__ __forceinline__ void fft_sync() {
  if constexpr (fft_wave_local<N, P>())
    wave_sync();
  else
    __syncthreads();
}

// One Stockham pass: radix 2^LOGR, sub-transform length Ns = 2^LOGNS.
// Element i of a slot is at fft_pad(i) = i + i/16.  When the stride between
// a butterfly's R inputs (or outputs) is a multiple of 16 the padded
// positions are a fixed stride apart, so every access is one base address
// plus a compile-time offset.  With TPS threads per slot, thread t computes
// butterflies j = t % TPS + it * TPS (it < P/R) of slot t / TPS, so
// consecutive lanes take consecutive j (consecutive twiddles, conflict-free
// LDS) and a slot never leaves its threads.
template <typename T, int LOGN, int LOGR, int LOGNS, int DIR, int NT, int P, bool KEYED = false>
__device__ __forceinline__ void fft_pass(cx<T>* __restrict__ lds, int active,
                                         const cx<T>* __restrict__ tw) {
  constexpr int N = 1 << LOGN;
  constexpr int R = 1 << LOGR;
  constexpr int NS = 1 << LOGNS;
  constexpr int BF = N / R;          // butterflies per transform
  constexpr int ITEMS = P / R;       // butterflies per thread
  constexpr int LDSN = fft_ldsn(N);
  constexpr int TPS = fft_tps<N, P>();
  constexpr bool RSEP = (BF % 16) == 0;
  constexpr int RSTR = BF + BF / 16;
  constexpr bool WSEP = (NS % 16) == 0 || (NS == 1 && R == 16);
  constexpr int WSTR = NS == 1 ? 1 : NS + NS / 16;
  constexpr int TOFF =
      LOGNS == 0 ? 0 : (KEYED ? fft_keyed_table(LOGNS, LOGR) : fft_pass_table(LOGN, LOGNS));
  const int tid = threadIdx.x;
  // slot and butterfly index of this thread's butterfly `it`
  auto slot = [&](int it) { return TPS ? tid / TPS : (tid + it * NT) / BF; };
  auto bfly = [&](int it) { return TPS ? tid % TPS + it * TPS : (tid + it * NT) % BF; };
  cx<T> v[ITEMS][R];
#pragma unroll
  for (int it = 0; it < ITEMS; ++it) {
    const int f = slot(it);
    const int j = bfly(it);
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
  fft_sync<N, P>();
#pragma unroll
  for (int it = 0; it < ITEMS; ++it) {
    const int f = slot(it);
    const int j = bfly(it);
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
  fft_sync<N, P>();
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
// (exp(-i...)), +1 inverse.  Requires NT*P/N >= active and, before the
// call, a barrier after the writes that filled the slots (a workgroup
// barrier, or wave_sync() when each wavefront filled its own slots and
// fft_wave_local holds).  On return the slots are complete for every
// thread of the block (SYNC_OUT = true) or, with SYNC_OUT = false and
// fft_wave_local, for the threads of each slot's own wavefront.
template <typename T, int LOGN, int DIR, int NT, int P, bool SYNC_OUT = true>
__device__ __forceinline__ void fft_lds(cx<T>* lds, int active, const cx<T>* tw) {
  static_assert(P == 16 || P == 32, "points per thread");
  static_assert((NT * P) % (1 << LOGN) == 0 || (1 << LOGN) > NT * P, "slots");
  FftPasses<T, LOGN, 0, DIR, NT, P>::run(lds, active, tw);
  if constexpr (SYNC_OUT && fft_wave_local<(1 << LOGN), P>()) __syncthreads();
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
