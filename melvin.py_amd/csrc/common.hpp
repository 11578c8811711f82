// Shared device helpers: complex arithmetic on float2/double2, the FP64-
// tabulated twiddle table, and compensated FP64 accumulation.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace mlv {

// Largest supported transform length (log2).  One twiddle table of this
// length serves every smaller power of two by striding.
constexpr int LOG_NMAX = 13;
constexpr int NMAX = 1 << LOG_NMAX;

template <typename T> struct CxOf;
template <> struct CxOf<float> { using type = float2; };
template <> struct CxOf<double> { using type = double2; };
template <typename T> using cx = typename CxOf<T>::type;

// A device pointer read out of a job table is generic to the compiler, so its
// loads and stores become flat instructions, which also count against the
// LDS counter and make every LDS wait wait for them.  glb() states that the
// pointer is global memory; code inlined after it then issues global_* ops.
// (The host pass of the same source sees plain pointers.)
#if defined(__HIP_DEVICE_COMPILE__)
template <typename T>
using gptr = __attribute__((address_space(1))) T*;
#else
template <typename T>
using gptr = T*;
#endif
template <typename T>
__device__ __forceinline__ gptr<T> glb(T* p) { return (gptr<T>)p; }

template <typename T>
__device__ __forceinline__ cx<T> mk(T re, T im) {
  cx<T> r;
  r.x = re;
  r.y = im;
  return r;
}
template <typename T>
__device__ __forceinline__ cx<T> cadd(cx<T> a, cx<T> b) { return mk<T>(a.x + b.x, a.y + b.y); }
template <typename T>
__device__ __forceinline__ cx<T> csub(cx<T> a, cx<T> b) { return mk<T>(a.x - b.x, a.y - b.y); }
template <typename T>
__device__ __forceinline__ cx<T> cmul(cx<T> a, cx<T> b) {
  return mk<T>(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
template <typename T>
__device__ __forceinline__ cx<T> cscale(cx<T> a, T s) { return mk<T>(a.x * s, a.y * s); }
template <typename T>
__device__ __forceinline__ cx<T> cconj(cx<T> a) { return mk<T>(a.x, -a.y); }
// i * a
template <typename T>
__device__ __forceinline__ cx<T> cmuli(cx<T> a) { return mk<T>(-a.y, a.x); }
template <typename T>
__device__ __forceinline__ cx<T> czero() { return mk<T>(T(0), T(0)); }

// W_N^k = exp(DIR * 2*pi*i*k/N) from the table exp(-2*pi*i*j/NMAX).
// DIR = -1 (forward) reads the table as is, DIR = +1 conjugates it.
template <typename T, int DIR>
__device__ __forceinline__ cx<T> twiddle(const cx<T>* __restrict__ tw, int k, int logn) {
  cx<T> w = tw[k << (LOG_NMAX - logn)];
  if (DIR > 0) w.y = -w.y;
  return w;
}

// sin(2*pi*k/N) for signed k, read from the same table.
template <typename T>
__device__ __forceinline__ T tsin(const cx<T>* __restrict__ tw, int k, int logn) {
  const int n = 1 << logn;
  int kk = k % n;
  if (kk < 0) kk += n;
  return -tw[kk << (LOG_NMAX - logn)].y;
}

// Neumaier-compensated FP64 sum.
struct KSum {
  double s = 0.0, c = 0.0;
  __device__ __forceinline__ void add(double x) {
    double t = s + x;
    if (fabs(s) >= fabs(x))
      c += (s - t) + x;
    else
      c += (x - t) + s;
    s = t;
  }
  __device__ __forceinline__ double value() const { return s + c; }
};

}  // namespace mlv
