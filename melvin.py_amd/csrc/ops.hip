// Operator-level element-wise and stencil kernels (mirror of
// SpatialDifferentiator, LaplacianSolver (fully spectral), Integrator and
// the FP64 reductions behind xp.max / xp.sum / xp.mean).
//
// Arithmetic follows the reference expression order so the FP64 results
// agree with NumPy to the last bit where the operation allows it:
//   diff_factor * n * var  ->  (-y*vi, y*vr), y = (2*pi/L) * n
//   (f[i+1] - f[i-1]) / (2*dx)     (division, not a reciprocal multiply)
#include "ops.hpp"

namespace mlv {

constexpr int kEW = 256;  // threads per block for element-wise kernels

__device__ __forceinline__ int mode_x(int r, int rows, int nn, int fully) {
  return fully ? (r <= nn ? r : r - rows) : r;
}

// --------------------------------------------------------------------------
// Spectral multiplies (SpatialDifferentiator.py:50-74, LaplacianSolver.py:58-68)
// --------------------------------------------------------------------------
template <typename T>
__global__ void __launch_bounds__(kEW) k_spec(const cx<T>* __restrict__ in, void* __restrict__ out,
                                              int rows, int cols, int nn, int fully, int kind,
                                              double fx, double fz, double fx2, double fz2,
                                              int real) {
  const size_t total = (size_t)rows * cols;
  for (size_t e = blockIdx.x * (size_t)kEW + threadIdx.x; e < total;
       e += (size_t)gridDim.x * kEW) {
    const int r = (int)(e / cols), c = (int)(e % cols);
    const int n = mode_x(r, rows, nn, fully);
    const int m = c;
    if (kind == MLV_SPEC_LAP) {
      // real output, always FP64 (the reference's lap is float64)
      reinterpret_cast<double*>(out)[e] = fx2 * (double)(n * n) + fz2 * (double)(m * m);
      continue;
    }
    const cx<T> v = in[e];
    cx<T> o;
    switch (kind) {
      case MLV_SPEC_DDX: {
        const T y = (T)(fx * n);
        o = (real & 1) ? mk<T>(y * v.x, y * v.y) : mk<T>(-(y * v.y), y * v.x);
        break;
      }
      case MLV_SPEC_DDZ: {
        const T y = (T)(fz * m);
        o = (real & 2) ? mk<T>(y * v.x, y * v.y) : mk<T>(-(y * v.y), y * v.x);
        break;
      }
      case MLV_SPEC_D2DX2: {
        const T a = (T)(fx2 * (double)(n * n));
        o = mk<T>(a * v.x, a * v.y);
        break;
      }
      case MLV_SPEC_D2DZ2: {
        const T a = (T)(fz2 * (double)(m * m));
        o = mk<T>(a * v.x, a * v.y);
        break;
      }
      case MLV_SPEC_SNABLA2: {
        const T a = (T)(fx2 * (double)(n * n));
        const T b = (T)(fz2 * (double)(m * m));
        o = mk<T>(a * v.x + b * v.x, a * v.y + b * v.y);
        break;
      }
      case kSpecAddD2dx2: {
        const T a = (T)(fx2 * (double)(n * n));
        const cx<T> b = reinterpret_cast<const cx<T>*>(out)[e];
        o = mk<T>(a * v.x + b.x, a * v.y + b.y);
        break;
      }
      case MLV_SPEC_POISSON: {
        double lap = fx2 * (double)(n * n) + fz2 * (double)(m * m);
        if (r == 0 && c == 0) lap = 1.0;
        // complex / real in NumPy is a complex division by (lap + 0i),
        // i.e. a multiply by the reciprocal
        const double rl = 1.0 / lap;
        o = mk<T>((T)(v.x * rl), (T)(v.y * rl));
        break;
      }
      default:
        o = v;
    }
    reinterpret_cast<cx<T>*>(out)[e] = o;
  }
}

// --------------------------------------------------------------------------
// Integrator update (Integrator.py:5-18, 53-63; TimeDerivative.py:35-45)
// --------------------------------------------------------------------------
template <typename T>
__device__ __forceinline__ cx<T> ab_combine(const cx<T>* __restrict__ hist, size_t slot,
                                            size_t e, int q, int curr, double h) {
  auto get = [&](int i) {
    int s = curr + i;
    if (s < 0) s += q;
    return hist[(size_t)s * slot + e];
  };
  const cx<T> d0 = get(0), d1 = get(-1);
  if (q == 2) {
    const T rx = T(3) * d0.x - d1.x, ry = T(3) * d0.y - d1.y;
    return mk<T>((T)h * rx, (T)h * ry);
  }
  const cx<T> d2 = get(-2), d3 = get(-3);
  const T rx = T(55) * d0.x - T(59) * d1.x + T(37) * d2.x - T(9) * d3.x;
  const T ry = T(55) * d0.y - T(59) * d1.y + T(37) * d2.y - T(9) * d3.y;
  return mk<T>((T)h * rx, (T)h * ry);
}

template <typename T>
__global__ void __launch_bounds__(kEW) k_update(cx<T>* __restrict__ var, cx<T>* __restrict__ hist,
                                                const void* __restrict__ lin, size_t n, int q,
                                                int curr, int scheme, double c1, double c2,
                                                double h) {
  for (size_t e = blockIdx.x * (size_t)kEW + threadIdx.x; e < n; e += (size_t)gridDim.x * kEW) {
    const cx<T> v = var[e];
    if (scheme == MLV_SCHEME_EXPLICIT) {
      // dvar[:] += diffusion (stored into the history), var += AB(dvar)
      const cx<T> dif = reinterpret_cast<const cx<T>*>(lin)[e];
      cx<T> d = hist[(size_t)curr * n + e];
      d = cadd<T>(d, dif);
      hist[(size_t)curr * n + e] = d;
      const cx<T> ab = ab_combine<T>(hist, n, e, q, curr, h);
      var[e] = cadd<T>(v, ab);
    } else {
      const double L = reinterpret_cast<const double*>(lin)[e];
      const double num = 1.0 + c1 * L;
      const double den = 1.0 - c2 * L;
      const cx<T> ab = ab_combine<T>(hist, n, e, q, curr, h);
      const T rx = (T)num * v.x + ab.x, ry = (T)num * v.y + ab.y;
      const T rd = T(1) / (T)den;
      var[e] = mk<T>(rx * rd, ry * rd);
    }
  }
}

// AB / AM combination alone (Integrator.predictor / corrector)
template <typename T>
__global__ void __launch_bounds__(kEW) k_combine(const cx<T>* __restrict__ hist, cx<T>* __restrict__ out,
                                                 size_t n, int q, int curr, int moulton, double h) {
  for (size_t e = blockIdx.x * (size_t)kEW + threadIdx.x; e < n; e += (size_t)gridDim.x * kEW) {
    if (!moulton) {
      out[e] = ab_combine<T>(hist, n, e, q, curr, h);
      continue;
    }
    auto get = [&](int i) {
      int s = curr + i;
      if (s < 0) s += q;
      return hist[(size_t)s * n + e];
    };
    const cx<T> d0 = get(0), d1 = get(-1);
    T rx, ry;
    if (q == 2) {
      rx = d0.x + d1.x;
      ry = d0.y + d1.y;
    } else {
      const cx<T> d2 = get(-2), d3 = get(-3);
      rx = T(9) * d0.x + T(19) * d1.x - T(5) * d2.x + d3.x;
      ry = T(9) * d0.y + T(19) * d1.y - T(5) * d2.y + d3.y;
    }
    out[e] = mk<T>((T)h * rx, (T)h * ry);
  }
}

// Velocity from vorticity, fully spectral (utility.py:62-79 with
// LaplacianSolver.py:58-68): psi = (-w)/lap, ux = -ddz(psi), uz = ddx(psi).
template <typename T>
__global__ void __launch_bounds__(kEW) k_velocity(const cx<T>* __restrict__ w, cx<T>* __restrict__ psi,
                                                  cx<T>* __restrict__ ux, cx<T>* __restrict__ uz,
                                                  int rows, int cols, int nn, double fx, double fz,
                                                  double fx2, double fz2) {
  const size_t total = (size_t)rows * cols;
  for (size_t e = blockIdx.x * (size_t)kEW + threadIdx.x; e < total;
       e += (size_t)gridDim.x * kEW) {
    const int r = (int)(e / cols), m = (int)(e % cols);
    const int n = mode_x(r, rows, nn, 1);
    double lap = fx2 * (double)(n * n) + fz2 * (double)(m * m);
    if (r == 0 && m == 0) lap = 1.0;
    const double rl = 1.0 / lap;
    const cx<T> v = w[e];
    const cx<T> p = mk<T>((T)(-(double)v.x * rl), (T)(-(double)v.y * rl));
    psi[e] = p;
    if (ux) {
      const T y = (T)(fz * m);
      ux[e] = mk<T>(y * p.y, -(y * p.x));
    }
    if (uz) {
      const T y = (T)(fx * n);
      uz[e] = mk<T>(-(y * p.y), y * p.x);
    }
  }
}

// --------------------------------------------------------------------------
// Physical-space central differences (SpatialDifferentiator.py:76-185)
// --------------------------------------------------------------------------
// One value of an FD stencil along a line with stride `st`, index i of n.
// Returns false when the point is a boundary point left at zero.
template <typename T>
__device__ __forceinline__ bool fd1(const T* __restrict__ f, int i, int n, size_t st, int order,
                                    int periodic, double h, T* res) {
  auto at = [&](int k) {
    if (k < 0) k += n;
    if (k >= n) k -= n;
    return f[(size_t)k * st];
  };
  if (order == 2) {
    if (!periodic && (i < 1 || i > n - 2)) return false;
    *res = (at(i + 1) - at(i - 1)) / (T)(2 * h);
  } else {
    if (!periodic && (i < 2 || i > n - 3)) return false;
    *res = (T(-0.25) * at(i + 2) + T(2) * at(i + 1) - T(2) * at(i - 1) + T(0.25) * at(i - 2)) /
           (T)(3 * h);
  }
  return true;
}

template <typename T>
__device__ __forceinline__ bool fd2(const T* __restrict__ f, int i, int n, size_t st, int order,
                                    double h, T* res) {
  auto at = [&](int k) { return f[(size_t)k * st]; };
  if (order == 2) {
    if (i < 1 || i > n - 2) return false;
    *res = (at(i + 1) - T(2) * at(i) + at(i - 1)) / (T)(h * h);
  } else {
    if (i < 2 || i > n - 3) return false;
    *res = (T(-1.0 / 12) * at(i + 2) + T(4.0 / 3) * at(i + 1) - T(5.0 / 2) * at(i) +
            T(4.0 / 3) * at(i - 1) - T(1.0 / 12) * at(i - 2)) /
           (T)(h * h);
  }
  return true;
}

// kind: 0 d/dx, 1 d/dz, 2 d2/dz2.  `comp` = number of interleaved real
// components (1 real, 2 complex).  rows x cols logical grid, z contiguous.
template <typename T>
__global__ void __launch_bounds__(kEW) k_fd(const T* __restrict__ in, T* __restrict__ out, int rows,
                                            int cols, int comp, int kind, int order, int periodic,
                                            double h) {
  const size_t total = (size_t)rows * cols * comp;
  for (size_t e = blockIdx.x * (size_t)kEW + threadIdx.x; e < total;
       e += (size_t)gridDim.x * kEW) {
    const int k = (int)(e % comp);
    const size_t rc = e / comp;
    const int r = (int)(rc / cols), c = (int)(rc % cols);
    T v = T(0);
    bool ok;
    if (kind == 0)
      ok = fd1<T>(in + (size_t)c * comp + k, r, rows, (size_t)cols * comp, order, periodic, h, &v);
    else if (kind == 1)
      ok = fd1<T>(in + (size_t)r * cols * comp + k, c, cols, comp, order, periodic, h, &v);
    else
      ok = fd2<T>(in + (size_t)r * cols * comp + k, c, cols, comp, order, h, &v);
    out[e] = ok ? v : T(0);
  }
}

// pddx(ux*f) + pddz(uz*f) (Variable.py:125-127)
template <typename T>
__global__ void __launch_bounds__(kEW) k_flux(const T* __restrict__ ux, const T* __restrict__ uz,
                                              const T* __restrict__ f, T* __restrict__ out, int nx,
                                              int nz, int order, int xper, int zper, double dx,
                                              double dz) {
  const size_t total = (size_t)nx * nz;
  for (size_t e = blockIdx.x * (size_t)kEW + threadIdx.x; e < total;
       e += (size_t)gridDim.x * kEW) {
    const int i = (int)(e / nz), j = (int)(e % nz);
    auto px = [&](int a) {
      if (a < 0) a += nx;
      if (a >= nx) a -= nx;
      const size_t q = (size_t)a * nz + j;
      return ux[q] * f[q];
    };
    auto pz = [&](int b) {
      if (b < 0) b += nz;
      if (b >= nz) b -= nz;
      const size_t q = (size_t)i * nz + b;
      return uz[q] * f[q];
    };
    T gx = T(0), gz = T(0);
    if (order == 2) {
      if (xper || (i >= 1 && i <= nx - 2)) gx = (px(i + 1) - px(i - 1)) / (T)(2 * dx);
      if (zper || (j >= 1 && j <= nz - 2)) gz = (pz(j + 1) - pz(j - 1)) / (T)(2 * dz);
    } else {
      if (xper || (i >= 2 && i <= nx - 3))
        gx = (T(-0.25) * px(i + 2) + T(2) * px(i + 1) - T(2) * px(i - 1) + T(0.25) * px(i - 2)) /
             (T)(3 * dx);
      if (zper || (j >= 2 && j <= nz - 3))
        gz = (T(-0.25) * pz(j + 2) + T(2) * pz(j + 1) - T(2) * pz(j - 1) + T(0.25) * pz(j - 2)) /
             (T)(3 * dz);
    }
    out[e] = gx + gz;
  }
}

// --------------------------------------------------------------------------
// FP64 reductions: max, sum, sum(a^2 + b^2), sum(a*b).  Two passes:
// per-block partials (Neumaier-compensated), then one block combines them.
// --------------------------------------------------------------------------
__device__ __forceinline__ double red_combine(double a, double b, int op) {
  return op == MLV_RED_MAX ? fmax(a, b) : a + b;
}

// Pairwise tree over the block's (sum, compensation) pairs in LDS: a
// NaN-propagating max, or Neumaier-compensated addition of the pairs.
// Leaves the result in slot 0.
__device__ __forceinline__ void red_tree(double* sh_s, double* sh_c, int op) {
  for (int w = kEW / 2; w > 0; w >>= 1) {
    if (threadIdx.x < w) {
      const int o = threadIdx.x + w;
      if (op == MLV_RED_MAX) {
        const double a = sh_s[threadIdx.x], b = sh_s[o];
        sh_s[threadIdx.x] = (a != a || b != b) ? NAN : fmax(a, b);
      } else {
        KSum k;
        k.s = sh_s[threadIdx.x];
        k.c = sh_c[threadIdx.x];
        k.add(sh_s[o]);
        k.add(sh_c[o]);
        sh_s[threadIdx.x] = k.s;
        sh_c[threadIdx.x] = k.c;
      }
    }
    __syncthreads();
  }
}

template <typename T>
__global__ void __launch_bounds__(kEW) k_reduce_part(const T* __restrict__ a, const T* __restrict__ b,
                                                     size_t n, int op, double* __restrict__ part) {
  __shared__ double sh_s[kEW], sh_c[kEW];
  KSum acc;
  double mx = -INFINITY;
  bool nan = false;
  for (size_t e = blockIdx.x * (size_t)kEW + threadIdx.x; e < n; e += (size_t)gridDim.x * kEW) {
    const double va = (double)a[e];
    if (op == MLV_RED_MAX) {
      if (va != va) nan = true;
      mx = fmax(mx, va);
    } else if (op == MLV_RED_SUM) {
      acc.add(va);
    } else if (op == MLV_RED_SUMSQ2) {
      const double vb = (double)b[e];
      acc.add(va * va + vb * vb);
    } else {
      acc.add(va * (double)b[e]);
    }
  }
  if (op == MLV_RED_MAX) {
    sh_s[threadIdx.x] = nan ? NAN : mx;
    sh_c[threadIdx.x] = 0.0;
  } else {
    sh_s[threadIdx.x] = acc.s;
    sh_c[threadIdx.x] = acc.c;
  }
  __syncthreads();
  red_tree(sh_s, sh_c, op);
  if (threadIdx.x == 0) {
    part[2 * blockIdx.x] = sh_s[0];
    part[2 * blockIdx.x + 1] = sh_c[0];
  }
}

__global__ void __launch_bounds__(kEW) k_reduce_final(const double* __restrict__ part, int nblk,
                                                      int op, double scale,
                                                      double* __restrict__ out) {
  __shared__ double sh_s[kEW], sh_c[kEW];
  // each thread folds a strided subset of the partials, then a tree
  if (op == MLV_RED_MAX) {
    double m = -INFINITY;
    bool nan = false;
    for (int t = threadIdx.x; t < nblk; t += kEW) {
      const double v = part[2 * t];
      if (v != v) nan = true;
      m = fmax(m, v);
    }
    sh_s[threadIdx.x] = nan ? NAN : m;
    sh_c[threadIdx.x] = 0.0;
  } else {
    KSum k;
    for (int t = threadIdx.x; t < nblk; t += kEW) {
      k.add(part[2 * t]);
      k.add(part[2 * t + 1]);
    }
    sh_s[threadIdx.x] = k.s;
    sh_c[threadIdx.x] = k.c;
  }
  __syncthreads();
  red_tree(sh_s, sh_c, op);
  if (threadIdx.x == 0) out[0] = op == MLV_RED_MAX ? sh_s[0] : (sh_s[0] + sh_c[0]) * scale;
}

// --------------------------------------------------------------------------
// Host launchers
// --------------------------------------------------------------------------
static inline unsigned ew_grid(size_t n) {
  size_t g = (n + kEW - 1) / kEW;
  if (g > 8192) g = 8192;
  if (g < 1) g = 1;
  return (unsigned)g;
}

template <typename T>
hipError_t launch_spec(const void* in, void* out, int rows, int cols, int nn, int fully, int kind,
                       double fx, double fz, double fx2, double fz2, hipStream_t s, int real) {
  hipLaunchKernelGGL(k_spec<T>, dim3(ew_grid((size_t)rows * cols)), dim3(kEW), 0, s,
                     (const cx<T>*)in, out, rows, cols, nn, fully, kind, fx, fz, fx2, fz2, real);
  return hipGetLastError();
}

// --------------------------------------------------------------------------
// Dense per-axis transforms (the cosine / sine bases): one thread per output
// element, the sum over j in the precision of T (FP64 for double plans).
// --------------------------------------------------------------------------
template <typename T>
__global__ void __launch_bounds__(kEW) k_dense(const cx<T>* __restrict__ M, int K, int J, int batch,
                                               const void* __restrict__ in, int in_cplx,
                                               long long in_sb, long long in_sj,
                                               void* __restrict__ out, int out_real,
                                               long long out_sb, long long out_sk) {
  const size_t total = (size_t)K * batch;
  for (size_t e = blockIdx.x * (size_t)kEW + threadIdx.x; e < total;
       e += (size_t)gridDim.x * kEW) {
    const int k = (int)(e % K), b = (int)(e / K);
    const cx<T>* __restrict__ row = M + (size_t)k * J;
    T re = T(0), im = T(0);
    if (in_cplx) {
      const cx<T>* __restrict__ x = reinterpret_cast<const cx<T>*>(in) + (size_t)b * in_sb;
      for (int j = 0; j < J; ++j) {
        const cx<T> m = row[j], v = x[(size_t)j * in_sj];
        re += m.x * v.x - m.y * v.y;
        im += m.x * v.y + m.y * v.x;
      }
    } else {
      const T* __restrict__ x = reinterpret_cast<const T*>(in) + (size_t)b * in_sb;
      for (int j = 0; j < J; ++j) {
        const cx<T> m = row[j];
        const T v = x[(size_t)j * in_sj];
        re += m.x * v;
        im += m.y * v;
      }
    }
    const size_t o = (size_t)b * out_sb + (size_t)k * out_sk;
    if (out_real)
      reinterpret_cast<T*>(out)[o] = re;
    else
      reinterpret_cast<cx<T>*>(out)[o] = mk<T>(re, im);
  }
}

template <typename T>
hipError_t launch_dense(const void* M, int K, int J, int batch, const void* in, int in_cplx,
                        long long in_sb, long long in_sj, void* out, int out_real,
                        long long out_sb, long long out_sk, hipStream_t s) {
  hipLaunchKernelGGL(k_dense<T>, dim3(ew_grid((size_t)K * batch)), dim3(kEW), 0, s,
                     (const cx<T>*)M, K, J, batch, in, in_cplx, in_sb, in_sj, out, out_real,
                     out_sb, out_sk);
  return hipGetLastError();
}

template <typename T>
hipError_t launch_update(void* var, void* hist, const void* lin, size_t n, int q, int curr,
                         int scheme, double c1, double c2, double h, hipStream_t s) {
  hipLaunchKernelGGL(k_update<T>, dim3(ew_grid(n)), dim3(kEW), 0, s, (cx<T>*)var, (cx<T>*)hist,
                     lin, n, q, curr, scheme, c1, c2, h);
  return hipGetLastError();
}

template <typename T>
hipError_t launch_fd(const void* in, void* out, int rows, int cols, int comp, int kind, int order,
                     int periodic, double h, hipStream_t s) {
  hipLaunchKernelGGL(k_fd<T>, dim3(ew_grid((size_t)rows * cols * comp)), dim3(kEW), 0, s,
                     (const T*)in, (T*)out, rows, cols, comp, kind, order, periodic, h);
  return hipGetLastError();
}

template <typename T>
hipError_t launch_flux(const void* ux, const void* uz, const void* f, void* out, int nx, int nz,
                       int order, int xper, int zper, double dx, double dz, hipStream_t s) {
  hipLaunchKernelGGL(k_flux<T>, dim3(ew_grid((size_t)nx * nz)), dim3(kEW), 0, s, (const T*)ux,
                     (const T*)uz, (const T*)f, (T*)out, nx, nz, order, xper, zper, dx, dz);
  return hipGetLastError();
}

template <typename T>
hipError_t launch_reduce(const void* a, const void* b, size_t n, int op, double scale,
                         double* part, double* out, hipStream_t s) {
  size_t g = (n + kEW - 1) / kEW;
  if (g > (size_t)kRedBlocks) g = kRedBlocks;
  if (g < 1) g = 1;
  hipLaunchKernelGGL(k_reduce_part<T>, dim3((unsigned)g), dim3(kEW), 0, s, (const T*)a,
                     (const T*)b, n, op, part);
  hipLaunchKernelGGL(k_reduce_final, dim3(1), dim3(kEW), 0, s, (const double*)part, (int)g, op,
                     scale, out);
  return hipGetLastError();
}

size_t reduce_partials_bytes() { return sizeof(double) * 2 * kRedBlocks; }

template <typename T>
hipError_t launch_combine(const void* hist, void* out, size_t n, int q, int curr, int moulton,
                          double h, hipStream_t s) {
  hipLaunchKernelGGL(k_combine<T>, dim3(ew_grid(n)), dim3(kEW), 0, s, (const cx<T>*)hist,
                     (cx<T>*)out, n, q, curr, moulton, h);
  return hipGetLastError();
}

template <typename T>
hipError_t launch_velocity(const void* w, void* psi, void* ux, void* uz, int rows, int cols, int nn,
                           double fx, double fz, double fx2, double fz2, hipStream_t s) {
  hipLaunchKernelGGL(k_velocity<T>, dim3(ew_grid((size_t)rows * cols)), dim3(kEW), 0, s,
                     (const cx<T>*)w, (cx<T>*)psi, (cx<T>*)ux, (cx<T>*)uz, rows, cols, nn, fx, fz,
                     fx2, fz2);
  return hipGetLastError();
}

#define MLV_INST(T)                                                                              \
  template hipError_t launch_spec<T>(const void*, void*, int, int, int, int, int, double, double, \
                                     double, double, hipStream_t, int);                          \
  template hipError_t launch_dense<T>(const void*, int, int, int, const void*, int, long long,    \
                                      long long, void*, int, long long, long long, hipStream_t); \
  template hipError_t launch_update<T>(void*, void*, const void*, size_t, int, int, int, double,  \
                                       double, double, hipStream_t);                             \
  template hipError_t launch_fd<T>(const void*, void*, int, int, int, int, int, int, double,      \
                                   hipStream_t);                                                 \
  template hipError_t launch_flux<T>(const void*, const void*, const void*, void*, int, int, int, \
                                     int, int, double, double, hipStream_t);                     \
  template hipError_t launch_reduce<T>(const void*, const void*, size_t, int, double, double*,    \
                                       double*, hipStream_t);                                    \
  template hipError_t launch_combine<T>(const void*, void*, size_t, int, int, int, double,        \
                                        hipStream_t);                                            \
  template hipError_t launch_velocity<T>(const void*, void*, void*, void*, int, int, int, double, \
                                         double, double, double, hipStream_t);
MLV_INST(float)
MLV_INST(double)

}  // namespace mlv
