// Fused Fourier-x / FD-z step kernels over a job table (see fdz.hpp).
#include <math.h>

#include "diag_dev.hpp"
#include "fdz.hpp"
#include "grid.hpp"

namespace mlv {

template <typename T>
__device__ __forceinline__ double2 d2(cx<T> v) {
  return make_double2((double)v.x, (double)v.y);
}
template <typename T>
__device__ __forceinline__ cx<T> t2(double2 v) {
  return mk<T>((T)v.x, (T)v.y);
}

// --------------------------------------------------------------------------
// Tridiagonal solve of one row by its WPR wavefronts, FP64.
//
// The system has identity first and last rows and constant interior bands
// (lower = upper = lo).  fac holds, for element j, (cp_j, 1/den_j) of the
// Thomas factorisation, stored chunk-major: entry i * LR + t is element
// j = t K + i, LR = 64 WPR threads per row.  Thread t owns elements
// [t K, t K + K).
//   forward   d_j = (r_j - lower_j d_{j-1}) / den_j
//   backward  x_j = d_j - cp_j x_{j+1},  x_{nz-1} = d_{nz-1}
// Each recurrence is an affine map per element; a thread composes the maps
// of its chunk, an inclusive shuffle scan composes them across a
// wavefront's lanes, and the thread then replays its chunk from the value
// entering it.  With two wavefronts per row the value leaving the first
// wavefront's last element (forward) or entering the second's first
// element (backward) goes through xch (4 doubles of LDS) between them.
// buf holds the right-hand side on entry and the solution on exit.
// --------------------------------------------------------------------------
template <int WPR>
__device__ __forceinline__ void row_sync() {
  if constexpr (WPR == 1)
    wave_sync();
  else
    __syncthreads();  // a block holds exactly one row
}

template <int KT, int WPR>
__device__ __forceinline__ void row_tridiag(double2* __restrict__ buf,
                                            gptr<const double2> fac, double lo, int nz, int K,
                                            int t, double* __restrict__ xch) {
  static_assert(WPR == 1 || WPR == 2, "one or two wavefronts per row");
  constexpr int LR = 64 * WPR;
  const int lane = t & 63, wv = t >> 6;
  const int j0 = t * K;
  double2 d[KT], fc[KT];
#pragma unroll
  for (int i = 0; i < KT; ++i) {
    const int j = j0 + i;
    const bool ok = i < K && j < nz;
    fc[i] = make_double2(0.0, 1.0);
    d[i] = make_double2(0.0, 0.0);
    if (ok) {
      fc[i] = fac[i * LR + t];
      d[i] = buf[j];
    }
  }
  // forward elimination
  double A = 1.0, Bx = 0.0, By = 0.0;
#pragma unroll
  for (int i = 0; i < KT; ++i) {
    const int j = j0 + i;
    if (i < K && j < nz) {
      const double l = (j == 0 || j == nz - 1) ? 0.0 : lo;
      const double m = -l * fc[i].y;
      A = m * A;
      Bx = m * Bx + d[i].x * fc[i].y;
      By = m * By + d[i].y * fc[i].y;
    }
  }
#pragma unroll
  for (int off = 1; off < 64; off <<= 1) {
    const double Ao = __shfl_up(A, off, 64);
    const double Xo = __shfl_up(Bx, off, 64);
    const double Yo = __shfl_up(By, off, 64);
    if (lane >= off) {
      Bx = A * Xo + Bx;
      By = A * Yo + By;
      A = A * Ao;
    }
  }
  double px = __shfl_up(Bx, 1, 64), py = __shfl_up(By, 1, 64);
  if (lane == 0) px = py = 0.0;
  if constexpr (WPR == 2) {
    // the second wavefront's lanes continue from the first's last value
    double ax = __shfl_up(A, 1, 64);
    if (lane == 0) ax = 1.0;
    if (t == 63) {
      xch[0] = Bx;
      xch[1] = By;
    }
    __syncthreads();
    if (wv == 1) {
      px = ax * xch[0] + px;
      py = ax * xch[1] + py;
    }
  }
#pragma unroll
  for (int i = 0; i < KT; ++i) {
    const int j = j0 + i;
    if (i < K && j < nz) {
      const double l = (j == 0 || j == nz - 1) ? 0.0 : lo;
      d[i].x = (d[i].x - l * px) * fc[i].y;
      d[i].y = (d[i].y - l * py) * fc[i].y;
      px = d[i].x;
      py = d[i].y;
    }
  }
  // back substitution
  A = 1.0;
  Bx = 0.0;
  By = 0.0;
#pragma unroll
  for (int i = KT - 1; i >= 0; --i) {
    const int j = j0 + i;
    if (i < K && j < nz) {
      const double c = fc[i].x;
      A = -c * A;
      Bx = d[i].x - c * Bx;
      By = d[i].y - c * By;
    }
  }
#pragma unroll
  for (int off = 1; off < 64; off <<= 1) {
    const double Ao = __shfl_down(A, off, 64);
    const double Xo = __shfl_down(Bx, off, 64);
    const double Yo = __shfl_down(By, off, 64);
    if (lane + off < 64) {
      Bx = A * Xo + Bx;
      By = A * Yo + By;
      A = A * Ao;
    }
  }
  double xx = __shfl_down(Bx, 1, 64), xy = __shfl_down(By, 1, 64);
  if (lane == 63) xx = xy = 0.0;
  if constexpr (WPR == 2) {
    // the first wavefront's lanes continue from the second's first value
    double ax = __shfl_down(A, 1, 64);
    if (lane == 63) ax = 1.0;
    if (t == 64) {
      xch[2] = Bx;
      xch[3] = By;
    }
    __syncthreads();
    if (wv == 0) {
      xx = ax * xch[2] + xx;
      xy = ax * xch[3] + xy;
    }
  }
#pragma unroll
  for (int i = KT - 1; i >= 0; --i) {
    const int j = j0 + i;
    if (i < K && j < nz) {
      xx = d[i].x - fc[i].x * xx;
      xy = d[i].y - fc[i].x * xy;
      d[i] = make_double2(xx, xy);
    }
  }
#pragma unroll
  for (int i = 0; i < KT; ++i) {
    const int j = j0 + i;
    if (i < K && j < nz) buf[j] = d[i];
  }
}

// First z difference of the row in buf at j (pddz, SpatialDifferentiator.py:
// 91-104, 159-185): zero on the end points the stencil does not reach.
__device__ __forceinline__ double2 ddz_row(const double2* __restrict__ b, int j, int nz, int order,
                                           double ih) {
  if (order == 2) {
    if (j < 1 || j > nz - 2) return make_double2(0.0, 0.0);
    return make_double2((b[j + 1].x - b[j - 1].x) * ih, (b[j + 1].y - b[j - 1].y) * ih);
  }
  if (j < 2 || j > nz - 3) return make_double2(0.0, 0.0);
  return make_double2(
      (-0.25 * b[j + 2].x + 2.0 * b[j + 1].x - 2.0 * b[j - 1].x + 0.25 * b[j - 2].x) * ih,
      (-0.25 * b[j + 2].y + 2.0 * b[j + 1].y - 2.0 * b[j - 1].y + 0.25 * b[j - 2].y) * ih);
}

// Second z difference (pd2dz2, SpatialDifferentiator.py:106-128).
__device__ __forceinline__ double2 d2dz2_row(const double2* __restrict__ b, int j, int nz,
                                             int order, double idz2) {
  if (order == 2) {
    if (j < 1 || j > nz - 2) return make_double2(0.0, 0.0);
    return make_double2((b[j + 1].x - 2.0 * b[j].x + b[j - 1].x) * idz2,
                        (b[j + 1].y - 2.0 * b[j].y + b[j - 1].y) * idz2);
  }
  if (j < 2 || j > nz - 3) return make_double2(0.0, 0.0);
  const double c1 = -1.0 / 12, c2 = 4.0 / 3, c3 = -5.0 / 2;
  return make_double2(
      (c1 * b[j + 2].x + c2 * b[j + 1].x + c3 * b[j].x + c2 * b[j - 1].x + c1 * b[j - 2].x) * idz2,
      (c1 * b[j + 2].y + c2 * b[j + 1].y + c3 * b[j].y + c2 * b[j - 1].y + c1 * b[j - 2].y) * idz2);
}

// --------------------------------------------------------------------------
// Row kernel body: mode n, field f (0 omega, 1 T), by the row's WPR
// wavefronts (thread t of 64 WPR).  Element-wise work uses the coalesced
// layout j = t + 64 WPR i.  The row owns two LDS rows of FP64 complex
// values: buf (operands with neighbours, and the tridiagonal solves) and
// vrow (the state v_n the update starts from).
// --------------------------------------------------------------------------
template <typename T, int KT, int WPR>
__device__ __forceinline__ void fdz_row(const FdzJob<T>& J, int stage, int n, int f, int step,
                                        double2* __restrict__ buf, double2* __restrict__ vrow,
                                        int t, double* __restrict__ xch) {
  constexpr int LR = 64 * WPR;  // threads per row
  const int Z = J.nz, K = J.K;
  const int par = (J.par0 + step) & 1;
  const size_t row = (size_t)n * Z;
  const int order = J.fd_order;
  const int kw = order == 2 ? 1 : 2;  // width of the wall band
  const double dz = J.dz;
  const double ih = order == 2 ? 1.0 / (2.0 * dz) : 1.0 / (3.0 * dz);
  if (stage != FDZ_PROLOGUE) {
    const int q = J.q;
    const int cur = ring(J.curr0, step, q);
    const cx<T>* __restrict__ qx = glb(J.qw[2 * f]) + row;
    const cx<T>* __restrict__ qz = glb(J.qw[2 * f + 1]) + row;
    const cx<T>* __restrict__ vin = glb(J.state[par][f]) + row;
    const cx<T>* __restrict__ cin = J.comp[f] ? glb(J.comp[f]) + row : nullptr;
#pragma unroll
    for (int i = 0; i < KT; ++i) {
      const int j = t + LR * i;
      if (i < K && j < Z) {
        buf[j] = d2<T>(qz[j]);
        double2 v = d2<T>(vin[j]);
        if (cin) {
          const double2 c = d2<T>(cin[j]);
          v.x -= c.x;
          v.y -= c.y;
        }
        vrow[j] = v;
      }
    }
    row_sync<WPR>();
    double sx;
    {
      const double th = 2.0 * M_PI * (double)n / (double)J.nx;
      sx = order == 2 ? sin(th) / J.dx : (8.0 * sin(th) - sin(2.0 * th)) / (6.0 * J.dx);
    }
    // -Pr Ra i (2 pi n / lx) T  (rayleigh_benard_convection.py:101-104)
    const double buoy = f == 0 ? J.Pr * J.Ra * (J.fx * (double)n) : 0.0;
    const cx<T>* __restrict__ tsrc = glb(stage == FDZ_CORR ? J.star[1] : J.state[par][1]) + row;
    const double coeff = f == 0 ? J.Pr : 1.0;
    const double k2 = J.fx2 * (double)n * (double)n;  // -(2 pi n / lx)^2
    const double idz2 = 1.0 / (dz * dz);
    const bool expl = J.scheme == MLV_FDZ_EXPLICIT;
    const int ws = stage == FDZ_CORR ? (cur + 1) % q : cur;  // ring slot written
    const int s1 = (ws + q - 1) % q, s2 = (ws + q - 2) % q, s3 = (ws + q - 3) % q;
    const size_t ssz = (size_t)J.nn * Z;
    cx<T>* __restrict__ hist = glb(J.hist[f]);
    const double dt = J.dt;
    double2 inc[KT];
#pragma unroll
    for (int i = 0; i < KT; ++i) {
      const int j = t + LR * i;
      inc[i] = make_double2(0.0, 0.0);
      if (i < K && j < Z) {
        const size_t e = row + j;
        // issue the global loads first
        const double2 a = d2<T>(qx[j]);
        const double2 t = f == 0 ? d2<T>(tsrc[j]) : make_double2(0.0, 0.0);
        const double2 h1 = d2<T>(hist[(size_t)s1 * ssz + e]);
        double2 h2 = make_double2(0.0, 0.0), h3 = make_double2(0.0, 0.0);
        if (q == 4) {
          h2 = d2<T>(hist[(size_t)s2 * ssz + e]);
          h3 = d2<T>(hist[(size_t)s3 * ssz + e]);
        }
        // advection N = -(i sx Qx + Dz Qz), buoyancy
        const double2 dd = ddz_row(buf, j, Z, order, ih);
        double2 h0 = make_double2(sx * a.y - dd.x, -(sx * a.x) - dd.y);
        h0.x += buoy * t.y;
        h0.y -= buoy * t.x;
        const double2 v = vrow[j];
        double2 lv = make_double2(0.0, 0.0);
        if (expl) {
          // diffusion added into the stored RHS (Integrator.py:53-56):
          // coeff * (sd2dx2 + pd2dz2) v
          const double2 dzz = d2dz2_row(vrow, j, Z, order, idz2);
          h0.x += coeff * (k2 * v.x + dzz.x);
          h0.y += coeff * (k2 * v.y + dzz.y);
        } else if (j > 0 && j < Z - 1) {
          // L v = -kx^2 v + second-order D2z v on interior points, 0 on walls
          lv = make_double2(k2 * v.x + (vrow[j + 1].x - 2.0 * v.x + vrow[j - 1].x) * idz2,
                            k2 * v.y + (vrow[j + 1].y - 2.0 * v.y + vrow[j - 1].y) * idz2);
        }
        hist[(size_t)ws * ssz + e] = t2<T>(h0);
        double2 E;
        if (stage == FDZ_CORR) {  // Adams-Moulton (Integrator.py:20-33)
          if (q == 2) {
            E = make_double2(dt / 2 * (h0.x + h1.x), dt / 2 * (h0.y + h1.y));
          } else {
            E = make_double2(dt / 24 * (9 * h0.x + 19 * h1.x - 5 * h2.x + h3.x),
                             dt / 24 * (9 * h0.y + 19 * h1.y - 5 * h2.y + h3.y));
          }
        } else {  // Adams-Bashforth (Integrator.py:5-18)
          if (q == 2) {
            E = make_double2(dt / 2 * (3 * h0.x - h1.x), dt / 2 * (3 * h0.y - h1.y));
          } else {
            E = make_double2(dt / 24 * (55 * h0.x - 59 * h1.x + 37 * h2.x - 9 * h3.x),
                             dt / 24 * (55 * h0.y - 59 * h1.y + 37 * h2.y - 9 * h3.y));
          }
        }
        inc[i] = expl ? E : make_double2(dt * coeff * lv.x + E.x, dt * coeff * lv.y + E.y);
      }
    }
    row_sync<WPR>();
#pragma unroll
    for (int i = 0; i < KT; ++i) {
      const int j = t + LR * i;
      if (i < K && j < Z) buf[j] = inc[i];
    }
    if (!expl) {
      // increment form of the theta scheme:
      //   (I - a dt c L) delta = dt c L v + E,  v_new = v + delta
      row_sync<WPR>();
      row_tridiag<KT, WPR>(buf, glb(J.fac[1 + f]) + (size_t)n * K * LR, J.lo[1 + f], Z, K,
                           t, xch);
    }
    row_sync<WPR>();
    // new state and wall writes (rayleigh_benard_convection.py:113-142);
    // buf[j] becomes -omega_new for the Poisson solve (same thread, same j)
    cx<T>* __restrict__ out = glb(stage == FDZ_PRED ? J.star[f] : J.state[par ^ 1][f]) + row;
    cx<T>* __restrict__ cout = (stage != FDZ_PRED && J.comp[f]) ? glb(J.comp[f]) + row : nullptr;
#pragma unroll
    for (int i = 0; i < KT; ++i) {
      const int j = t + LR * i;
      if (i < K && j < Z) {
        const double2 dl = buf[j], v0 = vrow[j];
        double2 v = make_double2(v0.x + dl.x, v0.y + dl.y);
        bool wall;
        if (f == 0) {
          wall = n == 0 || j < kw || j >= Z - kw;
          if (wall) v = make_double2(0.0, 0.0);
        } else {
          wall = j < kw || j >= Z - kw;
          if (wall) v = make_double2((j < kw && n == 0) ? 1.0 : 0.0, 0.0);
        }
        const cx<T> hi = t2<T>(v);
        out[j] = hi;
        if (cout) {
          // value = hi - residual, with the residual rounded once
          cout[j] = wall ? czero<T>() : mk<T>((T)((double)hi.x - v.x), (T)((double)hi.y - v.y));
        }
        // the Poisson solve sees the value the next evaluation reads
        const double2 wv = (cout || sizeof(T) == 8) ? v : d2<T>(hi);
        buf[j] = make_double2(-wv.x, -wv.y);
      }
    }
  } else {
    const cx<T>* __restrict__ vin = glb(J.state[par][0]) + row;
    const cx<T>* __restrict__ cin = J.comp[0] ? glb(J.comp[0]) + row : nullptr;
#pragma unroll
    for (int i = 0; i < KT; ++i) {
      const int j = t + LR * i;
      if (i < K && j < Z) {
        double2 w = d2<T>(vin[j]);
        if (cin) {
          const double2 c = d2<T>(cin[j]);
          w.x -= c.x;
          w.y -= c.y;
        }
        buf[j] = make_double2(-w.x, -w.y);
      }
    }
  }
  if (f != 0) return;
  // psi = solve(-omega) (utility.py:65, LaplacianSolver.py:70-78),
  // ux_s = -pddz(psi), uz_s = i kx psi (utility.py:66-79)
  row_sync<WPR>();
  row_tridiag<KT, WPR>(buf, glb(J.fac[0]) + (size_t)n * K * LR, J.lo[0], Z, K, t, xch);
  row_sync<WPR>();
  const int vs = stage == FDZ_PRED ? 2 : (stage == FDZ_PROLOGUE ? par : par ^ 1);
  cx<T>* __restrict__ ux = glb(J.vel[vs][0]) + row;
  cx<T>* __restrict__ uz = glb(J.vel[vs][1]) + row;
  cx<T>* __restrict__ ps = vs < 2 && J.psi[vs] ? glb(J.psi[vs]) + row : nullptr;
  const double kx = J.fx * (double)n;
#pragma unroll
  for (int i = 0; i < KT; ++i) {
    const int j = t + LR * i;
    if (i < K && j < Z) {
      const double2 p = buf[j];
      const double2 dd = ddz_row(buf, j, Z, order, ih);
      ux[j] = t2<T>(make_double2(-dd.x, -dd.y));
      uz[j] = t2<T>(make_double2(-(kx * p.y), kx * p.x));
      if (ps) ps[j] = (n == 0 || j < kw || j >= Z - kw) ? czero<T>() : t2<T>(p);
    }
  }
}

// Blocks of fdz_rows_per_block(WPR) rows of WPR wavefronts each.
template <typename T, int KT, int WPR>
__global__ void __launch_bounds__(fdz_row_threads(WPR))
    k_fdz_rows(const FdzJob<T>* __restrict__ jobs, FdzJob<T>* clock, const int* __restrict__ pref,
               int njobs, int phase, int step, int zstride) {
  extern __shared__ char smem_raw[];
  __shared__ double xch[4];
  const int lb = xcd_logical(pref[njobs]);
  if (lb < 0) return;
  const int jb = find_job(pref, njobs, lb);
  const FdzJob<T>& J = jobs[jb];
  constexpr int LR = 64 * WPR;
  const int rib = threadIdx.x / LR, t = threadIdx.x % LR;  // row in block, thread in row
  const int lw = (lb - pref[jb]) * fdz_rows_per_block(WPR) + rib;
  int stage, nwf;
  if (phase < 0) {
    stage = FDZ_PROLOGUE;
    nwf = 1;
  } else {
    stage = J.scheme == MLV_FDZ_PC ? (phase == 0 ? FDZ_PRED : FDZ_CORR) : FDZ_MAIN;
    nwf = 2;
  }
  if (phase == (J.scheme == MLV_FDZ_PC ? 1 : 0) && lb == pref[jb] && threadIdx.x == 0) {
    // t += dt after each loop (Simulation.py:199-209), through `clock` (the
    // same table as a plain pointer); this launch reads these two fields
    // through neither pointer
    clock[jb].t = clock[jb].t + J.dt;
    clock[jb].loops = clock[jb].loops + 1;
  }
  // (uniform over a block when WPR > 1: the block is one row)
  if (lw >= nwf * J.nn || J.kt != KT || J.wpr != WPR) return;
  const int n = lw / nwf, f = lw % nwf;
  double2* buf = reinterpret_cast<double2*>(smem_raw) + (size_t)(2 * rib) * zstride;
  fdz_row<T, KT, WPR>(J, stage, n, f, step, buf, buf + zstride, t, xch);
}

// --------------------------------------------------------------------------
// Column kernel: TC z columns per block (TC even), two LDS slots per column.
//
// Two real sequences share one complex transform (a + i b).  The rounding
// of a transform is relative to the whole complex sequence, so the pair
// must hold values of similar size: omega is ~Ra times T once convection
// starts, and packing omega with T would bury T under omega's rounding.
// Each slot therefore pairs the same field at two adjacent z columns.
// --------------------------------------------------------------------------
template <typename T>
__device__ __forceinline__ void herm_pack(cx<T>* __restrict__ z, int k, int N, cx<T> a, cx<T> b) {
  // Spectra of two real sequences (k < nn kept) packed into one complex
  // sequence a + i b; irfft ignores the imaginary part of the k = 0 term.
  if (k == 0) {
    z[0] = mk<T>(a.x, b.x);
  } else {
    z[fft_pad(k)] = mk<T>(a.x - b.y, a.y + b.x);
    z[fft_pad(N - k)] = mk<T>(a.x + b.y, b.x - a.y);
  }
}

template <typename T>
__host__ __device__ constexpr int fdz_tc(int slots) {
  return slots / 2 > kNT ? kNT : slots / 2;
}

template <typename T, int LOGN>
__device__ __forceinline__ void fdz_cols_body(const FdzJob<T>& J, int local, int phase, int flags,
                                              int step, const cx<T>* __restrict__ tw,
                                              cx<T>* lds) {
  constexpr int N = 1 << LOGN;
  constexpr int P = fdz_points<T>(LOGN);
  constexpr int SL = fft_slots<LOGN, kNT, P>();
  constexpr int TC = fdz_tc<T>(SL);
  if constexpr (TC < 2) {
    return;  // rejected on the host
  } else {
    constexpr int LDSN = fft_ldsn(N);
    const int Z = J.nz, nn = J.nn;
    const int par = (J.par0 + step) & 1;
    const bool corr = J.scheme == MLV_FDZ_PC && phase == 1;
    const int j0 = local * TC;
    const int ncol = min(TC, Z - j0);
    const int npair = (ncol + 1) / 2;
    const int vs = corr ? 2 : par;
    // field f of column pair p lives in slot 4 p + f: 0 ux, 1 uz, 2 omega, 3 T
    const gptr<const cx<T>> src[4] = {glb(J.vel[vs][0]), glb(J.vel[vs][1]),
                                      glb(corr ? J.star[0] : J.state[par][0]),
                                      glb(corr ? J.star[1] : J.state[par][1])};
    // The x mean of T, Tm(z) = Re T[kx=0], is O(1) (the conductive profile)
    // while the rest is small: transforming T whole in single precision would
    // spread Tm's rounding over every mode.  Tm is constant along x, so
    // F_x(u Tm) = Tm F_x(u) exactly; the transforms carry T - Tm and the
    // Tm u_s term is added per mode after the forward transform.
    __shared__ T tmean[TC];  // (every size's body has its own)
    for (int cc = threadIdx.x; cc < ncol; cc += kNT) tmean[cc] = src[3][j0 + cc].x;
    const int nslot = 4 * npair;
    const int nzero = N - 2 * nn + 1;
    for (int e = threadIdx.x; e < nslot * nzero; e += kNT) {
      const int sl = e / nzero, k = nn + e % nzero;
      lds[sl * LDSN + fft_pad(k)] = czero<T>();
    }
    // one work item = (field, k, pair): both columns' values of one field;
    // neighbouring threads read neighbouring column pairs
    const int nfill = npair * 4 * nn;
    constexpr int U = 4;
#pragma unroll 1
    for (int e0 = threadIdx.x; e0 < nfill; e0 += U * kNT) {
      cx<T> a[U], b[U];
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const int e = e0 + u * kNT;
        const int pr = e % npair, k = (e / npair) % nn, f = e / (npair * nn);
        const int c0 = 2 * pr;
        a[u] = czero<T>();
        b[u] = czero<T>();
        if (e < nfill) {
          const size_t idx = (size_t)k * Z + j0 + c0;
          const bool drop = f == 3 && k == 0;  // T without its x mean
          if (!drop) {
            a[u] = src[f][idx];
            if (c0 + 1 < ncol) b[u] = src[f][idx + 1];
          }
        }
      }
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const int e = e0 + u * kNT;
        if (e < nfill) {
          const int pr = e % npair, k = (e / npair) % nn, f = e / (npair * nn);
          herm_pack<T>(lds + (4 * pr + f) * LDSN, k, N, a[u], b[u]);
        }
      }
    }
    __syncthreads();
    fft_lds<T, LOGN, +1, kNT, P>(lds, nslot, tw);
    const bool wp = (flags & kFdzWritePhys) && !corr;
    const bool dg = (flags & kFdzDiag) && !corr && J.loops % (flags >> kFdzDiagShift) == 0;
    T mux = -INFINITY, muz = -INFINITY;  // NaN-propagating maxima, as np.max
    double ke = 0.0, ens = 0.0, flux = 0.0;
    for (int e = threadIdx.x; e < npair * N; e += kNT) {
      const int pr = e / N, x = e % N;
      cx<T>* z = lds + (4 * pr) * LDSN + fft_pad(x);
      const cx<T> vx = z[0], vz = z[LDSN], w = z[2 * LDSN], t = z[3 * LDSN];
      if (dg) {
        // the imaginary parts belong to column 2 pr + 1 when it exists
        const int c0 = 2 * pr;
        const double t0 = (double)(t.x + tmean[c0]);
        mux = nan_max(mux, vx.x);
        muz = nan_max(muz, vz.x);
        ke += (double)vz.x * (double)vz.x + (double)vx.x * (double)vx.x;
        ens += (double)w.x * (double)w.x;
        flux += t0 * (double)vz.x;
        if (c0 + 1 < ncol) {
          const double t1 = (double)(t.y + tmean[c0 + 1]);
          mux = nan_max(mux, vx.y);
          muz = nan_max(muz, vz.y);
          ke += (double)vz.y * (double)vz.y + (double)vx.y * (double)vx.y;
          ens += (double)w.y * (double)w.y;
          flux += t1 * (double)vz.y;
        }
      }
      // real parts: column 2 pr, imaginary parts: column 2 pr + 1
      if (wp) {
        const int c0 = 2 * pr;
        const size_t o = (size_t)x * Z + j0 + c0;
        const bool two = c0 + 1 < ncol;
        const gptr<T> p0 = glb(J.phys[0]), p1 = glb(J.phys[1]), p2 = glb(J.phys[2]),
                      p3 = glb(J.phys[3]);
        if (p0) { p0[o] = vx.x; if (two) p0[o + 1] = vx.y; }
        if (p1) { p1[o] = vz.x; if (two) p1[o + 1] = vz.y; }
        if (p2) { p2[o] = w.x; if (two) p2[o + 1] = w.y; }
        if (p3) {
          p3[o] = t.x + tmean[c0];
          if (two) p3[o + 1] = t.y + tmean[c0 + 1];
        }
      }
      z[0] = mk<T>(vx.x * w.x, vx.y * w.y);              // ux omega
      z[LDSN] = mk<T>(vz.x * w.x, vz.y * w.y);           // uz omega
      z[2 * LDSN] = mk<T>(vx.x * t.x, vx.y * t.y);       // ux T'
      z[3 * LDSN] = mk<T>(vz.x * t.x, vz.y * t.y);       // uz T'
    }
    if (dg) block_diag<kNT>((double)mux, (double)muz, ke, ens, flux, glb(J.diag) + (size_t)kDiagPart * local);
    __syncthreads();
    fft_lds<T, LOGN, -1, kNT, P>(lds, nslot, tw);
    const T hs = (T)(0.5 / (double)N);
    // one work item = (column, k): the four outputs of that column
    for (int e = threadIdx.x; e < TC * nn; e += kNT) {
      const int cc = e % TC, k = e / TC;
      if (cc >= ncol) continue;
      const int pr = cc >> 1;
      const bool hi = cc & 1;
      const size_t idx = (size_t)k * Z + j0 + cc;
      cx<T> ov[4];
#pragma unroll
      for (int m = 0; m < 4; ++m) {
        const cx<T>* z = lds + (4 * pr + m) * LDSN;
        const cx<T> y = z[fft_pad(k)];
        const cx<T> ym = z[fft_pad((N - k) & (N - 1))];
        // (y + conj ym) / 2 is the first sequence's spectrum, (y - conj ym) / 2i the second's
        ov[m] = hi ? mk<T>(hs * (y.y + ym.y), hs * (ym.x - y.x))
                   : mk<T>(hs * (y.x + ym.x), hs * (y.y - ym.y));
      }
      cx<T> vx = src[0][idx], vz = src[1][idx];
      if (k == 0) {  // irfft keeps the real part of the kx = 0 term
        vx.y = T(0);
        vz.y = T(0);
      }
      const T tm = tmean[cc];
      glb(J.qw[0])[idx] = ov[0];
      glb(J.qw[1])[idx] = ov[1];
      glb(J.qw[2])[idx] = mk<T>(ov[2].x + tm * vx.x, ov[2].y + tm * vx.y);
      glb(J.qw[3])[idx] = mk<T>(ov[3].x + tm * vz.x, ov[3].y + tm * vz.y);
    }
  }
}

template <typename T>
__global__ void __launch_bounds__(kNT)
    k_fdz_cols(const FdzJob<T>* __restrict__ jobs, const int* __restrict__ pref, int njobs,
               int phase, int flags, int step, const cx<T>* __restrict__ tw) {
  extern __shared__ char smem_raw[];
  cx<T>* lds = reinterpret_cast<cx<T>*>(smem_raw);
  const int lb = xcd_logical(pref[njobs]);
  if (lb < 0) return;
  const int jb = find_job(pref, njobs, lb);
  const FdzJob<T>& J = jobs[jb];
  const int local = lb - pref[jb];
  switch (J.lognx) {
#define X(L)                                                          \
  case L:                                                             \
    fdz_cols_body<T, L>(J, local, phase, flags, step, tw, lds);       \
    break;
    X(3) X(4) X(5) X(6) X(7) X(8) X(9) X(10) X(11) X(12)
#undef X
    default:
      break;
  }
}

// Thomas factors of one job's operator for mode n, by lr = 64 wpr threads
// (FP64), in the row kernel's layout.
//   which = 0: Poisson, interior bands [1/dz^2, -(kx^2 + 2/dz^2), 1/dz^2]
//              (LaplacianSolver.py:21-52)
//   which > 0: theta diffusion, [-a/dz^2, 1 + a (kx^2 + 2/dz^2), -a/dz^2]
// First and last rows (and the padding up to lr K) are identity rows.  The
// table is chunk-major (entry i * lr + t is element t K + i): thread t
// computes its chunk [t K, t K + K) and stores entry i of it with the
// other threads' entries i, so each store instruction writes consecutive
// entries.  The interior recurrence cp_j = lo / (dg - lo cp_{j-1}) has
// constant coefficients: a thread runs it from the first row up to its
// chunk, stopping once it maps a value to itself in floating point (every
// later factor is then that value; diffusion gets there within a few dozen
// rows), and then through its chunk -- the same operations in the same
// order as a single sequential pass, so the table is bit-identical to one.
__device__ void fdz_factor_row(double2* __restrict__ fac, int n, int nz, int K, int lr,
                               double fxabs, double dz, int which, double a, int t) {
  gptr<double2> F = glb(fac) + (size_t)n * K * lr;
  const double k = (double)n * fxabs;
  const double k2 = k * k;
  const double idz2 = 1.0 / (dz * dz);
  double lo, dg;
  if (which == 0) {
    lo = idz2;
    dg = -(k2 + 2.0 / (dz * dz));
  } else {
    lo = -a * idz2;
    dg = 1.0 + a * (k2 + 2.0 / (dz * dz));
  }
  const int j0 = t * K;
  // factors of element j - 1 before element j (row 0: cp = 0)
  double cp = 0.0, rp = 0.0;
  for (int j = 1; j < j0 && j < nz - 1; ++j) {
    const double rd = 1.0 / (dg - lo * cp);
    const double c = lo * rd;
    const bool fixed = c == cp && rd == rp;
    cp = c;
    rp = rd;
    if (fixed) break;
  }
  for (int i = 0; i < K; ++i) {
    const int j = j0 + i;
    double2 v = make_double2(0.0, 1.0);  // identity rows and padding
    if (j >= 1 && j < nz - 1) {
      const double rd = 1.0 / (dg - lo * cp);
      const double c = lo * rd;
      cp = c;
      rp = rd;
      v = make_double2(c, rd);
    }
    F[i * lr + t] = v;
  }
}

// Grid: one block of lr threads per mode.
__global__ void __launch_bounds__(128) k_fdz_factors(double2* __restrict__ fac, int nn, int nz,
                                                     int K, int lr, double fxabs, double dz,
                                                     int which, double a) {
  if ((int)blockIdx.x < nn && (int)threadIdx.x < lr)
    fdz_factor_row(fac, blockIdx.x, nz, K, lr, fxabs, dz, which, a, threadIdx.x);
}

// One block per job: tracker record and CFL rule (diag_apply) from the
// column kernel's partials; a changed dt also moves the theta operators'
// interior bands and flags the job for k_fdz_refactor.
template <typename T>
__global__ void __launch_bounds__(kNT) k_fdz_diag(FdzJob<T>* jobs, double cutoff, int cadence) {
  FdzJob<T>& J = jobs[blockIdx.x];
  // loops has been advanced by this step's last row kernel
  if ((J.loops - 1) % cadence != 0) {
    if (threadIdx.x == 0) J.refactor = 0;
    return;
  }
  double dt = J.dt;
  const bool changed = diag_apply<kNT>(J.diag, J.nbC, J.dx, J.dz, (double)J.nx * (double)J.nz,
                                       true, cutoff, J.t, J.loops, dt);
  if (threadIdx.x == 0) {
    J.refactor = changed ? 1 : 0;
    if (changed) {
      const double idz2 = 1.0 / (J.dz * J.dz);
      J.dt = dt;
      J.lo[1] = -(J.alpha * dt * J.Pr) * idz2;
      J.lo[2] = -(J.alpha * dt * 1.0) * idz2;
    }
  }
}

// Theta tables of the jobs whose dt the last CFL update changed: grid
// (modes, jobs, 2 operators) of 128-thread blocks, 64 wpr of them used.
template <typename T>
__global__ void __launch_bounds__(128) k_fdz_refactor(const FdzJob<T>* __restrict__ jobs) {
  const FdzJob<T>& J = jobs[blockIdx.y];
  if (!J.refactor || J.scheme == MLV_FDZ_EXPLICIT || (int)threadIdx.x >= 64 * J.wpr) return;
  const int n = blockIdx.x;
  if (n >= J.nn) return;
  const int which = 1 + (int)blockIdx.z;
  const double a = J.alpha * J.dt * (which == 1 ? J.Pr : 1.0);
  fdz_factor_row(J.fac[which], n, J.nz, J.K, 64 * J.wpr, J.fx, J.dz, which, a, threadIdx.x);
}

template <typename T>
__global__ void k_fdz_set_dt(FdzJob<T>* jobs, int j, double dt, double lo_w, double lo_t) {
  if (threadIdx.x == 0 && blockIdx.x == 0) {
    jobs[j].dt = dt;
    jobs[j].lo[1] = lo_w;
    jobs[j].lo[2] = lo_t;
  }
}

// --------------------------------------------------------------------------
// Host side
// --------------------------------------------------------------------------
template <typename T, int V>
static void launch_rows(const FdzJob<T>* jobs, const FdzLaunch& fl, const int* pref, int nblocks,
                        int phase, int step, hipStream_t s) {
  if (nblocks <= 0) return;
  constexpr int KT = kFdzKts[V], WPR = kFdzWprs[V];
  MLV_SET_LDS_ONCE((k_fdz_rows<T, KT, WPR>), fl.ldsrow[V]);
  hipLaunchKernelGGL((k_fdz_rows<T, KT, WPR>), dim3(xcd_grid(nblocks)), dim3(fdz_row_threads(WPR)),
                     fl.ldsrow[V], s, jobs, const_cast<FdzJob<T>*>(jobs), pref, fl.njobs, phase,
                     step, fl.zstride);
}

template <typename T>
static void launch_rows_all(const FdzJob<T>* jobs, const FdzLaunch& fl, int phase, int step,
                            hipStream_t s) {
  const int p = phase < 0 ? 2 : phase;
  launch_rows<T, 0>(jobs, fl, fl.prefrow[p][0], fl.nrow[p][0], phase, step, s);
  launch_rows<T, 1>(jobs, fl, fl.prefrow[p][1], fl.nrow[p][1], phase, step, s);
  launch_rows<T, 2>(jobs, fl, fl.prefrow[p][2], fl.nrow[p][2], phase, step, s);
  launch_rows<T, 3>(jobs, fl, fl.prefrow[p][3], fl.nrow[p][3], phase, step, s);
  launch_rows<T, 4>(jobs, fl, fl.prefrow[p][4], fl.nrow[p][4], phase, step, s);
  static_assert(kFdzVariants == 5, "row kernel variants");
}

template <typename T>
hipError_t fdz_phase(const FdzJob<T>* jobs, const FdzLaunch& fl, int phase, int flags, int step,
                     const cx<T>* tw, hipStream_t s, hipEvent_t* ev) {
  if (ev) (void)hipEventRecord(ev[0], s);
  if (fl.ncol[phase] > 0) {
    MLV_SET_LDS_ONCE(k_fdz_cols<T>, fl.ldscol);
    hipLaunchKernelGGL(k_fdz_cols<T>, dim3(xcd_grid(fl.ncol[phase])), dim3(kNT), fl.ldscol, s,
                       jobs, (const int*)fl.prefcol[phase], fl.njobs, phase, flags, step, tw);
  }
  if (ev) (void)hipEventRecord(ev[1], s);
  launch_rows_all<T>(jobs, fl, phase, step, s);
  if (ev) (void)hipEventRecord(ev[2], s);
  return hipGetLastError();
}

template <typename T>
hipError_t fdz_prologue(const FdzJob<T>* jobs, const FdzLaunch& fl, int step, hipStream_t s) {
  launch_rows_all<T>(jobs, fl, -1, step, s);
  return hipGetLastError();
}

hipError_t fdz_factors(double2* fac, int nn, int nz, int K, int wpr, double fxabs, double dz,
                       int which, double a, hipStream_t s) {
  hipLaunchKernelGGL(k_fdz_factors, dim3(nn), dim3(64 * wpr), 0, s, fac, nn, nz, K, 64 * wpr,
                     fxabs, dz, which, a);
  return hipGetLastError();
}

template <typename T>
hipError_t fdz_set_dt(FdzJob<T>* jobs, int job, double dt, double lo_w, double lo_t,
                      hipStream_t s) {
  hipLaunchKernelGGL(k_fdz_set_dt<T>, dim3(1), dim3(64), 0, s, jobs, job, dt, lo_w, lo_t);
  return hipGetLastError();
}

template <typename T>
hipError_t fdz_diag(FdzJob<T>* jobs, int njobs, int max_nn, double cutoff, int cadence,
                    hipStream_t s) {
  hipLaunchKernelGGL(k_fdz_diag<T>, dim3(njobs), dim3(kNT), 0, s, jobs, cutoff, cadence);
  hipLaunchKernelGGL(k_fdz_refactor<T>, dim3(max_nn, njobs, 2), dim3(128), 0, s,
                     (const FdzJob<T>*)jobs);
  return hipGetLastError();
}

template <typename T>
hipError_t fdz_prepare(const FdzLaunch& fl) {
  if (fl.ldscol) MLV_SET_LDS_ONCE(k_fdz_cols<T>, fl.ldscol);
#define MLV_ROWS_LDS(V) \
  if (fl.ldsrow[V]) MLV_SET_LDS_ONCE((k_fdz_rows<T, kFdzKts[V], kFdzWprs[V]>), fl.ldsrow[V]);
  MLV_ROWS_LDS(0) MLV_ROWS_LDS(1) MLV_ROWS_LDS(2) MLV_ROWS_LDS(3) MLV_ROWS_LDS(4)
#undef MLV_ROWS_LDS
  return hipGetLastError();
}

template hipError_t fdz_phase<float>(const FdzJob<float>*, const FdzLaunch&, int, int, int,
                                     const cx<float>*, hipStream_t, hipEvent_t*);
template hipError_t fdz_phase<double>(const FdzJob<double>*, const FdzLaunch&, int, int, int,
                                      const cx<double>*, hipStream_t, hipEvent_t*);
template hipError_t fdz_prologue<float>(const FdzJob<float>*, const FdzLaunch&, int, hipStream_t);
template hipError_t fdz_prologue<double>(const FdzJob<double>*, const FdzLaunch&, int, hipStream_t);
template hipError_t fdz_set_dt<float>(FdzJob<float>*, int, double, double, double, hipStream_t);
template hipError_t fdz_set_dt<double>(FdzJob<double>*, int, double, double, double, hipStream_t);
template hipError_t fdz_diag<float>(FdzJob<float>*, int, int, double, int, hipStream_t);
template hipError_t fdz_diag<double>(FdzJob<double>*, int, int, double, int, hipStream_t);
template hipError_t fdz_prepare<float>(const FdzLaunch&);
template hipError_t fdz_prepare<double>(const FdzLaunch&);

}  // namespace mlv
