// Device body of the fused step's update kernel C (forward x pass, FD
// operator, sources, Adams-Bashforth + semi-implicit update), launched by
// fused_c.hip in its generic and size-specialised forms.
#pragma once
#include "fused.hpp"
#include "diag_dev.hpp"
#include "grid.hpp"

namespace mlv {
// --------------------------------------------------------------------------
// Kernel C: forward x pass, FD operator, sources, AB + semi-implicit update
// --------------------------------------------------------------------------
template <typename T>
__device__ __forceinline__ T modk(const cx<T>* __restrict__ tw, int k, int logn, int order,
                                  double h) {
  const T s1 = tsin<T>(tw, k, logn);
  if (order == 2) return s1 / (T)h;
  const T s2 = tsin<T>(tw, 2 * k, logn);
  return (T(8) * s1 - s2) / (T)(6.0 * h);
}

template <typename T, int LOGN, int WIDE, int NT = kNT>
__device__ __forceinline__ void stepC_body(const JobDev<T>& J, const StageDev& S, int local, int cur,
                                           int par, const cx<T>* __restrict__ tw, cx<T>* lds) {
  constexpr int N = 1 << LOGN;
  constexpr int P = fused_points<T>(LOGN, WIDE);
  constexpr int TC = fused_cols<LOGN, P, NT>();
  constexpr int LDSN = fft_ldsn(N);
  const int nn = J.nn, nm = J.nm, rows = J.rows;
  const size_t chs = work_chs(N, nm);
  const size_t ssz = (size_t)rows * nm;
  const int ntile = (nm + TC - 1) / TC;
  const int eqi = local / ntile, tile = local % ntile;
  const int field = S.eq_field[eqi];
  const int c0 = tile * TC;
  const int ncol = min(TC, nm - c0);
  const cx<T>* __restrict__ pxc = glb(J.work) + (size_t)pchan(J.pbase, eqi, 0) * chs;
  const cx<T>* __restrict__ pzc = glb(J.work) + (size_t)pchan(J.pbase, eqi, 1) * chs;
  // One transform per column: the periodic x difference acts on the
  // intermediate rows (it commutes with the z transform) and the z
  // difference is its modified wavenumber, so
  //   D = F_x[ FDx(Px) + i sz(kz) Pz ].
  const int order = J.fd_order;
  const T ihx = order == 2 ? (T)(1.0 / (2.0 * J.dx)) : (T)(1.0 / (3.0 * J.dx));
  constexpr int ITER = (N * TC) / NT;
  // loads in flight per thread: Px phase UP, Pz phase U (held beside the
  // stencil's registers)
  constexpr int UP = 8;
  constexpr int U = (sizeof(T) == 4 && !WIDE) ? 4 : 8;
  static_assert(ITER % U == 0 && ITER % UP == 0, "load batching");
  // sz depends only on the column; every thread keeps one column (e % TC
  // is fixed because NT is a multiple of TC)
  const int cme = threadIdx.x % TC;
  const T szc = modk<T>(tw, c0 + cme, J.lognz, order, J.dz);
  const bool cok = cme < ncol;
  const int col = c0 + cme;
  cx<T>* __restrict__ slot = lds + cme * LDSN;
  // Single precision with at most 16 points per thread: Px is loaded once
  // into the slot and the stencil reads its neighbours from LDS (4 % faster
  // at 1024^2 than four extra global loads per point, which hit L2).
  if constexpr (sizeof(T) == 4 && (WIDE || ITER <= 16)) {
    // 1. Px column into the slot, and the thread's Pz points into registers:
    //    both issued UP at a time, so the column arrives in ITER / UP memory
    //    round trips and the stencil phase needs no loads of its own
    cx<T> zv[ITER];
#pragma unroll
    for (int i0 = 0; i0 < ITER; i0 += UP) {
      cx<T> pv[UP];
#pragma unroll
      for (int u = 0; u < UP; ++u) {
        const int x = (threadIdx.x + (i0 + u) * NT) / TC;
        pv[u] = cok ? pxc[pidx(x, col, N)] : czero<T>();
        zv[i0 + u] = cok ? pzc[pidx(x, col, N)] : czero<T>();
      }
#pragma unroll
      for (int u = 0; u < UP; ++u) {
        const int x = (threadIdx.x + (i0 + u) * NT) / TC;
        slot[fft_pad(x)] = pv[u];
      }
    }
    __syncthreads();
    // 2. x stencil from LDS plus i sz Pz, held in registers, then 3. stored
    cx<T> dv[ITER];
#pragma unroll
    for (int i = 0; i < ITER; ++i) {
      const int x = (threadIdx.x + i * NT) / TC;
      auto at = [&](int d) { return slot[fft_pad((x + d) & (N - 1))]; };
      cx<T> d;
      if (order == 2) {
        const cx<T> a = at(1), b = at(-1);
        d = mk<T>((a.x - b.x) * ihx, (a.y - b.y) * ihx);
      } else {
        const cx<T> a2 = at(2), a1 = at(1), b1 = at(-1), b2 = at(-2);
        d = mk<T>((T(-0.25) * a2.x + T(2) * a1.x - T(2) * b1.x + T(0.25) * b2.x) * ihx,
                  (T(-0.25) * a2.y + T(2) * a1.y - T(2) * b1.y + T(0.25) * b2.y) * ihx);
      }
      dv[i] = cok ? mk<T>(d.x - szc * zv[i].y, d.y + szc * zv[i].x) : czero<T>();
    }
    __syncthreads();
#pragma unroll
    for (int i = 0; i < ITER; ++i) {
      const int x = (threadIdx.x + i * NT) / TC;
      slot[fft_pad(x)] = dv[i];
    }
  } else {
    // double precision and 32 points per thread: the register file is the
    // limit, so the stencil reads its neighbours straight from global
    // memory (L2 hits)
#pragma unroll 1
    for (int i0 = 0; i0 < ITER; i0 += U) {
      cx<T> pa[U], pb[U], pz[U], pa2[U], pb2[U];
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const int x = (threadIdx.x + (i0 + u) * NT) / TC;
        if (cok) {
          pa[u] = pxc[pidx((x + 1) & (N - 1), col, N)];
          pb[u] = pxc[pidx((x - 1) & (N - 1), col, N)];
          pz[u] = pzc[pidx(x, col, N)];
          if (order != 2) {
            pa2[u] = pxc[pidx((x + 2) & (N - 1), col, N)];
            pb2[u] = pxc[pidx((x - 2) & (N - 1), col, N)];
          }
        }
      }
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const int x = (threadIdx.x + (i0 + u) * NT) / TC;
        cx<T> v = czero<T>();
        if (cok) {
          cx<T> d;
          if (order == 2) {
            d = mk<T>((pa[u].x - pb[u].x) * ihx, (pa[u].y - pb[u].y) * ihx);
          } else {
            d = mk<T>((T(-0.25) * pa2[u].x + T(2) * pa[u].x - T(2) * pb[u].x + T(0.25) * pb2[u].x) * ihx,
                      (T(-0.25) * pa2[u].y + T(2) * pa[u].y - T(2) * pb[u].y + T(0.25) * pb2[u].y) * ihx);
          }
          v = mk<T>(d.x - szc * pz[u].y, d.y + szc * pz[u].x);
        }
        slot[fft_pad(x)] = v;
      }
    }
  }
  __syncthreads();
  // the epilogue gives column cc to wavefront cc when the block has as many
  // columns as wavefronts; with 64 threads per transform that wavefront
  // transformed the column itself
  constexpr bool WL_OUT = fft_tps<N, P>() == 64 && TC == NT / 64;
  fft_lds<T, LOGN, -1, NT, P, !WL_OUT>(lds, ncol, tw);

  const int q = J.q;
  const double dt = J.dt, alpha = J.alpha;
  const double c1 = (1 - alpha) * dt, c2 = alpha * dt;
  const double hab = q == 2 ? dt / 2 : dt / 24;
  const double coeff = J.coeff[field];
  const T scale = (T)J.scale;
  const cx<T>* __restrict__ sin_ = glb(J.state[par][field]);
  cx<T>* __restrict__ sout = glb(J.state[par ^ 1][field]);
  cx<T>* __restrict__ hist = glb(J.hist[field]);
  // Epilogue: one wavefront per column, lanes on consecutive rows (the
  // spectral state is kz-major, so a column is contiguous).  Loads of U2
  // row groups are issued before any is used.
  constexpr int KR = (2 * ((N - 1) / 3) + 1 + 63) / 64;  // bound on ceil(rows / 64)
  constexpr int U2 = 4;
  const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const bool ddc = J.problem == MLV_PROB_DDC;
  uint32_t* __restrict__ carry = sizeof(T) == 4 ? glb(J.comp[field]) : nullptr;
  const int prev1 = (cur + q - 1) % q, prev2 = (cur + q - 2) % q, prev3 = (cur + q - 3) % q;
  const float fcoeff = (float)coeff, ffx2 = (float)J.fx2, ffz2 = (float)J.fz2;
  const float fdt = (float)dt, fc2 = (float)c2;
  // one wavefront per column (WPC per column when the block has fewer
  // columns than wavefronts, splitting its row groups)
  constexpr int NW = NT / 64;
  constexpr int WPC = TC >= NW ? 1 : NW / TC;
  const int wsub = wv % WPC;
#pragma unroll 1
  for (int cc = wv / WPC; cc < ncol; cc += NW / WPC) {
    const int m = c0 + cc;
    const size_t cbase = (size_t)m * rows;
    const cx<T>* __restrict__ dcol = lds + cc * LDSN;
#pragma unroll 1
    for (int i0 = wsub * U2; i0 < KR; i0 += WPC * U2) {
      cx<T> vv[U2], h1[U2], h2[U2], h3[U2], lo[U2], e1[U2], e2[U2];
#pragma unroll
      for (int u = 0; u < U2; ++u) {
        const int r = lane + 64 * (i0 + u);
        if (r < rows) {
          const size_t idx = cbase + r;
          vv[u] = sin_[idx];
          h1[u] = hist[(size_t)prev1 * ssz + idx];
          if (q == 4) {
            h2[u] = hist[(size_t)prev2 * ssz + idx];
            h3[u] = hist[(size_t)prev3 * ssz + idx];
          }
          if (carry) {
            const float2 b = bf2_load(carry[idx]);
            lo[u] = mk<T>((T)b.x, (T)b.y);
          }
          if (ddc) {
            if (field == 0) {
              e1[u] = glb(J.state[par][1])[idx];
              e2[u] = glb(J.state[par][2])[idx];
            } else {
              e1[u] = glb(J.state[par][0])[idx];
            }
          }
        }
      }
#pragma unroll
      for (int u = 0; u < U2; ++u) {
        const int r = lane + 64 * (i0 + u);
        if (r >= rows) continue;
        const int x = r <= nn ? r : r + N - rows;
        const int n = r <= nn ? r : r - rows;
        const size_t idx = cbase + r;
        const cx<T> D = dcol[fft_pad(x)];
        cx<T> rhs = mk<T>(D.x * scale, D.y * scale);
        if (ddc) {
          const double Pr = J.param[0], R0 = J.param[1];
          if (field == 0) {
            const T y = (T)(J.fx * n);
            const cx<T> t = e1[u], sv = e2[u];
            const T pr = (T)Pr;
            rhs = cadd<T>(rhs, mk<T>(pr * -(y * sv.y), pr * (y * sv.x)));
            rhs = csub<T>(rhs, mk<T>(pr * -(y * t.y), pr * (y * t.x)));
          } else {
            double lap = J.fx2 * (double)(n * n) + J.fz2 * (double)(m * m);
            if (r == 0 && m == 0) lap = 1.0;
            const T rl = (T)(1.0 / lap);
            const cx<T> w = e1[u];
            const cx<T> ps = mk<T>(-w.x * rl, -w.y * rl);
            const T y = (T)(J.fx * n);
            cx<T> uz = mk<T>(-(y * ps.y), y * ps.x);
            if (field == 2) {
              const T ir0 = (T)(1.0 / R0);
              uz = mk<T>(uz.x * ir0, uz.y * ir0);
            }
            rhs = csub<T>(rhs, uz);
          }
        }
        hist[(size_t)cur * ssz + idx] = rhs;
        T abx, aby;
        if (q == 2) {
          abx = T(3) * rhs.x - h1[u].x;
          aby = T(3) * rhs.y - h1[u].y;
        } else {
          abx = T(55) * rhs.x - T(59) * h1[u].x + T(37) * h2[u].x - T(9) * h3[u].x;
          aby = T(55) * rhs.y - T(59) * h1[u].y + T(37) * h2[u].y - T(9) * h3[u].y;
        }
        const cx<T> v = vv[u];
        cx<T> nv;
        const bool masked = ddc && field != 0 && m == 0;
        if constexpr (sizeof(T) == 8) {
          // (1 + (1-a) dt L) v + AB, divided by (1 - a dt L)  (Integrator.py:58-63)
          const double lin = coeff * (J.fx2 * (double)(n * n) + J.fz2 * (double)(m * m));
          const T num = (T)(1.0 + c1 * lin);
          const T rden = (T)(1.0 / (1.0 - c2 * lin));
          nv = mk<T>((num * v.x + (T)hab * abx) * rden, (num * v.y + (T)hab * aby) * rden);
        } else {
          // The same update written as an increment, v + (dt L v + AB)/(1 - a dt L),
          // so the rounding falls on the small change, not on (1 + (1-a) dt L);
          // the factors therefore only need single precision.  With a carry
          // array the addition is Kahan-compensated: increments below half an
          // ulp of v are kept instead of being dropped every step.
          const float lin = fcoeff * (ffx2 * (float)(n * n) + ffz2 * (float)(m * m));
          const T a = fdt * lin;
          const T rden = __builtin_amdgcn_rcpf(1.0f - fc2 * lin);
          const T ix = (a * v.x + (T)hab * abx) * rden, iy = (a * v.y + (T)hab * aby) * rden;
          if (carry) {
            const T yx = ix - lo[u].x, yy = iy - lo[u].y;
            const T tx = v.x + yx, ty = v.y + yy;
            carry[idx] = masked ? 0u : bf2_pack((float)((tx - v.x) - yx), (float)((ty - v.y) - yy));
            nv = mk<T>(tx, ty);
          } else {
            nv = mk<T>(v.x + ix, v.y + iy);
          }
        }
        if (masked) nv = czero<T>();
        sout[idx] = nv;
      }
    }
  }
}

// Kernel C's work for logical block lb of a launch over the job table's
// update blocks (prefix array pref).  The block of local index 0 of a job
// advances the job's device-side t and loop count after its last stage
// (Simulation.py:199-209), through `clock` -- the same table, as a plain
// pointer: the launch reads those two fields through neither pointer, so
// the read-only, non-aliased `jobs` stays valid for every field it reads.
template <typename T, int WIDE>
__device__ __forceinline__ void roleC(const JobDev<T>* __restrict__ jobs, JobDev<T>* clock,
                                      const int* __restrict__ pref, int njobs, int lb, int stage,
                                      int step, const cx<T>* __restrict__ tw, cx<T>* lds) {
  const int j = find_job(pref, njobs, lb);
  const JobDev<T>& J = jobs[j];
  const int local = lb - pref[j];
  const int par = (J.par0 + step) & 1;
  const int cur = ring(J.curr0, step, J.q);
  if (stage == J.nstage - 1 && local == 0 && threadIdx.x == 0) {
    clock[j].t = clock[j].t + J.dt;
    clock[j].loops = clock[j].loops + 1;
  }
  switch (J.lognx) {
#define X(L)                                                            \
  case L:                                                               \
    stepC_body<T, L, WIDE>(J, J.st[stage], local, cur, par, tw, lds);   \
    break;
    X(3) X(4) X(5) X(6) X(7) X(8) X(9) X(10) X(11) X(12)
#undef X
    default:
      break;
  }
}

// Kernel C's work for a batch whose jobs all have 2^LOGN points along x.
template <typename T, int SHAPE, int LOGN>
__device__ __forceinline__ void roleC_fixed(const JobDev<T>* __restrict__ jobs, JobDev<T>* clock,
                                            const int* __restrict__ pref, int njobs, int lb,
                                            int stage, int step, const cx<T>* __restrict__ tw,
                                            cx<T>* lds) {
  const int j = find_job(pref, njobs, lb);
  const JobDev<T>& J = jobs[j];
  const int local = lb - pref[j];
  const int par = (J.par0 + step) & 1;
  const int cur = ring(J.curr0, step, J.q);
  if (stage == J.nstage - 1 && local == 0 && threadIdx.x == 0) {
    clock[j].t = clock[j].t + J.dt;
    clock[j].loops = clock[j].loops + 1;
  }
  stepC_body<T, LOGN, cshape_wide(SHAPE), cshape_threads(SHAPE)>(J, J.st[stage], local, cur, par,
                                                                 tw, lds);
}

}  // namespace mlv
