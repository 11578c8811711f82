// Device bodies of the fused step's x-inverse kernels: A (spectral
// prologue + inverse x pass, xinv_threads<T, SHAPE>() threads per block) and M
// (single-precision x-mean profiles, kNT threads per block).
#pragma once
#include "fused.hpp"
#include "diag_dev.hpp"
#include "grid.hpp"

namespace mlv {
// --------------------------------------------------------------------------
// Kernel A: spectral prologue + inverse x pass
// --------------------------------------------------------------------------
template <typename T, int LOGN, int SHAPE>
__device__ __forceinline__ void stepA_body(const JobDev<T>& J, const StageDev& S, int local,
                                           int par, const cx<T>* __restrict__ tw, cx<T>* lds) {
  constexpr int N = 1 << LOGN;
  constexpr int P = fused_points<T>(LOGN, 0);
  constexpr int NT = xinv_threads<T, SHAPE>();
  constexpr int TC = xinv_cols<T, LOGN, P, SHAPE>();
  constexpr int LDSN = fft_ldsn(N);
  const int nn = J.nn, nm = J.nm, rows = J.rows;
  // ops outermost: consecutive blocks write neighbouring tiles of one
  // channel (measured at 64 x 1024^2: 293 us per launch against 329 us
  // with a tile's channel ops side by side, although those share its reads)
  const int ntile = (nm + TC - 1) / TC;
  const int opi = local / ntile, tile = local % ntile;
  const XinvOp op = S.a[opi];
  const int c0 = tile * TC;
  const int ncol = min(TC, nm - c0);
  const cx<T>* __restrict__ src = glb(J.state[op.from_new ? (par ^ 1) : par][op.src]);
  cx<T>* __restrict__ dst = glb(J.work) + (size_t)op.ch * work_chs(N, nm);
  cx<T>* __restrict__ psi_out = (op.op == XOP_UX && op.src == 0 && !op.from_new) ? glb(J.psi) : nullptr;
  const double fx = J.fx, fz = J.fz, fx2 = J.fx2, fz2 = J.fz2;
  const float ffx = (float)fx, ffz = (float)fz, ffx2 = (float)fx2, ffz2 = (float)fz2;
  // Each thread owns ITER = P elements; loads are issued U at a time so
  // several are in flight before the first is used.  When a column's
  // transform is private to one wavefront (WL), the thread loads the
  // elements its own slot's threads transform: column tid / TPS, rows
  // tid % TPS + i TPS (a column of the kz-major state is contiguous, so the
  // wavefront reads consecutive addresses) and no workgroup barrier is
  // needed before the transform.  Otherwise rows are fastest over the
  // whole block.
  constexpr int ITER = (N * TC) / NT;
  // every load of the thread in flight at once where the registers allow
  // (single precision, up to 16 points: one memory round trip)
  constexpr int U = (sizeof(T) == 4 && ITER <= 16) ? ITER : 8;
  static_assert(ITER % U == 0, "load batching");
  constexpr int TPS = fft_tps<N, P>();
  constexpr bool WL = fft_wave_local<N, P>() && TC * TPS == NT;
  auto cx_of = [&](int i, int& c, int& x) {
    if constexpr (WL) {
      c = threadIdx.x / TPS;
      x = threadIdx.x % TPS + i * TPS;
    } else {
      const int e = threadIdx.x + i * NT;
      c = e >> LOGN;
      x = e & (N - 1);
    }
  };
#pragma unroll 1
  for (int i0 = 0; i0 < ITER; i0 += U) {
    cx<T> fv[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      int c, x;
      cx_of(i0 + u, c, x);
      const int r = x <= nn ? x : (x >= N - nn ? x - N + rows : -1);
      fv[u] = (c < ncol && r >= 0) ? src[(size_t)(c0 + c) * rows + r] : czero<T>();
    }
#pragma unroll
    for (int u = 0; u < U; ++u) {
    int c, x;
    cx_of(i0 + u, c, x);
    cx<T> v = czero<T>();
    if (c < ncol) {
      const int r = x <= nn ? x : (x >= N - nn ? x - N + rows : -1);
      if (r >= 0) {
        const int m = c0 + c;
        const size_t idx = (size_t)m * rows + r;
        const cx<T> f = fv[u];
        if (op.op == XOP_ID) {
          v = f;
        } else {
          const int n = r <= nn ? r : r - rows;
          T rl, yz, yx;
          if constexpr (sizeof(T) == 8) {
            double lap = fx2 * (double)(n * n) + fz2 * (double)(m * m);
            if (r == 0 && m == 0) lap = 1.0;
            rl = 1.0 / lap;
            yz = fz * m;
            yx = fx * n;
          } else {
            // single precision: the symbol and its reciprocal in FP32
            // (hardware reciprocal, 1 ulp) instead of an FP64 division
            float lap = ffx2 * (float)(n * n) + ffz2 * (float)(m * m);
            if (r == 0 && m == 0) lap = 1.0f;
            rl = __builtin_amdgcn_rcpf(lap);
            yz = ffz * (float)m;
            yx = ffx * (float)n;
          }
          // psi = solve(-f) (utility.py:65, LaplacianSolver.py:64-66)
          const cx<T> ps = mk<T>(-f.x * rl, -f.y * rl);
          if (psi_out) psi_out[idx] = ps;
          if (op.op == XOP_UX) {  // ux = -d(psi)/dz
            v = mk<T>(yz * ps.y, -(yz * ps.x));
          } else {  // uz = d(psi)/dx
            v = mk<T>(-(yx * ps.y), yx * ps.x);
          }
        }
        // single precision: the channel's x mean (its kx = 0 row) is carried
        // by the mean profiles (kernel M), so the transforms see only the
        // fluctuation and round relative to it
        if (sizeof(T) == 4 && J.split && r == 0) v = czero<T>();
      }
    }
    lds[c * LDSN + fft_pad(x)] = v;
    }
  }
  if constexpr (WL)
    wave_sync();
  else
    __syncthreads();
  fft_lds<T, LOGN, +1, NT, P>(lds, ncol, tw);
  if constexpr (sizeof(T) == 4 && kXil == 2) {
    // a lane stores both rows of a row pair of one column, side by side in
    // the interleaved layout, as one 16-byte store; consecutive lanes take
    // consecutive columns (the channel base and the row pairs are 16-byte
    // aligned: work_chs and x nm are even)
    static_assert(ITER % 2 == 0, "row pairs");
#pragma unroll 8
    for (int i = 0; i < ITER / 2; ++i) {
      const int g = threadIdx.x + i * NT;
      const int c = g % TC, x = (g / TC) * 2;
      if (c < ncol) {
        const cx<T> a = lds[c * LDSN + fft_pad(x)], b = lds[c * LDSN + fft_pad(x + 1)];
        *reinterpret_cast<float4*>(dst + xrow_base(x, nm) + (size_t)(c0 + c) * 2) =
            make_float4(a.x, a.y, b.x, b.y);
      }
    }
  } else {
    // x % kXil fastest, then the column: consecutive lanes store one row
    // group's TC columns, contiguous in the interleaved layout
#pragma unroll 8
    for (int i = 0; i < ITER; ++i) {
      const int e = threadIdx.x + i * NT;
      const int g = e / kXil;
      const int c = g % TC, x = (g / TC) * kXil + e % kXil;
      if (c < ncol) dst[xrow_base(x, nm) + (size_t)(c0 + c) * kXil] = lds[c * LDSN + fft_pad(x)];
    }
  }
}

// Kernel A with mixed tiles (block shape 2): a block holds kMixCols
// columns of every op of the stage -- slot s is op s % nA of column
// c0 + s / nA -- so the ops derived from one field (its identity, ux, uz)
// read its columns once between them (from the cache after the first),
// where one op per block reads each column once per op.  Each op's rows
// are then stored in runs of kMixCols columns, and the neighbouring blocks
// of the same XCD complete the lines.  Single precision at 2^9 and 2^10
// points, 512 threads, when 2 nA fits the block's slots (xmix_ok).
template <typename T, int LOGN>
__host__ __device__ constexpr bool xmix_ok() {
  return sizeof(T) == 4 && (LOGN == 9 || LOGN == 10) &&
         fft_wave_local<(1 << LOGN), fused_points<T>(LOGN, 0)>() &&
         xinv_cols<T, LOGN, fused_points<T>(LOGN, 0), 0>() *
                 fft_tps<(1 << LOGN), fused_points<T>(LOGN, 0)>() ==
             xinv_threads<T, 0>();
}
template <typename T, int LOGN>
__device__ __forceinline__ void stepA_mixed(const JobDev<T>& J, const StageDev& S, int local,
                                            int par, const cx<T>* __restrict__ tw, cx<T>* lds) {
  constexpr int N = 1 << LOGN;
  constexpr int P = fused_points<T>(LOGN, 0);
  constexpr int NT = xinv_threads<T, 0>();
  constexpr int TC = xinv_cols<T, LOGN, P, 0>();
  constexpr int LDSN = fft_ldsn(N);
  constexpr int TPS = fft_tps<N, P>();
  constexpr int ITER = N / TPS;
  constexpr int U = 8;
  static_assert(xmix_ok<T, LOGN>() && ITER % U == 0, "mixed x-inverse tile shape");
  const int nn = J.nn, nm = J.nm, rows = J.rows;
  const int nA = S.nA;
  const int c0 = local * kMixCols;
  const int ncol = min(kMixCols, nm - c0);
  const int nslot = ncol * nA;
  // this thread's slot: op sl % nA of column c0 + sl / nA
  const int sl = threadIdx.x / TPS;
  const int xl = threadIdx.x % TPS;
  const bool live = sl < nslot;
  const XinvOp op = S.a[live ? sl % nA : 0];
  const int m = c0 + sl / nA;
  const cx<T>* __restrict__ src = glb(J.state[op.from_new ? (par ^ 1) : par][op.src]);
  cx<T>* __restrict__ psi_out = (op.op == XOP_UX && op.src == 0 && !op.from_new) ? glb(J.psi) : nullptr;
  const float ffx = (float)J.fx, ffz = (float)J.fz, ffx2 = (float)J.fx2, ffz2 = (float)J.fz2;
#pragma unroll 1
  for (int i0 = 0; i0 < ITER; i0 += U) {
    cx<T> fv[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int x = xl + (i0 + u) * TPS;
      const int r = x <= nn ? x : (x >= N - nn ? x - N + rows : -1);
      fv[u] = (live && r >= 0) ? src[(size_t)m * rows + r] : czero<T>();
    }
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int x = xl + (i0 + u) * TPS;
      cx<T> v = czero<T>();
      const int r = x <= nn ? x : (x >= N - nn ? x - N + rows : -1);
      if (live && r >= 0) {
        const cx<T> f = fv[u];
        if (op.op == XOP_ID) {
          v = f;
        } else {
          const int n = r <= nn ? r : r - rows;
          // the symbol and its reciprocal in FP32, as stepA_body
          float lap = ffx2 * (float)(n * n) + ffz2 * (float)(m * m);
          if (r == 0 && m == 0) lap = 1.0f;
          const T rl = __builtin_amdgcn_rcpf(lap);
          const cx<T> ps = mk<T>(-f.x * rl, -f.y * rl);
          if (psi_out) psi_out[(size_t)m * rows + r] = ps;
          if (op.op == XOP_UX) {
            const T yz = ffz * (float)m;
            v = mk<T>(yz * ps.y, -(yz * ps.x));
          } else {
            const T yx = ffx * (float)n;
            v = mk<T>(-(yx * ps.y), yx * ps.x);
          }
        }
        if (J.split && r == 0) v = czero<T>();
      }
      lds[sl * LDSN + fft_pad(x)] = v;
    }
  }
  wave_sync();
  fft_lds<T, LOGN, +1, NT, P>(lds, nslot, tw);
  // per op: the row pairs' kMixCols columns, x % kXil fastest, then the
  // column, so consecutive lanes store one contiguous run
  for (int o = 0; o < nA; ++o) {
    cx<T>* __restrict__ dst = glb(J.work) + (size_t)S.a[o].ch * work_chs(N, nm);
#pragma unroll 4
    for (int e = threadIdx.x; e < kMixCols * N; e += NT) {
      const int x2 = e % kXil, cc = (e / kXil) % kMixCols, x = (e / (kXil * kMixCols)) * kXil + x2;
      if (cc < ncol)
        dst[xrow_base(x, nm) + (size_t)(c0 + cc) * kXil] = lds[(cc * nA + o) * LDSN + fft_pad(x)];
    }
  }
}

// Kernel A's work for logical block lb of a launch over the job table's
// x-inverse blocks (prefix array pref), in block shape SHAPE.  The host
// selects a shape only when every job's transform fits its block
// (xinv_cols_host >= 1).
template <typename T, int SHAPE>
__device__ __forceinline__ void roleA(const JobDev<T>* __restrict__ jobs, const int* __restrict__ pref,
                                      int njobs, int lb, int stage, int step,
                                      const cx<T>* __restrict__ tw, cx<T>* lds) {
  const int j = find_job(pref, njobs, lb);
  const JobDev<T>& J = jobs[j];
  const int local = lb - pref[j];
  const int par = (J.par0 + step) & 1;
  switch (J.lognx) {
#define X(L)                                                                       \
  case L:                                                                          \
    if constexpr (xinv_cols<T, L, fused_points<T>(L, 0), SHAPE>() >= 1)            \
      stepA_body<T, L, SHAPE>(J, J.st[stage], local, par, tw, lds);                \
    break;
    X(3) X(4) X(5) X(6) X(7) X(8) X(9) X(10) X(11) X(12)
#undef X
    default:
      break;
  }
}

// Kernel A's work for a batch whose jobs all have 2^LOGN points along x.
template <typename T, int SHAPE, int LOGN>
__device__ __forceinline__ void roleA_fixed(const JobDev<T>* __restrict__ jobs,
                                            const int* __restrict__ pref, int njobs, int lb,
                                            int stage, int step, const cx<T>* __restrict__ tw,
                                            cx<T>* lds) {
  const int j = find_job(pref, njobs, lb);
  const JobDev<T>& J = jobs[j];
  const int local = lb - pref[j];
  const int par = (J.par0 + step) & 1;
  if constexpr (SHAPE == 2) {
    if constexpr (xmix_ok<T, LOGN>()) stepA_mixed<T, LOGN>(J, J.st[stage], local, par, tw, lds);
  } else if constexpr (xinv_cols<T, LOGN, fused_points<T>(LOGN, 0), SHAPE>() >= 1) {
    stepA_body<T, LOGN, SHAPE>(J, J.st[stage], local, par, tw, lds);
  }
}

// --------------------------------------------------------------------------
// Kernel M (single precision): the x mean of each channel an x-inverse op
// of the stage writes, as a function of z -- the z c2r of the channel's
// kx = 0 coefficients -- into J.mean[op.ch * nz + n].  One block per job.
// Kernel A leaves the kx = 0 row out of the channels, so kernel B's
// products are formed from mean + fluctuation with the mean-times-mean part
// left out: its x difference is zero, and in z it vanishes because uz and bz
// have no x mean.  Without the split, an O(1) x-independent profile (a shear
// layer, a current sheet) puts FP32 rounding of its own size into products
// whose x differences should cancel exactly.
// --------------------------------------------------------------------------
template <typename T, int LOGN>
__device__ __forceinline__ void stepM_body(const JobDev<T>& J, const StageDev& S, int pair,
                                           int par, const cx<T>* __restrict__ tw, cx<T>* lds) {
  constexpr int N = 1 << LOGN;
  constexpr int P = 16;  // one slot per block: 256 threads x 16 points >= N
  const int o0 = 2 * pair;
  if (o0 >= S.nA) return;
  const bool hb = o0 + 1 < S.nA;
  const XinvOp& oa = S.a[o0];
  const XinvOp& ob = S.a[hb ? o0 + 1 : o0];
  const int nm = J.nm, rows = J.rows;
  const float ffz = (float)J.fz, ffz2 = (float)J.fz2;
  // the kx = 0 coefficient of channel op at kz = m (the values kernel A
  // would put in row 0)
  auto coef = [&](const XinvOp& op, int m) -> cx<T> {
    const cx<T> f = glb(J.state[op.from_new ? (par ^ 1) : par][op.src])[(size_t)m * rows];
    if (op.op == XOP_ID) return f;
    if (op.op == XOP_UZ) return czero<T>();  // i kx psi with kx = 0
    float lap = ffz2 * (float)(m * m);
    if (m == 0) lap = 1.0f;
    const T rl = __builtin_amdgcn_rcpf(lap);
    const cx<T> ps = mk<T>(-f.x * rl, -f.y * rl);
    const T yz = ffz * (float)m;
    return mk<T>(yz * ps.y, -(yz * ps.x));
  };
  cx<T>* z = lds;
  for (int k = threadIdx.x; k < nm; k += kNT) {
    const cx<T> a = coef(oa, k);
    const cx<T> b = hb ? coef(ob, k) : czero<T>();
    if (k == 0) {
      z[0] = mk<T>(a.x, b.x);
    } else {
      z[fft_pad(k)] = mk<T>(a.x - b.y, a.y + b.x);
      z[fft_pad(N - k)] = mk<T>(a.x + b.y, b.x - a.y);
    }
  }
  for (int k = nm + threadIdx.x; k <= N - nm; k += kNT) z[fft_pad(k)] = czero<T>();
  __syncthreads();
  fft_lds<T, LOGN, +1, kNT, P>(lds, 1, tw);
  T* __restrict__ ma = glb(J.mean) + (size_t)oa.ch * N;
  T* __restrict__ mb = glb(J.mean) + (size_t)ob.ch * N;
  for (int n = threadIdx.x; n < N; n += kNT) {
    const cx<T> zv = z[fft_pad(n)];
    ma[n] = zv.x;
    if (hb) mb[n] = zv.y;
  }
}

// Kernel M's work for op pair `pair` of job J.
template <typename T>
__device__ __forceinline__ void roleM(const JobDev<T>& J, int pair, int stage, int step,
                                      const cx<T>* __restrict__ tw, cx<T>* lds) {
  if (stage >= J.nstage || !J.split) return;
  const int par = (J.par0 + step) & 1;
  switch (J.lognz) {
#define X(L)                                                       \
  case L:                                                          \
    stepM_body<T, L>(J, J.st[stage], pair, par, tw, lds);          \
    break;
    X(3) X(4) X(5) X(6) X(7) X(8) X(9) X(10) X(11) X(12)
#undef X
    default:
      break;
  }
}

}  // namespace mlv
