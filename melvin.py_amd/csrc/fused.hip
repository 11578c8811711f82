// Fused fully-spectral step kernels over a job table (see fused.hpp).
#include "fused.hpp"
#include "diag_dev.hpp"
#include "grid.hpp"

namespace mlv {

// --------------------------------------------------------------------------
// Kernel A: spectral prologue + inverse x pass
// --------------------------------------------------------------------------
template <typename T, int LOGN>
__device__ __forceinline__ void stepA_body(const JobDev<T>& J, const StageDev& S, int local,
                                           int par, const cx<T>* __restrict__ tw, cx<T>* lds) {
  constexpr int N = 1 << LOGN;
  constexpr int P = fused_points<T>(LOGN, 0);
  constexpr int TC = fused_cols<LOGN, P>();
  constexpr int LDSN = fft_ldsn(N);
  const int nn = J.nn, nm = J.nm, rows = J.rows;
  const int ntile = (nm + TC - 1) / TC;
  const int opi = local / ntile, tile = local % ntile;
  const XinvOp op = S.a[opi];
  const int c0 = tile * TC;
  const int ncol = min(TC, nm - c0);
  const cx<T>* __restrict__ src = J.state[op.from_new ? (par ^ 1) : par][op.src];
  cx<T>* __restrict__ dst = J.work + (size_t)op.ch * N * nm;
  cx<T>* __restrict__ psi_out = (op.op == XOP_UX && op.src == 0 && !op.from_new) ? J.psi : nullptr;
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
  constexpr int ITER = (N * TC) / kNT;
  constexpr int U = 8;
  static_assert(ITER % U == 0, "load batching");
  constexpr int TPS = fft_tps<N, P>();
  constexpr bool WL = fft_wave_local<N, P>() && TC * TPS == kNT;
  auto cx_of = [&](int i, int& c, int& x) {
    if constexpr (WL) {
      c = threadIdx.x / TPS;
      x = threadIdx.x % TPS + i * TPS;
    } else {
      const int e = threadIdx.x + i * kNT;
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
  fft_lds<T, LOGN, +1, kNT, P>(lds, ncol, tw);
  // x % kXil fastest, then the column: consecutive lanes store one row
  // group's TC columns, contiguous in the interleaved layout
#pragma unroll 8
  for (int i = 0; i < ITER; ++i) {
    const int e = threadIdx.x + i * kNT;
    const int g = e / kXil;
    const int c = g % TC, x = (g / TC) * kXil + e % kXil;
    if (c < ncol) dst[xrow_base(x, nm) + (size_t)(c0 + c) * kXil] = lds[c * LDSN + fft_pad(x)];
  }
}

template <typename T>
__global__ void __launch_bounds__(kNT) k_stepA(const JobDev<T>* __restrict__ jobs,
                                               const int* __restrict__ pref, int njobs, int stage,
                                               int step, const cx<T>* __restrict__ tw) {
  extern __shared__ char smem_raw[];
  cx<T>* lds = reinterpret_cast<cx<T>*>(smem_raw);
  const int lb = xcd_logical(pref[njobs]);
  if (lb < 0) return;
  const int j = find_job(pref, njobs, lb);
  const JobDev<T>& J = jobs[j];
  const int local = lb - pref[j];
  const int par = (J.par0 + step) & 1;
  switch (J.lognx) {
#define X(L)                                                       \
  case L:                                                          \
    stepA_body<T, L>(J, J.st[stage], local, par, tw, lds);         \
    break;
    X(3) X(4) X(5) X(6) X(7) X(8) X(9) X(10) X(11) X(12)
#undef X
    default:
      break;
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
    const cx<T> f = J.state[op.from_new ? (par ^ 1) : par][op.src][(size_t)m * rows];
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
  T* __restrict__ ma = J.mean + (size_t)oa.ch * N;
  T* __restrict__ mb = J.mean + (size_t)ob.ch * N;
  for (int n = threadIdx.x; n < N; n += kNT) {
    const cx<T> zv = z[fft_pad(n)];
    ma[n] = zv.x;
    if (hb) mb[n] = zv.y;
  }
}

// Grid: jobs x (kMeanCh / 2) op pairs.
template <typename T>
__global__ void __launch_bounds__(kNT) k_stepM(const JobDev<T>* __restrict__ jobs, int stage,
                                               int step, const cx<T>* __restrict__ tw) {
  extern __shared__ char smem_raw[];
  cx<T>* lds = reinterpret_cast<cx<T>*>(smem_raw);
  const JobDev<T>& J = jobs[blockIdx.x / (kMeanCh / 2)];
  const int pair = blockIdx.x % (kMeanCh / 2);
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

// --------------------------------------------------------------------------
// Kernel B: z pass, c2r of the channels, products, r2c of the products
// --------------------------------------------------------------------------
// SPLIT (single precision): jobs with split carry mean profiles (kernel M)
// and mean-free channels; a batch without such jobs runs SPLIT = 0.
// Shapes of a z-pass block of program V at 2^LOGN points with P points per
// thread: RB rows, each of NIN input channels (packed two real sequences
// per LDS slot) and NEQ product slots.
template <typename T, int LOGN, int V, int P>
struct BShape {
  using Pg = BProg<V>;
  static constexpr int NIN = Pg::NIN, NEQ = Pg::NEQ;
  static constexpr int N = 1 << LOGN;
  static constexpr int SL = fft_slots<LOGN, kNT, P>();
  static constexpr int LDSN = fft_ldsn(N);
  static constexpr int RB = brows(NIN, NEQ, SL);
  static constexpr int NW = kNT / 64;
  static constexpr int KIT = ((N - 1) / 3 + 63) / 64;  // bound on ceil(nm / 64)
  static constexpr int SLN = (RB * NIN + 1) / 2;       // input slots
  static constexpr int WPS = SLN >= NW ? 1 : NW / SLN;  // wavefronts per input slot
  static constexpr int NO = RB * NEQ;                  // product slots
  static constexpr int WPO = NO >= NW ? 1 : NW / NO;   // wavefronts per product slot
};

// Load phase: the block's input channel rows into the LDS slots, two real
// sequences per slot packed as one Hermitian spectrum (zero-padded from nm
// to N/2).  One wavefront per slot (WPS per slot when the block has fewer
// slots than wavefronts, splitting its k range): lane l handles k = l + 64 i,
// so no index needs a division by the runtime row length nm.
template <typename T, typename Sh>
__device__ __forceinline__ void bload(const cx<T>* __restrict__ work, size_t chs, int nm, int x0,
                                      int nseq, int nslot, cx<T>* lds) {
  constexpr int N = Sh::N, NIN = Sh::NIN, KIT = Sh::KIT, WPS = Sh::WPS, NW = Sh::NW;
  constexpr int KB = KIT < 6 ? KIT : 6;
  const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int ssub = wv % WPS;
  for (int s = wv / WPS; s < nslot; s += NW / WPS) {
    cx<T>* z = lds + s * Sh::LDSN;
    const int qa = 2 * s, qb = qa + 1;
    const cx<T>* __restrict__ pa = work + (size_t)(qa % NIN) * chs + xrow_base(x0 + qa / NIN, nm);
    const cx<T>* __restrict__ pb =
        qb < nseq ? work + (size_t)(qb % NIN) * chs + xrow_base(x0 + qb / NIN, nm) : nullptr;
    // loads of up to KB iterations in flight before any is used
#pragma unroll 1
    for (int i0 = ssub * KB; i0 < KIT; i0 += WPS * KB) {
      cx<T> av[KB], bv[KB];
#pragma unroll
      for (int i = 0; i < KB; ++i) {
        const int k = lane + 64 * (i0 + i);
        av[i] = k < nm ? pa[(size_t)k * kXil] : czero<T>();
        bv[i] = (pb && k < nm) ? pb[(size_t)k * kXil] : czero<T>();
      }
#pragma unroll
      for (int i = 0; i < KB; ++i) {
        const int k = lane + 64 * (i0 + i);
        if (k < nm) {
          const cx<T> a = av[i], b = bv[i];
          if (k == 0) {
            z[0] = mk<T>(a.x, b.x);
          } else {
            z[fft_pad(k)] = mk<T>(a.x - b.y, a.y + b.x);
            z[fft_pad(N - k)] = mk<T>(a.x + b.y, b.x - a.y);
          }
        }
      }
    }
    for (int k = nm + lane + 64 * ssub; k <= N - nm; k += 64 * WPS) z[fft_pad(k)] = czero<T>();
  }
}

// Running diagnostics partials of a z-pass thread.
template <typename T>
struct BDiag {
  T mux = -INFINITY, muz = -INFINITY;  // NaN-propagating maxima, as np.max
  double ke = 0.0, ens = 0.0, flux = 0.0;
};

// The physical work at one point n of the block's rows, v[q] the value of
// sequence q = row * NIN + channel (a fluctuation when kSplit): the
// diagnostics partials, the physical outputs, and each row's products into
// px / pz[row * NEQ + e].
template <typename T, typename Sh>
__device__ __forceinline__ void bpoint(const JobDev<T>& J, const StageDev& S, bool kSplit,
                                       bool diag, int flags, int x0, int nrows, int nseq, int n,
                                       const T* v, BDiag<T>& d, T* px, T* pz) {
  using Pg = typename Sh::Pg;
  constexpr int NIN = Sh::NIN, NEQ = Sh::NEQ, RB = Sh::RB, N = Sh::N;
  T mv[NIN];
#pragma unroll
  for (int q = 0; q < NIN; ++q)
    mv[q] = (kSplit && !Pg::zero_mean(q)) ? J.mean[(size_t)q * N + n] : T(0);
  // the whole value of sequence q (row q / NIN, channel q % NIN)
  auto whole = [&](int q) { return kSplit ? mv[q % NIN] + v[q] : v[q]; };
  if (diag) {
#pragma unroll
    for (int r = 0; r < RB; ++r) {
      if (r < nrows) {
        const T ux = whole(r * NIN + Pg::UX), uz = whole(r * NIN + Pg::UZ);
        d.mux = nan_max(d.mux, ux);
        d.muz = nan_max(d.muz, uz);
        d.ke += (double)uz * (double)uz + (double)ux * (double)ux;
        const double w = (double)whole(r * NIN);
        d.ens += w * w;
        if constexpr (Pg::FLUX >= 0) d.flux += (double)whole(r * NIN + Pg::FLUX) * (double)uz;
      }
    }
  }
  if (flags & kWritePhys) {
#pragma unroll
    for (int q = 0; q < RB * NIN; ++q) {
      if (q < nseq) {
#pragma unroll
        for (int i = 0; i < kMaxPhys; ++i)
          if (S.phys_ch[i] == q % NIN) J.phys[i][(size_t)(x0 + q / NIN) * N + n] = whole(q);
      }
    }
  }
#pragma unroll
  for (int r = 0; r < RB; ++r) {
    if (kSplit)
      Pg::template products_split<T>(mv, &v[r * NIN], &px[r * NEQ], &pz[r * NEQ]);
    else
      Pg::template products<T>(&v[r * NIN], &px[r * NEQ], &pz[r * NEQ]);
  }
}

// Output phase: the kz < nm coefficients of each product row, split from
// the packed (Px + i Pz) spectrum, into the Px / Pz channels (row-major).
// One wavefront per product slot (WPO per slot when there are fewer).
template <typename T, typename Sh>
__device__ __forceinline__ void bstore(cx<T>* __restrict__ work, size_t chs, int nm, int x0,
                                       int nout, const cx<T>* lds) {
  constexpr int N = Sh::N, NIN = Sh::NIN, NEQ = Sh::NEQ, KIT = Sh::KIT, WPO = Sh::WPO;
  constexpr int NW = Sh::NW;
  const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int osub = wv % WPO;
  for (int o = wv / WPO; o < nout; o += NW / WPO) {
    const int r = o / NEQ, eq = o % NEQ;
    const cx<T>* z = lds + o * Sh::LDSN;
    cx<T>* __restrict__ ox = work + (size_t)pchan(NIN, eq, 0) * chs + (size_t)(x0 + r) * nm;
    cx<T>* __restrict__ oz = work + (size_t)pchan(NIN, eq, 1) * chs + (size_t)(x0 + r) * nm;
#pragma unroll 6
    for (int i = osub; i < KIT; i += WPO) {
      const int k = lane + 64 * i;
      if (k < nm) {
        const cx<T> y = z[fft_pad(k)];
        const cx<T> ym = z[fft_pad((N - k) & (N - 1))];
        ox[k] = mk<T>(T(0.5) * (y.x + ym.x), T(0.5) * (y.y - ym.y));
        oz[k] = mk<T>(T(0.5) * (y.y + ym.y), T(0.5) * (ym.x - y.x));
      }
    }
  }
}

// Generic z pass: load, inverse transforms, products point by point through
// the LDS, forward transforms, store.
template <typename T, int LOGN, int V, int SPLIT>
__device__ __forceinline__ void stepB_body(const JobDev<T>& J, const StageDev& S, int local,
                                           int flags, const cx<T>* __restrict__ tw, cx<T>* lds) {
  using Pg = BProg<V>;
  constexpr int P = bpoints<T>(LOGN, Pg::NIN, Pg::NEQ);
  using Sh = BShape<T, LOGN, V, P>;
  constexpr int NIN = Sh::NIN, NEQ = Sh::NEQ, N = Sh::N, LDSN = Sh::LDSN, RB = Sh::RB;
  if constexpr (!bfits(NIN, NEQ, Sh::SL)) {
    return;  // rejected on the host
  } else {
    const int nx = J.nx, nm = J.nm;
    const size_t chs = (size_t)nx * nm;
    const int x0 = local * RB;
    const int nrows = min(RB, nx - x0);
    const int nseq = nrows * NIN;
    const int nslot = (nseq + 1) / 2;
    // slot s is loaded by wavefront s and transformed by the same
    // wavefront: no workgroup barrier before the inverse transform
    constexpr bool WL_IN = fft_tps<N, P>() == 64 && Sh::WPS == 1 && Sh::SL == Sh::NW;
    bload<T, Sh>(J.work, chs, nm, x0, nseq, nslot, lds);
    if constexpr (WL_IN)
      wave_sync();
    else
      __syncthreads();
    fft_lds<T, LOGN, +1, kNT, P>(lds, nslot, tw);
    // this job's check loop (J.loops counts the loops before this one)
    const bool diag = (flags & kDiag) && J.loops % (flags >> kDiagShift) == 0;
    BDiag<T> d;
    // single precision: the channels are fluctuations about their x means,
    // whose z profiles kernel M wrote
    const bool kSplit = SPLIT && J.split;
    for (int n = threadIdx.x; n < N; n += kNT) {
      T v[RB * NIN];
#pragma unroll
      for (int q = 0; q < RB * NIN; ++q) {
        if (q < nseq) {
          const cx<T> z = lds[(q >> 1) * LDSN + fft_pad(n)];
          v[q] = (q & 1) ? z.y : z.x;
        } else {
          v[q] = T(0);
        }
      }
      T px[RB * NEQ], pz[RB * NEQ];
      bpoint<T, Sh>(J, S, kSplit, diag, flags, x0, nrows, nseq, n, v, d, px, pz);
#pragma unroll
      for (int o = 0; o < RB * NEQ; ++o)
        if (o < nrows * NEQ) lds[o * LDSN + fft_pad(n)] = mk<T>(px[o], pz[o]);
    }
    if (diag) block_diag<kNT>((double)d.mux, (double)d.muz, d.ke, d.ens, d.flux,
                              J.diag + (size_t)kDiagPart * local);
    __syncthreads();
    // when wavefront o transformed product slot o itself the outputs need no
    // workgroup barrier after the transform
    constexpr bool WL_OUT = fft_tps<N, P>() == 64 && Sh::NO == Sh::NW;
    fft_lds<T, LOGN, -1, kNT, P, !WL_OUT>(lds, nrows * NEQ, tw);
    bstore<T, Sh>(J.work, chs, nm, x0, nrows * NEQ, lds);
  }
}

// z pass with the physical products between two fused passes, for
// N = 512, 1024, 2048 at 16 points per thread (256 threads, 4096 / N slots).
// The inverse transforms run the standard plan 16, 16, R (R = N / 256); its
// last pass (Ns = 256) gives thread t the butterfly t of every slot, whose
// outputs are the points t + 256 r, r < R.  The forward transforms run the
// remainder-first plan R, 16, 16, whose first pass (Ns = 1) takes exactly
// those points.  So thread t reads the last inverse pass's inputs, finishes
// the inverse, forms every row's products at its R points, and does the
// first forward pass in registers: the products never make an LDS round
// trip of their own, and the twiddles of the last inverse pass, which
// depend only on t, are loaded once for all slots.
template <typename T, int LOGN, int V, int SPLIT>
__device__ __forceinline__ void stepB_split(const JobDev<T>& J, const StageDev& S, int local,
                                            int flags, const cx<T>* __restrict__ tw, cx<T>* lds) {
  constexpr int P = 16;
  using Sh = BShape<T, LOGN, V, P>;
  constexpr int NIN = Sh::NIN, NEQ = Sh::NEQ, N = Sh::N, LDSN = Sh::LDSN, RB = Sh::RB;
  constexpr int LOGR = LOGN - 8, R = 1 << LOGR, NSL = 256;
  static_assert(LOGR >= 1 && LOGR <= 3 && kNT == NSL, "split pass shape");
  static_assert(fft_pass_logr(LOGN, 8) == LOGR, "standard plan ends with radix N/256");
  const int nx = J.nx, nm = J.nm;
  const size_t chs = (size_t)nx * nm;
  const int x0 = local * RB;
  const int nrows = min(RB, nx - x0);
  const int nseq = nrows * NIN;
  const int nslot = (nseq + 1) / 2;
  constexpr bool WL_IN = fft_tps<N, P>() == 64 && Sh::WPS == 1 && Sh::SL == Sh::NW;
  bload<T, Sh>(J.work, chs, nm, x0, nseq, nslot, lds);
  if constexpr (WL_IN)
    wave_sync();
  else
    __syncthreads();
  // inverse: the two radix-16 passes
  fft_pass<T, LOGN, 4, 0, +1, kNT, P>(lds, nslot, tw);
  fft_pass<T, LOGN, 4, 4, +1, kNT, P>(lds, nslot, tw);
  __syncthreads();
  // the last inverse pass: butterfly j of every slot
  const int j = threadIdx.x;
  cx<T> w[R];
  {
    const cx<T>* t = tw + fft_pass_table(LOGN, 8) + j;
#pragma unroll
    for (int r = 1; r < R; ++r) {
      w[r] = t[(r - 1) * NSL];
      w[r].y = -w[r].y;  // inverse
    }
  }
  T val[R][2 * Sh::SLN];  // val[r][q]: sequence q at the point j + 256 r
#pragma unroll
  for (int s = 0; s < Sh::SLN; ++s) {
    cx<T> v[R];
    if (s < nslot) {
      const cx<T>* p = lds + s * LDSN + fft_pad(j);
#pragma unroll
      for (int r = 0; r < R; ++r) v[r] = p[r * (NSL + NSL / 16)];
#pragma unroll
      for (int r = 1; r < R; ++r) v[r] = cmul<T>(v[r], w[r]);
      Dft<R, +1, T>::run(v);
    } else {
#pragma unroll
      for (int r = 0; r < R; ++r) v[r] = czero<T>();
    }
#pragma unroll
    for (int r = 0; r < R; ++r) {
      val[r][2 * s] = v[r].x;
      val[r][2 * s + 1] = v[r].y;
    }
  }
  const bool diag = (flags & kDiag) && J.loops % (flags >> kDiagShift) == 0;
  BDiag<T> d;
  const bool kSplit = SPLIT && J.split;
  cx<T> pr[Sh::NO][R];  // product slot o = row * NEQ + e at the point j + 256 r
#pragma unroll
  for (int r = 0; r < R; ++r) {
    T v[RB * NIN];
#pragma unroll
    for (int q = 0; q < RB * NIN; ++q) v[q] = q < nseq ? val[r][q] : T(0);
    T px[Sh::NO], pz[Sh::NO];
    bpoint<T, Sh>(J, S, kSplit, diag, flags, x0, nrows, nseq, j + NSL * r, v, d, px, pz);
#pragma unroll
    for (int o = 0; o < Sh::NO; ++o) pr[o][r] = mk<T>(px[o], pz[o]);
  }
  if (diag) block_diag<kNT>((double)d.mux, (double)d.muz, d.ke, d.ens, d.flux,
                            J.diag + (size_t)kDiagPart * local);
  // every thread has read its inputs before any product slot is written
  // over them
  __syncthreads();
  // the first forward pass (radix R, Ns = 1): outputs at R j + r
#pragma unroll
  for (int o = 0; o < Sh::NO; ++o) {
    if (o < nrows * NEQ) {
      Dft<R, -1, T>::run(pr[o]);
      cx<T>* p = lds + o * LDSN;
#pragma unroll
      for (int r = 0; r < R; ++r) p[fft_pad(R * j + r)] = pr[o][r];
    }
  }
  __syncthreads();
  // the remaining forward passes: radix 16 at Ns = R and Ns = 16 R
  fft_pass<T, LOGN, 4, LOGR, -1, kNT, P, true>(lds, nrows * NEQ, tw);
  fft_pass<T, LOGN, 4, LOGR + 4, -1, kNT, P, true>(lds, nrows * NEQ, tw);
  constexpr bool WL_OUT = fft_tps<N, P>() == 64 && Sh::NO == Sh::NW;
  if constexpr (!WL_OUT) __syncthreads();
  bstore<T, Sh>(J.work, chs, nm, x0, nrows * NEQ, lds);
}

// The split-pass z pass applies at 2^9..2^11 points when the program runs
// at 16 points per thread and fits the slots.
template <typename T, int LOGN, int V>
__host__ __device__ constexpr bool bsplit_ok() {
  return LOGN >= 9 && LOGN <= 11 && bpoints<T>(LOGN, bvar_nin(V), bvar_neq(V)) == 16 &&
         bfits(bvar_nin(V), bvar_neq(V), fft_slots<LOGN, kNT, 16>());
}

// Single precision is held to 128 VGPRs: four wavefronts per SIMD, as many
// as the four blocks per CU its LDS allows at 1024 points.
template <typename T, int V, int SPLIT>
__global__ void __launch_bounds__(kNT)
    __attribute__((amdgpu_waves_per_eu(sizeof(T) == 4 ? 4 : 1))) k_stepB(const JobDev<T>* __restrict__ jobs,
                                               const int* __restrict__ pref, int njobs, int stage,
                                               int flags, const cx<T>* __restrict__ tw) {
  extern __shared__ char smem_raw[];
  cx<T>* lds = reinterpret_cast<cx<T>*>(smem_raw);
  const int lb = xcd_logical(pref[njobs]);
  if (lb < 0) return;
  const int j = find_job(pref, njobs, lb);
  const JobDev<T>& J = jobs[j];
  const int local = lb - pref[j];
  switch (J.lognz) {
#define X(L)                                                          \
  case L:                                                             \
    if constexpr (bsplit_ok<T, L, V>())                               \
      stepB_split<T, L, V, SPLIT>(J, J.st[stage], local, flags, tw, lds); \
    else                                                              \
      stepB_body<T, L, V, SPLIT>(J, J.st[stage], local, flags, tw, lds);  \
    break;
    X(3) X(4) X(5) X(6) X(7) X(8) X(9) X(10) X(11) X(12)
#undef X
    default:
      break;
  }
}

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

template <typename T, int LOGN, int WIDE>
__device__ __forceinline__ void stepC_body(const JobDev<T>& J, const StageDev& S, int local, int cur,
                                           int par, const cx<T>* __restrict__ tw, cx<T>* lds) {
  constexpr int N = 1 << LOGN;
  constexpr int P = fused_points<T>(LOGN, WIDE);
  constexpr int TC = fused_cols<LOGN, P>();
  constexpr int LDSN = fft_ldsn(N);
  const int nn = J.nn, nm = J.nm, rows = J.rows;
  const size_t chs = (size_t)N * nm;
  const size_t ssz = (size_t)rows * nm;
  const int ntile = (nm + TC - 1) / TC;
  const int eqi = local / ntile, tile = local % ntile;
  const int field = S.eq_field[eqi];
  const int c0 = tile * TC;
  const int ncol = min(TC, nm - c0);
  const cx<T>* __restrict__ pxc = J.work + (size_t)pchan(J.pbase, eqi, 0) * chs;
  const cx<T>* __restrict__ pzc = J.work + (size_t)pchan(J.pbase, eqi, 1) * chs;
  // One transform per column: the periodic x difference acts on the
  // intermediate rows (it commutes with the z transform) and the z
  // difference is its modified wavenumber, so
  //   D = F_x[ FDx(Px) + i sz(kz) Pz ].
  const int order = J.fd_order;
  const T ihx = order == 2 ? (T)(1.0 / (2.0 * J.dx)) : (T)(1.0 / (3.0 * J.dx));
  constexpr int ITER = (N * TC) / kNT;
  constexpr int U = (sizeof(T) == 4 && !WIDE) ? 4 : 8;
  static_assert(ITER % U == 0, "load batching");
  // sz depends only on the column; every thread keeps one column (e % TC
  // is fixed because kNT is a multiple of TC)
  const int cme = threadIdx.x % TC;
  const T szc = modk<T>(tw, c0 + cme, J.lognz, order, J.dz);
  const bool cok = cme < ncol;
  const size_t col = (size_t)(c0 + cme);
  cx<T>* __restrict__ slot = lds + cme * LDSN;
  // Single precision with at most 16 points per thread: Px is loaded once
  // into the slot and the stencil reads its neighbours from LDS (4 % faster
  // at 1024^2 than four extra global loads per point, which hit L2).
  if constexpr (sizeof(T) == 4 && (WIDE || ITER <= 16)) {
    // 1. Px column into the slot
#pragma unroll 1
    for (int i0 = 0; i0 < ITER; i0 += U) {
      cx<T> pv[U];
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const int x = (threadIdx.x + (i0 + u) * kNT) / TC;
        pv[u] = cok ? pxc[(size_t)x * nm + col] : czero<T>();
      }
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const int x = (threadIdx.x + (i0 + u) * kNT) / TC;
        slot[fft_pad(x)] = pv[u];
      }
    }
    __syncthreads();
    // 2. x stencil from LDS plus i sz Pz, held in registers, then 3. stored
    cx<T> dv[ITER];
#pragma unroll
    for (int i0 = 0; i0 < ITER; i0 += U) {
      cx<T> zv[U];
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const int x = (threadIdx.x + (i0 + u) * kNT) / TC;
        zv[u] = cok ? pzc[(size_t)x * nm + col] : czero<T>();
      }
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const int x = (threadIdx.x + (i0 + u) * kNT) / TC;
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
        dv[i0 + u] = cok ? mk<T>(d.x - szc * zv[u].y, d.y + szc * zv[u].x) : czero<T>();
      }
    }
    __syncthreads();
#pragma unroll
    for (int i = 0; i < ITER; ++i) {
      const int x = (threadIdx.x + i * kNT) / TC;
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
        const int x = (threadIdx.x + (i0 + u) * kNT) / TC;
        if (cok) {
          pa[u] = pxc[(size_t)((x + 1) & (N - 1)) * nm + col];
          pb[u] = pxc[(size_t)((x - 1) & (N - 1)) * nm + col];
          pz[u] = pzc[(size_t)x * nm + col];
          if (order != 2) {
            pa2[u] = pxc[(size_t)((x + 2) & (N - 1)) * nm + col];
            pb2[u] = pxc[(size_t)((x - 2) & (N - 1)) * nm + col];
          }
        }
      }
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const int x = (threadIdx.x + (i0 + u) * kNT) / TC;
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
  constexpr bool WL_OUT = fft_tps<N, P>() == 64 && TC == kNT / 64;
  fft_lds<T, LOGN, -1, kNT, P, !WL_OUT>(lds, ncol, tw);

  const int q = J.q;
  const double dt = J.dt, alpha = J.alpha;
  const double c1 = (1 - alpha) * dt, c2 = alpha * dt;
  const double hab = q == 2 ? dt / 2 : dt / 24;
  const double coeff = J.coeff[field];
  const T scale = (T)J.scale;
  const cx<T>* __restrict__ sin_ = J.state[par][field];
  cx<T>* __restrict__ sout = J.state[par ^ 1][field];
  cx<T>* __restrict__ hist = J.hist[field];
  // Epilogue: one wavefront per column, lanes on consecutive rows (the
  // spectral state is kz-major, so a column is contiguous).  Loads of U2
  // row groups are issued before any is used.
  constexpr int KR = (2 * ((N - 1) / 3) + 1 + 63) / 64;  // bound on ceil(rows / 64)
  constexpr int U2 = 4;
  const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const bool ddc = J.problem == MLV_PROB_DDC;
  uint32_t* __restrict__ carry = sizeof(T) == 4 ? J.comp[field] : nullptr;
  const int prev1 = (cur + q - 1) % q, prev2 = (cur + q - 2) % q, prev3 = (cur + q - 3) % q;
  const float fcoeff = (float)coeff, ffx2 = (float)J.fx2, ffz2 = (float)J.fz2;
  const float fdt = (float)dt, fc2 = (float)c2;
  // one wavefront per column (WPC per column when the block has fewer
  // columns than wavefronts, splitting its row groups)
  constexpr int NW = kNT / 64;
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
              e1[u] = J.state[par][1][idx];
              e2[u] = J.state[par][2][idx];
            } else {
              e1[u] = J.state[par][0][idx];
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

// Narrow single-precision blocks are held to 128 VGPRs (four wavefronts per
// SIMD, matching the four blocks per CU their LDS allows).
template <typename T, int WIDE>
__global__ void __launch_bounds__(kNT)
    __attribute__((amdgpu_waves_per_eu((WIDE || sizeof(T) == 8) ? 1 : 4)))
    k_stepC(const JobDev<T>* __restrict__ jobs, const int* __restrict__ pref, int njobs,
            int stage, int step, const cx<T>* __restrict__ tw) {
  extern __shared__ char smem_raw[];
  cx<T>* lds = reinterpret_cast<cx<T>*>(smem_raw);
  const int lb = xcd_logical(pref[njobs]);
  if (lb < 0) return;
  const int j = find_job(pref, njobs, lb);
  const JobDev<T>& J = jobs[j];
  const int local = lb - pref[j];
  const int par = (J.par0 + step) & 1;
  const int cur = ring(J.curr0, step, J.q);
  if (stage == J.nstage - 1 && local == 0 && threadIdx.x == 0) {
    // t += dt after each loop (Simulation.py:199-209); nothing else in this
    // launch reads these two fields
    JobDev<T>* Jw = const_cast<JobDev<T>*>(jobs) + j;
    Jw->t += J.dt;
    Jw->loops += 1;
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

template <typename T>
__global__ void k_set_dt(JobDev<T>* jobs, int j, double dt) {
  if (threadIdx.x == 0 && blockIdx.x == 0) jobs[j].dt = dt;
}

// One block per job, after a step whose z pass wrote diagnostics partials:
// tracker record and CFL rule (diag_apply).  The first advected scalar
// exists in vorticity and DDC jobs of two or more fields.
template <typename T>
__global__ void __launch_bounds__(kNT) k_diag(JobDev<T>* jobs, double cutoff, int cadence) {
  JobDev<T>& J = jobs[blockIdx.x];
  // loops has been advanced by this step's update kernel: the step was this
  // job's check loop when the count before it was a multiple of cadence
  if ((J.loops - 1) % cadence != 0) return;
  double dt = J.dt;
  const bool flux = J.problem != MLV_PROB_MHD && J.nfield >= 2;
  if (diag_apply<kNT>(J.diag, J.nbB, J.dx, J.dz, (double)J.nx * (double)J.nz, flux, cutoff, J.t,
                      J.loops, dt))
    J.dt = dt;
}

// --------------------------------------------------------------------------
// Host side
// --------------------------------------------------------------------------
template <typename T, int V>
static void launchB(const JobDev<T>* jobs, const BatchLaunch& bl, int stage, int flags,
                    const cx<T>* tw, hipStream_t s) {
  if (bl.nB[stage][V] == 0) return;
  const dim3 grid(xcd_grid(bl.nB[stage][V]));
  const int* pref = bl.prefB[stage][V];
  if constexpr (sizeof(T) == 4) {
    if (bl.ldsM) {  // some job carries mean profiles
      MLV_SET_LDS_ONCE((k_stepB<T, V, 1>), bl.ldsB[V]);
      hipLaunchKernelGGL((k_stepB<T, V, 1>), grid, dim3(kNT), bl.ldsB[V], s, jobs, pref,
                         bl.njobs, stage, flags, tw);
      return;
    }
  }
  MLV_SET_LDS_ONCE((k_stepB<T, V, 0>), bl.ldsB[V]);
  hipLaunchKernelGGL((k_stepB<T, V, 0>), grid, dim3(kNT), bl.ldsB[V], s, jobs, pref, bl.njobs,
                     stage, flags, tw);
}

template <typename T>
hipError_t fused_stage(const JobDev<T>* jobs, const BatchLaunch& bl, int stage, int flags,
                       int step, const cx<T>* tw, hipStream_t s, hipEvent_t* ev) {
  if (ev) (void)hipEventRecord(ev[0], s);
  if (bl.nA[stage] > 0) {
    MLV_SET_LDS_ONCE(k_stepA<T>, bl.ldsA);
    hipLaunchKernelGGL(k_stepA<T>, dim3(xcd_grid(bl.nA[stage])), dim3(kNT), bl.ldsA, s, jobs,
                       (const int*)bl.prefA[stage], bl.njobs, stage, step, tw);
  }
  if constexpr (sizeof(T) == 4) {
    if (bl.ldsM) {
      MLV_SET_LDS_ONCE(k_stepM<T>, bl.ldsM);
      hipLaunchKernelGGL(k_stepM<T>, dim3(bl.njobs * (kMeanCh / 2)), dim3(kNT), bl.ldsM, s, jobs,
                         stage, step, tw);
    }
  }
  if (ev) (void)hipEventRecord(ev[1], s);
  launchB<T, BV_VORT1>(jobs, bl, stage, flags, tw, s);
  launchB<T, BV_VORT2>(jobs, bl, stage, flags, tw, s);
  launchB<T, BV_VORT3>(jobs, bl, stage, flags, tw, s);
  launchB<T, BV_MHD1>(jobs, bl, stage, flags, tw, s);
  launchB<T, BV_MHD2>(jobs, bl, stage, flags, tw, s);
  if (ev) (void)hipEventRecord(ev[2], s);
  if (bl.nC[stage] > 0) {
    if (bl.wideC) {
      MLV_SET_LDS_ONCE((k_stepC<T, 1>), bl.ldsC);
      hipLaunchKernelGGL((k_stepC<T, 1>), dim3(xcd_grid(bl.nC[stage])), dim3(kNT), bl.ldsC, s,
                         jobs, (const int*)bl.prefC[stage], bl.njobs, stage, step, tw);
    } else {
      MLV_SET_LDS_ONCE((k_stepC<T, 0>), bl.ldsC);
      hipLaunchKernelGGL((k_stepC<T, 0>), dim3(xcd_grid(bl.nC[stage])), dim3(kNT), bl.ldsC, s,
                         jobs, (const int*)bl.prefC[stage], bl.njobs, stage, step, tw);
    }
  }
  if (ev) (void)hipEventRecord(ev[3], s);
  return hipGetLastError();
}

// Raise the dynamic-LDS limit of every fused kernel once, outside any
// stream capture.
template <typename T>
hipError_t fused_prepare(const BatchLaunch& bl) {
  size_t b = 0;
  for (int v = 0; v < BV_COUNT; ++v) b = bl.ldsB[v] > b ? bl.ldsB[v] : b;
  if (bl.ldsA) MLV_SET_LDS_ONCE(k_stepA<T>, bl.ldsA);
  if constexpr (sizeof(T) == 4) {
    if (bl.ldsM) MLV_SET_LDS_ONCE(k_stepM<T>, bl.ldsM);
  }

  if (bl.ldsCp[1]) MLV_SET_LDS_ONCE((k_stepC<T, 1>), bl.ldsCp[1]);
  if (bl.ldsCp[0]) MLV_SET_LDS_ONCE((k_stepC<T, 0>), bl.ldsCp[0]);
#define MLV_B_LDS(V)                                                           \
  if (bl.ldsB[V]) {                                                            \
    MLV_SET_LDS_ONCE((k_stepB<T, V, 0>), bl.ldsB[V]);                          \
    if constexpr (sizeof(T) == 4) MLV_SET_LDS_ONCE((k_stepB<T, V, 1>), bl.ldsB[V]); \
  }
  MLV_B_LDS(BV_VORT1)
  MLV_B_LDS(BV_VORT2)
  MLV_B_LDS(BV_VORT3)
  MLV_B_LDS(BV_MHD1)
  MLV_B_LDS(BV_MHD2)
#undef MLV_B_LDS
  (void)b;
  return hipGetLastError();
}
template hipError_t fused_prepare<float>(const BatchLaunch&);
template hipError_t fused_prepare<double>(const BatchLaunch&);

template <typename T>
hipError_t fused_set_dt(JobDev<T>* jobs, int job, double dt, hipStream_t s) {
  hipLaunchKernelGGL(k_set_dt<T>, dim3(1), dim3(64), 0, s, jobs, job, dt);
  return hipGetLastError();
}

template <typename T>
hipError_t fused_diag(JobDev<T>* jobs, int njobs, double cutoff, int cadence, hipStream_t s) {
  hipLaunchKernelGGL(k_diag<T>, dim3(njobs), dim3(kNT), 0, s, jobs, cutoff, cadence);
  return hipGetLastError();
}
template hipError_t fused_diag<float>(JobDev<float>*, int, double, int, hipStream_t);
template hipError_t fused_diag<double>(JobDev<double>*, int, double, int, hipStream_t);
template hipError_t fused_set_dt<float>(JobDev<float>*, int, double, hipStream_t);
template hipError_t fused_set_dt<double>(JobDev<double>*, int, double, hipStream_t);
template hipError_t fused_stage<float>(const JobDev<float>*, const BatchLaunch&, int, int, int,
                                       const cx<float>*, hipStream_t, hipEvent_t*);
template hipError_t fused_stage<double>(const JobDev<double>*, const BatchLaunch&, int, int, int,
                                        const cx<double>*, hipStream_t, hipEvent_t*);

}  // namespace mlv
