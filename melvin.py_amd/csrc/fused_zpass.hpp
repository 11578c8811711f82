// Device bodies of the fused step's z-pass kernel B (c2r of the channel
// rows, physical products, r2c of the products), launched by fused_b.hip
// (generic) and fused_b_fixed.hip (one z size).
#pragma once
#include "fused.hpp"
#include "diag_dev.hpp"
#include "grid.hpp"

namespace mlv {
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
  // Row-pair slots (an even number of rows): slot s holds channel s % NIN of
  // the rows x0 + 2 (s / NIN) and x0 + 2 (s / NIN) + 1, which the
  // interleaved channel layout stores side by side, so one load of two
  // elements per k fills both halves of the slot and every load instruction
  // uses whole cache lines.  Otherwise consecutive sequences share a slot.
  static constexpr bool PAIR = RB % 2 == 0 && kXil == 2;
  static constexpr int SLN = PAIR ? NIN * (RB / 2) : (RB * NIN + 1) / 2;  // input slots
  static constexpr int WPS = SLN >= NW ? 1 : NW / SLN;  // wavefronts per input slot
  // slot, and half (0 real, 1 imaginary part), of sequence q = row * NIN +
  // channel
  static constexpr int slot_of(int q) { return PAIR ? (q / NIN / 2) * NIN + q % NIN : q / 2; }
  static constexpr int half_of(int q) { return PAIR ? (q / NIN) % 2 : q % 2; }
  // slots in use for a block of nrows rows (even when PAIR: RB is even and
  // the grid sizes are powers of two)
  __device__ __forceinline__ static int nslots(int nrows) {
    return PAIR ? NIN * (nrows / 2) : (nrows * NIN + 1) / 2;
  }
  static constexpr int NO = RB * NEQ;                  // product slots
  static constexpr int WPO = NO >= NW ? 1 : NW / NO;   // wavefronts per product slot
  static_assert(SLN <= bslots_used(NIN, NEQ, SL) && NO <= bslots_used(NIN, NEQ, SL),
                "the launch's LDS holds the slots in use");
};

// Two consecutive elements (16 bytes in single precision, one load).
template <typename T>
__device__ __forceinline__ void load_two(const cx<T>* __restrict__ p, cx<T>& a, cx<T>& b) {
  if constexpr (sizeof(T) == 4) {
    const float4 v = *reinterpret_cast<const float4*>(p);
    a = mk<T>(v.x, v.y);
    b = mk<T>(v.z, v.w);
  } else {
    a = p[0];
    b = p[1];
  }
}

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
    const cx<T>* __restrict__ pa;
    const cx<T>* __restrict__ pb = nullptr;
    if constexpr (Sh::PAIR) {
      // channel s % NIN, rows x0 + 2 (s / NIN) + {0, 1}: element k of the
      // pair at 2k and 2k + 1
      pa = work + (size_t)(s % NIN) * chs + xrow_base(x0 + 2 * (s / NIN), nm);
    } else {
      const int qa = 2 * s, qb = qa + 1;
      pa = work + (size_t)(qa % NIN) * chs + xrow_base(x0 + qa / NIN, nm);
      if (qb < nseq) pb = work + (size_t)(qb % NIN) * chs + xrow_base(x0 + qb / NIN, nm);
    }
    // loads of up to KB iterations in flight before any is used
#pragma unroll 1
    for (int i0 = ssub * KB; i0 < KIT; i0 += WPS * KB) {
      cx<T> av[KB], bv[KB];
#pragma unroll
      for (int i = 0; i < KB; ++i) {
        const int k = lane + 64 * (i0 + i);
        if constexpr (Sh::PAIR) {
          if (k < nm) {
            load_two<T>(pa + (size_t)k * kXil, av[i], bv[i]);
          } else {
            av[i] = czero<T>();
            bv[i] = czero<T>();
          }
        } else {
          av[i] = k < nm ? pa[(size_t)k * kXil] : czero<T>();
          bv[i] = (pb && k < nm) ? pb[(size_t)k * kXil] : czero<T>();
        }
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
    mv[q] = (kSplit && !Pg::zero_mean(q)) ? glb(J.mean)[(size_t)q * N + n] : T(0);
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
          if (S.phys_ch[i] == q % NIN) glb(J.phys[i])[(size_t)(x0 + q / NIN) * N + n] = whole(q);
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
// the packed (Px + i Pz) spectrum, into the Px / Pz channels (pidx layout).
// A work item is (equation, row pair, four kz groups); its lanes take row
// (lane >> 3) & 1 of the pair, group (lane >> 4) of the four and kz % 8 =
// lane & 7, so each 16 lanes store one group of both rows -- one whole line
// per channel.  Every product slot must be complete for the whole block.
template <typename T, typename Sh>
__device__ __forceinline__ void bstore(cx<T>* __restrict__ work, size_t chs, int nm, int nx,
                                       int x0, int nrows, const cx<T>* lds) {
  constexpr int N = Sh::N, NIN = Sh::NIN, NEQ = Sh::NEQ, RB = Sh::RB, NW = Sh::NW;
  static_assert(kPGroup == 8, "lane mapping: 8 kz per group");
  constexpr int NPAIR = (RB + 1) / 2;
  const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int r2 = (lane >> 3) & 1, gsub = lane >> 4, kk = lane & 7;
  const int nquad = (nm + 4 * kPGroup - 1) / (4 * kPGroup);
  const int nitem = NEQ * NPAIR * nquad;
  for (int it = wv; it < nitem; it += NW) {
    const int eq = it % NEQ, rest = it / NEQ;
    const int pr = rest % NPAIR, quad = rest / NPAIR;
    const int r = 2 * pr + r2;
    const int k = (quad * 4 + gsub) * kPGroup + kk;
    if (r < nrows && k < nm) {
      const cx<T>* z = lds + (r * NEQ + eq) * Sh::LDSN;
      const cx<T> y = z[fft_pad(k)];
      const cx<T> ym = z[fft_pad((N - k) & (N - 1))];
      const size_t at = pidx(x0 + r, k, nx);
      work[(size_t)pchan(NIN, eq, 0) * chs + at] = mk<T>(T(0.5) * (y.x + ym.x), T(0.5) * (y.y - ym.y));
      work[(size_t)pchan(NIN, eq, 1) * chs + at] = mk<T>(T(0.5) * (y.y + ym.y), T(0.5) * (ym.x - y.x));
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
    const size_t chs = work_chs(nx, nm);
    const int x0 = local * RB;
    const int nrows = min(RB, nx - x0);
    const int nseq = nrows * NIN;
    const int nslot = Sh::nslots(nrows);
    // slot s is loaded by wavefront s and transformed by the same
    // wavefront: no workgroup barrier before the inverse transform
    constexpr bool WL_IN = fft_tps<N, P>() == 64 && Sh::WPS == 1 && Sh::SL == Sh::NW;
    // a wavefront issuing its loads goes ahead of other blocks' arithmetic
    __builtin_amdgcn_s_setprio(1);
    bload<T, Sh>(glb(J.work), chs, nm, x0, nseq, nslot, lds);
    __builtin_amdgcn_s_setprio(0);
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
          const cx<T> z = lds[Sh::slot_of(q) * LDSN + fft_pad(n)];
          v[q] = Sh::half_of(q) ? z.y : z.x;
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
                              glb(J.diag) + (size_t)kDiagPart * local);
    __syncthreads();
    fft_lds<T, LOGN, -1, kNT, P, true>(lds, nrows * NEQ, tw);
    bstore<T, Sh>(glb(J.work), chs, nm, nx, x0, nrows, lds);
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
  const size_t chs = work_chs(nx, nm);
  const int x0 = local * RB;
  const int nrows = min(RB, nx - x0);
  const int nseq = nrows * NIN;
  const int nslot = Sh::nslots(nrows);
  constexpr bool WL_IN = fft_tps<N, P>() == 64 && Sh::WPS == 1 && Sh::SL == Sh::NW;
  // a wavefront issuing its loads goes ahead of other blocks' arithmetic
  __builtin_amdgcn_s_setprio(1);
  bload<T, Sh>(glb(J.work), chs, nm, x0, nseq, nslot, lds);
  __builtin_amdgcn_s_setprio(0);
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
    for (int q = 0; q < RB * NIN; ++q)
      v[q] = q < nseq ? val[r][2 * Sh::slot_of(q) + Sh::half_of(q)] : T(0);
    T px[Sh::NO], pz[Sh::NO];
    bpoint<T, Sh>(J, S, kSplit, diag, flags, x0, nrows, nseq, j + NSL * r, v, d, px, pz);
#pragma unroll
    for (int o = 0; o < Sh::NO; ++o) pr[o][r] = mk<T>(px[o], pz[o]);
  }
  if (diag) block_diag<kNT>((double)d.mux, (double)d.muz, d.ke, d.ens, d.flux,
                            glb(J.diag) + (size_t)kDiagPart * local);
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
  __syncthreads();
  bstore<T, Sh>(glb(J.work), chs, nm, nx, x0, nrows, lds);
}

template <typename T, int LOGN, int V>
__host__ __device__ constexpr bool bsplit_ok() {
  static_assert(fft_slots<LOGN, kNT, 16>() == ((kNT * 16) >> LOGN), "slots at 16 points");
  return bsplit_fits<T>(LOGN, V);
}

// Kernel B's work for logical block lb of a launch over the job table's
// z-pass blocks of program V (prefix array pref), in plan PLAN.
template <typename T, int V, int SPLIT, int PLAN>
__device__ __forceinline__ void roleB(const JobDev<T>* __restrict__ jobs, const int* __restrict__ pref,
                                      int njobs, int lb, int stage, int flags,
                                      const cx<T>* __restrict__ tw, cx<T>* lds) {
  const int j = find_job(pref, njobs, lb);
  const JobDev<T>& J = jobs[j];
  const int local = lb - pref[j];
  switch (J.lognz) {
#define X(L)                                                          \
  case L:                                                             \
    if constexpr (PLAN == kPlanSplit && bsplit_ok<T, L, V>())         \
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

// Kernel B's work for a batch whose jobs all have 2^LOGN points along z.
template <typename T, int V, int SPLIT, int PLAN, int LOGN>
__device__ __forceinline__ void roleB_fixed(const JobDev<T>* __restrict__ jobs,
                                            const int* __restrict__ pref, int njobs, int lb,
                                            int stage, int flags, const cx<T>* __restrict__ tw,
                                            cx<T>* lds) {
  const int j = find_job(pref, njobs, lb);
  const JobDev<T>& J = jobs[j];
  const int local = lb - pref[j];
  if constexpr (PLAN == kPlanSplit && bsplit_ok<T, LOGN, V>())
    stepB_split<T, LOGN, V, SPLIT>(J, J.st[stage], local, flags, tw, lds);
  else
    stepB_body<T, LOGN, V, SPLIT>(J, J.st[stage], local, flags, tw, lds);
}

}  // namespace mlv
