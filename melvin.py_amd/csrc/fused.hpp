// Fused fully-spectral timestep over a table of jobs (simulations).
//
// One stage of a step is three kernels:
//   A  x pass inverse: per output channel, a spectral prologue on a state
//      field (identity, or velocity from the streamfunction psi = -f/lap),
//      zero-pad |kx| <= nn to nx, inverse FFT along x -> channel (nx, nm).
//   B  z pass per row block: c2r of every channel row, physical products
//      (the flux terms sign * u * f of each equation), r2c of the products,
//      keep kz < nm -> Px_e, Pz_e channels (nx, nm).  Optionally writes the
//      physical rows of the input channels (for CFL and diagnostics).
//   C  x pass forward per equation: FFT along x of Px_e and Pz_e, keep
//      |kx| <= nn, apply the periodic finite-difference operator as its
//      exact modified wavenumber, add the equation's spectral source terms,
//      store the RHS into the Adams-Bashforth ring, and apply the
//      semi-implicit update, writing the new state to the other buffer of
//      a ping-pong pair.
//
// The periodic central difference of a periodic grid function is a
// circulant operator, so on the DFT grid it is exactly multiplication by
// i*sin(k*h)/h (order 2) or i*(8 sin(k h) - sin(2 k h))/(6 h) (order 4).
// Applying it after the forward transform is therefore the same discrete
// operator as the reference's physical-space stencil followed by rfft2
// (Variable.py:125-127, SpatialDifferentiator.py:76-104,130-185), without
// a halo exchange between rows.
#pragma once
#include <hip/hip_runtime.h>
#include "../../include/melvin_hip.h"
#include "diag.hpp"
#include "transforms.hpp"

namespace mlv {

constexpr int kMaxFields = 4;
constexpr int kMaxChan = 8;   // input channels of the z pass
constexpr int kMaxEq = 3;
constexpr int kMaxA = 8;
constexpr int kMaxPhys = 8;

enum XinvOpKind : int { XOP_ID = 0, XOP_UX = 1, XOP_UZ = 2 };

struct XinvOp {
  int src;       // field index
  int op;        // XinvOpKind
  int ch;        // destination channel
  int from_new;  // read the field from the freshly written state buffer
};

// z-pass variants (compile-time product programs)
enum BVariant : int { BV_VORT1 = 0, BV_VORT2 = 1, BV_VORT3 = 2, BV_MHD1 = 3, BV_MHD2 = 4, BV_COUNT = 5 };

struct StageDev {
  int nA;
  XinvOp a[kMaxA];
  int bvar;      // BVariant
  int neq;       // equations updated by kernel C
  int eq_field[kMaxEq];
  int phys_ch[kMaxPhys];  // z-pass input index written to phys[i] (-1 = none)
};

template <typename T>
struct JobDev {
  int problem, nstage;
  int nx, nz, lognx, lognz, nn, nm, rows;
  int q, curr0, par0, fd_order;
  int pbase;                 // first workspace channel of the Px/Pz outputs
  double fx, fz, fx2, fz2;   // 2*pi/L and -(2*pi)^2/L^2, as the reference forms them
  double dx, dz, scale;      // grid spacings, 1/(nx*nz)
  double alpha, dt;
  double coeff[kMaxFields];  // diffusion coefficient of each field's equation
  double param[4];           // problem parameters (DDC: Pr, R0)
  int nfield;
  cx<T>* state[2][kMaxFields];
  cx<T>* hist[kMaxFields];
  // single precision, optional: the Kahan residual of the state (value =
  // state - residual) as bfloat16 (re, im) pairs packed in 32 bits; it only
  // has to resolve a fraction of one FP32 ulp, so 8 bits of it suffice
  uint32_t* comp[kMaxFields];
  cx<T>* work;
  // single precision with split = 1: each z-pass input channel's x mean as
  // a function of z (kMeanCh rows of nz values, in the workspace after the
  // channels); the x passes then carry only the fluctuations
  int split;
  T* mean;
  cx<T>* psi;
  T* phys[kMaxPhys];
  StageDev st[2];
  // device-side time keeping and diagnostics (see k_diag): the update kernel
  // advances t and loops every step; diag holds the z pass's per-block
  // partials (kDiagPart doubles per block, nbB blocks), then the record
  // count, the CFL breach flag and the series of kDiagRec-double records
  double t;
  long long loops;
  int nbB;
  double* diag;
};

// bfloat16 pair <-> float2 (round to nearest even; the residuals are finite)
__device__ __forceinline__ float2 bf2_load(uint32_t v) {
  return make_float2(__uint_as_float(v << 16), __uint_as_float(v & 0xffff0000u));
}
__device__ __forceinline__ uint32_t bf16_bits(float f) {
  const uint32_t u = __float_as_uint(f);
  return (u + 0x7fffu + ((u >> 16) & 1u)) >> 16;
}
__device__ __forceinline__ uint32_t bf2_pack(float re, float im) {
  return bf16_bits(re) | (bf16_bits(im) << 16);
}

// a*b without its mean-times-mean part: (am + af)(bm + bf) - am bm, with
// means m and fluctuations f (single-precision mean split, see kernel M)
template <typename T>
__device__ __forceinline__ T bil(T am, T af, T bm, T bf) {
  return am * bf + af * bm + af * bf;
}

// Per-variant channel/product description used by kernel B.  products()
// takes whole values; products_split() takes the x means m and the
// fluctuations f of the channels and leaves out the mean-times-mean part.
template <int V> struct BProg;
template <int NF> struct BVortProg {
  static constexpr int NIN = NF + 2;  // fields, ux, uz
  static constexpr int NEQ = NF;
  static constexpr int UX = NF, UZ = NF + 1;  // velocity channels
  static constexpr int FLUX = NF >= 2 ? 1 : -1;  // scalar whose flux f*uz is tracked
  // channels whose x mean is zero (d/dx of a potential)
  static constexpr bool zero_mean(int q) { return q == UZ; }
  template <typename T>
  __device__ __forceinline__ static void products(const T* v, T* px, T* pz) {
#pragma unroll
    for (int e = 0; e < NF; ++e) {
      px[e] = -(v[NF] * v[e]);
      pz[e] = -(v[NF + 1] * v[e]);
    }
  }
  template <typename T>
  __device__ __forceinline__ static void products_split(const T* m, const T* f, T* px, T* pz) {
#pragma unroll
    for (int e = 0; e < NF; ++e) {
      px[e] = -bil(m[NF], f[NF], m[e], f[e]);
      pz[e] = -bil(m[NF + 1], f[NF + 1], m[e], f[e]);
    }
  }
};
template <> struct BProg<BV_VORT1> : BVortProg<1> {};
template <> struct BProg<BV_VORT2> : BVortProg<2> {};
template <> struct BProg<BV_VORT3> : BVortProg<3> {};
// MHD channels: 0 w, 1 j, 2 ux, 3 uz, 4 bx, 5 bz
template <> struct BProg<BV_MHD1> {
  static constexpr int NIN = 6;
  static constexpr int UX = 2, UZ = 3;
  static constexpr int FLUX = -1;
  static constexpr bool zero_mean(int q) { return q == 3 || q == 5; }  // uz, bz
  static constexpr int NEQ = 1;
  template <typename T>
  __device__ __forceinline__ static void products(const T* v, T* px, T* pz) {
    px[0] = -(v[2] * v[0]) + v[4] * v[1];
    pz[0] = -(v[3] * v[0]) + v[5] * v[1];
  }
  template <typename T>
  __device__ __forceinline__ static void products_split(const T* m, const T* f, T* px, T* pz) {
    px[0] = -bil(m[2], f[2], m[0], f[0]) + bil(m[4], f[4], m[1], f[1]);
    pz[0] = -bil(m[3], f[3], m[0], f[0]) + bil(m[5], f[5], m[1], f[1]);
  }
};
template <> struct BProg<BV_MHD2> {
  static constexpr int NIN = 6;
  static constexpr int UX = 2, UZ = 3;
  static constexpr int FLUX = -1;
  static constexpr bool zero_mean(int q) { return q == 3 || q == 5; }  // uz, bz
  static constexpr int NEQ = 1;
  template <typename T>
  __device__ __forceinline__ static void products(const T* v, T* px, T* pz) {
    px[0] = -(v[2] * v[1]) + v[4] * v[0];
    pz[0] = -(v[3] * v[1]) + v[5] * v[0];
  }
  template <typename T>
  __device__ __forceinline__ static void products_split(const T* m, const T* f, T* px, T* pz) {
    px[0] = -bil(m[2], f[2], m[1], f[1]) + bil(m[4], f[4], m[0], f[0]);
    pz[0] = -bil(m[3], f[3], m[1], f[1]) + bil(m[5], f[5], m[0], f[0]);
  }
};

// Columns per block of the x-pass kernels: every LDS slot, but at most one
// per thread, so that each thread keeps a single column (kNT % cols == 0).
template <int LOGN, int P, int NT = kNT>
__host__ __device__ constexpr int fused_cols() {
  return fft_slots<LOGN, NT, P>() < NT ? fft_slots<LOGN, NT, P>() : NT;
}
// Kernel C's tile shapes (BatchLaunch::wideC): 0 narrow and 1 wide at kNT
// threads (fused_points), 2 narrow at 512 threads -- twice the columns per
// block, so each row of Px / Pz is read in runs of 8 (single precision,
// batches of one x size, fixed_logn).
__host__ __device__ constexpr int cshape_threads(int shape) { return shape == 2 ? 512 : kNT; }
__host__ __device__ constexpr int cshape_wide(int shape) { return shape == 1 ? 1 : 0; }
// Points per thread of the x passes.  wide = 1 is the column-pass policy
// of fft_points_col (32 points from N = 1024 in single precision: more
// columns per block); wide = 0 keeps 16 points up to N = 2048 (half the LDS
// per block, so twice the blocks per CU, and a batch that mixes 2048 with
// smaller grids keeps the smaller grids' occupancy: the launch's LDS is the
// largest job's).  The x-inverse always runs narrow;
// the x-forward/update kernel runs narrow in single precision and, in
// double precision, wide when its grid fills the device several times over.
template <typename T>
__host__ __device__ constexpr int fused_points(int logn, int wide) {
  return wide ? fft_points_col<T>(logn) : (logn >= 12 ? 32 : 16);
}
template <typename T>
inline int fused_cols_host(int logn, int wide) {
  const int s = (kNT * fused_points<T>(logn, wide)) >> logn;
  return s < kNT ? s : kNT;
}
template <typename T>
inline size_t fused_lds_bytes(int logn, int wide) {
  const int s = (kNT * fused_points<T>(logn, wide)) >> logn;
  return (size_t)s * fft_ldsn(1 << logn) * sizeof(cx<T>);
}

// Threads per block of the x-inverse kernel A, by shape (BatchLaunch::
// shapeA).  Shape 0, the default: single precision takes 512, so that a
// block holds 8 columns at 1024 points and each interleaved row pair's store
// is a whole 128-byte line (two such blocks per CU give as many wavefronts
// as four blocks of 256); with 64-byte runs the stores of neighbouring tiles
// only merged in L2 when their blocks ran side by side.  Double precision
// takes 256 (4 columns of 16-byte values already fill a line, and 512 would
// not fit the LDS at 4096 points).  Shape 1 halves the block (half the
// columns per block, twice the blocks); shape 2 keeps shape 0's block with
// every op of the stage side by side (stepA_mixed); the autotuner times
// them.
template <typename T, int SHAPE = 0>
__host__ __device__ constexpr int xinv_threads() {
  return (sizeof(T) == 4 ? 512 : 256) >> (SHAPE == 1 ? 1 : 0);
}
template <typename T, int LOGN, int P, int SHAPE = 0>
__host__ __device__ constexpr int xinv_cols() {
  return ((xinv_threads<T, SHAPE>() * P) >> LOGN) < xinv_threads<T, SHAPE>()
             ? ((xinv_threads<T, SHAPE>() * P) >> LOGN)
             : xinv_threads<T, SHAPE>();
}
template <typename T>
inline int xinv_cols_host(int logn, int shape = 0) {
  const int nt = (sizeof(T) == 4 ? 512 : 256) >> (shape == 1 ? 1 : 0);
  const int s = (nt * fused_points<T>(logn, 0)) >> logn;
  return s < nt ? s : nt;
}
template <typename T>
inline size_t xinv_lds_bytes(int logn, int shape = 0) {
  return (size_t)xinv_cols_host<T>(logn, shape) * fft_ldsn(1 << logn) * sizeof(cx<T>);
}
// Columns per block of every op in kernel A's mixed tiles (shape 2), and
// the host check that a job's stages fit them: single precision at 2^9 and
// 2^10 points, kMixCols x (the most ops of any stage) <= the block's slots.
constexpr int kMixCols = 2;
template <typename T>
inline bool xmix_fits(int logn, int max_nA) {
  return sizeof(T) == 4 && (logn == 9 || logn == 10) &&
         kMixCols * max_nA <= xinv_cols_host<T>(logn, 0);
}

// Rows per z-pass block: as many as fit the LDS slots, at most 8.
__host__ __device__ constexpr int brows(int nin, int neq, int slots) {
  int best = 1;
  for (int r = 1; r <= 8; ++r) {
    const int need_in = (r * nin + 1) / 2;
    const int need_out = r * neq;
    const int need = need_in > need_out ? need_in : need_out;
    if (need <= slots) best = r;
  }
  return best;
}
__host__ __device__ constexpr bool bfits(int nin, int neq, int slots) {
  return ((nin + 1) / 2 <= slots) && (neq <= slots);
}
// Points per thread of the z pass of a program with nin inputs and neq
// equations: 16 up to N = 1024, and at N = 2048 when the program fits the
// two slots that leaves (for the reason given at fused_points); otherwise
// the row policy's 32.
template <typename T>
__host__ __device__ constexpr int bpoints(int logn, int nin, int neq) {
  return (logn <= 10 || (logn == 11 && bfits(nin, neq, (kNT * 16) >> 11)))
             ? 16
             : fft_points_row<T>(logn);
}
__host__ __device__ constexpr int bvar_nin(int v) { return v <= BV_VORT3 ? v + 3 : 6; }
__host__ __device__ constexpr int bvar_neq(int v) { return v <= BV_VORT3 ? v + 1 : 1; }

// The split-pass z pass (inverse 16,16,R with the products in its last
// pass, forward R,16,16) applies at 2^9..2^11 points when the program runs
// at 16 points per thread and fits the slots.
template <typename T>
__host__ __device__ constexpr bool bsplit_fits(int logn, int v) {
  return logn >= 9 && logn <= 11 && bpoints<T>(logn, bvar_nin(v), bvar_neq(v)) == 16 &&
         bfits(bvar_nin(v), bvar_neq(v), (kNT * 16) >> logn);
}
// Plans of the z pass (BatchLaunch::planB): the split-pass plan where it
// applies, or the two-sweep plan everywhere (16,16,R both ways, the
// products in a sweep of their own between the transforms).
constexpr int kPlanSplit = 0, kPlanTwoSweep = 1;

// workspace channel of equation e's Px (k=0) / Pz (k=1) output; the
// z-pass inputs occupy channels [0, base)
__host__ __device__ constexpr int pchan(int base, int e, int k) { return base + 2 * e + k; }

// Layout of the z-pass input channels (kernel A writes them, kernel B reads
// them): x rows in groups of kXil, interleaved by kz, so that one of A's
// column tiles stores runs of kXil x TC elements rather than TC.  Element
// (x, k) of a channel is at xrow_base(x, nm) + kXil k.  Groups of two beat
// four: A's runs are 64 bytes instead of 128 (+10 us at 64 x 1024^2), but a
// z-pass block reads both rows of each group (-32 us).
constexpr int kXil = 2;
__host__ __device__ constexpr size_t xrow_base(int x, int nm) {
  return (size_t)(x / kXil) * (kXil * (size_t)nm) + (size_t)(x % kXil);
}

// LDS slots a z-pass block of `slots` available slots touches: its input
// slots (a row pair's channel per slot when the row count is even, as
// BShape lays them out) or its product slots, whichever is more.  The
// launch sizes the LDS to these alone, so a slot the block never uses does
// not cost blocks per CU (3 of 4 slots for a one-field vortex at 1024 z
// points: 5 blocks per CU instead of 4).
__host__ __device__ constexpr int bslots_used(int nin, int neq, int slots) {
  const int rb = brows(nin, neq, slots);
  const int in = (rb % 2 == 0 && kXil == 2) ? nin * (rb / 2) : (rb * nin + 1) / 2;
  const int out = rb * neq;
  return in > out ? in : out;
}

// Layout of the Px / Pz channels (kernel B writes them, kernel C reads
// them): kz in groups of kPGroup, each group x-major, so element (x, k) is
// at pidx(x, k, nx).  A column tile of kernel C (kPGroup columns, or half
// that) is then one contiguous run, and a z-pass block, which holds a row
// pair, stores each group of both rows as one whole 128-byte line (single
// precision).  Row-major rows cost the update kernel 64-byte runs at a
// stride of a row.
constexpr int kPGroup = 8;
__host__ __device__ constexpr size_t pidx(int x, int k, int nx) {
  return (size_t)(k / kPGroup) * ((size_t)nx * kPGroup) + (size_t)x * kPGroup + (size_t)(k % kPGroup);
}
// Elements per workspace channel: (nx, nm), nm rounded up to whole groups.
__host__ __device__ constexpr size_t work_chs(int nx, int nm) {
  return (size_t)nx * (size_t)((nm + kPGroup - 1) / kPGroup * kPGroup);
}

// Rows of the mean-profile area (the most z-pass input channels).
constexpr int kMeanCh = 6;
// Workspace of a job, in complex elements: the z-pass inputs and the Px/Pz
// outputs, (nin + 2 neq) channels of work_chs, then the mean profiles.
inline size_t fused_work_elems(int nin, int neq, int nx, int nz) {
  const int nm = (nz - 1) / 3;
  return (size_t)(nin + 2 * neq) * work_chs(nx, nm) + (size_t)(kMeanCh / 2) * nz;
}

struct BatchLaunch {
  int njobs;
  int shapeA;                          // block shape of kernel A (xinv_threads)
  int planB;                           // plan of kernel B (kPlanSplit / kPlanTwoSweep)
  int wideC;                           // points policy of kernel C (fused_points)
  int nA[2], nB[2][BV_COUNT], nC[2];  // total blocks per launch
  size_t ldsA, ldsB[BV_COUNT], ldsC;
  size_t ldsM;                         // single precision mean profiles: one slot; 0 = none
  int* prefA[2];                       // device prefix arrays (njobs + 1)
  int* prefB[2][BV_COUNT];
  int* prefC[2];
  // the tile shapes of kernel C (index = shape, cshape_*; okC[2] = 0 outside
  // single-precision batches of one x size); wideC / nC / ldsC / prefC
  // above are the selected one
  int okC[3];
  int nCp[3][2];
  size_t ldsCp[3];
  int* prefCp[3][2];
  // the block shapes of kernel A (index = shape; okA[s] = 0 where shape s
  // does not apply to every job: the smaller block where a job's transform
  // does not fit it, the mixed tiles outside xmix_fits or in a batch of
  // several x sizes); nA / ldsA / prefA above are the selected one
  int okA[3];
  int nAp[3][2];
  size_t ldsAp[3];
  int* prefAp[3][2];
  // whether some job's z pass takes the split-pass plan under kPlanSplit
  int splitB;
  // log2 of the x / z grid size shared by every job of the batch, -1 when
  // the jobs differ: a batch of one grid size launches kernels compiled for
  // that size alone (kFixedLogn), whose registers are their own rather than
  // the largest over every size the generic kernels handle
  int lognx_all, lognz_all;
};

// Grid sizes with kernels compiled for them alone: 2^9 .. 2^11.
__host__ __device__ constexpr bool fixed_logn(int logn) { return logn >= 9 && logn <= 11; }

// Select kernel A's block shape.
inline void select_a_shape(BatchLaunch& bl, int shape) {
  bl.shapeA = shape;
  bl.ldsA = bl.ldsAp[shape];
  for (int s = 0; s < 2; ++s) {
    bl.nA[s] = bl.nAp[shape][s];
    bl.prefA[s] = bl.prefAp[shape][s];
  }
}

// Select kernel C's tile shape (wide = 1: fused_points' wide policy).
inline void select_c_policy(BatchLaunch& bl, int shape) {
  bl.wideC = shape;
  bl.ldsC = bl.ldsCp[shape];
  for (int s = 0; s < 2; ++s) {
    bl.nC[s] = bl.nCp[shape][s];
    bl.prefC[s] = bl.prefCp[shape][s];
  }
}

// `step` is the batch's step count modulo 4 (a multiple of both ring
// lengths and of the ping-pong period), so a captured sequence of 4k steps
// replays correctly.
template <typename T>
hipError_t fused_stage(const JobDev<T>* jobs, const BatchLaunch& bl, int stage, int flags,
                       int step, const cx<T>* tw, hipStream_t s, hipEvent_t* ev = nullptr);
template <typename T>
hipError_t fused_set_dt(JobDev<T>* jobs, int job, double dt, hipStream_t s);
// flags of a stage's z pass: physical outputs, diagnostics partials; with
// kDiag the CFL cadence sits in the bits from kDiagShift up, and a job
// writes partials on the loops whose count before the loop is a multiple of
// it (the reference's ticker fires after loops 1, 1 + cadence, ...)
constexpr int kWritePhys = 1, kDiag = 2, kDiagShift = 8;
// After a step whose z pass ran with kDiag: for each job whose check loop
// it was, reduce its partials, append a tracker record, and apply the CFL
// rule to its dt.
template <typename T>
hipError_t fused_diag(JobDev<T>* jobs, int njobs, double cutoff, int cadence, hipStream_t s);
template <typename T>
hipError_t fused_prepare(const BatchLaunch& bl);

// Launchers and one-time LDS set-up of the three kernel families (one
// translation unit each: fused_a.hip, fused_b.hip, fused_c.hip).
template <typename T>
void launch_xinv(const JobDev<T>* jobs, const BatchLaunch& bl, int stage, int step,
                 const cx<T>* tw, hipStream_t s);
template <typename T>
void launch_zpass(const JobDev<T>* jobs, const BatchLaunch& bl, int stage, int flags,
                  const cx<T>* tw, hipStream_t s);
template <typename T>
void launch_update(const JobDev<T>* jobs, const BatchLaunch& bl, int stage, int step,
                   const cx<T>* tw, hipStream_t s);
template <typename T>
void prepare_xinv(const BatchLaunch& bl);
template <typename T>
void prepare_zpass(const BatchLaunch& bl);
template <typename T>
void prepare_update(const BatchLaunch& bl);
// The z pass of program v for a batch of one z size (fixed_logn), compiled
// in its own translation unit (fused_b_fixed.hip).
template <typename T>
void launch_zpass_fixed(int v, const JobDev<T>* jobs, const BatchLaunch& bl, int stage, int flags,
                        const cx<T>* tw, hipStream_t s);
template <typename T>
void prepare_zpass_fixed(const BatchLaunch& bl);

}  // namespace mlv
