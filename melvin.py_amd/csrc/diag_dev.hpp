// Device side of the on-device diagnostics and CFL rule (see diag.hpp).
#pragma once
#include <hip/hip_runtime.h>
#include <math.h>

#include "common.hpp"
#include "diag.hpp"

namespace mlv {

// NaN-propagating maximum (np.max semantics; fmax would drop a NaN).
template <typename T>
__device__ __forceinline__ T nan_max(T a, T b) {
  return (a != a || b != b) ? T(NAN) : (a > b ? a : b);
}

// Neumaier-compensated sum of the (sum, compensation) pairs of one
// wavefront, by a butterfly over its lanes; every lane ends with the total.
__device__ __forceinline__ KSum wave_ksum(KSum k) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    const double s = __shfl_xor(k.s, off, 64), c = __shfl_xor(k.c, off, 64);
    k.add(s);
    k.add(c);
  }
  return k;
}

// Block reduction (NT threads) of one pass's partials; thread 0 writes
// (max ux, max uz, then sum(ux^2 + uz^2), sum(w^2), sum(f1 uz) as
// Neumaier (sum, compensation) pairs: the sums first, then their
// compensations) to out.  Each thread's own values are short plain sums.
template <int NT>
__device__ void block_diag(double a, double b, double ke, double ens, double flux, double* out) {
  __shared__ double red[NT / 64][kDiagPart];
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    a = nan_max(a, __shfl_xor(a, off, 64));
    b = nan_max(b, __shfl_xor(b, off, 64));
  }
  KSum k[3];
  k[0].s = ke;
  k[1].s = ens;
  k[2].s = flux;
#pragma unroll
  for (int q = 0; q < 3; ++q) k[q] = wave_ksum(k[q]);
  const int wv = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0) {
    red[wv][0] = a;
    red[wv][1] = b;
#pragma unroll
    for (int q = 0; q < 3; ++q) {
      red[wv][2 + q] = k[q].s;
      red[wv][5 + q] = k[q].c;
    }
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int w = 1; w < NT / 64; ++w) {
      a = nan_max(a, red[w][0]);
      b = nan_max(b, red[w][1]);
#pragma unroll
      for (int q = 0; q < 3; ++q) {
        k[q].add(red[w][2 + q]);
        k[q].add(red[w][5 + q]);
      }
    }
    out[0] = a;
    out[1] = b;
#pragma unroll
    for (int q = 0; q < 3; ++q) {
      out[2 + q] = k[q].s;
      out[5 + q] = k[q].c;
    }
  }
}

// Called by every thread of an NT block, once per job after a pass that
// wrote nb partials at diag: reduces them, then thread 0 appends a tracker
// record (loop, t, dt, kinetic energy 0.5 sum(ux^2+uz^2)/npts as utility.py:
// 42-59, max ux, max uz, next dt, enstrophy, scalar flux) and applies the
// CFL rule of Integrator.py:35-44: cfl_dt = min(dx / max ux, dz / max uz)
// with Python's min (a NaN in the second operand is ignored, in the first
// it is returned); a breach or a NaN sets the breach flag and leaves dt,
// else dt shrinks by 0.9 until dt <= cutoff * cfl_dt.  Returns true on
// thread 0 when dt changed (written back through dt).
template <int NT>
__device__ bool diag_apply(double* diag, int nb, double dx, double dz, double npts,
                           bool has_flux, double cutoff, double t, long long loops, double& dt) {
  __shared__ double red[NT / 64][kDiagPart];
  double a = -INFINITY, b = -INFINITY;
  KSum k[3];
  for (int i = threadIdx.x; i < nb; i += NT) {
    const double* p = diag + (size_t)kDiagPart * i;
    a = nan_max(a, p[0]);
    b = nan_max(b, p[1]);
#pragma unroll
    for (int q = 0; q < 3; ++q) {
      k[q].add(p[2 + q]);
      k[q].add(p[5 + q]);
    }
  }
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    a = nan_max(a, __shfl_xor(a, off, 64));
    b = nan_max(b, __shfl_xor(b, off, 64));
  }
#pragma unroll
  for (int q = 0; q < 3; ++q) k[q] = wave_ksum(k[q]);
  const int wv = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0) {
    red[wv][0] = a;
    red[wv][1] = b;
#pragma unroll
    for (int q = 0; q < 3; ++q) {
      red[wv][2 + q] = k[q].s;
      red[wv][5 + q] = k[q].c;
    }
  }
  __syncthreads();
  if (threadIdx.x != 0) return false;
  for (int w = 1; w < NT / 64; ++w) {
    a = nan_max(a, red[w][0]);
    b = nan_max(b, red[w][1]);
#pragma unroll
    for (int q = 0; q < 3; ++q) {
      k[q].add(red[w][2 + q]);
      k[q].add(red[w][5 + q]);
    }
  }
  const double ke = k[0].value(), ens = k[1].value(), flux = k[2].value();
  const double c1 = dx / a, c2 = dz / b;
  const double cfl = c2 < c1 ? c2 : c1;
  const double used = dt;
  double* hdr = diag + (size_t)kDiagPart * nb;
  double next = dt;
  if (dt > cfl || cfl != cfl) {
    hdr[1] = 1.0;  // breach
  } else {
    while (next > cutoff * cfl) next = next * 0.9;
  }
  const long long nrec = (long long)hdr[0];
  double* rec = hdr + kDiagHdr + (size_t)kDiagRec * (nrec % kDiagCap);
  rec[0] = (double)loops;
  rec[1] = t;
  rec[2] = used;
  rec[3] = 0.5 * ke / npts;
  rec[4] = a;
  rec[5] = b;
  rec[6] = next;
  rec[7] = 0.5 * ens / npts;
  rec[8] = has_flux ? flux / npts : (double)NAN;
  hdr[0] = (double)(nrec + 1);
  const bool changed = next != used;
  dt = next;
  return changed;
}

}  // namespace mlv
