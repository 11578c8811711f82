// Host-side launchers and enums of the element-wise operator kernels.
#pragma once
#include <hip/hip_runtime.h>
#include "common.hpp"
#include "../../include/melvin_hip.h"

namespace mlv {

// internal spectral kind: out += d2/dx2(in) (FD-z snabla2)
constexpr int kSpecAddD2dx2 = 100;

// real: bit 0 / bit 1 make the x / z first-derivative factor real (sine and
// cosine bases) instead of i times the factor (complex exponentials)
template <typename T>
hipError_t launch_spec(const void* in, void* out, int rows, int cols, int nn, int fully, int kind,
                       double fx, double fz, double fx2, double fz2, hipStream_t s,
                       int real = 0);
// Batched dense complex matrix-vector products, out[b, k] = sum_j M[k, j] in[b, j]
// (M row-major K x J); the input may be real, the output reduced to its real
// part.  Strides in elements.
template <typename T>
hipError_t launch_dense(const void* M, int K, int J, int batch, const void* in, int in_cplx,
                        long long in_sb, long long in_sj, void* out, int out_real,
                        long long out_sb, long long out_sk, hipStream_t s);
template <typename T>
hipError_t launch_update(void* var, void* hist, const void* lin, size_t n, int q, int curr,
                         int scheme, double c1, double c2, double h, hipStream_t s);
template <typename T>
hipError_t launch_fd(const void* in, void* out, int rows, int cols, int comp, int kind, int order,
                     int periodic, double h, hipStream_t s);
template <typename T>
hipError_t launch_flux(const void* ux, const void* uz, const void* f, void* out, int nx, int nz,
                       int order, int xper, int zper, double dx, double dz, hipStream_t s);
// Blocks of the first reduction stage at most (one (sum, compensation) pair
// of partials each); the partials buffer holds 2 * kRedBlocks doubles.
constexpr int kRedBlocks = 1024;
template <typename T>
hipError_t launch_reduce(const void* a, const void* b, size_t n, int op, double scale,
                         double* part, double* out, hipStream_t s);
size_t reduce_partials_bytes();
template <typename T>
hipError_t launch_combine(const void* hist, void* out, size_t n, int q, int curr, int moulton,
                          double h, hipStream_t s);
template <typename T>
hipError_t launch_velocity(const void* w, void* psi, void* ux, void* uz, int rows, int cols, int nn,
                           double fx, double fz, double fx2, double fz2, hipStream_t s);
template <typename T>
hipError_t launch_thomas(const void* rhs, void* out, void* dwork, const double* cp,
                         const double* rden, double lower, double sign, int nn, int nz,
                         hipStream_t s);

}  // namespace mlv
