// extern "C" entry points of libmelvin_hip.so (declared in
// include/melvin_hip.h): plans, operator-level calls, fused job batches.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <complex>
#include <cstdio>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/melvin_hip.h"
#include "capi_common.hpp"
#include "fused.hpp"
#include "ops.hpp"
#include "transforms.hpp"

using namespace mlv;

namespace mlv {
namespace abi {

thread_local std::string g_err = "";

int fail(int code, const std::string& msg) {
  g_err = msg;
  return code;
}

int hip_fail(hipError_t e, const char* what) {
  g_err = std::string(what) + ": " + hipGetErrorString(e);
  return MLV_ERR_HIP;
}


int ilog2_exact(int n) {
  if (n <= 0 || (n & (n - 1))) return -1;
  int l = 0;
  while ((1 << l) < n) ++l;
  return l;
}

// Reduction partials for calls without a plan: one buffer per device.
// Every plan or batch creation on a device allocates it (see get_twiddles),
// so a plan-less reduction allocates only when nothing was created on its
// device before.
std::mutex g_red_mu;
std::map<int, double*> g_red;

int global_partials(int device, double** out) {
  std::lock_guard<std::mutex> lk(g_red_mu);
  double*& p = g_red[device];
  if (!p) HIP_TRY(hipMalloc((void**)&p, reduce_partials_bytes()), "hipMalloc partials");
  *out = p;
  return MLV_OK;
}

// First-stage partials of mlv_reduce_noalloc: a file-scope device array of
// kRedBlocks (sum, compensation) pairs, 16 KiB on each device the library
// runs on, held until the process ends, so that the call never allocates.
// launch_reduce uses at most kRedBlocks blocks for any n, so the array fits
// every input size.  In host code `scratch` only names the symbol: the
// device address of each device's copy comes from hipGetSymbolAddress (once
// per device, with that device current), and only that address is passed
// to kernels.  Every call on a device shares its copy; the partials a
// reduction reads are all written by its own first stage, so nothing needs
// clearing between calls.
static __device__ __align__(256) double scratch[2 * kRedBlocks];
static_assert(sizeof(scratch) == 2 * kRedBlocks * sizeof(double), "partials size");
std::mutex g_scratch_mu;
std::map<int, double*> g_scratch;

int scratch_partials(int device, double** out) {
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  auto it = g_scratch.find(device);
  if (it == g_scratch.end()) {
    void* ptr = nullptr;
    HIP_TRY(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)), "hipGetSymbolAddress(scratch)");
    it = g_scratch.emplace(device, static_cast<double*>(ptr)).first;
  }
  *out = it->second;
  return MLV_OK;
}

// FP64-tabulated twiddles exp(-2 pi i k / NMAX), built once per device.
struct Twiddles {
  void* f32 = nullptr;
  void* f64 = nullptr;
};
std::mutex g_tw_mu;
std::map<int, Twiddles> g_tw;

// Per-device set-up shared by every create call: the twiddle tables and the
// plan-less reduction partials.
int get_twiddles(int device, int dtype, const void** out) {
  std::lock_guard<std::mutex> lk(g_tw_mu);
  Twiddles& t = g_tw[device];
  if (!t.f64) {
    double* part = nullptr;
    int rc = global_partials(device, &part);
    if (!rc) rc = scratch_partials(device, &part);
    if (rc) return rc;
    // main table exp(-2 pi i k / NMAX), then the per-pass tables of every
    // length (fft.hpp: fft_pass_table), all from long-double angles
    const int total = fft_twiddle_entries();
    std::vector<double> d(2 * (size_t)total);
    std::vector<float> f(2 * (size_t)total);
    const long double pi = 3.141592653589793238462643383279502884L;
    auto put = [&](int idx, long double num, long double den) {
      const long double a = 2.0L * pi * num / den;
      const long double c = cosl(a), sn = -sinl(a);
      d[2 * (size_t)idx] = (double)c;
      d[2 * (size_t)idx + 1] = (double)sn;
      f[2 * (size_t)idx] = (float)c;
      f[2 * (size_t)idx + 1] = (float)sn;
    };
    for (int k = 0; k < NMAX; ++k) put(k, k, NMAX);
    // tables keyed by (Ns, R) (fft_keyed_table)
    for (int lr = 1; lr <= 4; ++lr)
      for (int ln = 1; ln + lr <= LOG_NMAX; ++ln) {
        const int off = fft_keyed_table(ln, lr);
        const int NS = 1 << ln, R = 1 << lr;
        for (int r = 1; r < R; ++r)
          for (int jm = 0; jm < NS; ++jm)
            put(off + (r - 1) * NS + jm, (long double)((long long)jm * r), (long double)(NS * R));
      }
    for (int logn = 3; logn <= LOG_NMAX; ++logn) {
      int ns = 0;
      while (ns < logn) {
        const int lr = fft_pass_logr(logn, ns);
        if (ns > 0) {
          const int off = fft_pass_table(logn, ns);
          const int NS = 1 << ns, R = 1 << lr;
          for (int r = 1; r < R; ++r)
            for (int jm = 0; jm < NS; ++jm)
              put(off + (r - 1) * NS + jm, (long double)((long long)jm * r),
                  (long double)(NS * R));
        }
        ns += lr;
      }
    }
    HIP_TRY(hipMalloc(&t.f64, d.size() * sizeof(double)), "hipMalloc twiddles");
    HIP_TRY(hipMalloc(&t.f32, f.size() * sizeof(float)), "hipMalloc twiddles");
    HIP_TRY(hipMemcpy(t.f64, d.data(), d.size() * sizeof(double), hipMemcpyHostToDevice),
            "upload twiddles");
    HIP_TRY(hipMemcpy(t.f32, f.data(), f.size() * sizeof(float), hipMemcpyHostToDevice),
            "upload twiddles");
  }
  *out = dtype == MLV_F64 ? t.f64 : t.f32;
  return MLV_OK;
}

size_t max_lds_per_block(int device) {
  int v = 0;
  if (hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerBlock, device) != hipSuccess)
    return 65536;
  return (size_t)v;
}

size_t cplx_bytes(int dtype) { return dtype == MLV_F64 ? 16 : 8; }

int copy_tracker(const double* hdr, double* out, int max_records, int* nrec) {
  double h[kDiagHdr];
  HIP_TRY(hipMemcpy(h, hdr, sizeof(h), hipMemcpyDeviceToHost), "read diagnostics");
  const long long total = (long long)h[0];
  const long long kept = total < kDiagCap ? total : kDiagCap;
  const long long n = kept < max_records ? kept : max_records;
  if (n > 0 && out) {
    // oldest kept record first
    std::vector<double> all((size_t)kDiagRec * kDiagCap);
    HIP_TRY(hipMemcpy(all.data(), hdr + kDiagHdr, all.size() * sizeof(double),
                      hipMemcpyDeviceToHost),
            "read tracker");
    for (long long i = 0; i < n; ++i) {
      const long long r = (total - kept + i) % kDiagCap;
      std::memcpy(out + (size_t)kDiagRec * i, all.data() + (size_t)kDiagRec * r,
                  kDiagRec * sizeof(double));
    }
  }
  *nrec = (int)n;
  return MLV_OK;
}

}  // namespace abi
}  // namespace mlv

using namespace mlv::abi;

struct mlv_plan {
  int device;
  int nx, nz, lognx, lognz, nn, nm, rows, cols;
  int fully, dtype, fd_order;
  double lx, lz, dx, dz, fx, fz, fx2, fz2;
  const void* tw;
  void* work;     // 2-D: (nx, nm) complex intermediate; FD-z: Thomas d' (nn, nz) complex128
  double* cp;     // FD-z Thomas factors (nn, nz)
  double* rden;
  double lower;
  double* red_part;
  std::map<int, void*> basis_mats;  // dense per-axis transforms, built on first use
};

extern "C" {

const char* mlv_last_error(void) { return g_err.c_str(); }
int mlv_abi_version(void) { return 2; }

int mlv_plan_create(mlv_plan** out, int nx, int nz, double lx, double lz, int disc, int dtype,
                    int fd_order) {
  if (!out) return fail(MLV_ERR_ARG, "plan pointer is NULL");
  *out = nullptr;
  if (dtype != MLV_F32 && dtype != MLV_F64) return fail(MLV_ERR_ARG, "dtype must be F32 or F64");
  if (fd_order != 2 && fd_order != 4) return fail(MLV_ERR_ARG, "fd_order must be 2 or 4");
  if (disc != MLV_FULLY_SPECTRAL && disc != MLV_FOURIER_X_FD_Z)
    return fail(MLV_ERR_ARG, "unknown discretisation");
  if (!(lx > 0) || !(lz > 0)) return fail(MLV_ERR_ARG, "domain lengths must be positive");
  const int lnx = ilog2_exact(nx);
  if (lnx < 3 || lnx > 12) return fail(MLV_ERR_UNSUPPORTED, "nx must be a power of two in [8, 4096]");
  int lnz = -1;
  if (disc == MLV_FULLY_SPECTRAL) {
    lnz = ilog2_exact(nz);
    if (lnz < 3 || lnz > 12)
      return fail(MLV_ERR_UNSUPPORTED, "nz must be a power of two in [8, 4096]");
  } else if (nz < 5) {
    return fail(MLV_ERR_UNSUPPORTED, "FD-z needs nz >= 5");
  }
  mlv_plan* p = new mlv_plan();  // value-initialised: scalars and pointers zero
  HIP_TRY(hipGetDevice(&p->device), "hipGetDevice");
  p->nx = nx;
  p->nz = nz;
  p->lognx = lnx;
  p->lognz = lnz;
  p->fully = disc == MLV_FULLY_SPECTRAL;
  p->dtype = dtype;
  p->fd_order = fd_order;
  p->lx = lx;
  p->lz = lz;
  p->nn = (nx - 1) / 3;
  if (p->fully) {
    p->nm = (nz - 1) / 3;
    p->rows = 2 * p->nn + 1;
    p->cols = p->nm;
  } else {
    p->nm = nz;
    p->rows = p->nn;
    p->cols = nz;
  }
  // Parameters.py:81-82 and BasisFunctions.py:26-59, in the same operation order
  p->dx = lx / nx;
  p->dz = lz / nz;
  const double two_pi = 2.0 * M_PI;
  p->fx = two_pi / lx;
  p->fz = two_pi / lz;
  p->fx2 = -(two_pi * two_pi) / (lx * lx);
  p->fz2 = -(two_pi * two_pi) / (lz * lz);
  int rc = get_twiddles(p->device, dtype, &p->tw);
  if (rc) {
    delete p;
    return rc;
  }
  if (p->fully) {
    HIP_TRY(hipMalloc(&p->work, (size_t)nx * p->nm * cplx_bytes(dtype)), "hipMalloc workspace");
  } else {
    // Thomas factors of LaplacianSolver.py:21-52, built as a NumPy sweep would
    const int nn = p->nn;
    std::vector<double> cp((size_t)nn * nz), rden((size_t)nn * nz);
    const double inv_dz2 = 1.0 / (p->dz * p->dz);
    const double ddx = two_pi / lx;
    for (int n = 0; n < nn; ++n) {
      const double diag = -((n * ddx) * (n * ddx) + 2.0 / (p->dz * p->dz));
      double cprev = 0.0;
      for (int j = 0; j < nz; ++j) {
        const double dj = (j == 0 || j == nz - 1) ? 1.0 : diag;
        const double lj = (j == 0 || j == nz - 1) ? 0.0 : inv_dz2;
        const double uj = (j == 0 || j == nz - 1) ? 0.0 : inv_dz2;
        const double den = j == 0 ? dj : dj - lj * cprev;
        const double c = uj / den;
        cp[(size_t)n * nz + j] = c;
        rden[(size_t)n * nz + j] = 1.0 / den;
        cprev = c;
      }
    }
    p->lower = inv_dz2;
    HIP_TRY(hipMalloc(&p->work, (size_t)nn * nz * 16), "hipMalloc workspace");
    HIP_TRY(hipMalloc((void**)&p->cp, cp.size() * sizeof(double)), "hipMalloc thomas");
    HIP_TRY(hipMalloc((void**)&p->rden, rden.size() * sizeof(double)), "hipMalloc thomas");
    HIP_TRY(hipMemcpy(p->cp, cp.data(), cp.size() * sizeof(double), hipMemcpyHostToDevice),
            "upload thomas");
    HIP_TRY(hipMemcpy(p->rden, rden.data(), rden.size() * sizeof(double), hipMemcpyHostToDevice),
            "upload thomas");
  }
  HIP_TRY(hipMalloc((void**)&p->red_part, reduce_partials_bytes()), "hipMalloc partials");
  size_t lds_need = dtype == MLV_F64 ? fft_lds_bytes<double>(lnx) : fft_lds_bytes<float>(lnx);
  if (lnz > 0)
    lds_need = std::max(lds_need, dtype == MLV_F64 ? fft_lds_bytes_row<double>(lnz)
                                                   : fft_lds_bytes_row<float>(lnz));
  if (lds_need > max_lds_per_block(p->device)) {
    mlv_plan_destroy(p);
    return fail(MLV_ERR_UNSUPPORTED, "transform does not fit the LDS of this device");
  }
  *out = p;
  return MLV_OK;
}

int mlv_plan_destroy(mlv_plan* p) {
  if (!p) return MLV_OK;
  if (p->work) (void)hipFree(p->work);
  if (p->cp) (void)hipFree(p->cp);
  if (p->rden) (void)hipFree(p->rden);
  if (p->red_part) (void)hipFree(p->red_part);
  for (auto& kv : p->basis_mats) (void)hipFree(kv.second);
  delete p;
  return MLV_OK;
}

int mlv_plan_info(const mlv_plan* p, int* nn, int* nm, int* rows, int* cols) {
  if (!p) return fail(MLV_ERR_ARG, "plan is NULL");
  if (nn) *nn = p->nn;
  if (nm) *nm = p->nm;
  if (rows) *rows = p->rows;
  if (cols) *cols = p->cols;
  return MLV_OK;
}

#define CHECK_PLAN(p) \
  if (!(p)) return fail(MLV_ERR_ARG, "plan is NULL")
#define CHECK_PTR(x) \
  if (!(x)) return fail(MLV_ERR_ARG, #x " is NULL")
#define DISPATCH(p, CALL_F32, CALL_F64) ((p)->dtype == MLV_F64 ? (CALL_F64) : (CALL_F32))

int mlv_to_physical(mlv_plan* p, const void* spec, void* phys, void* stream) {
  CHECK_PLAN(p);
  CHECK_PTR(spec);
  CHECK_PTR(phys);
  hipStream_t s = (hipStream_t)stream;
  hipError_t e;
  if (p->fully) {
    e = DISPATCH(p, launch_xinv<float>(p->lognx, spec, p->work, p->nn, p->nm, p->tw, s),
                 launch_xinv<double>(p->lognx, spec, p->work, p->nn, p->nm, p->tw, s));
    if (e != hipSuccess) return hip_fail(e, "to_physical x pass");
    e = DISPATCH(p, launch_zc2r<float>(p->lognz, p->work, phys, p->nx, p->nm, p->tw, s),
                 launch_zc2r<double>(p->lognz, p->work, phys, p->nx, p->nm, p->tw, s));
    if (e != hipSuccess) return hip_fail(e, "to_physical z pass");
  } else {
    e = DISPATCH(p, launch_xc2r_cols<float>(p->lognx, spec, phys, p->nn, p->nz, p->tw, s),
                 launch_xc2r_cols<double>(p->lognx, spec, phys, p->nn, p->nz, p->tw, s));
    if (e != hipSuccess) return hip_fail(e, "to_physical x pass");
  }
  return MLV_OK;
}

int mlv_to_spectral(mlv_plan* p, const void* phys, void* spec, void* stream) {
  CHECK_PLAN(p);
  CHECK_PTR(spec);
  CHECK_PTR(phys);
  hipStream_t s = (hipStream_t)stream;
  hipError_t e;
  if (p->fully) {
    const double scale = 1.0 / ((double)p->nx * (double)p->nz);
    e = DISPATCH(p, launch_zr2c<float>(p->lognz, phys, p->work, p->nx, p->nm, p->tw, s),
                 launch_zr2c<double>(p->lognz, phys, p->work, p->nx, p->nm, p->tw, s));
    if (e != hipSuccess) return hip_fail(e, "to_spectral z pass");
    e = DISPATCH(p, launch_xfwd<float>(p->lognx, p->work, spec, p->nn, p->nm, scale, p->tw, s),
                 launch_xfwd<double>(p->lognx, p->work, spec, p->nn, p->nm, scale, p->tw, s));
    if (e != hipSuccess) return hip_fail(e, "to_spectral x pass");
  } else {
    const double scale = 1.0 / (double)p->nx;
    e = DISPATCH(p,
                 launch_xr2c_cols<float>(p->lognx, phys, spec, p->nn, p->nz, scale, p->tw, s),
                 launch_xr2c_cols<double>(p->lognx, phys, spec, p->nn, p->nz, scale, p->tw, s));
    if (e != hipSuccess) return hip_fail(e, "to_spectral x pass");
  }
  return MLV_OK;
}

// numpy.fft's real transforms on a row-major (nx, nz) array (the xp.fft
// calls of SpectralTransformer.py:58,85,132,191): the transformed axes'
// lengths must be powers of two from 8 to 4096 (the plans' range).
static int fft_axis_log(int n, const char* what, int* logn) {
  int l = 0;
  while (l < 30 && (1 << l) < n) ++l;
  if (n < 8 || (1 << l) != n || l > 12)
    return fail(MLV_ERR_UNSUPPORTED, std::string(what) + " must be a power of two in [8, 4096], got " +
                                         std::to_string(n));
  *logn = l;
  return MLV_OK;
}

static int fft_setup(int dtype, int axes, int nx, int nz, int* lognx, int* lognz,
                     const void** tw) {
  if (dtype != MLV_F32 && dtype != MLV_F64) return fail(MLV_ERR_ARG, "bad dtype");
  if (axes < MLV_FFT_AXIS0 || axes > MLV_FFT_AXES01) return fail(MLV_ERR_ARG, "bad axes");
  if (nx < 1 || nz < 1) return fail(MLV_ERR_ARG, "empty array");
  *lognx = *lognz = 0;
  int rc = MLV_OK;
  if (axes != MLV_FFT_AXIS1) rc = fft_axis_log(nx, "nx", lognx);
  if (!rc && axes != MLV_FFT_AXIS0) rc = fft_axis_log(nz, "nz", lognz);
  if (rc) return rc;
  int device = 0;
  HIP_TRY(hipGetDevice(&device), "hipGetDevice");
  const bool f64 = dtype == MLV_F64;
  size_t lds = 0;
  if (*lognx) lds = f64 ? fft_lds_bytes<double>(*lognx) : fft_lds_bytes<float>(*lognx);
  if (*lognz)
    lds = std::max(lds, f64 ? fft_lds_bytes_row<double>(*lognz) : fft_lds_bytes_row<float>(*lognz));
  if (lds > max_lds_per_block(device)) return fail(MLV_ERR_UNSUPPORTED, "transform too long for LDS");
  return get_twiddles(device, dtype, tw);
}

int mlv_rfft(int dtype, int axes, int nx, int nz, const void* in, void* out, void* work,
             void* stream) {
  CHECK_PTR(in);
  CHECK_PTR(out);
  int lx, lz;
  const void* tw = nullptr;
  int rc = fft_setup(dtype, axes, nx, nz, &lx, &lz, &tw);
  if (rc) return rc;
  hipStream_t s = (hipStream_t)stream;
  const bool f64 = dtype == MLV_F64;
  hipError_t e = hipSuccess;
  if (axes == MLV_FFT_AXIS0) {
    e = f64 ? launch_xr2c_cols<double>(lx, in, out, nx / 2 + 1, nz, 1.0, tw, s)
            : launch_xr2c_cols<float>(lx, in, out, nx / 2 + 1, nz, 1.0, tw, s);
  } else if (axes == MLV_FFT_AXIS1) {
    e = f64 ? launch_zr2c<double>(lz, in, out, nx, nz / 2 + 1, tw, s)
            : launch_zr2c<float>(lz, in, out, nx, nz / 2 + 1, tw, s);
  } else {
    CHECK_PTR(work);
    const int nm = nz / 2 + 1;
    e = f64 ? launch_zr2c<double>(lz, in, work, nx, nm, tw, s)
            : launch_zr2c<float>(lz, in, work, nx, nm, tw, s);
    if (e == hipSuccess)
      e = f64 ? launch_xfwd<double>(lx, work, out, nx / 2, nm, 1.0, tw, s, nx)
              : launch_xfwd<float>(lx, work, out, nx / 2, nm, 1.0, tw, s, nx);
  }
  if (e != hipSuccess) return hip_fail(e, "rfft");
  return MLV_OK;
}

int mlv_irfft(int dtype, int axes, int nx, int nz, const void* in, void* out, void* work,
              void* stream) {
  CHECK_PTR(in);
  CHECK_PTR(out);
  int lx, lz;
  const void* tw = nullptr;
  int rc = fft_setup(dtype, axes, nx, nz, &lx, &lz, &tw);
  if (rc) return rc;
  hipStream_t s = (hipStream_t)stream;
  const bool f64 = dtype == MLV_F64;
  hipError_t e = hipSuccess;
  if (axes == MLV_FFT_AXIS0) {
    const double sc = 1.0 / nx;
    e = f64 ? launch_xc2r_cols<double>(lx, in, out, nx / 2 + 1, nz, tw, s, sc)
            : launch_xc2r_cols<float>(lx, in, out, nx / 2 + 1, nz, tw, s, sc);
  } else if (axes == MLV_FFT_AXIS1) {
    const double sc = 1.0 / nz;
    e = f64 ? launch_zc2r<double>(lz, in, out, nx, nz / 2 + 1, tw, s, sc)
            : launch_zc2r<float>(lz, in, out, nx, nz / 2 + 1, tw, s, sc);
  } else {
    CHECK_PTR(work);
    const int nm = nz / 2 + 1;
    const double sc = 1.0 / ((double)nx * (double)nz);
    e = f64 ? launch_xinv<double>(lx, in, work, nx / 2, nm, tw, s, nx, sc)
            : launch_xinv<float>(lx, in, work, nx / 2, nm, tw, s, nx, sc);
    if (e == hipSuccess)
      e = f64 ? launch_zc2r<double>(lz, work, out, nx, nm, tw, s)
              : launch_zc2r<float>(lz, work, out, nx, nm, tw, s);
  }
  if (e != hipSuccess) return hip_fail(e, "irfft");
  return MLV_OK;
}

int mlv_spec_op(mlv_plan* p, int kind, const void* in, void* out, void* stream) {
  CHECK_PLAN(p);
  CHECK_PTR(out);
  if (kind != MLV_SPEC_LAP) CHECK_PTR(in);
  hipStream_t s = (hipStream_t)stream;
  if (kind < MLV_SPEC_DDX || kind > MLV_SPEC_POISSON) return fail(MLV_ERR_ARG, "unknown spectral op");
  if (!p->fully) {
    // FD-z: z derivatives are finite differences on the complex rows
    // (SpatialDifferentiator.py:35-40); the FD-z Laplacian has no symbol.
    if (kind == MLV_SPEC_DDZ || kind == MLV_SPEC_POISSON)
      return fail(MLV_ERR_UNSUPPORTED,
                  "operator not defined for Fourier-x/FD-z (use mlv_fd_op / mlv_poisson)");
    if (kind == MLV_SPEC_LAP) {
      // the FDM basis has a zero second-derivative factor (BasisFunctions.py:40-43)
      hipError_t e = DISPATCH(p,
                              launch_spec<float>(in, out, p->rows, p->cols, p->nn, 0, kind, p->fx,
                                                 p->fz, p->fx2, 0.0, s),
                              launch_spec<double>(in, out, p->rows, p->cols, p->nn, 0, kind, p->fx,
                                                  p->fz, p->fx2, 0.0, s));
      if (e != hipSuccess) return hip_fail(e, "lap");
      return MLV_OK;
    }
    if (kind == MLV_SPEC_D2DZ2)
      return mlv_fd_op(p, MLV_FD_D2DZ2, 1, p->rows, p->cols, in, out, stream);
    if (kind == MLV_SPEC_SNABLA2) {
      // sd2dx2 + FD d2/dz2: out = d2z(in); then out += d2x(in)
      int rc = mlv_fd_op(p, MLV_FD_D2DZ2, 1, p->rows, p->cols, in, out, stream);
      if (rc) return rc;
      hipError_t e =
          DISPATCH(p,
                   launch_spec<float>(in, out, p->rows, p->cols, p->nn, 0, kSpecAddD2dx2, p->fx,
                                      p->fz, p->fx2, p->fz2, s),
                   launch_spec<double>(in, out, p->rows, p->cols, p->nn, 0, kSpecAddD2dx2, p->fx,
                                       p->fz, p->fx2, p->fz2, s));
      if (e != hipSuccess) return hip_fail(e, "snabla2");
      return MLV_OK;
    }
  }
  hipError_t e = DISPATCH(p,
                          launch_spec<float>(in, out, p->rows, p->cols, p->nn, p->fully, kind,
                                             p->fx, p->fz, p->fx2, p->fz2, s),
                          launch_spec<double>(in, out, p->rows, p->cols, p->nn, p->fully, kind,
                                              p->fx, p->fz, p->fx2, p->fz2, s));
  if (e != hipSuccess) return hip_fail(e, "spectral op");
  return MLV_OK;
}

int mlv_fd_op(mlv_plan* p, int kind, int is_complex, int rows, int cols, const void* in, void* out,
              void* stream) {
  CHECK_PLAN(p);
  CHECK_PTR(in);
  CHECK_PTR(out);
  if (rows <= 0 || cols <= 0) return fail(MLV_ERR_ARG, "bad array shape");
  if (kind < MLV_FD_DDX || kind > MLV_FD_D2DZ2) return fail(MLV_ERR_ARG, "unknown FD op");
  // the stencils wrap at most once, so the differentiated axis needs 5 points
  if ((kind == MLV_FD_DDX ? rows : cols) < 5)
    return fail(MLV_ERR_ARG, "FD axis shorter than 5 points");
  int periodic = 0;
  double h = p->dz;
  if (kind == MLV_FD_DDX) {
    periodic = 1;  // x is spectral in every supported discretisation
    h = p->dx;
  } else if (kind == MLV_FD_DDZ) {
    periodic = p->fully;
  }
  hipError_t e = DISPATCH(
      p,
      launch_fd<float>(in, out, rows, cols, is_complex ? 2 : 1, kind, p->fd_order, periodic, h,
                       (hipStream_t)stream),
      launch_fd<double>(in, out, rows, cols, is_complex ? 2 : 1, kind, p->fd_order, periodic, h,
                        (hipStream_t)stream));
  if (e != hipSuccess) return hip_fail(e, "fd op");
  return MLV_OK;
}

int mlv_flux_divergence(mlv_plan* p, const void* ux, const void* uz, const void* f, void* out,
                        void* stream) {
  CHECK_PLAN(p);
  CHECK_PTR(ux);
  CHECK_PTR(uz);
  CHECK_PTR(f);
  CHECK_PTR(out);
  hipError_t e = DISPATCH(
      p,
      launch_flux<float>(ux, uz, f, out, p->nx, p->nz, p->fd_order, 1, p->fully, p->dx, p->dz,
                         (hipStream_t)stream),
      launch_flux<double>(ux, uz, f, out, p->nx, p->nz, p->fd_order, 1, p->fully, p->dx, p->dz,
                          (hipStream_t)stream));
  if (e != hipSuccess) return hip_fail(e, "flux divergence");
  return MLV_OK;
}

int mlv_poisson(mlv_plan* p, const void* rhs, void* out, void* stream) {
  CHECK_PLAN(p);
  CHECK_PTR(rhs);
  CHECK_PTR(out);
  if (p->fully) return mlv_spec_op(p, MLV_SPEC_POISSON, rhs, out, stream);
  hipError_t e = DISPATCH(p,
                          launch_thomas<float>(rhs, out, p->work, p->cp, p->rden, p->lower, 1.0,
                                               p->nn, p->nz, (hipStream_t)stream),
                          launch_thomas<double>(rhs, out, p->work, p->cp, p->rden, p->lower, 1.0,
                                                p->nn, p->nz, (hipStream_t)stream));
  if (e != hipSuccess) return hip_fail(e, "poisson");
  return MLV_OK;
}

int mlv_update(mlv_plan* p, int scheme, int order, double alpha, double dt, const void* lin,
               void* var, void* hist, int curr, void* stream) {
  CHECK_PLAN(p);
  CHECK_PTR(lin);
  CHECK_PTR(var);
  CHECK_PTR(hist);
  if (order != 2 && order != 4) return fail(MLV_ERR_ARG, "integrator order must be 2 or 4");
  if (curr < 0 || curr >= order) return fail(MLV_ERR_ARG, "curr out of range");
  if (scheme != MLV_SCHEME_SEMI_IMPLICIT && scheme != MLV_SCHEME_EXPLICIT)
    return fail(MLV_ERR_ARG, "unknown scheme");
  const size_t n = (size_t)p->rows * p->cols;
  const double c1 = (1 - alpha) * dt, c2 = alpha * dt;
  const double h = order == 2 ? dt / 2 : dt / 24;
  hipError_t e = DISPATCH(
      p, launch_update<float>(var, hist, lin, n, order, curr, scheme, c1, c2, h, (hipStream_t)stream),
      launch_update<double>(var, hist, lin, n, order, curr, scheme, c1, c2, h, (hipStream_t)stream));
  if (e != hipSuccess) return hip_fail(e, "update");
  return MLV_OK;
}

int mlv_reduce(mlv_plan* p, int op, int dtype, const void* a, const void* b, size_t n,
               double scale, void* out, void* stream) {
  CHECK_PTR(a);
  CHECK_PTR(out);
  if (op < MLV_RED_MAX || op > MLV_RED_DOT) return fail(MLV_ERR_ARG, "unknown reduction");
  if ((op == MLV_RED_SUMSQ2 || op == MLV_RED_DOT) && !b) return fail(MLV_ERR_ARG, "b is NULL");
  if (n == 0) return fail(MLV_ERR_ARG, "empty reduction");
  double* part = p ? p->red_part : nullptr;
  if (!part) {
    int dev = 0;
    HIP_TRY(hipGetDevice(&dev), "hipGetDevice");
    int rc = global_partials(dev, &part);
    if (rc) return rc;
  }
  hipError_t e =
      dtype == MLV_F64
          ? launch_reduce<double>(a, b, n, op, scale, part, (double*)out, (hipStream_t)stream)
          : launch_reduce<float>(a, b, n, op, scale, part, (double*)out, (hipStream_t)stream);
  if (e != hipSuccess) return hip_fail(e, "reduce");
  return MLV_OK;
}

int mlv_reduce_noalloc(int op, int dtype, const void* a, const void* b, size_t n, double scale,
                       void* out, void* stream) {
  CHECK_PTR(a);
  CHECK_PTR(out);
  if (op < MLV_RED_MAX || op > MLV_RED_DOT) return fail(MLV_ERR_ARG, "unknown reduction");
  if ((op == MLV_RED_SUMSQ2 || op == MLV_RED_DOT) && !b) return fail(MLV_ERR_ARG, "b is NULL");
  if (dtype != MLV_F32 && dtype != MLV_F64) return fail(MLV_ERR_ARG, "dtype must be F32 or F64");
  if (n == 0) return fail(MLV_ERR_ARG, "empty reduction");
  // the first stage runs min(ceil(n / 256), kRedBlocks) blocks, one pair of
  // partials each: never more than the scratch array holds
  int dev = 0;
  HIP_TRY(hipGetDevice(&dev), "hipGetDevice");
  double* part = nullptr;
  const int rc = scratch_partials(dev, &part);
  if (rc) return rc;
  hipError_t e =
      dtype == MLV_F64
          ? launch_reduce<double>(a, b, n, op, scale, part, (double*)out, (hipStream_t)stream)
          : launch_reduce<float>(a, b, n, op, scale, part, (double*)out, (hipStream_t)stream);
  if (e != hipSuccess) return hip_fail(e, "reduce");
  return MLV_OK;
}

int mlv_combine(mlv_plan* p, int moulton, int order, double dt, const void* hist, int curr,
                void* out, void* stream) {
  CHECK_PLAN(p);
  CHECK_PTR(hist);
  CHECK_PTR(out);
  if (order != 2 && order != 4) return fail(MLV_ERR_ARG, "integrator order must be 2 or 4");
  if (curr < 0 || curr >= order) return fail(MLV_ERR_ARG, "curr out of range");
  const size_t n = (size_t)p->rows * p->cols;
  const double h = order == 2 ? dt / 2 : dt / 24;
  hipError_t e = DISPATCH(
      p, launch_combine<float>(hist, out, n, order, curr, moulton ? 1 : 0, h, (hipStream_t)stream),
      launch_combine<double>(hist, out, n, order, curr, moulton ? 1 : 0, h, (hipStream_t)stream));
  if (e != hipSuccess) return hip_fail(e, "combine");
  return MLV_OK;
}

int mlv_velocity(mlv_plan* p, const void* w, void* psi, void* ux_spec, void* uz_spec,
                 void* stream) {
  CHECK_PLAN(p);
  CHECK_PTR(w);
  CHECK_PTR(psi);
  hipStream_t s = (hipStream_t)stream;
  hipError_t e;
  if (p->fully) {
    e = DISPATCH(p,
                 launch_velocity<float>(w, psi, ux_spec, uz_spec, p->rows, p->cols, p->nn, p->fx,
                                        p->fz, p->fx2, p->fz2, s),
                 launch_velocity<double>(w, psi, ux_spec, uz_spec, p->rows, p->cols, p->nn, p->fx,
                                         p->fz, p->fx2, p->fz2, s));
    if (e != hipSuccess) return hip_fail(e, "velocity");
    return MLV_OK;
  }
  if (ux_spec) return fail(MLV_ERR_ARG, "FD-z: ux comes from the physical z-derivative of psi");
  e = DISPATCH(p,
               launch_thomas<float>(w, psi, p->work, p->cp, p->rden, p->lower, -1.0, p->nn, p->nz, s),
               launch_thomas<double>(w, psi, p->work, p->cp, p->rden, p->lower, -1.0, p->nn, p->nz, s));
  if (e != hipSuccess) return hip_fail(e, "velocity poisson");
  if (uz_spec) return mlv_spec_op(p, MLV_SPEC_DDX, psi, uz_spec, stream);
  return MLV_OK;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Batches of fused jobs
// ---------------------------------------------------------------------------
namespace {

struct BatchHost {
  int dtype = 0, njobs = 0, device = 0, nstage = 1;
  void* jobs_dev = nullptr;
  BatchLaunch bl{};
  long long steps = 0;
  const void* tw = nullptr;
  // optional kernel timing: 4 events per stage and step (before A, B, C, after C)
  // captured sequence of graph_steps steps (see mlv_batch_graph_step)
  hipStream_t cap_stream = nullptr;
  hipGraph_t graph = nullptr;
  hipGraphExec_t graph_exec = nullptr;
  int graph_steps = 0;
  int graph_cadence = 0;
  double graph_cutoff = 0;
  bool prof = false;
  std::vector<hipEvent_t> events;
  size_t used = 0;
  double ms[3] = {0, 0, 0};
  long long launches[3] = {0, 0, 0};
  std::vector<char> host;          // host copy of the job table
  std::vector<double*> diag;       // per job diagnostics buffer (in diag_buf)
  std::vector<int> nbB;
  std::vector<long long> loops0;   // per job loop count at creation
  std::string signature;           // the job set's shapes (batch_signature)
  // device memory: the job table and every prefix array in one allocation,
  // the jobs' diagnostics in one of two buffers (a rebuild copies the
  // records of the jobs that stay into the other); each is reused while a
  // rebuilt job set fits it
  void* tables = nullptr;
  size_t tables_cap = 0;
  double* diag_buf[2] = {nullptr, nullptr};
  size_t diag_cap[2] = {0, 0};
  int diag_cur = 0;
  std::vector<const void*> keys;   // a job's identity across rebuilds: its workspace
};

// Kernel choices picked by mlv_batch_autotune, by job-set signature; a
// batch created later for the same set of shapes starts with them.
struct TunedPlan {
  int shapeA, planB, wideC;
};
std::mutex g_tuned_mu;
std::map<std::string, TunedPlan> g_tuned;

// Precision and the sorted (problem, fields, nx, nz, order) of every job:
// what the kernels' block grids depend on.
std::string batch_signature(int dtype, const mlv_job_desc* jobs, int njobs) {
  std::vector<std::string> parts;
  for (int j = 0; j < njobs; ++j) {
    char buf[64];
    std::snprintf(buf, sizeof(buf), "%d:%d:%d:%d:%d", jobs[j].problem, jobs[j].nfield, jobs[j].nx,
                  jobs[j].nz, jobs[j].ab_order);
    parts.push_back(buf);
  }
  std::sort(parts.begin(), parts.end());
  std::string s = std::to_string(dtype);
  for (const auto& p : parts) s += "|" + p;
  return s;
}

hipEvent_t* prof_slots(BatchHost* b) {
  if (!b->prof) return nullptr;
  while (b->events.size() < b->used + 4) {
    hipEvent_t e;
    if (hipEventCreate(&e) != hipSuccess) return nullptr;
    b->events.push_back(e);
  }
  hipEvent_t* r = &b->events[b->used];
  b->used += 4;
  return r;
}

int prof_collect(BatchHost* b) {
  if (b->used == 0) return MLV_OK;
  HIP_TRY(hipEventSynchronize(b->events[b->used - 1]), "hipEventSynchronize");
  for (size_t i = 0; i + 3 < b->used; i += 4) {
    for (int k = 0; k < 3; ++k) {
      float t = 0.f;
      HIP_TRY(hipEventElapsedTime(&t, b->events[i + k], b->events[i + k + 1]),
              "hipEventElapsedTime");
      b->ms[k] += t;
      b->launches[k] += 1;
    }
  }
  b->used = 0;
  return MLV_OK;
}

int problem_fields(int prob) {
  switch (prob) {
    case MLV_PROB_VORTICITY:
      return -1;  // 1..3, given by nfield
    case MLV_PROB_VORT_SCALAR:
      return 2;
    case MLV_PROB_DDC:
      return 3;
    case MLV_PROB_MHD:
      return 2;
    default:
      return 0;
  }
}

template <typename T>
int build_job(const mlv_job_desc& d, JobDev<T>& J, int j) {
  char msg[160];
  auto bad = [&](const char* what) {
    std::snprintf(msg, sizeof(msg), "job %d: %s", j, what);
    return fail(MLV_ERR_ARG, msg);
  };
  std::memset(&J, 0, sizeof(J));
  const int lnx = ilog2_exact(d.nx), lnz = ilog2_exact(d.nz);
  if (lnx < 3 || lnx > 12 || lnz < 3 || lnz > 12) {
    std::snprintf(msg, sizeof(msg), "job %d: nx, nz must be powers of two in [8, 4096]", j);
    return fail(MLV_ERR_UNSUPPORTED, msg);
  }
  int nf = problem_fields(d.problem);
  if (nf == 0) return bad("unknown problem");
  if (nf < 0) {
    if (d.nfield < 1 || d.nfield > 3) return bad("vorticity problem needs 1..3 fields");
    nf = d.nfield;
  } else if (d.nfield != nf) {
    return bad("nfield does not match the problem");
  }
  if (d.ab_order != 2 && d.ab_order != 4) return bad("ab_order must be 2 or 4");
  if (d.fd_order != 2 && d.fd_order != 4) return bad("fd_order must be 2 or 4");
  if (!(d.lx > 0) || !(d.lz > 0) || !(d.dt > 0)) return bad("lengths and dt must be positive");
  if (d.curr < 0 || d.curr >= d.ab_order || (d.parity & ~1)) return bad("curr/parity out of range");
  if (!d.work) return bad("work is NULL");
  // the z pass loads and the x-inverse stores row pairs as 16-byte vectors
  if (reinterpret_cast<uintptr_t>(d.work) % 16) return bad("work must be 16-byte aligned");
  for (int f = 0; f < nf; ++f)
    if (!d.state[0][f] || !d.state[1][f] || !d.hist[f]) return bad("state/hist pointer is NULL");
  J.problem = d.problem;
  J.nx = d.nx;
  J.nz = d.nz;
  J.lognx = lnx;
  J.lognz = lnz;
  J.nn = (d.nx - 1) / 3;
  J.nm = (d.nz - 1) / 3;
  J.rows = 2 * J.nn + 1;
  J.q = d.ab_order;
  J.curr0 = d.curr;
  J.par0 = d.parity;
  J.fd_order = d.fd_order;
  const double two_pi = 2.0 * M_PI;
  J.fx = two_pi / d.lx;
  J.fz = two_pi / d.lz;
  J.fx2 = -(two_pi * two_pi) / (d.lx * d.lx);
  J.fz2 = -(two_pi * two_pi) / (d.lz * d.lz);
  J.dx = d.lx / d.nx;
  J.dz = d.lz / d.nz;
  J.scale = 1.0 / ((double)d.nx * (double)d.nz);
  J.alpha = d.alpha;
  J.dt = d.dt;
  J.nfield = nf;
  for (int i = 0; i < 4; ++i) {
    J.coeff[i] = d.coeff[i];
    J.param[i] = d.param[i];
  }
  for (int b = 0; b < 2; ++b)
    for (int f = 0; f < nf; ++f) J.state[b][f] = (cx<T>*)d.state[b][f];
  for (int f = 0; f < nf; ++f) J.hist[f] = (cx<T>*)d.hist[f];
  for (int f = 0; f < nf; ++f) J.comp[f] = sizeof(T) == 4 ? (uint32_t*)d.comp[f] : nullptr;
  J.work = (cx<T>*)d.work;
  J.psi = (cx<T>*)d.psi;
  J.t = d.t;
  J.loops = d.loops;
  J.split = (sizeof(T) == 4 && d.mean_split) ? 1 : 0;
  for (int i = 0; i < kMaxPhys; ++i) J.phys[i] = (T*)d.phys[i];
  for (int s = 0; s < 2; ++s)
    for (int i = 0; i < kMaxPhys; ++i) J.st[s].phys_ch[i] = -1;
  auto addA = [&](StageDev& S, int src, int op, int ch, int from_new) {
    S.a[S.nA++] = XinvOp{src, op, ch, from_new};
  };
  if (d.problem == MLV_PROB_MHD) {
    J.nstage = 2;
    StageDev& s0 = J.st[0];
    addA(s0, 0, XOP_ID, 0, 0);
    addA(s0, 1, XOP_ID, 1, 0);
    addA(s0, 0, XOP_UX, 2, 0);
    addA(s0, 0, XOP_UZ, 3, 0);
    addA(s0, 1, XOP_UX, 4, 0);
    addA(s0, 1, XOP_UZ, 5, 0);
    s0.bvar = BV_MHD1;
    s0.neq = 1;
    s0.eq_field[0] = 0;
    for (int i = 0; i < 6; ++i)
      if (J.phys[i]) s0.phys_ch[i] = i;
    StageDev& s1 = J.st[1];
    addA(s1, 0, XOP_ID, 0, 1);
    s1.bvar = BV_MHD2;
    s1.neq = 1;
    s1.eq_field[0] = 1;
  } else {
    J.nstage = 1;
    StageDev& s0 = J.st[0];
    for (int f = 0; f < nf; ++f) addA(s0, f, XOP_ID, f, 0);
    addA(s0, 0, XOP_UX, nf, 0);
    addA(s0, 0, XOP_UZ, nf + 1, 0);
    s0.bvar = BV_VORT1 + (nf - 1);
    s0.neq = nf;
    for (int f = 0; f < nf; ++f) s0.eq_field[f] = f;
    for (int i = 0; i < nf + 2; ++i)
      if (J.phys[i]) s0.phys_ch[i] = i;
  }
  J.pbase = bvar_nin(J.st[0].bvar);
  {
    const int nin = bvar_nin(J.st[0].bvar), neq = bvar_neq(J.st[0].bvar);
    J.mean = (T*)(J.work + (fused_work_elems(nin, neq, J.nx, J.nz) - (size_t)(kMeanCh / 2) * J.nz));
  }
  return MLV_OK;
}

}  // namespace

extern "C" {

int mlv_job_work_bytes(int problem, int nfield, int nx, int nz, int dtype, size_t* bytes) {
  if (!bytes) return fail(MLV_ERR_ARG, "bytes is NULL");
  if (nx < 8 || nz < 8) return fail(MLV_ERR_ARG, "grid too small");
  int nf = problem_fields(problem);
  if (nf == 0) return fail(MLV_ERR_ARG, "unknown problem");
  if (nf < 0) {
    if (nfield < 1 || nfield > 3) return fail(MLV_ERR_ARG, "vorticity problem needs 1..3 fields");
    nf = nfield;
  }
  const int bv = problem == MLV_PROB_MHD ? BV_MHD1 : BV_VORT1 + (nf - 1);
  *bytes = fused_work_elems(bvar_nin(bv), bvar_neq(bv), nx, nz) * cplx_bytes(dtype);
  return MLV_OK;
}

// Host plan of a batch: the device job table and every launch table of
// the job set, ready for one upload (mlv_batch_create, mlv_batch_rebuild).
struct BatchPlan {
  std::vector<char> host;  // JobDev table
  std::vector<int> pA[3][2], pB[2][BV_COUNT], pC[2], pCn[2], pCm[2];  // pA[shape][stage]
  BatchLaunch bl{};        // block counts and LDS sizes; pointers set by upload_batch
  size_t ldsCn = 0;
  std::vector<int> nbB;
  std::vector<long long> loops0;
  std::string signature;
  int nstage = 1;
};

static int plan_batch(int dtype, const mlv_job_desc* jobs, int njobs, int device, BatchPlan& P) {
  int rc = MLV_OK;
  const size_t lds_max = max_lds_per_block(device);
  const size_t jsz = dtype == MLV_F64 ? sizeof(JobDev<double>) : sizeof(JobDev<float>);
  std::vector<char>& host = P.host;
  host.assign(jsz * njobs, 0);
  auto& pA = P.pA;
  auto& pB = P.pB;
  auto& pC = P.pC;
  auto& pCn = P.pCn;
  size_t& ldsCn = P.ldsCn;
  for (int s = 0; s < 2; ++s) {
    for (int a = 0; a < 3; ++a) pA[a][s].assign(njobs + 1, 0);
    pC[s].assign(njobs + 1, 0);
    pCn[s].assign(njobs + 1, 0);
    P.pCm[s].assign(njobs + 1, 0);
    for (int v = 0; v < BV_COUNT; ++v) pB[s][v].assign(njobs + 1, 0);
  }
  BatchLaunch& bl = P.bl;
  std::memset(&bl, 0, sizeof(bl));
  bl.njobs = njobs;
  bl.okA[0] = bl.okA[1] = bl.okA[2] = 1;
  bl.okC[0] = bl.okC[1] = bl.okC[2] = 1;
  size_t ldsCm = 0;
  bl.lognx_all = bl.lognz_all = 0;
  for (int j = 0; j < njobs; ++j) {
    int nstage, lognx, lognz, nm, nx;
    int nA[2], neq[2], bvar[2];
    if (dtype == MLV_F64) {
      JobDev<double>& J = reinterpret_cast<JobDev<double>*>(host.data())[j];
      rc = build_job<double>(jobs[j], J, j);
      if (rc) break;
      nstage = J.nstage;
      lognx = J.lognx;
      lognz = J.lognz;
      nm = J.nm;
      nx = J.nx;
      for (int s = 0; s < nstage; ++s) {
        nA[s] = J.st[s].nA;
        neq[s] = J.st[s].neq;
        bvar[s] = J.st[s].bvar;
      }
    } else {
      JobDev<float>& J = reinterpret_cast<JobDev<float>*>(host.data())[j];
      rc = build_job<float>(jobs[j], J, j);
      if (rc) break;
      nstage = J.nstage;
      lognx = J.lognx;
      lognz = J.lognz;
      nm = J.nm;
      nx = J.nx;
      for (int s = 0; s < nstage; ++s) {
        nA[s] = J.st[s].nA;
        neq[s] = J.st[s].neq;
        bvar[s] = J.st[s].bvar;
      }
    }
    const bool f64 = dtype == MLV_F64;
    bl.lognx_all = (j == 0 || bl.lognx_all == lognx) ? lognx : -1;
    bl.lognz_all = (j == 0 || bl.lognz_all == lognz) ? lognz : -1;
    // x passes: narrow (n) and wide (w) points policies, see fused_points
    const int sla = f64 ? xinv_cols_host<double>(lognx) : xinv_cols_host<float>(lognx);
    const size_t ldsa = f64 ? xinv_lds_bytes<double>(lognx) : xinv_lds_bytes<float>(lognx);
    // kernel A's smaller block shape, where the job's transform fits it
    const int sla1 = f64 ? xinv_cols_host<double>(lognx, 1) : xinv_cols_host<float>(lognx, 1);
    const size_t ldsa1 = f64 ? xinv_lds_bytes<double>(lognx, 1) : xinv_lds_bytes<float>(lognx, 1);
    if (sla1 < 1) bl.okA[1] = 0;
    // kernel A's mixed tiles (stepA_mixed): every stage's ops, kMixCols
    // columns of each, in shape 0's block
    {
      int max_nA = 0;
      for (int s = 0; s < nstage; ++s) max_nA = std::max(max_nA, nA[s]);
      if (f64 || !xmix_fits<float>(lognx, max_nA)) bl.okA[2] = 0;
    }
    const int sln = f64 ? fused_cols_host<double>(lognx, 0) : fused_cols_host<float>(lognx, 0);
    const int slw = f64 ? fused_cols_host<double>(lognx, 1) : fused_cols_host<float>(lognx, 1);
    const size_t ldsn = f64 ? fused_lds_bytes<double>(lognx, 0) : fused_lds_bytes<float>(lognx, 0);
    const size_t ldsw = f64 ? fused_lds_bytes<double>(lognx, 1) : fused_lds_bytes<float>(lognx, 1);
    // kernel C's 512-thread narrow shape (cshape_threads): single precision
    const int slm = (cshape_threads(2) * 16) >> lognx;
    const size_t ldsm = (size_t)slm * fft_ldsn(1 << lognx) * 8;
    if (f64 || slm < 1 || slm > cshape_threads(2) || ldsm > lds_max) bl.okC[2] = 0;
    // z pass: points per thread by program (bpoints); the largest LDS any
    // of the job's programs takes is checked here
    auto zslots = [&](int v) {
      const int p = f64 ? bpoints<double>(lognz, bvar_nin(v), bvar_neq(v))
                        : bpoints<float>(lognz, bvar_nin(v), bvar_neq(v));
      return (kNT * p) >> lognz;
    };
    auto zlds = [&](int v) {
      return (size_t)bslots_used(bvar_nin(v), bvar_neq(v), zslots(v)) * fft_ldsn(1 << lognz) *
             (f64 ? 16 : 8);
    };
    size_t ldsz = 0;
    for (int s = 0; s < nstage; ++s) ldsz = std::max(ldsz, zlds(bvar[s]));
    if (ldsa > lds_max || ldsn > lds_max || ldsw > lds_max || ldsz > lds_max || sla < 1 ||
        sln < 1 || slw < 1) {
      rc = fail(MLV_ERR_UNSUPPORTED, "job " + std::to_string(j) + ": grid too large for LDS");
      break;
    }
    if (nstage > P.nstage) P.nstage = nstage;
    for (int s = 0; s < 2; ++s) {
      int a = 0, a1 = 0, a2 = 0, c = 0, cn = 0, cm = 0, bb = 0, bv = 0;
      if (s < nstage) {
        const int nin_v = bvar_nin(bvar[s]), neq_v = bvar_neq(bvar[s]);
        const int slz = zslots(bvar[s]);
        if (!bfits(nin_v, neq_v, slz)) {
          rc = fail(MLV_ERR_UNSUPPORTED,
                    "job " + std::to_string(j) + ": problem does not fit the z-pass LDS at this nz");
          break;
        }
        const int rb = brows(nin_v, neq_v, slz);
        a = nA[s] * ((nm + sla - 1) / sla);
        a1 = sla1 >= 1 ? nA[s] * ((nm + sla1 - 1) / sla1) : 0;
        a2 = (nm + kMixCols - 1) / kMixCols;
        if (f64 ? bsplit_fits<double>(lognz, bvar[s]) : bsplit_fits<float>(lognz, bvar[s]))
          bl.splitB = 1;
        c = neq[s] * ((nm + slw - 1) / slw);
        cn = neq[s] * ((nm + sln - 1) / sln);
        cm = slm >= 1 ? neq[s] * ((nm + slm - 1) / slm) : 0;
        bb = (nx + rb - 1) / rb;
        bv = bvar[s];
        bl.ldsB[bv] = std::max(bl.ldsB[bv], zlds(bv));
      }
      if (s == 0) {
        P.nbB.push_back(bb);
        P.loops0.push_back(jobs[j].loops);
      }
      pA[0][s][j + 1] = pA[0][s][j] + a;
      pA[1][s][j + 1] = pA[1][s][j] + a1;
      pA[2][s][j + 1] = pA[2][s][j] + a2;
      pC[s][j + 1] = pC[s][j] + c;
      pCn[s][j + 1] = pCn[s][j] + cn;
      P.pCm[s][j + 1] = P.pCm[s][j] + cm;
      for (int v = 0; v < BV_COUNT; ++v) pB[s][v][j + 1] = pB[s][v][j] + (v == bv ? bb : 0);
    }
    if (rc) break;
    bl.ldsAp[0] = std::max(bl.ldsAp[0], ldsa);
    if (sla1 >= 1) bl.ldsAp[1] = std::max(bl.ldsAp[1], ldsa1);
    ldsCm = std::max(ldsCm, ldsm);
    bl.ldsAp[2] = bl.ldsAp[0];
    // kernel M (single precision mean profiles): one LDS slot of nz
    if (!f64 && jobs[j].mean_split)
      bl.ldsM = std::max(bl.ldsM, (size_t)fft_ldsn(1 << lognz) * 8);
    bl.ldsC = std::max(bl.ldsC, ldsw);
    ldsCn = std::max(ldsCn, ldsn);
  }
  if (!rc) {
    // Kernel C: single precision always runs narrow -- 16 points per thread,
    // the stencil's neighbours read through the cache, 112 VGPRs, so four
    // blocks (16 wavefronts) share a CU and their loads overlap the others'
    // transforms (measured on MI355X, KH 1024^2: -7 % kernel time at 8 and
    // at 32 jobs against the wide form).  Double precision needs the
    // register file either way and runs wide when its grid covers the
    // device's block slots twice over at two blocks per CU.
    int cus = 0;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device) != hipSuccess ||
        cus <= 0)
      cus = 256;
    bl.wideC = (dtype == MLV_F64 && pC[0][njobs] >= 4 * cus) ? 1 : 0;
    // the mixed tiles and the 512-thread update run in the kernels compiled
    // for one x size; where it applies, the 512-thread update is the default
    // (64 x 1024^2 FP32: 556 against 578 us per launch for the 256-thread
    // narrow shape -- Px and Pz rows read in runs of 8 columns, not 4)
    if (!fixed_logn(bl.lognx_all)) bl.okA[2] = bl.okC[2] = 0;
    if (bl.okC[2]) bl.wideC = 2;
    // kernel A: the default block shape; kernel B: the split-pass plan
    bl.shapeA = 0;
    bl.planB = kPlanSplit;
    // choices picked by mlv_batch_autotune for the same set of jobs win
    P.signature = batch_signature(dtype, jobs, njobs);
    {
      std::lock_guard<std::mutex> lk(g_tuned_mu);
      auto it = g_tuned.find(P.signature);
      if (it != g_tuned.end()) {
        bl.shapeA = bl.okA[it->second.shapeA] ? it->second.shapeA : 0;
        bl.planB = it->second.planB;
        bl.wideC = bl.okC[it->second.wideC] ? it->second.wideC : bl.wideC;
      }
    }
    bl.ldsCp[1] = bl.ldsC;
    bl.ldsCp[0] = ldsCn;
    bl.ldsCp[2] = bl.okC[2] ? ldsCm : 0;
  }
  return rc;
}

// The tables of plan P on the device: the job table (with each job's
// diagnostics buffer) and the prefix arrays in one allocation, grown only
// when P needs more; the diagnostics buffers in the other of the two
// diagnostics allocations, with the records of jobs that were in the batch
// before (matched by workspace pointer) copied over and the others
// cleared.  The device must be idle (the old tables may still be read by
// queued kernels otherwise).
static int upload_batch(BatchHost* b, BatchPlan& P, const mlv_job_desc* jobs, int njobs) {
  const size_t jsz = b->dtype == MLV_F64 ? sizeof(JobDev<double>) : sizeof(JobDev<float>);
  auto align = [](size_t x) { return (x + 255) & ~(size_t)255; };
  // layout of the tables allocation
  size_t off = align(P.host.size());
  std::vector<std::pair<const std::vector<int>*, int**>> arrays;
  BatchLaunch& bl = P.bl;
  for (int s = 0; s < 2; ++s) {
    arrays.push_back({&P.pA[0][s], &bl.prefAp[0][s]});
    arrays.push_back({&P.pA[1][s], &bl.prefAp[1][s]});
    arrays.push_back({&P.pA[2][s], &bl.prefAp[2][s]});
    arrays.push_back({&P.pC[s], &bl.prefCp[1][s]});
    arrays.push_back({&P.pCn[s], &bl.prefCp[0][s]});
    arrays.push_back({&P.pCm[s], &bl.prefCp[2][s]});
    for (int v = 0; v < BV_COUNT; ++v) arrays.push_back({&P.pB[s][v], &bl.prefB[s][v]});
  }
  std::vector<size_t> aoff;
  for (auto& a : arrays) {
    aoff.push_back(off);
    off = align(off + a.first->size() * sizeof(int));
  }
  const size_t tbytes = off;
  if (tbytes > b->tables_cap) {
    if (b->tables) (void)hipFree(b->tables);
    b->tables = nullptr;
    b->tables_cap = 0;
    HIP_TRY(hipMalloc(&b->tables, tbytes), "hipMalloc batch tables");
    b->tables_cap = tbytes;
  }
  char* base = static_cast<char*>(b->tables);
  for (size_t k = 0; k < arrays.size(); ++k)
    *arrays[k].second = reinterpret_cast<int*>(base + aoff[k]);
  for (int s = 0; s < 2; ++s) {
    bl.nAp[0][s] = P.pA[0][s][njobs];
    bl.nAp[1][s] = P.pA[1][s][njobs];
    bl.nAp[2][s] = P.pA[2][s][njobs];
    bl.nCp[1][s] = P.pC[s][njobs];
    bl.nCp[0][s] = P.pCn[s][njobs];
    bl.nCp[2][s] = P.pCm[s][njobs];
    for (int v = 0; v < BV_COUNT; ++v) bl.nB[s][v] = P.pB[s][v][njobs];
  }
  // diagnostics: the buffer not in use now, grown when needed
  std::vector<size_t> doff(njobs);
  size_t dtotal = 0;
  for (int j = 0; j < njobs; ++j) {
    doff[j] = dtotal;
    dtotal += align(diag_doubles(P.nbB[j]) * sizeof(double));
  }
  const int nd = b->diag_buf[b->diag_cur] ? 1 - b->diag_cur : b->diag_cur;
  if (dtotal > b->diag_cap[nd]) {
    if (b->diag_buf[nd]) (void)hipFree(b->diag_buf[nd]);
    b->diag_buf[nd] = nullptr;
    b->diag_cap[nd] = 0;
    HIP_TRY(hipMalloc((void**)&b->diag_buf[nd], dtotal), "hipMalloc diagnostics");
    b->diag_cap[nd] = dtotal;
  }
  std::map<const void*, int> old;
  for (size_t j = 0; j < b->keys.size(); ++j) old[b->keys[j]] = (int)j;
  std::vector<double*> diag(njobs);
  std::vector<const void*> keys(njobs);
  for (int j = 0; j < njobs; ++j) {
    diag[j] = reinterpret_cast<double*>(reinterpret_cast<char*>(b->diag_buf[nd]) + doff[j]);
    keys[j] = jobs[j].work;
    const size_t bytes = diag_doubles(P.nbB[j]) * sizeof(double);
    auto it = old.find(keys[j]);
    if (it != old.end() && b->nbB[it->second] == P.nbB[j]) {
      HIP_TRY(hipMemcpy(diag[j], b->diag[it->second], bytes, hipMemcpyDeviceToDevice),
              "carry diagnostics");
    } else {
      HIP_TRY(hipMemset(diag[j], 0, bytes), "clear diagnostics");
    }
    if (b->dtype == MLV_F64) {
      reinterpret_cast<JobDev<double>*>(P.host.data())[j].diag = diag[j];
      reinterpret_cast<JobDev<double>*>(P.host.data())[j].nbB = P.nbB[j];
    } else {
      reinterpret_cast<JobDev<float>*>(P.host.data())[j].diag = diag[j];
      reinterpret_cast<JobDev<float>*>(P.host.data())[j].nbB = P.nbB[j];
    }
  }
  // one upload of the job table and the prefix arrays
  std::vector<char> stage(tbytes, 0);
  std::memcpy(stage.data(), P.host.data(), P.host.size());
  for (size_t k = 0; k < arrays.size(); ++k)
    std::memcpy(stage.data() + aoff[k], arrays[k].first->data(), arrays[k].first->size() * sizeof(int));
  HIP_TRY(hipMemcpy(b->tables, stage.data(), tbytes, hipMemcpyHostToDevice), "upload batch tables");
  (void)jsz;
  b->diag_cur = nd;
  b->diag = diag;
  b->keys = keys;
  b->jobs_dev = b->tables;
  b->njobs = njobs;
  b->nstage = P.nstage;
  b->nbB = P.nbB;
  b->loops0 = P.loops0;
  b->signature = P.signature;
  b->host = P.host;
  b->steps = 0;
  b->bl = bl;
  select_a_shape(b->bl, b->bl.shapeA);
  select_c_policy(b->bl, b->bl.wideC);
  hipError_t e = b->dtype == MLV_F64 ? fused_prepare<double>(b->bl) : fused_prepare<float>(b->bl);
  if (e != hipSuccess) return hip_fail(e, "kernel attributes");
  return MLV_OK;
}

static void free_batch_memory(BatchHost* b) {
  if (b->tables) (void)hipFree(b->tables);
  for (int k = 0; k < 2; ++k)
    if (b->diag_buf[k]) (void)hipFree(b->diag_buf[k]);
  b->tables = nullptr;
  b->diag_buf[0] = b->diag_buf[1] = nullptr;
}

int mlv_batch_create(mlv_batch** out, int dtype, const mlv_job_desc* jobs, int njobs) {
  if (!out || !jobs || njobs <= 0) return fail(MLV_ERR_ARG, "bad batch arguments");
  *out = nullptr;
  if (dtype != MLV_F32 && dtype != MLV_F64) return fail(MLV_ERR_ARG, "dtype must be F32 or F64");
  BatchHost* b = new BatchHost();
  b->dtype = dtype;
  HIP_TRY(hipGetDevice(&b->device), "hipGetDevice");
  int rc = get_twiddles(b->device, dtype, &b->tw);
  BatchPlan P;
  if (!rc) rc = plan_batch(dtype, jobs, njobs, b->device, P);
  if (!rc) rc = upload_batch(b, P, jobs, njobs);
  if (rc) {
    free_batch_memory(b);
    delete b;
    return rc;
  }
  *out = reinterpret_cast<mlv_batch*>(b);
  return MLV_OK;
}

int mlv_batch_rebuild(mlv_batch* bp, const mlv_job_desc* jobs, int njobs) {
  if (!bp || !jobs || njobs <= 0) return fail(MLV_ERR_ARG, "bad batch arguments");
  BatchHost* b = reinterpret_cast<BatchHost*>(bp);
  BatchPlan P;
  int rc = plan_batch(b->dtype, jobs, njobs, b->device, P);
  if (rc) return rc;  // the batch is unchanged
  // queued steps may still read the tables about to be overwritten
  HIP_TRY(hipDeviceSynchronize(), "hipDeviceSynchronize");
  int rc2 = prof_collect(b);
  if (rc2) return rc2;
  if (b->graph_exec) (void)hipGraphExecDestroy(b->graph_exec);
  if (b->graph) (void)hipGraphDestroy(b->graph);
  b->graph_exec = nullptr;
  b->graph = nullptr;
  return upload_batch(b, P, jobs, njobs);
}

int mlv_batch_destroy(mlv_batch* bp) {
  if (!bp) return MLV_OK;
  BatchHost* b = reinterpret_cast<BatchHost*>(bp);
  (void)hipDeviceSynchronize();
  free_batch_memory(b);
  for (hipEvent_t e : b->events) (void)hipEventDestroy(e);
  if (b->graph_exec) (void)hipGraphExecDestroy(b->graph_exec);
  if (b->graph) (void)hipGraphDestroy(b->graph);
  if (b->cap_stream) (void)hipStreamDestroy(b->cap_stream);
  delete b;
  return MLV_OK;
}

int mlv_batch_profile(mlv_batch* bp, int enable) {
  if (!bp) return fail(MLV_ERR_ARG, "batch is NULL");
  BatchHost* b = reinterpret_cast<BatchHost*>(bp);
  int rc = prof_collect(b);
  if (rc) return rc;
  b->prof = enable != 0;
  for (int k = 0; k < 3; ++k) {
    b->ms[k] = 0;
    b->launches[k] = 0;
  }
  return MLV_OK;
}

int mlv_batch_profile_read(mlv_batch* bp, double* ms, long long* launches) {
  if (!bp || !ms || !launches) return fail(MLV_ERR_ARG, "NULL argument");
  BatchHost* b = reinterpret_cast<BatchHost*>(bp);
  int rc = prof_collect(b);
  if (rc) return rc;
  for (int k = 0; k < 3; ++k) {
    ms[k] = b->ms[k];
    launches[k] = b->launches[k];
  }
  return MLV_OK;
}

static hipError_t enqueue_steps(BatchHost* b, int nsteps, int write_phys, long long first,
                                hipStream_t s, bool timed, int cadence = 0, double cutoff = 0.5);

int mlv_batch_step(mlv_batch* bp, int nsteps, int write_phys, void* stream) {
  if (!bp) return fail(MLV_ERR_ARG, "batch is NULL");
  if (nsteps < 0) return fail(MLV_ERR_ARG, "nsteps < 0");
  BatchHost* b = reinterpret_cast<BatchHost*>(bp);
  hipError_t e = enqueue_steps(b, nsteps, write_phys, b->steps, (hipStream_t)stream, b->prof);
  if (e != hipSuccess) return hip_fail(e, "fused step");
  b->steps += nsteps;
  return MLV_OK;
}

// A job's CFL check loops: the loops whose count before them (the job's
// loops at creation plus the batch's steps) is a multiple of cadence --
// loops 1, 1+cadence, ... of Simulation.end_loop's ticker, counted from
// the start of the simulation, so a resumed job keeps the phase its
// uninterrupted run had.  Returns whether step `step` of the batch is a
// check loop of any job.
static bool any_check(const BatchHost* b, long long step, int cadence) {
  if (cadence <= 0) return false;
  for (long long l0 : b->loops0)
    if ((l0 + step) % cadence == 0) return true;
  return false;
}

// Steps first .. first+nsteps-1 (counted from the batch's creation).  With
// cadence > 0 the steps that are some job's check loop also produce that
// job's diagnostics partials and are followed by the on-device CFL update.
static hipError_t enqueue_steps(BatchHost* b, int nsteps, int write_phys, long long first,
                                hipStream_t s, bool timed, int cadence, double cutoff) {
  for (int k = 0; k < nsteps; ++k) {
    const int wp = (write_phys && k == nsteps - 1) ? kWritePhys : 0;
    const bool dg = any_check(b, first + k, cadence);
    const int flags0 = wp | (dg ? (kDiag | (cadence << kDiagShift)) : 0);
    const int step = (int)((first + k) & 3);
    for (int st = 0; st < b->nstage; ++st) {
      hipEvent_t* ev = timed ? prof_slots(b) : nullptr;
      hipError_t e =
          b->dtype == MLV_F64
              ? fused_stage<double>((const JobDev<double>*)b->jobs_dev, b->bl, st,
                                    st == 0 ? flags0 : 0, step, (const cx<double>*)b->tw, s, ev)
              : fused_stage<float>((const JobDev<float>*)b->jobs_dev, b->bl, st,
                                   st == 0 ? flags0 : 0, step, (const cx<float>*)b->tw, s, ev);
      if (e != hipSuccess) return e;
    }
    if (dg) {
      hipError_t e = b->dtype == MLV_F64
                         ? fused_diag<double>((JobDev<double>*)b->jobs_dev, b->njobs, cutoff,
                                              cadence, s)
                         : fused_diag<float>((JobDev<float>*)b->jobs_dev, b->njobs, cutoff,
                                             cadence, s);
      if (e != hipSuccess) return e;
    }
  }
  return hipSuccess;
}

static int graph_run(BatchHost* b, int nsteps, int cadence, double cutoff, void* stream) {
  if (nsteps <= 0 || (nsteps & 3)) return fail(MLV_ERR_ARG, "graph steps must be a positive multiple of 4");
  if (b->steps & 3)
    return fail(MLV_ERR_ARG, "graph replay needs the batch's step count at a multiple of 4");
  if (cadence > 0 && ((nsteps % cadence) || (b->steps % cadence)))
    return fail(MLV_ERR_ARG, "graph runs with CFL checks need nsteps and the step count at "
                             "multiples of the cadence");
  if (!b->graph_exec || b->graph_steps != nsteps || b->graph_cadence != cadence ||
      b->graph_cutoff != cutoff) {
    if (b->graph_exec) (void)hipGraphExecDestroy(b->graph_exec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    b->graph_exec = nullptr;
    b->graph = nullptr;
    if (!b->cap_stream)
      HIP_TRY(hipStreamCreateWithFlags(&b->cap_stream, hipStreamNonBlocking), "capture stream");
    HIP_TRY(hipStreamBeginCapture(b->cap_stream, hipStreamCaptureModeThreadLocal), "begin capture");
    hipError_t e = enqueue_steps(b, nsteps, 0, 0, b->cap_stream, false, cadence, cutoff);
    hipGraph_t g = nullptr;
    hipError_t e2 = hipStreamEndCapture(b->cap_stream, &g);
    if (e != hipSuccess) return hip_fail(e, "capture");
    if (e2 != hipSuccess) return hip_fail(e2, "end capture");
    b->graph = g;
    HIP_TRY(hipGraphInstantiate(&b->graph_exec, g, nullptr, nullptr, 0), "instantiate");
    b->graph_steps = nsteps;
    b->graph_cadence = cadence;
    b->graph_cutoff = cutoff;
  }
  HIP_TRY(hipGraphLaunch(b->graph_exec, (hipStream_t)stream), "graph launch");
  b->steps += nsteps;
  return MLV_OK;
}

int mlv_batch_graph_step(mlv_batch* bp, int nsteps, void* stream) {
  if (!bp) return fail(MLV_ERR_ARG, "batch is NULL");
  return graph_run(reinterpret_cast<BatchHost*>(bp), nsteps, 0, 0.0, stream);
}

// Drop a captured graph: it holds the launches of the choices it was
// captured with.
static void drop_graph(BatchHost* b) {
  if (b->graph_exec) (void)hipGraphExecDestroy(b->graph_exec);
  if (b->graph) (void)hipGraphDestroy(b->graph);
  b->graph_exec = nullptr;
  b->graph = nullptr;
}

// Apply kernel choices (validated by the caller).
static void apply_plan(BatchHost* b, int shapeA, int planB, int wideC) {
  select_a_shape(b->bl, shapeA);
  b->bl.planB = planB;
  select_c_policy(b->bl, wideC);
}

int mlv_batch_autotune(mlv_batch* bp, int nsteps, void* stream, int* chosen) {
  if (!bp) return fail(MLV_ERR_ARG, "batch is NULL");
  if (nsteps < 1) return fail(MLV_ERR_ARG, "nsteps must be positive");
  BatchHost* b = reinterpret_cast<BatchHost*>(bp);
  hipStream_t s = (hipStream_t)stream;
  // Options per kernel: A's block shapes (the smaller where every job fits
  // it), B's plans (two where some job takes the split-pass plan), C's tile
  // shapes.  Segments run every kernel at option 0, then at option 1 (a
  // kernel with one option stays at 0), then 1 and 0 again -- half of
  // nsteps each; one segment per option when nsteps is 1 -- with events
  // around each kernel of every step, so each kernel's options are timed in
  // the same steps.  A kernel's option figure is its faster segment per
  // step: the mirrored order and the minimum keep a clock ramp after an
  // idle period, and a kernel's first launch, from deciding the choice.
  // The kernels read and write whole buffers in HBM between launches, so
  // one kernel's choice does not change another's time.
  // the values each kernel can take
  int opts[3][3], nopt[3] = {0, 0, 0};
  for (int a = 0; a < 3; ++a)
    if (b->bl.okA[a]) opts[0][nopt[0]++] = a;
  opts[1][nopt[1]++] = kPlanSplit;
  if (b->bl.splitB) opts[1][nopt[1]++] = kPlanTwoSweep;
  for (int c = 0; c < 3; ++c)
    if (b->bl.okC[c]) opts[2][nopt[2]++] = c;
  const int maxopt = std::max(nopt[0], std::max(nopt[1], nopt[2]));
  const int h1 = nsteps >= 2 ? nsteps / 2 : nsteps, h2 = nsteps - h1;
  // option indices 0, 1, .., maxopt - 1, then back down
  int seg_opt[6], seg_steps[6];
  const int nseg = 2 * maxopt;
  for (int k = 0; k < maxopt; ++k) {
    seg_opt[k] = k;
    seg_steps[k] = h1;
    seg_opt[nseg - 1 - k] = k;
    seg_steps[nseg - 1 - k] = h2;
  }
  float best_ms[3][3];
  for (auto& r : best_ms)
    for (float& v : r) v = INFINITY;
  const int per_step = 4 * b->nstage;
  std::vector<hipEvent_t> ev((size_t)std::max(h1, h2) * per_step, nullptr);
  hipError_t e = hipSuccess;
  for (auto& x : ev)
    if (e == hipSuccess) e = hipEventCreate(&x);
  for (int k = 0; k < nseg && e == hipSuccess; ++k) {
    const int n = seg_steps[k];
    if (n == 0) continue;
    int opt[3];
    for (int q = 0; q < 3; ++q) opt[q] = opts[q][seg_opt[k] < nopt[q] ? seg_opt[k] : 0];
    apply_plan(b, opt[0], opt[1], opt[2]);
    for (int t = 0; t < n && e == hipSuccess; ++t) {
      const int step = (int)((b->steps + t) & 3);
      for (int st = 0; st < b->nstage && e == hipSuccess; ++st) {
        hipEvent_t* evs = &ev[(size_t)t * per_step + 4 * st];
        e = b->dtype == MLV_F64
                ? fused_stage<double>((const JobDev<double>*)b->jobs_dev, b->bl, st, 0, step,
                                      (const cx<double>*)b->tw, s, evs)
                : fused_stage<float>((const JobDev<float>*)b->jobs_dev, b->bl, st, 0, step,
                                     (const cx<float>*)b->tw, s, evs);
      }
    }
    if (e == hipSuccess) e = hipEventSynchronize(ev[(size_t)n * per_step - 1]);
    if (e != hipSuccess) break;
    b->steps += n;
    float sum[3] = {0.f, 0.f, 0.f};
    for (int t = 0; t < n * b->nstage && e == hipSuccess; ++t)
      for (int q = 0; q < 3 && e == hipSuccess; ++q) {
        float ms = 0.f;
        e = hipEventElapsedTime(&ms, ev[4 * t + q], ev[4 * t + q + 1]);
        sum[q] += ms;
      }
    for (int q = 0; q < 3; ++q)
      if (seg_opt[k] < nopt[q]) best_ms[q][seg_opt[k]] = std::min(best_ms[q][seg_opt[k]], sum[q] / n);
  }
  for (auto& x : ev)
    if (x) (void)hipEventDestroy(x);
  if (e != hipSuccess) return hip_fail(e, "autotune");
  int best[3];
  for (int q = 0; q < 3; ++q) {
    int bi = 0;
    for (int o = 1; o < nopt[q]; ++o)
      if (best_ms[q][o] < best_ms[q][bi]) bi = o;
    best[q] = opts[q][bi];
  }
  apply_plan(b, best[0], best[1], best[2]);
  drop_graph(b);
  {
    std::lock_guard<std::mutex> lk(g_tuned_mu);
    g_tuned[b->signature] = TunedPlan{best[0], best[1], best[2]};
  }
  if (chosen)
    for (int q = 0; q < 3; ++q) chosen[q] = best[q];
  return MLV_OK;
}

int mlv_batch_plan(const mlv_batch* bp, int* plan) {
  if (!bp || !plan) return fail(MLV_ERR_ARG, "batch or plan is NULL");
  const BatchHost* b = reinterpret_cast<const BatchHost*>(bp);
  plan[0] = b->bl.shapeA;
  plan[1] = b->bl.planB;
  plan[2] = b->bl.wideC;
  plan[3] = b->bl.okA[1];
  plan[4] = b->bl.splitB;
  plan[5] = b->bl.okA[2];
  plan[6] = b->bl.okC[2];
  return MLV_OK;
}

int mlv_batch_set_plan(mlv_batch* bp, int shape_xinv, int plan_zpass, int wide_update) {
  if (!bp) return fail(MLV_ERR_ARG, "batch is NULL");
  BatchHost* b = reinterpret_cast<BatchHost*>(bp);
  if (shape_xinv < 0 || shape_xinv > 2 || plan_zpass < 0 || plan_zpass > 1 || wide_update < 0 ||
      wide_update > 2)
    return fail(MLV_ERR_ARG, "kernel choices: x-inverse and update 0, 1 or 2, z pass 0 or 1");
  if (wide_update == 2 && !b->bl.okC[2])
    return fail(MLV_ERR_UNSUPPORTED,
                "the 512-thread update shape needs single-precision jobs of one x size in [512, 2048]");
  if (shape_xinv == 1 && !b->bl.okA[1])
    return fail(MLV_ERR_UNSUPPORTED, "some job's x transform does not fit the smaller x-inverse block");
  if (shape_xinv == 2 && !b->bl.okA[2])
    return fail(MLV_ERR_UNSUPPORTED,
                "the mixed x-inverse tiles need single-precision jobs of one x size, 512 or 1024, "
                "whose stages' ops fit two columns each in a block");
  apply_plan(b, shape_xinv, plan_zpass, wide_update);
  drop_graph(b);
  return MLV_OK;
}

int mlv_batch_run(mlv_batch* bp, int nsteps, int cadence, double cutoff, int graph, void* stream) {
  if (!bp) return fail(MLV_ERR_ARG, "batch is NULL");
  if (nsteps < 0 || cadence < 0 || !(cutoff > 0)) return fail(MLV_ERR_ARG, "bad run arguments");
  BatchHost* b = reinterpret_cast<BatchHost*>(bp);
  if (graph) return graph_run(b, nsteps, cadence, cutoff, stream);
  hipError_t e =
      enqueue_steps(b, nsteps, 0, b->steps, (hipStream_t)stream, b->prof, cadence, cutoff);
  if (e != hipSuccess) return hip_fail(e, "fused run");
  b->steps += nsteps;
  return MLV_OK;
}

}  // extern "C"

template <typename T>
static int read_job(BatchHost* b, int job, double* dt, double* t, long long* loops) {
  JobDev<T> J;
  HIP_TRY(hipMemcpy(&J, (const char*)b->jobs_dev + job * sizeof(JobDev<T>), sizeof(J),
                    hipMemcpyDeviceToHost),
          "read job");
  if (dt) *dt = J.dt;
  if (t) *t = J.t;
  if (loops) *loops = J.loops;
  return MLV_OK;
}

extern "C" {

int mlv_batch_job_time(mlv_batch* bp, int job, double* dt, double* t, long long* loops,
                       int* breach) {
  if (!bp) return fail(MLV_ERR_ARG, "batch is NULL");
  BatchHost* b = reinterpret_cast<BatchHost*>(bp);
  if (job < 0 || job >= b->njobs) return fail(MLV_ERR_ARG, "job index out of range");
  HIP_TRY(hipDeviceSynchronize(), "synchronize");
  int rc = b->dtype == MLV_F64 ? read_job<double>(b, job, dt, t, loops)
                               : read_job<float>(b, job, dt, t, loops);
  if (rc) return rc;
  if (breach) {
    double h[2];
    HIP_TRY(hipMemcpy(h, b->diag[job] + (size_t)kDiagPart * b->nbB[job], sizeof(h),
                      hipMemcpyDeviceToHost),
            "read diagnostics");
    *breach = h[1] != 0.0;
  }
  return MLV_OK;
}

int mlv_batch_tracker(mlv_batch* bp, int job, double* out, int max_records, int* nrec) {
  if (!bp || !nrec) return fail(MLV_ERR_ARG, "NULL argument");
  BatchHost* b = reinterpret_cast<BatchHost*>(bp);
  if (job < 0 || job >= b->njobs) return fail(MLV_ERR_ARG, "job index out of range");
  HIP_TRY(hipDeviceSynchronize(), "synchronize");
  return copy_tracker(b->diag[job] + (size_t)kDiagPart * b->nbB[job], out, max_records, nrec);
}

int mlv_batch_set_dt(mlv_batch* bp, int job, double dt, void* stream) {
  if (!bp) return fail(MLV_ERR_ARG, "batch is NULL");
  BatchHost* b = reinterpret_cast<BatchHost*>(bp);
  if (job < 0 || job >= b->njobs) return fail(MLV_ERR_ARG, "job index out of range");
  if (!(dt > 0)) return fail(MLV_ERR_ARG, "dt must be positive");
  hipError_t e = b->dtype == MLV_F64
                     ? fused_set_dt<double>((JobDev<double>*)b->jobs_dev, job, dt, (hipStream_t)stream)
                     : fused_set_dt<float>((JobDev<float>*)b->jobs_dev, job, dt, (hipStream_t)stream);
  if (e != hipSuccess) return hip_fail(e, "set dt");
  return MLV_OK;
}

long long mlv_batch_steps(const mlv_batch* bp) {
  if (!bp) return -1;
  return reinterpret_cast<const BatchHost*>(bp)->steps;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Cosine and sine bases (SpectralTransformer.py:90-199, BasisFunctions.py:
// 26-59).  The reference mirrors the signal about its end point to length
// 2(n-1) (sine: with a sign flip) and takes an FFT of that length, dividing
// by n-1 (sine: -i(n-1)) and halving the k = 0 cosine term; the inverse
// doubles that term, multiplies by the same factor, inverse-transforms and
// crops to n points.  2(n-1) is rarely a power of two, so each axis is a
// dense transform whose matrix folds the mirror images, the factors and the
// crop together; the matrices are built once per plan in long double.
// ---------------------------------------------------------------------------
namespace {

using cld = std::complex<long double>;
constexpr int kExp = MLV_BASIS_COMPLEX_EXP, kCos = MLV_BASIS_COSINE, kSin = MLV_BASIS_SINE;
const long double kPiL = 3.141592653589793238462643383279502884L;

cld cis(long double a) { return cld(cosl(a), sinl(a)); }

// length of the periodic sequence the reference transforms along one axis
int axis_period(int b, int n) { return b == kExp ? n : 2 * (n - 1); }
// the reference's axis factor: n, n-1 or -i(n-1)
cld axis_factor(int b, int n) {
  if (b == kExp) return cld((long double)n, 0);
  if (b == kCos) return cld((long double)(n - 1), 0);
  return cld(0, -(long double)(n - 1));
}

// Forward: spectral value of mode k (signed) from sample j of n.
cld fwd_entry(int b, int n, int k, int j) {
  const int L = axis_period(b, n);
  cld v = 0;
  if (b == kExp) {
    v = cis(-2 * kPiL * (long double)j * k / L);
  } else {
    if (j <= n - 2) v += cis(-2 * kPiL * (long double)j * k / L);
    if (j >= 1) v += (b == kCos ? 1.0L : -1.0L) * cis(-2 * kPiL * (long double)(L - j) * k / L);
  }
  v /= axis_factor(b, n);
  if (b == kCos && k == 0) v /= 2.0L;
  return v;
}

// Inverse along the full complex axis (x): sample j from mode k (signed).
cld inv_entry(int b, int n, int k, int j) {
  const int L = axis_period(b, n);
  cld v = cis(2 * kPiL * (long double)j * k / L) / (long double)L * axis_factor(b, n);
  if (b == kCos && k == 0) v *= 2.0L;
  return v;
}

// Inverse along the half-spectrum axis (z, real output): the weight of
// mode m in sample j, so that sample = Re(sum_m W U_m) as irfft gives it.
cld inv_half_entry(int b, int n, int m, int j) {
  const int L = axis_period(b, n);
  cld v = cis(2 * kPiL * (long double)j * m / L) / (long double)L * axis_factor(b, n);
  if (m > 0) v *= 2.0L;
  if (b == kCos && m == 0) v *= 2.0L;
  return v;
}

int mode_of_row(const mlv_plan* p, int r) { return r <= p->nn ? r : r - p->rows; }

// Matrix key: direction (0 forward, 1 inverse), axis (0 x, 1 z), basis.
int get_basis_mat(mlv_plan* p, int dir, int axis, int b, const void** out) {
  const int key = dir * 16 + axis * 4 + b;
  auto it = p->basis_mats.find(key);
  if (it != p->basis_mats.end()) {
    *out = it->second;
    return MLV_OK;
  }
  int K, J;
  if (dir == 0 && axis == 1) { K = p->nm; J = p->nz; }
  else if (dir == 0) { K = p->rows; J = p->nx; }
  else if (axis == 0) { K = p->nx; J = p->rows; }
  else { K = p->nz; J = p->nm; }
  std::vector<cld> m((size_t)K * J);
  for (int k = 0; k < K; ++k)
    for (int j = 0; j < J; ++j) {
      cld v;
      if (dir == 0 && axis == 1) v = fwd_entry(b, p->nz, k, j);
      else if (dir == 0) v = fwd_entry(b, p->nx, mode_of_row(p, k), j);
      else if (axis == 0) v = inv_entry(b, p->nx, mode_of_row(p, j), k);
      else v = inv_half_entry(b, p->nz, j, k);
      m[(size_t)k * J + j] = v;
    }
  const size_t nel = (size_t)K * J;
  void* d = nullptr;
  if (p->dtype == MLV_F64) {
    std::vector<double> h(2 * nel);
    for (size_t i = 0; i < nel; ++i) { h[2 * i] = (double)m[i].real(); h[2 * i + 1] = (double)m[i].imag(); }
    HIP_TRY(hipMalloc(&d, h.size() * sizeof(double)), "hipMalloc basis matrix");
    HIP_TRY(hipMemcpy(d, h.data(), h.size() * sizeof(double), hipMemcpyHostToDevice), "upload");
  } else {
    std::vector<float> h(2 * nel);
    for (size_t i = 0; i < nel; ++i) { h[2 * i] = (float)m[i].real(); h[2 * i + 1] = (float)m[i].imag(); }
    HIP_TRY(hipMalloc(&d, h.size() * sizeof(float)), "hipMalloc basis matrix");
    HIP_TRY(hipMemcpy(d, h.data(), h.size() * sizeof(float), hipMemcpyHostToDevice), "upload");
  }
  p->basis_mats[key] = d;
  *out = d;
  return MLV_OK;
}

int check_bases(const mlv_plan* p, int bx, int bz) {
  if (!p->fully) return fail(MLV_ERR_UNSUPPORTED, "cosine/sine bases need a fully spectral plan");
  if (bx < kExp || bx > kSin || bz < kExp || bz > kSin) return fail(MLV_ERR_ARG, "unknown basis");
  return MLV_OK;
}

}  // namespace

extern "C" {

int mlv_to_spectral_basis(mlv_plan* p, int bx, int bz, const void* phys, void* spec, void* stream) {
  CHECK_PLAN(p);
  CHECK_PTR(phys);
  CHECK_PTR(spec);
  int rc = check_bases(p, bx, bz);
  if (rc) return rc;
  if (bx == kExp && bz == kExp) return mlv_to_spectral(p, phys, spec, stream);
  const void *mz, *mx;
  if ((rc = get_basis_mat(p, 0, 1, bz, &mz)) || (rc = get_basis_mat(p, 0, 0, bx, &mx))) return rc;
  hipStream_t s = (hipStream_t)stream;
  // z: (nx, nz) real -> (nx, nm) complex in the plan workspace
  hipError_t e = DISPATCH(p,
      launch_dense<float>(mz, p->nm, p->nz, p->nx, phys, 0, p->nz, 1, p->work, 0, p->nm, 1, s),
      launch_dense<double>(mz, p->nm, p->nz, p->nx, phys, 0, p->nz, 1, p->work, 0, p->nm, 1, s));
  // x: per kz column, (nx) -> (rows) complex
  if (e == hipSuccess)
    e = DISPATCH(p,
        launch_dense<float>(mx, p->rows, p->nx, p->nm, p->work, 1, 1, p->nm, spec, 0, 1, p->nm, s),
        launch_dense<double>(mx, p->rows, p->nx, p->nm, p->work, 1, 1, p->nm, spec, 0, 1, p->nm, s));
  if (e != hipSuccess) return hip_fail(e, "to_spectral (basis)");
  return MLV_OK;
}

int mlv_to_physical_basis(mlv_plan* p, int bx, int bz, const void* spec, void* phys, void* stream) {
  CHECK_PLAN(p);
  CHECK_PTR(phys);
  CHECK_PTR(spec);
  int rc = check_bases(p, bx, bz);
  if (rc) return rc;
  if (bx == kExp && bz == kExp) return mlv_to_physical(p, spec, phys, stream);
  const void *mx, *mz;
  if ((rc = get_basis_mat(p, 1, 0, bx, &mx)) || (rc = get_basis_mat(p, 1, 1, bz, &mz))) return rc;
  hipStream_t s = (hipStream_t)stream;
  // x: per kz column, (rows) -> (nx) complex
  hipError_t e = DISPATCH(p,
      launch_dense<float>(mx, p->nx, p->rows, p->nm, spec, 1, 1, p->nm, p->work, 0, 1, p->nm, s),
      launch_dense<double>(mx, p->nx, p->rows, p->nm, spec, 1, 1, p->nm, p->work, 0, 1, p->nm, s));
  // z: per x row, (nm) complex -> (nz) real
  if (e == hipSuccess)
    e = DISPATCH(p,
        launch_dense<float>(mz, p->nz, p->nm, p->nx, p->work, 1, p->nm, 1, phys, 1, p->nz, 1, s),
        launch_dense<double>(mz, p->nz, p->nm, p->nx, p->work, 1, p->nm, 1, phys, 1, p->nz, 1, s));
  if (e != hipSuccess) return hip_fail(e, "to_physical (basis)");
  return MLV_OK;
}

int mlv_spec_op_basis(mlv_plan* p, int kind, int bx, int bz, const void* in, void* out,
                      void* stream) {
  CHECK_PLAN(p);
  CHECK_PTR(out);
  if (kind != MLV_SPEC_LAP) CHECK_PTR(in);
  int rc = check_bases(p, bx, bz);
  if (rc) return rc;
  if (kind < MLV_SPEC_DDX || kind > MLV_SPEC_POISSON) return fail(MLV_ERR_ARG, "unknown spectral op");
  // BasisFunctions.py:26-59: first-derivative wavelength i 2 pi (exp), pi
  // (sine), -pi (cosine); second derivative -|wavelength|^2; over L, L^2
  auto w1 = [](int b) { return b == kExp ? 2.0 * M_PI : (b == kSin ? M_PI : -M_PI); };
  auto w2 = [](int b) { const double a = b == kExp ? 2.0 * M_PI : M_PI; return -(a * a); };
  const double fx = w1(bx) / p->lx, fz = w1(bz) / p->lz;
  const double fx2 = w2(bx) / (p->lx * p->lx), fz2 = w2(bz) / (p->lz * p->lz);
  const int real = (bx != kExp ? 1 : 0) | (bz != kExp ? 2 : 0);
  hipStream_t s = (hipStream_t)stream;
  hipError_t e = DISPATCH(p,
      launch_spec<float>(in, out, p->rows, p->cols, p->nn, 1, kind, fx, fz, fx2, fz2, s, real),
      launch_spec<double>(in, out, p->rows, p->cols, p->nn, 1, kind, fx, fz, fx2, fz2, s, real));
  if (e != hipSuccess) return hip_fail(e, "spectral op (basis)");
  return MLV_OK;
}

}  // extern "C"
