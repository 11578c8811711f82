// extern "C" entry points of the fused Fourier-x / FD-z batch (declared in
// include/melvin_hip.h).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/melvin_hip.h"
#include "capi_common.hpp"
#include "fdz.hpp"

using namespace mlv;
using namespace mlv::abi;

namespace {

// Row kernel variant of a job with wpr wavefronts per row and K elements
// per thread: the smallest compiled size >= K of that wpr.
int fdz_variant(int K, int wpr) {
  for (int v = 0; v < kFdzVariants; ++v)
    if (kFdzWprs[v] == wpr && kFdzKts[v] >= K) return v;
  return -1;
}

struct FdzHost {
  int dtype = 0, njobs = 0, device = 0, nphase = 1;
  void* jobs_dev = nullptr;
  FdzLaunch fl{};
  long long steps = 0;
  bool primed = false;
  const void* tw = nullptr;
  std::vector<void*> allocs;
  std::vector<char> host;       // host copy of the job table
  hipStream_t cap_stream = nullptr;
  hipGraph_t graph = nullptr;
  hipGraphExec_t graph_exec = nullptr;
  int graph_steps = 0, graph_cadence = 0;
  double graph_cutoff = 0.0;
  int max_nn = 0;
  std::vector<double*> diag;    // per job: partials, header, tracker records
  std::vector<int> nbC;         // per job: column blocks of phase 0
  std::vector<long long> loops0;  // per job: loop count at creation
  bool prof = false;
  std::vector<hipEvent_t> events;
  size_t used = 0;
  double ms[2] = {0, 0};
  long long launches[2] = {0, 0};
};

hipEvent_t* prof_slots(FdzHost* b) {
  if (!b->prof) return nullptr;
  while (b->events.size() < b->used + 3) {
    hipEvent_t e;
    if (hipEventCreate(&e) != hipSuccess) return nullptr;
    b->events.push_back(e);
  }
  hipEvent_t* r = &b->events[b->used];
  b->used += 3;
  return r;
}

int prof_collect(FdzHost* b) {
  if (b->used == 0) return MLV_OK;
  HIP_TRY(hipEventSynchronize(b->events[b->used - 1]), "hipEventSynchronize");
  for (size_t i = 0; i + 2 < b->used; i += 3) {
    for (int k = 0; k < 2; ++k) {
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

size_t fdz_work_elems(int nx, int nz) {
  const size_t s = (size_t)((nx - 1) / 3) * nz;
  return 12 * s;  // Qx/Qz per field (4), velocity of 3 states (6), predicted state (2)
}

template <typename T>
int build_fdz_job(const mlv_fdz_job_desc& d, FdzJob<T>& J, int j) {
  char msg[160];
  auto bad = [&](const char* what) {
    std::snprintf(msg, sizeof(msg), "job %d: %s", j, what);
    return fail(MLV_ERR_ARG, msg);
  };
  std::memset(&J, 0, sizeof(J));
  const int lnx = ilog2_exact(d.nx);
  if (lnx < 3 || lnx > 11) return bad("nx must be a power of two in [8, 2048]");
  if (d.nz < 5 || d.nz > kFdzMaxNz) return bad("nz must be in [5, 768]");
  if (d.fd_order != 2 && d.fd_order != 4) return bad("fd_order must be 2 or 4");
  if (d.ab_order != 2 && d.ab_order != 4) return bad("ab_order must be 2 or 4");
  if (d.scheme != MLV_FDZ_EXPLICIT && d.scheme != MLV_FDZ_SEMI_IMPLICIT && d.scheme != MLV_FDZ_PC)
    return bad("unknown scheme");
  if (!(d.lx > 0) || !(d.lz > 0) || !(d.dt > 0)) return bad("lengths and dt must be positive");
  if (d.curr < 0 || d.curr >= d.ab_order || (d.parity & ~1)) return bad("curr/parity out of range");
  if (!d.work) return bad("work is NULL");
  for (int f = 0; f < 2; ++f)
    if (!d.state[0][f] || !d.state[1][f] || !d.hist[f]) return bad("state/hist pointer is NULL");
  J.nx = d.nx;
  J.nz = d.nz;
  J.lognx = lnx;
  J.nn = (d.nx - 1) / 3;
  J.wpr = fdz_wpr(d.nz);
  J.K = (d.nz + 64 * J.wpr - 1) / (64 * J.wpr);
  J.kt = kFdzKts[fdz_variant(J.K, J.wpr)];
  J.fd_order = d.fd_order;
  J.scheme = d.scheme;
  J.q = d.ab_order;
  J.curr0 = d.curr;
  J.par0 = d.parity;
  const double two_pi = 2.0 * M_PI;
  J.fx = two_pi / d.lx;
  J.fx2 = -(two_pi * two_pi) / (d.lx * d.lx);
  J.dx = d.lx / d.nx;
  J.dz = d.lz / d.nz;
  J.dt = d.dt;
  J.alpha = d.alpha;
  J.Pr = d.Pr;
  J.Ra = d.Ra;
  for (int b = 0; b < 2; ++b)
    for (int f = 0; f < 2; ++f) J.state[b][f] = (cx<T>*)d.state[b][f];
  for (int f = 0; f < 2; ++f) {
    J.hist[f] = (cx<T>*)d.hist[f];
    J.comp[f] = sizeof(T) == 4 ? (cx<T>*)d.comp[f] : nullptr;
    J.psi[f] = (cx<T>*)d.psi[f];
  }
  const size_t s = (size_t)J.nn * d.nz;
  cx<T>* w = (cx<T>*)d.work;
  for (int i = 0; i < 4; ++i) J.qw[i] = w + i * s;
  for (int v = 0; v < 3; ++v)
    for (int c = 0; c < 2; ++c) J.vel[v][c] = w + (4 + 2 * v + c) * s;
  for (int f = 0; f < 2; ++f) J.star[f] = w + (10 + f) * s;
  for (int i = 0; i < 4; ++i) J.phys[i] = (T*)d.phys[i];
  const double idz2 = 1.0 / (J.dz * J.dz);
  J.lo[0] = idz2;
  J.lo[1] = -(d.alpha * d.dt * d.Pr) * idz2;
  J.lo[2] = -(d.alpha * d.dt * 1.0) * idz2;
  J.t = d.t;
  J.loops = d.loops;
  return MLV_OK;
}

template <typename T>
FdzJob<T>& host_job(FdzHost* b, int j) {
  return reinterpret_cast<FdzJob<T>*>(b->host.data())[j];
}

// Theta tables of job j for its current dt (Poisson too when all = true).
template <typename T>
hipError_t job_factors(FdzHost* b, int j, bool all, hipStream_t s) {
  const FdzJob<T>& J = host_job<T>(b, j);
  const double fxabs = J.fx;
  hipError_t e = hipSuccess;
  if (all) e = fdz_factors(J.fac[0], J.nn, J.nz, J.K, J.wpr, fxabs, J.dz, 0, 0.0, s);
  if (e == hipSuccess && J.scheme != MLV_FDZ_EXPLICIT) {
    e = fdz_factors(J.fac[1], J.nn, J.nz, J.K, J.wpr, fxabs, J.dz, 1, J.alpha * J.dt * J.Pr, s);
    if (e == hipSuccess)
      e = fdz_factors(J.fac[2], J.nn, J.nz, J.K, J.wpr, fxabs, J.dz, 2, J.alpha * J.dt * 1.0, s);
  }
  return e;
}

template <typename T>
int create_impl(FdzHost* b, const mlv_fdz_job_desc* jobs, int njobs) {
  const int dtype = b->dtype;
  b->host.assign(sizeof(FdzJob<T>) * njobs, 0);
  const size_t lds_max = max_lds_per_block(b->device);
  std::vector<int> prow[3][kFdzVariants], pcol[2];
  for (int p = 0; p < 3; ++p)
    for (int v = 0; v < kFdzVariants; ++v) prow[p][v].assign(njobs + 1, 0);
  for (int p = 0; p < 2; ++p) pcol[p].assign(njobs + 1, 0);
  FdzLaunch& fl = b->fl;
  std::memset(&fl, 0, sizeof(fl));
  fl.njobs = njobs;
  int zmax = 0;
  for (int j = 0; j < njobs; ++j) {
    FdzJob<T>& J = host_job<T>(b, j);
    int rc = build_fdz_job<T>(jobs[j], J, j);
    if (rc) return rc;
    const int tc = fdz_cols_per_block<T>(J.lognx);
    const size_t ldsc = fdz_cols_lds_bytes<T>(J.lognx);
    if (tc < 1 || ldsc > lds_max)
      return fail(MLV_ERR_UNSUPPORTED, "job " + std::to_string(j) + ": nx too large for LDS");
    fl.ldscol = std::max(fl.ldscol, ldsc);
    zmax = std::max(zmax, J.nz);
    const int nph = J.scheme == MLV_FDZ_PC ? 2 : 1;
    b->nphase = std::max(b->nphase, nph);
    const int var = fdz_variant(J.K, J.wpr);
    const int cb = (J.nz + tc - 1) / tc;
    const int rpb = fdz_rows_per_block(J.wpr);
    const int rb2 = (2 * J.nn + rpb - 1) / rpb;
    const int rb1 = (J.nn + rpb - 1) / rpb;
    for (int p = 0; p < 2; ++p) pcol[p][j + 1] = pcol[p][j] + (p < nph ? cb : 0);
    J.nbC = cb;
    b->max_nn = std::max(b->max_nn, J.nn);
    for (int p = 0; p < 3; ++p) {
      const int nb = p == 2 ? rb1 : (p < nph ? rb2 : 0);
      for (int v = 0; v < kFdzVariants; ++v) prow[p][v][j + 1] = prow[p][v][j] + (v == var ? nb : 0);
    }
  }
  fl.zstride = zmax | 1;  // odd stride: the waves' rows start on different banks
  // two LDS rows (buf, vrow) per row of a block
  for (int v = 0; v < kFdzVariants; ++v) {
    fl.ldsrow[v] = (size_t)fdz_rows_per_block(kFdzWprs[v]) * 2 * fl.zstride * sizeof(double2);
    if (fl.ldsrow[v] > lds_max)
      return fail(MLV_ERR_UNSUPPORTED, "nz too large for the row kernel LDS");
  }
  auto dalloc = [&](size_t bytes, void** p) -> int {
    hipError_t e = hipMalloc(p, bytes);
    if (e != hipSuccess) return hip_fail(e, "hipMalloc fdz batch");
    b->allocs.push_back(*p);
    return MLV_OK;
  };
  auto upload = [&](const std::vector<int>& v, int** dst) -> int {
    int r = dalloc(v.size() * sizeof(int), (void**)dst);
    if (r) return r;
    HIP_TRY(hipMemcpy(*dst, v.data(), v.size() * sizeof(int), hipMemcpyHostToDevice), "upload");
    return MLV_OK;
  };
  // diagnostics: partials of phase 0's column blocks, header, records
  b->diag.assign(njobs, nullptr);
  b->nbC.assign(njobs, 0);
  b->loops0.assign(njobs, 0);
  for (int j = 0; j < njobs; ++j) {
    FdzJob<T>& J = host_job<T>(b, j);
    const size_t bytes = diag_doubles(J.nbC) * sizeof(double);
    int r = dalloc(bytes, (void**)&J.diag);
    if (r) return r;
    HIP_TRY(hipMemset(J.diag, 0, bytes), "clear diagnostics");
    b->diag[j] = J.diag;
    b->nbC[j] = J.nbC;
    b->loops0[j] = J.loops;
  }
  // factor tables, chunk-major, one (cp, 1/den) pair per element
  for (int j = 0; j < njobs; ++j) {
    FdzJob<T>& J = host_job<T>(b, j);
    const size_t per = (size_t)J.nn * J.K * 64 * J.wpr * sizeof(double2);
    const int nops = J.scheme == MLV_FDZ_EXPLICIT ? 1 : 3;
    for (int o = 0; o < nops; ++o) {
      int r = dalloc(per, (void**)&J.fac[o]);
      if (r) return r;
    }
  }
  for (int p = 0; p < 2; ++p) {
    fl.ncol[p] = pcol[p][njobs];
    int r = upload(pcol[p], &fl.prefcol[p]);
    if (r) return r;
  }
  for (int p = 0; p < 3; ++p)
    for (int v = 0; v < kFdzVariants; ++v) {
      fl.nrow[p][v] = prow[p][v][njobs];
      if (fl.nrow[p][v] == 0) continue;
      int r = upload(prow[p][v], &fl.prefrow[p][v]);
      if (r) return r;
    }
  int r = MLV_OK;
  r = dalloc(b->host.size(), &b->jobs_dev);
  if (r) return r;
  HIP_TRY(hipMemcpy(b->jobs_dev, b->host.data(), b->host.size(), hipMemcpyHostToDevice),
          "upload jobs");
  for (int j = 0; j < njobs; ++j) HIP_TRY(job_factors<T>(b, j, true, nullptr), "factor tables");
  HIP_TRY(hipDeviceSynchronize(), "factor tables");
  HIP_TRY(fdz_prepare<T>(fl), "kernel attributes");
  (void)dtype;
  return MLV_OK;
}

// A job's CFL check loops: the loops whose count before them (the job's
// loops at creation plus the batch's steps) is a multiple of cadence --
// loops 1, 1+cadence, ... of Simulation.end_loop's ticker, counted from
// the start of the simulation, so a resumed job keeps its phase.
bool any_check(const FdzHost* b, long long step, int cadence) {
  if (cadence <= 0) return false;
  for (long long l0 : b->loops0)
    if ((l0 + step) % cadence == 0) return true;
  return false;
}

// Steps first .. first+nsteps-1 (counted from the batch's creation).  With
// cadence > 0 the steps that are some job's check loop also produce that
// job's diagnostics partials and are followed by the on-device CFL update.
template <typename T>
hipError_t enqueue(FdzHost* b, int nsteps, int write_phys, long long first, hipStream_t s,
                   bool timed, int cadence, double cutoff) {
  FdzJob<T>* jd = (FdzJob<T>*)b->jobs_dev;
  const cx<T>* tw = (const cx<T>*)b->tw;
  for (int k = 0; k < nsteps; ++k) {
    const int wp = (write_phys && k == nsteps - 1) ? kFdzWritePhys : 0;
    const bool dg = any_check(b, first + k, cadence);
    const int flags0 = wp | (dg ? (kFdzDiag | (cadence << kFdzDiagShift)) : 0);
    const int step = (int)((first + k) & 3);
    for (int ph = 0; ph < b->nphase; ++ph) {
      hipEvent_t* ev = timed ? prof_slots(b) : nullptr;
      hipError_t e = fdz_phase<T>(jd, b->fl, ph, ph == 0 ? flags0 : 0, step, tw, s, ev);
      if (e != hipSuccess) return e;
    }
    if (dg) {
      hipError_t e = fdz_diag<T>(jd, b->njobs, b->max_nn, cutoff, cadence, s);
      if (e != hipSuccess) return e;
    }
  }
  return hipSuccess;
}

// dt and the theta operators' bands of job j, on the host copy and the device.
template <typename T>
hipError_t set_dt_impl(FdzHost* b, int j, double dt, hipStream_t s) {
  FdzJob<T>& J = host_job<T>(b, j);
  const double idz2 = 1.0 / (J.dz * J.dz);
  J.dt = dt;
  J.lo[1] = -(J.alpha * dt * J.Pr) * idz2;
  J.lo[2] = -(J.alpha * dt * 1.0) * idz2;
  hipError_t e = fdz_set_dt<T>((FdzJob<T>*)b->jobs_dev, j, dt, J.lo[1], J.lo[2], s);
  if (e == hipSuccess) e = job_factors<T>(b, j, false, s);
  return e;
}

hipError_t enqueue_steps(FdzHost* b, int nsteps, int write_phys, long long first, hipStream_t s,
                         bool timed, int cadence = 0, double cutoff = 0.5) {
  return b->dtype == MLV_F64
             ? enqueue<double>(b, nsteps, write_phys, first, s, timed, cadence, cutoff)
             : enqueue<float>(b, nsteps, write_phys, first, s, timed, cadence, cutoff);
}

int prime(FdzHost* b, hipStream_t s);

int graph_run(FdzHost* b, int nsteps, int cadence, double cutoff, void* stream) {
  if (nsteps <= 0 || (nsteps & 3)) return fail(MLV_ERR_ARG, "graph steps must be a positive multiple of 4");
  if (b->steps & 3)
    return fail(MLV_ERR_ARG, "graph replay needs the batch's step count at a multiple of 4");
  if (cadence > 0 && ((nsteps % cadence) || (b->steps % cadence)))
    return fail(MLV_ERR_ARG, "graph runs with CFL checks need nsteps and the step count at "
                             "multiples of the cadence");
  if (!b->primed) {
    int rc = prime(b, (hipStream_t)stream);
    if (rc) return rc;
  }
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

template <typename T>
int read_fdz_job(FdzHost* b, int job, double* dt, double* t, long long* loops) {
  FdzJob<T> J;
  HIP_TRY(hipMemcpy(&J, (const char*)b->jobs_dev + job * sizeof(FdzJob<T>), sizeof(J),
                    hipMemcpyDeviceToHost),
          "read job");
  if (dt) *dt = J.dt;
  if (t) *t = J.t;
  if (loops) *loops = J.loops;
  // keep the host copy's dt (and theta bands) in step for set_dt
  FdzJob<T>& H = host_job<T>(b, job);
  H.dt = J.dt;
  H.lo[1] = J.lo[1];
  H.lo[2] = J.lo[2];
  return MLV_OK;
}

int prime(FdzHost* b, hipStream_t s) {
  const int step = (int)(b->steps & 3);
  hipError_t e = b->dtype == MLV_F64
                     ? fdz_prologue<double>((const FdzJob<double>*)b->jobs_dev, b->fl, step, s)
                     : fdz_prologue<float>((const FdzJob<float>*)b->jobs_dev, b->fl, step, s);
  if (e != hipSuccess) return hip_fail(e, "fdz prologue");
  b->primed = true;
  return MLV_OK;
}

}  // namespace

extern "C" {

int mlv_fdz_work_bytes(int nx, int nz, int dtype, size_t* bytes) {
  if (!bytes) return fail(MLV_ERR_ARG, "bytes is NULL");
  if (nx < 8 || nz < 5) return fail(MLV_ERR_ARG, "grid too small");
  if (dtype != MLV_F32 && dtype != MLV_F64) return fail(MLV_ERR_ARG, "dtype must be F32 or F64");
  *bytes = fdz_work_elems(nx, nz) * cplx_bytes(dtype);
  return MLV_OK;
}

int mlv_fdz_batch_create(mlv_fdz_batch** out, int dtype, const mlv_fdz_job_desc* jobs, int njobs) {
  if (!out || !jobs || njobs <= 0) return fail(MLV_ERR_ARG, "bad batch arguments");
  *out = nullptr;
  if (dtype != MLV_F32 && dtype != MLV_F64) return fail(MLV_ERR_ARG, "dtype must be F32 or F64");
  // validate every descriptor before touching the device
  for (int j = 0; j < njobs; ++j) {
    FdzJob<double> probe;
    int rc = build_fdz_job<double>(jobs[j], probe, j);
    if (rc) return rc;
  }
  FdzHost* b = new FdzHost();
  b->dtype = dtype;
  b->njobs = njobs;
  int rc = MLV_OK;
  if (hipGetDevice(&b->device) != hipSuccess) rc = fail(MLV_ERR_HIP, "hipGetDevice");
  if (!rc) rc = get_twiddles(b->device, dtype, &b->tw);
  if (!rc) rc = dtype == MLV_F64 ? create_impl<double>(b, jobs, njobs)
                                 : create_impl<float>(b, jobs, njobs);
  if (rc) {
    for (void* p : b->allocs) (void)hipFree(p);
    delete b;
    return rc;
  }
  *out = reinterpret_cast<mlv_fdz_batch*>(b);
  return MLV_OK;
}

int mlv_fdz_batch_destroy(mlv_fdz_batch* bp) {
  if (!bp) return MLV_OK;
  FdzHost* b = reinterpret_cast<FdzHost*>(bp);
  (void)hipDeviceSynchronize();
  for (void* p : b->allocs) (void)hipFree(p);
  for (hipEvent_t e : b->events) (void)hipEventDestroy(e);
  if (b->graph_exec) (void)hipGraphExecDestroy(b->graph_exec);
  if (b->graph) (void)hipGraphDestroy(b->graph);
  if (b->cap_stream) (void)hipStreamDestroy(b->cap_stream);
  delete b;
  return MLV_OK;
}

int mlv_fdz_batch_prime(mlv_fdz_batch* bp, void* stream) {
  if (!bp) return fail(MLV_ERR_ARG, "batch is NULL");
  return prime(reinterpret_cast<FdzHost*>(bp), (hipStream_t)stream);
}

int mlv_fdz_batch_step(mlv_fdz_batch* bp, int nsteps, int write_phys, void* stream) {
  if (!bp) return fail(MLV_ERR_ARG, "batch is NULL");
  if (nsteps < 0) return fail(MLV_ERR_ARG, "nsteps < 0");
  FdzHost* b = reinterpret_cast<FdzHost*>(bp);
  if (!b->primed) {
    int rc = prime(b, (hipStream_t)stream);
    if (rc) return rc;
  }
  hipError_t e = enqueue_steps(b, nsteps, write_phys, b->steps, (hipStream_t)stream, b->prof);
  if (e != hipSuccess) return hip_fail(e, "fdz step");
  b->steps += nsteps;
  return MLV_OK;
}

int mlv_fdz_batch_graph_step(mlv_fdz_batch* bp, int nsteps, void* stream) {
  if (!bp) return fail(MLV_ERR_ARG, "batch is NULL");
  return graph_run(reinterpret_cast<FdzHost*>(bp), nsteps, 0, 0.0, stream);
}

int mlv_fdz_batch_run(mlv_fdz_batch* bp, int nsteps, int cadence, double cutoff, int graph,
                      void* stream) {
  if (!bp) return fail(MLV_ERR_ARG, "batch is NULL");
  if (nsteps < 0 || cadence < 0 || !(cutoff > 0)) return fail(MLV_ERR_ARG, "bad run arguments");
  FdzHost* b = reinterpret_cast<FdzHost*>(bp);
  if (graph) return graph_run(b, nsteps, cadence, cutoff, stream);
  if (!b->primed) {
    int rc = prime(b, (hipStream_t)stream);
    if (rc) return rc;
  }
  hipError_t e =
      enqueue_steps(b, nsteps, 0, b->steps, (hipStream_t)stream, b->prof, cadence, cutoff);
  if (e != hipSuccess) return hip_fail(e, "fdz run");
  b->steps += nsteps;
  return MLV_OK;
}

int mlv_fdz_batch_job_time(mlv_fdz_batch* bp, int job, double* dt, double* t, long long* loops,
                           int* breach) {
  if (!bp) return fail(MLV_ERR_ARG, "batch is NULL");
  FdzHost* b = reinterpret_cast<FdzHost*>(bp);
  if (job < 0 || job >= b->njobs) return fail(MLV_ERR_ARG, "job index out of range");
  HIP_TRY(hipDeviceSynchronize(), "synchronize");
  int rc = b->dtype == MLV_F64 ? read_fdz_job<double>(b, job, dt, t, loops)
                               : read_fdz_job<float>(b, job, dt, t, loops);
  if (rc) return rc;
  if (breach) {
    double h[2];
    HIP_TRY(hipMemcpy(h, b->diag[job] + (size_t)kDiagPart * b->nbC[job], sizeof(h),
                      hipMemcpyDeviceToHost),
            "read diagnostics");
    *breach = h[1] != 0.0;
  }
  return MLV_OK;
}

int mlv_fdz_batch_tracker(mlv_fdz_batch* bp, int job, double* out, int max_records, int* nrec) {
  if (!bp || !nrec) return fail(MLV_ERR_ARG, "NULL argument");
  FdzHost* b = reinterpret_cast<FdzHost*>(bp);
  if (job < 0 || job >= b->njobs) return fail(MLV_ERR_ARG, "job index out of range");
  HIP_TRY(hipDeviceSynchronize(), "synchronize");
  return copy_tracker(b->diag[job] + (size_t)kDiagPart * b->nbC[job], out, max_records, nrec);
}

int mlv_fdz_batch_set_dt(mlv_fdz_batch* bp, int job, double dt, void* stream) {
  if (!bp) return fail(MLV_ERR_ARG, "batch is NULL");
  FdzHost* b = reinterpret_cast<FdzHost*>(bp);
  if (job < 0 || job >= b->njobs) return fail(MLV_ERR_ARG, "job index out of range");
  if (!(dt > 0)) return fail(MLV_ERR_ARG, "dt must be positive");
  hipStream_t s = (hipStream_t)stream;
  hipError_t e = b->dtype == MLV_F64 ? set_dt_impl<double>(b, job, dt, s)
                                     : set_dt_impl<float>(b, job, dt, s);
  if (e != hipSuccess) return hip_fail(e, "set dt");
  return MLV_OK;
}

long long mlv_fdz_batch_steps(const mlv_fdz_batch* bp) {
  if (!bp) return -1;
  return reinterpret_cast<const FdzHost*>(bp)->steps;
}

int mlv_fdz_batch_profile(mlv_fdz_batch* bp, int enable) {
  if (!bp) return fail(MLV_ERR_ARG, "batch is NULL");
  FdzHost* b = reinterpret_cast<FdzHost*>(bp);
  int rc = prof_collect(b);
  if (rc) return rc;
  b->prof = enable != 0;
  b->ms[0] = b->ms[1] = 0;
  b->launches[0] = b->launches[1] = 0;
  return MLV_OK;
}

int mlv_fdz_batch_profile_read(mlv_fdz_batch* bp, double* ms, long long* launches) {
  if (!bp || !ms || !launches) return fail(MLV_ERR_ARG, "NULL argument");
  FdzHost* b = reinterpret_cast<FdzHost*>(bp);
  int rc = prof_collect(b);
  if (rc) return rc;
  for (int k = 0; k < 2; ++k) {
    ms[k] = b->ms[k];
    launches[k] = b->launches[k];
  }
  return MLV_OK;
}

}  // extern "C"
