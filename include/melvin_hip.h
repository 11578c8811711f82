/*
 * melvin_hip.h -- C ABI of libmelvin_hip.so, the MI355X (gfx950) kernels
 * behind the melvin_amd array backend.
 *
 * Every entry point takes raw device pointers and a HIP stream (passed as
 * void*, a hipStream_t; NULL = the null stream), returns an int status
 * (MLV_OK = 0) and never allocates device memory except in the *_create
 * calls.  On failure mlv_last_error() returns a thread-local message.
 * Calls are asynchronous on the stream.  Plans and batches are not
 * thread-safe.
 *
 * Layouts (C order, z contiguous):
 *   fully spectral  spectral (2nn+1, nm) complex, nn = (nx-1)/3, nm = (nz-1)/3;
 *                   row r <-> kx = r (r <= nn) or r - (2nn+1); column <-> kz
 *   Fourier-x/FD-z  spectral (nn, nz) complex, row <-> kx in [0, nn)
 *   physical        (nx, nz) real
 * "complex" is interleaved (re, im) of the plan's real type (float or
 * double).  The Laplacian symbol (MLV_SPEC_LAP) and semi-implicit linear
 * operators are always float64, as in the reference.
 *
 * Reference interfaces cited as path:line in JamieJQuinn/Melvin.py.
 */
#ifndef MELVIN_HIP_H
#define MELVIN_HIP_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* status codes */
#define MLV_OK 0
#define MLV_ERR_ARG 1
#define MLV_ERR_HIP 2
#define MLV_ERR_UNSUPPORTED 3

/* discretisation (Parameters.py:22, 57-61) */
#define MLV_FULLY_SPECTRAL 0
#define MLV_FOURIER_X_FD_Z 1

/* real type of the plan (Parameters.py:84-89) */
#define MLV_F32 0
#define MLV_F64 1

/* spectral operators (SpatialDifferentiator.py:50-74, Variable.py:111-113,
 * LaplacianSolver.py:58-68) */
#define MLV_SPEC_DDX 0
#define MLV_SPEC_DDZ 1
#define MLV_SPEC_D2DX2 2
#define MLV_SPEC_D2DZ2 3
#define MLV_SPEC_LAP 4
#define MLV_SPEC_SNABLA2 5
#define MLV_SPEC_POISSON 6

/* bases of the spectral axes (BasisFunctions.py:5-9) */
#define MLV_BASIS_COMPLEX_EXP 0
#define MLV_BASIS_COSINE 1
#define MLV_BASIS_SINE 2

/* physical finite differences (SpatialDifferentiator.py:76-185) */
#define MLV_FD_DDX 0
#define MLV_FD_DDZ 1
#define MLV_FD_D2DZ2 2

/* time integration (Integrator.py:53-63, 78-83) */
#define MLV_SCHEME_SEMI_IMPLICIT 0
#define MLV_SCHEME_EXPLICIT 1

/* FP64 reductions (Integrator.py:38-39 xp.max, utility.py:58 xp.sum,
 * double_diffusive_convection.py:22 xp.mean) */
#define MLV_RED_MAX 0
#define MLV_RED_SUM 1
#define MLV_RED_SUMSQ2 2 /* sum(a*a + b*b) */
#define MLV_RED_DOT 3    /* sum(a*b) */

/* fused step problems: the example loop bodies */
#define MLV_PROB_VORTICITY 0 /* taylor_green_vortex.py:87-93, vortex_pairs.py:122-128 */
#define MLV_PROB_VORT_SCALAR 1 /* kelvin_helmholtz_instability.py:117-129, ink enabled */
#define MLV_PROB_DDC 2       /* double_diffusive_convection.py:102-124 */
#define MLV_PROB_MHD 3       /* resistive_tearing_instability.py:127-146 */

typedef struct mlv_plan mlv_plan;
typedef struct mlv_batch mlv_batch;

/* Thread-local message of the last failure; never NULL. */
const char* mlv_last_error(void);
/* ABI version, bumped on any signature change. */
int mlv_abi_version(void);

/* ---------------------------------------------------------------------
 * Plans: one grid, discretisation and precision.  Replaces the state the
 * reference keeps in Parameters (Parameters.py:63-92), ArrayFactory
 * (ArrayFactory.py:8-26), SpatialDifferentiator (SpatialDifferentiator.py:
 * 8-48) and the FD-z LaplacianSolver factorisation (LaplacianSolver.py:
 * 21-56).  nx must be a power of two in [8, 4096]; nz too when fully
 * spectral, else any nz >= 3.
 * ------------------------------------------------------------------- */
int mlv_plan_create(mlv_plan** plan, int nx, int nz, double lx, double lz, int discretisation,
                    int dtype, int fd_order);
int mlv_plan_destroy(mlv_plan* plan);
/* nn, nm (nm = nz for FD-z), spectral rows and columns */
int mlv_plan_info(const mlv_plan* plan, int* nn, int* nm, int* spec_rows, int* spec_cols);

/* SpectralTransformer.to_physical (SpectralTransformer.py:33-61, 90-150) */
int mlv_to_physical(mlv_plan* plan, const void* spec, void* phys, void* stream);
/* SpectralTransformer.to_spectral (SpectralTransformer.py:63-88, 152-199) */
int mlv_to_spectral(mlv_plan* plan, const void* phys, void* spec, void* stream);
/* numpy.fft real transforms of a row-major (nx, nz) array, the xp.fft
 * calls the reference's transformer makes (SpectralTransformer.py:58 irfft,
 * 85 rfft, 132 irfft2, 191 rfft2).  axes: MLV_FFT_AXIS0 along x (rfft:
 * real (nx, nz) -> complex (nx/2+1, nz)), MLV_FFT_AXIS1 along z (real
 * (nx, nz) -> complex (nx, nz/2+1)), MLV_FFT_AXES01 both (rfft2: -> complex
 * (nx, nz/2+1), kx in FFT order; `work` holds nx * (nz/2+1) complex
 * values).  Forward unscaled; mlv_irfft is the inverse with numpy's 1/n
 * scaling, nx and nz the REAL array's dimensions, and ignores the
 * imaginary parts of the DC and Nyquist bins along the real axis as numpy
 * does.  Transformed lengths: powers of two in [8, 4096]. */
#define MLV_FFT_AXIS0 0
#define MLV_FFT_AXIS1 1
#define MLV_FFT_AXES01 2
int mlv_rfft(int dtype, int axes, int nx, int nz, const void* in, void* out, void* work,
             void* stream);
int mlv_irfft(int dtype, int axes, int nx, int nz, const void* in, void* out, void* work,
              void* stream);
/* SpatialDifferentiator sddx/sddz/sd2dx2/sd2dz2/calc_lap, Variable.snabla2
 * and the fully spectral LaplacianSolver.solve; `in` is spectral, `out` is
 * spectral (complex) or, for MLV_SPEC_LAP, float64 with `in` ignored.  In
 * FD-z, MLV_SPEC_D2DZ2 / MLV_SPEC_SNABLA2 use the FD z-derivative
 * (SpatialDifferentiator.py:35-40). */
int mlv_spec_op(mlv_plan* plan, int kind, const void* in, void* out, void* stream);
/* The same three calls with a basis per axis (MLV_BASIS_COMPLEX_EXP, _COSINE
 * or _SINE; fully spectral plans only): SpectralTransformer.to_spectral /
 * to_physical with basis_functions (SpectralTransformer.py:90-199) and the
 * spectral operators with the factors of BasisFunctions.py:26-59 (bx for the
 * x factors, bz for the z factors; MLV_SPEC_POISSON divides by that basis's
 * Laplacian symbol, LaplacianSolver.py:58-68).  Cosine / sine axes are dense
 * per-axis transforms of the mirrored signal, built once per plan and basis.
 * to_physical does not double the caller's k = 0 cosine terms in place as
 * the reference does (SpectralTransformer.py:117-120); its result is that of
 * the reference. */
int mlv_to_spectral_basis(mlv_plan* plan, int bx, int bz, const void* phys, void* spec,
                          void* stream);
int mlv_to_physical_basis(mlv_plan* plan, int bx, int bz, const void* spec, void* phys,
                          void* stream);
int mlv_spec_op_basis(mlv_plan* plan, int kind, int bx, int bz, const void* in, void* out,
                      void* stream);
/* SpatialDifferentiator.pddx/pddz/pd2dz2 on a rows x cols array of real
 * (is_complex = 0) or complex values (SpatialDifferentiator.py:76-185). */
int mlv_fd_op(mlv_plan* plan, int kind, int is_complex, int rows, int cols, const void* in,
              void* out, void* stream);
/* pddx(ux*f) + pddz(uz*f), physical (Variable.py:125-127) */
int mlv_flux_divergence(mlv_plan* plan, const void* ux, const void* uz, const void* f, void* out,
                        void* stream);
/* LaplacianSolver.solve: fully spectral divide (LaplacianSolver.py:58-68) or
 * the per-mode tridiagonal FD-z solve (LaplacianSolver.py:70-78), FP64
 * internally. */
int mlv_poisson(mlv_plan* plan, const void* rhs, void* out, void* stream);
/* Integrator.integrate for one variable: store nothing new, combine the
 * history ring (order 2 or 4, current slot `curr`) with Adams-Bashforth and
 * apply the semi-implicit (lin = float64 linear operator) or explicit
 * (lin = complex diffusion term, added into the current history slot)
 * update in place (Integrator.py:5-18, 53-63). */
int mlv_update(mlv_plan* plan, int scheme, int order, double alpha, double dt, const void* lin,
               void* var, void* hist, int curr, void* stream);
/* Integrator.predictor (moulton = 0: AB2/AB4, Integrator.py:5-18) or
 * Integrator.corrector (moulton = 1: AM2/AM4, Integrator.py:20-33) of the
 * history ring into out. */
int mlv_combine(mlv_plan* plan, int moulton, int order, double dt, const void* hist, int curr,
                void* out, void* stream);
/* calc_velocity_from_vorticity's spectral part (utility.py:62-79):
 * psi = solve(-w); fully spectral also ux_spec = -ddz(psi) and
 * uz_spec = ddx(psi); FD-z gives psi (tridiagonal solve) and uz_spec, and
 * ux_spec must be NULL (ux is the physical -pddz of psi there).  Outputs
 * may be NULL except psi. */
int mlv_velocity(mlv_plan* plan, const void* w, void* psi, void* ux_spec, void* uz_spec,
                 void* stream);
/* FP64 reduction of n values of type dtype (b used by SUMSQ2/DOT); writes
 * scale * result to out_f64 (device).  plan may be NULL: the call then uses
 * its device's shared partials buffer, which the first *_create call on the
 * device allocates (a plan-less call on a device where nothing was created
 * yet allocates it).  Reductions that share a partials buffer (the same
 * plan, or NULL on one device) must not overlap on different streams.
 * Replaces the reference's xp.max / xp.sum / xp.mean calls
 * (Integrator.py:35-44, utility.py:42-59). */
int mlv_reduce(mlv_plan* plan, int op, int dtype, const void* a, const void* b, size_t n,
               double scale, void* out_f64, void* stream);
/* mlv_reduce without a plan that never allocates device memory: no
 * counterpart in the reference; it is for callers that cannot allocate
 * (code under HIP graph capture, serving loops and harnesses that forbid
 * hipMalloc).  Same kernels and results as mlv_reduce, bit for bit.  The
 * first-stage partials live in a fixed array of the library, 1024
 * (sum, compensation) pairs = 16 KiB on each device it runs on, held until
 * the process ends; at most that many partials are used for any n >= 1.
 * Every call on one device shares that array: calls that overlap on one
 * device (other streams, other host threads) race, and the caller must
 * serialise them. */
int mlv_reduce_noalloc(int op, int dtype, const void* a, const void* b, size_t n, double scale,
                       void* out_f64, void* stream);

/* ---------------------------------------------------------------------
 * Fused steps over a table of jobs (independent simulations).  One call
 * advances every job one full loop body of its problem: the velocity solve
 * (utility.py:62-79), every advected field's vec_dot_nabla (Variable.py:
 * 119-128) and source terms, and Integrator.integrate (Integrator.py:5-18,
 * 58-63), i.e. the bodies cited at MLV_PROB_*.  Fully spectral only.
 * ------------------------------------------------------------------- */
typedef struct mlv_job_desc {
  int problem;         /* MLV_PROB_* */
  int nx, nz;          /* powers of two in [8, 4096] */
  double lx, lz;
  int fd_order;        /* 2 or 4 */
  int ab_order;        /* 2 or 4 */
  double alpha;        /* semi-implicit weight (Parameters.py:13) */
  double dt;
  double coeff[4];     /* diffusion coefficient per field equation */
  double param[4];     /* DDC: Pr, R0 */
  int nfield;
  /* The fused batch keeps its spectral arrays kz-major: (nm, 2nn+1), the
   * element of row r (kx) and column m (kz) at m * (2nn+1) + r, so that a kz
   * column is contiguous for both x passes.  This applies to state, hist,
   * comp and psi. */
  void* state[2][4];   /* ping-pong spectral state per field */
  void* hist[4];       /* AB ring per field: ab_order x spectral */
  void* comp[4];       /* F32 only, optional: per-field Kahan residual of the
                          state (value = state - residual) as bfloat16 (re, im)
                          pairs, 4 bytes per spectral element,
                          zero-initialised; or NULL */
  void* work;          /* mlv_job_work_bytes() bytes, 16-byte aligned */
  void* psi;           /* optional spectral psi of the step, or NULL */
  void* phys[8];       /* optional physical outputs (problem order), or NULL */
  int curr;            /* ring slot the next step writes */
  int parity;          /* state buffer holding the current state */
  double t;            /* simulated time so far (advanced on the device) */
  long long loops;     /* loops so far (advanced on the device) */
  int mean_split;      /* F32 only: carry each z-pass channel's x mean (kx = 0
                          row) apart from its fluctuation, so that products
                          with an O(1) x-independent profile keep FP32
                          accuracy relative to the fluctuation; costs a few
                          per cent of the step.  Ignored for F64. */
} mlv_job_desc;

int mlv_job_work_bytes(int problem, int nfield, int nx, int nz, int dtype, size_t* bytes);
int mlv_batch_create(mlv_batch** batch, int dtype, const mlv_job_desc* jobs, int njobs);
int mlv_batch_destroy(mlv_batch* batch);
/* Replace the batch's job set (jobs leave, new jobs join, the others keep
 * going): the same tables mlv_batch_create builds, written into the
 * batch's device memory, which grows only when the new set needs more.
 * Each job that was already in the batch (same workspace pointer) keeps its
 * tracker records; the jobs' clocks, ring slots and dt come from the
 * descriptors as at creation, and the batch's step count restarts at 0.
 * Synchronises the device first (queued steps may still read the old
 * tables) and drops a captured graph.  The kernel choices become those of
 * a new batch of the same jobs (the autotuner's cached choice for the job
 * set, else the defaults), and the kernels compiled for one grid size are
 * used when the new set shares one.  On an invalid descriptor the batch is
 * left unchanged.  Serves continuous batching (the reference has no
 * batches). */
int mlv_batch_rebuild(mlv_batch* batch, const mlv_job_desc* jobs, int njobs);
/* Run nsteps steps; write_phys = 1 also writes the physical fields of the
 * last of them.  Each step reads dt from the job table. */
int mlv_batch_step(mlv_batch* batch, int nsteps, int write_phys, void* stream);
/* Same as mlv_batch_step without physical outputs, replayed from a HIP
 * graph captured on first use (re-captured when nsteps changes).  nsteps
 * must be a multiple of 4 and the batch's step count a multiple of 4 on
 * entry. */
int mlv_batch_graph_step(mlv_batch* batch, int nsteps, void* stream);
/* Set one job's dt for the following steps (stream-ordered). */
int mlv_batch_set_dt(mlv_batch* batch, int job, double dt, void* stream);
/* Pick each fused kernel's variant for this batch by measurement (the
 * tuning the reference leaves to its array library's FFT plans,
 * SpectralTransformer.py:58,85,132,191): kernel A's block shape (0: 512 threads in
 * single precision, 256 in double; 1: half that, where every job's x
 * transform fits; 2: shape 0's block holding two columns of every op of
 * the stage side by side, so the ops derived from one field read it once --
 * single precision, batches of one x size, 512 or 1024), kernel B's plan
 * (0: split-pass -- inverse 16,16,R with
 * the products in its last pass, forward R,16,16 -- where a job's z
 * transform takes it, else the two-sweep plan; 1: the two-sweep plan
 * everywhere) and kernel C's tile shape (0 narrow, 1 wide: 32 points per
 * thread where it applies, 2 narrow in 512-thread blocks: single
 * precision, batches of one x size in [512, 2048]).  Runs nsteps real steps with every kernel at
 * its first option, its second, .. up to the most options any kernel has,
 * then back down (half of nsteps per segment; a kernel with fewer options
 * stays at its first in the extra segments), with events around each
 * kernel, so the steps advance the simulations as mlv_batch_step would; a
 * kernel's option figure is its fastest segment per step.  Keeps each kernel's faster option and remembers the choice for
 * batches created later for the same set of job shapes (process-wide).
 * chosen (3 ints, or NULL) receives {shape A, plan B, shape C}.  A's shapes
 * and C's shapes give bit-identical results; B's plans agree to rounding.
 * Synchronises. */
int mlv_batch_autotune(mlv_batch* batch, int nsteps, void* stream, int* chosen);
/* The batch's kernel choices: plan[0..2] = {shape A, plan B, shape C} as
 * in mlv_batch_autotune, plan[3] = 1 if A's smaller shape is available,
 * plan[4] = 1 if some job's z pass takes the split-pass plan, plan[5] = 1
 * if A's shape 2 is available, plan[6] = 1 if C's shape 2 is: plan has 7
 * ints.  No counterpart in the reference (its kernels are library calls). */
int mlv_batch_plan(const mlv_batch* batch, int* plan);
/* Set the batch's kernel choices (values as in mlv_batch_autotune);
 * MLV_ERR_UNSUPPORTED for an x-inverse shape that does not apply.
 * Drops a captured graph.  No counterpart in the reference. */
int mlv_batch_set_plan(mlv_batch* batch, int shape_xinv, int plan_zpass, int wide_update);
/* nsteps loops with Simulation.end_loop's CFL ticker on the device
 * (Simulation.py:199-209, Ticker.py:14-25, Integrator.py:35-44): after each
 * job's loops 1, 1+cadence, ... -- counted from the job descriptor's loops,
 * so a resumed simulation keeps its phase -- the job's maxima of its
 * physical ux, uz (FP64, NaN-propagating, as xp.max) give
 * cfl_dt = min(dx / max ux, dz / max uz); a breach or NaN raises the job's
 * breach flag, else dt shrinks by 0.9 until dt <= cutoff * cfl_dt.  No host
 * synchronisation; graph = 1 replays a captured graph (nsteps and the step
 * count multiples of 4 and of cadence).  Each check also appends a tracker
 * record (loop, t, dt, kinetic energy as utility.py:42-59, max ux, max uz,
 * next dt). */
int mlv_batch_run(mlv_batch* batch, int nsteps, int cadence, double cutoff, int graph,
                  void* stream);
/* One job's device-side dt, t, loop count and CFL breach flag
 * (synchronises the device). */
int mlv_batch_job_time(mlv_batch* batch, int job, double* dt, double* t, long long* loops,
                       int* breach);
/* Up to max_records tracker records of one job, oldest first, 9 doubles
 * each: loop, t, dt of the step, kinetic energy, max ux, max uz, next dt,
 * enstrophy 0.5 mean(w^2), and mean(f uz) of the job's first advected scalar
 * f (NaN without one; double_diffusive_convection.py:20-23's Nusselt number
 * is 1 minus it).  The last 4096 records are kept (synchronises). */
int mlv_batch_tracker(mlv_batch* batch, int job, double* out, int max_records, int* n_records);
/* Steps completed since creation (host count). */
long long mlv_batch_steps(const mlv_batch* batch);
/* Kernel timing (no reference counterpart; measurement only): while enabled,
 * every step brackets its three kernels with HIP events on the step's
 * stream.  Enabling or disabling resets the totals. */
int mlv_batch_profile(mlv_batch* batch, int enable);
/* Milliseconds and launch counts of the x-inverse, z and x-forward/update
 * kernels since profiling was enabled (synchronises with the events). */
int mlv_batch_profile_read(mlv_batch* batch, double* ms, long long* launches);

/* ---------------------------------------------------------------------
 * Fused Fourier-x / finite-difference-z steps over a table of jobs: the
 * Rayleigh-Benard loop body (rayleigh_benard_convection.py:95-145) --
 * LaplacianSolver.solve in FD-z (LaplacianSolver.py:70-78),
 * calc_velocity_from_vorticity (utility.py:62-79), vec_dot_nabla of omega
 * and T (Variable.py:119-128), buoyancy, Integrator.integrate and the
 * spectral wall writes -- for omega and T on (nn, nz) spectral rows.
 * ------------------------------------------------------------------- */
/* Integrator._explicit, the reference's FD-z scheme (Integrator.py:53-56,
 * 78-83) */
#define MLV_FDZ_EXPLICIT 0
/* theta-implicit second-order z diffusion solved per x mode (FP64
 * tridiagonal), Adams-Bashforth advection and buoyancy; no reference
 * counterpart (the reference has no semi-implicit FD-z path) */
#define MLV_FDZ_SEMI_IMPLICIT 1
/* Adams-Bashforth predictor and Adams-Moulton corrector (the formulas of
 * Integrator.py:5-33; the reference never calls the corrector) around the
 * same theta-implicit diffusion */
#define MLV_FDZ_PC 2

typedef struct mlv_fdz_batch mlv_fdz_batch;

typedef struct mlv_fdz_job_desc {
  int nx, nz;          /* nx a power of two in [8, 2048]; nz in [5, 768] */
  double lx, lz;       /* dz = lz / nz (Parameters.py:82) */
  int fd_order;        /* 2 or 4: z differences of advection and explicit
                          diffusion, the x difference, ux = -pddz(psi) */
  int ab_order;        /* 2 or 4 */
  int scheme;          /* MLV_FDZ_* */
  double alpha, dt, Pr, Ra;
  void* state[2][2];   /* [buffer][omega, T] spectral (nn, nz) */
  void* hist[2];       /* ring per field: ab_order x (nn, nz) */
  void* comp[2];       /* F32 only, optional: per-field residual of the state
                          (value = state - comp), zero-initialised, or NULL */
  void* psi[2];        /* optional, per buffer: the walled psi of that
                          buffer's omega, written by the step that produced
                          it (psi of a step = psi[buffer the step read]) */
  void* work;          /* mlv_fdz_work_bytes() bytes */
  void* phys[4];       /* optional physical ux, uz, omega, T (nx, nz) of the
                          state a step starts from, or NULL */
  int curr;            /* ring slot the next step writes */
  int parity;          /* state buffer holding the current state */
  double t;            /* simulation time (advanced on the device) */
  long long loops;     /* loop count (Simulation._loop_counter) */
} mlv_fdz_job_desc;

int mlv_fdz_work_bytes(int nx, int nz, int dtype, size_t* bytes);
int mlv_fdz_batch_create(mlv_fdz_batch** batch, int dtype, const mlv_fdz_job_desc* jobs,
                         int njobs);
int mlv_fdz_batch_destroy(mlv_fdz_batch* batch);
/* Velocity (and psi) of every job's current state; run by the first step
 * and needed again only after a state is replaced from outside. */
int mlv_fdz_batch_prime(mlv_fdz_batch* batch, void* stream);
/* Run nsteps steps; write_phys = 1 also writes the physical fields of the
 * last of them.  Each step reads dt from the job table. */
int mlv_fdz_batch_step(mlv_fdz_batch* batch, int nsteps, int write_phys, void* stream);
/* As mlv_batch_graph_step, for FD-z batches. */
int mlv_fdz_batch_graph_step(mlv_fdz_batch* batch, int nsteps, void* stream);
/* Set one job's dt (and its implicit operators) for the following steps
 * (Integrator.override_dt, Integrator.py:46-48). */
int mlv_fdz_batch_set_dt(mlv_fdz_batch* batch, int job, double dt, void* stream);
/* As mlv_batch_run, for FD-z batches: the CFL ticker on the device from the
 * physical ux, uz of each check loop's starting state (as
 * rayleigh_benard_convection.py's Simulation.end_loop sees them); a changed
 * dt also rebuilds the job's theta operators on the device. */
int mlv_fdz_batch_run(mlv_fdz_batch* batch, int nsteps, int cadence, double cutoff, int graph,
                      void* stream);
/* As mlv_batch_job_time and mlv_batch_tracker, for FD-z batches (the
 * scalar flux column is mean(T uz)). */
int mlv_fdz_batch_job_time(mlv_fdz_batch* batch, int job, double* dt, double* t,
                           long long* loops, int* breach);
int mlv_fdz_batch_tracker(mlv_fdz_batch* batch, int job, double* out, int max_records,
                          int* n_records);
long long mlv_fdz_batch_steps(const mlv_fdz_batch* batch);
/* Kernel timing (measurement only): milliseconds and launch counts of the
 * column (x transform) and row (z solve / update) kernels. */
int mlv_fdz_batch_profile(mlv_fdz_batch* batch, int enable);
int mlv_fdz_batch_profile_read(mlv_fdz_batch* batch, double* ms, long long* launches);

#ifdef __cplusplus
}
#endif
#endif /* MELVIN_HIP_H */
