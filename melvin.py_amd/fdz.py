"""Fused Fourier-x / finite-difference-z steps over batches of simulations.

An ``FdzJob`` is one Rayleigh-Benard simulation -- omega and T periodic in
x (Fourier, spectral rows kx in [0, nn)) and on nz finite-difference points
in z with the walls of rayleigh_benard_convection.py (T = 1 at the bottom,
T = 0 at the top, psi = omega = 0) -- whose state, Adams-Bashforth ring and
workspace live in GPU memory.  An ``FdzBatch`` advances any set of such jobs
(mixed sizes, parameters and schemes, one precision) with two HIP kernels
per evaluation of the explicit terms.

Schemes:

* ``"explicit"``: the reference's FD-z integrator (Integrator.py:53-56):
  the diffusion term is added into the stored RHS, AB2/AB4 update.  One step
  is the loop body of rayleigh_benard_convection.py:95-145.
* ``"semi-implicit"``: theta-implicit z diffusion (a tridiagonal solve per x
  mode) with Adams-Bashforth advection and buoyancy.  No reference
  counterpart: the reference has no semi-implicit FD-z path.
* ``"pc"``: Adams-Bashforth predictor, Adams-Moulton corrector (the
  formulas of Integrator.py:5-33, whose corrector the reference never
  calls) around the same theta-implicit diffusion.

The tridiagonal solves (Poisson and diffusion) run in FP64 in both
precisions.
"""
from __future__ import annotations

import ctypes
import math

import numpy as np
import torch

from . import _lib, fused

SCHEMES = {
    "explicit": _lib.FDZ_EXPLICIT,
    "semi-implicit": _lib.FDZ_SEMI_IMPLICIT,
    "pc": _lib.FDZ_PC,
}


class FdzJob:
    """One Rayleigh-Benard simulation's device state (omega, T)."""

    def __init__(self, nx, nz, lx, lz, Pr, Ra, dtype="f32", dt=1e-5, scheme="pc",
                 ab_order=2, fd_order=2, alpha=0.51, want_psi=False, want_phys=False,
                 compensated=True, device=None):
        if scheme not in SCHEMES:
            raise ValueError(f"unknown scheme {scheme!r}")
        if dtype not in ("f32", "f64"):
            raise ValueError("dtype must be 'f32' or 'f64'")
        self.nx, self.nz, self.lx, self.lz = int(nx), int(nz), float(lx), float(lz)
        self.Pr, self.Ra = float(Pr), float(Ra)
        self.dtype = dtype
        self.double = dtype == "f64"
        self.real = torch.float64 if self.double else torch.float32
        self.cplx = torch.complex128 if self.double else torch.complex64
        self.dt = float(dt)
        self.scheme = scheme
        self.ab_order = int(ab_order)
        self.fd_order = int(fd_order)
        self.alpha = float(alpha)
        self.nn = (self.nx - 1) // 3
        dev = device if device is not None else torch.device("cuda", torch.cuda.current_device())
        shape = (self.nn, self.nz)
        self.buf = [[torch.zeros(shape, dtype=self.cplx, device=dev) for _ in range(2)]
                    for _ in range(2)]
        self.hist = [torch.zeros((self.ab_order,) + shape, dtype=self.cplx, device=dev)
                     for _ in range(2)]
        # single precision: the value held is buf - carry (carry = rounding residual)
        self.carry = ([torch.zeros(shape, dtype=self.cplx, device=dev) for _ in range(2)]
                      if (compensated and not self.double) else None)
        nbytes = ctypes.c_size_t()
        _lib.call("mlv_fdz_work_bytes", self.nx, self.nz, _lib.F64 if self.double else _lib.F32,
                  ctypes.byref(nbytes))
        self.work = torch.zeros(nbytes.value, dtype=torch.uint8, device=dev)
        self.psi = ([torch.zeros(shape, dtype=self.cplx, device=dev) for _ in range(2)]
                    if want_psi else None)
        self.phys = ([torch.zeros((self.nx, self.nz), dtype=self.real, device=dev)
                      for _ in range(4)] if want_phys else None)
        self.curr = 0
        self.parity = 0
        self.steps = 0
        self.t = 0.0
        # the batch whose device job table holds dt and t after a run_device
        self.time_owner = None

    @property
    def spectral_shape(self):
        return (self.nn, self.nz)

    def hbm_bytes(self):
        ts = [b for row in self.buf for b in row] + self.hist + [self.work]
        ts += self.carry or []
        ts += self.psi or []
        ts += self.phys or []
        return sum(t.numel() * t.element_size() for t in ts)

    def set_state(self, fields):
        """Load (omega, T) spectra (host or device arrays) into the current
        buffer; a batch holding the job must be primed again afterwards."""
        if len(fields) != 2:
            raise ValueError("expected (omega, T)")
        for i, f in enumerate(fields):
            t = f if isinstance(f, torch.Tensor) else torch.as_tensor(np.asarray(f))
            if tuple(t.shape) != self.spectral_shape:
                raise ValueError(f"field {i} has shape {tuple(t.shape)}, expected "
                                 f"{self.spectral_shape}")
            t = t.to(self.buf[0][0].device, torch.complex128)
            hi = t.to(self.cplx)
            self.buf[self.parity][i].copy_(hi)
            if self.carry is not None:
                self.carry[i].copy_((hi.to(torch.complex128) - t).to(self.cplx))

    def state(self):
        """Current (omega, T) as complex128 host arrays."""
        out = []
        for i, b in enumerate(self.buf[self.parity]):
            v = b.detach().cpu().numpy().astype(np.complex128)
            if self.carry is not None:
                v = v - self.carry[i].detach().cpu().numpy().astype(np.complex128)
            out.append(v)
        return out

    def step_psi(self):
        """Walled streamfunction of the last step (of the state it started
        from), as the reference's psi variable after its loop body."""
        if self.psi is None:
            raise RuntimeError("job was created without psi outputs")
        b = self.parity ^ 1 if self.steps else self.parity
        return self.psi[b].detach().cpu().numpy().astype(np.complex128)

    def physical(self):
        """Physical ux, uz, omega, T of the state the last step with
        write_phys started from."""
        if self.phys is None:
            raise RuntimeError("job was created without physical outputs")
        return [p.detach().cpu().numpy().astype(np.float64) for p in self.phys]

    def velocity_phys(self):
        """Device tensors of the physical ux, uz (for the CFL check)."""
        return self.phys[0], self.phys[1]

    def desc(self):
        d = _lib.FdzJobDesc()
        d.nx, d.nz, d.lx, d.lz = self.nx, self.nz, self.lx, self.lz
        d.fd_order, d.ab_order, d.scheme = self.fd_order, self.ab_order, SCHEMES[self.scheme]
        d.alpha, d.dt, d.Pr, d.Ra = self.alpha, self.dt, self.Pr, self.Ra
        for b in range(2):
            for f in range(2):
                d.state[b][f] = self.buf[b][f].data_ptr()
        for f in range(2):
            d.hist[f] = self.hist[f].data_ptr()
            if self.carry is not None:
                d.comp[f] = self.carry[f].data_ptr()
            if self.psi is not None:
                d.psi[f] = self.psi[f].data_ptr()
        d.work = self.work.data_ptr()
        if self.phys is not None:
            for i, p in enumerate(self.phys):
                d.phys[i] = p.data_ptr()
        d.curr, d.parity = self.curr, self.parity
        d.t, d.loops = self.t, self.steps
        return d


class FdzBatch(fused.DeviceRun):
    """A set of FD-z jobs advanced together; all jobs share one precision.
    ``run_device`` / ``sync`` / ``tracker`` / ``save_trackers`` run the CFL
    ticker and diagnostics on the device (mlv_fdz_batch_run); the tracker's
    scalar flux is mean(T uz)."""

    _C = "mlv_fdz_batch"

    def __init__(self, jobs):
        if not jobs:
            raise ValueError("empty batch")
        if len({j.dtype for j in jobs}) != 1:
            raise ValueError("all jobs of a batch must share one precision")
        self.jobs = list(jobs)
        self.double = self.jobs[0].double
        arr = (_lib.FdzJobDesc * len(self.jobs))(*[j.desc() for j in self.jobs])
        h = ctypes.c_void_p()
        _lib.call("mlv_fdz_batch_create", ctypes.byref(h), _lib.F64 if self.double else _lib.F32,
                  arr, len(self.jobs))
        self.handle = h
        self._steps0 = [j.steps for j in self.jobs]

    def _advance(self, nsteps):
        for j in self.jobs:
            j.steps += nsteps
            j.curr = (j.curr + nsteps) % j.ab_order
            j.parity = (j.parity + nsteps) & 1
            j.t += nsteps * j.dt

    def prime(self, stream=None):
        """Recompute every job's velocity from its current state (after
        set_state on a job of the batch)."""
        _lib.call("mlv_fdz_batch_prime", self.handle, _lib.stream_handle(stream))

    def step(self, nsteps=1, write_phys=False, stream=None):
        """Queue ``nsteps`` steps of every job (asynchronous)."""
        if nsteps <= 0:
            return
        _lib.call("mlv_fdz_batch_step", self.handle, int(nsteps), 1 if write_phys else 0,
                  _lib.stream_handle(stream))
        self._advance(nsteps)

    def graph_step(self, nsteps, stream=None):
        """``nsteps`` (a multiple of 4) steps as one HIP graph launch."""
        if nsteps % 4:
            raise ValueError("graph steps must be a multiple of 4")
        if (self.jobs[0].steps - self._steps0[0]) % 4:
            raise ValueError("graph replay needs the batch step count at a multiple of 4")
        _lib.call("mlv_fdz_batch_graph_step", self.handle, int(nsteps),
                  _lib.stream_handle(stream))
        self._advance(nsteps)

    def set_dt(self, index, dt, stream=None):
        self.jobs[index].dt = float(dt)
        _lib.call("mlv_fdz_batch_set_dt", self.handle, int(index), float(dt),
                  _lib.stream_handle(stream))

    def profile(self, enable=True):
        _lib.call("mlv_fdz_batch_profile", self.handle, 1 if enable else 0)

    def profile_read(self):
        """(column ms, row ms), (column launches, row launches) since
        profiling was enabled."""
        ms = (ctypes.c_double * 2)()
        n = (ctypes.c_longlong * 2)()
        _lib.call("mlv_fdz_batch_profile_read", self.handle, ms, n)
        return [ms[0], ms[1]], [n[0], n[1]]

    def close(self):
        if getattr(self, "handle", None) is not None and self.handle.value:
            if any(getattr(j, "time_owner", None) is self for j in self.jobs):
                self.sync()  # keep the device-side dt and t of run_device
            _lib.lib().mlv_fdz_batch_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def grid_points(self):
        return sum(j.nx * j.nz for j in self.jobs)


def cfl_dt(job, dt, cutoff=0.5):
    """The reference's CFL rule (Integrator.py:35-44) on the job's physical
    velocities (needs ``want_phys`` and a step run with write_phys)."""
    from . import hip

    ux, uz = job.velocity_phys()
    dx, dz = job.lx / job.nx, job.lz / job.nz
    with np.errstate(divide="ignore", invalid="ignore"):
        c = min(np.float64(dx) / np.float64(hip.max(ux)), np.float64(dz) / np.float64(hip.max(uz)))
    if dt > c or math.isnan(c):
        raise Exception("CFL condition breached")
    while dt > cutoff * c:
        dt = dt * 0.9
    return dt


def run(batch, nsteps, cfl_cadence=10, cutoff=0.5, on_cfl=None):
    """``fused.run`` for FD-z batches: the reference's CFL ticker around
    the fused steps."""
    return fused.run(batch, nsteps, cfl_cadence, cutoff, on_cfl, cfl=cfl_dt)


def run_device(batch, nsteps, cfl_cadence=10, cutoff=0.5, chunk=None):
    """``fused.run_device`` for FD-z batches: the same CFL ticker evaluated
    on the device (a changed dt rebuilds the theta operators there too),
    captured graphs of whole cadence periods, one synchronisation."""
    return fused.run_device(batch, nsteps, cfl_cadence, cutoff, chunk)
