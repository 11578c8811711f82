"""Fused whole-step kernels over batches of simulations.

A ``Job`` is one fully spectral simulation (its state, Adams-Bashforth
ring and workspace live in GPU memory); a ``Batch`` advances any set of
jobs -- of mixed grid sizes and problems, one precision -- with three HIP
kernels per stage and step (two stages for MHD).  One step of a job is the
whole loop body of the matching reference example:

* ``"vorticity"``: omega plus up to two passive scalars sharing the
  velocity (taylor_green_vortex.py:87-93, vortex_pairs.py:122-128,
  kelvin_helmholtz_instability.py:117-129 with the ink block enabled);
* ``"ddc"``: omega, T, S of double_diffusive_convection.py:102-124;
* ``"mhd"``: omega and j of resistive_tearing_instability.py:127-146.

Every step reads dt from the device job table, so steps can be queued
back to back (and captured) without host round trips.
"""
from __future__ import annotations

import ctypes
import math

import numpy as np
import torch

from . import _lib

PROBLEMS = {
    "vorticity": _lib.PROB_VORTICITY,
    "ddc": _lib.PROB_DDC,
    "mhd": _lib.PROB_MHD,
}
FIELDS = {"ddc": 3, "mhd": 2}


def _phys_count(problem, nfield):
    return {"vorticity": nfield + 2, "ddc": 5, "mhd": 6}[problem]


class Job:
    """One simulation's device state.

    ``coeff[i]`` is the diffusion coefficient of field i's equation (the
    ``lin_op = coeff * lap()`` factor of the examples); ``param`` holds the
    problem constants (DDC: Pr, R0).
    """

    def __init__(self, problem, nx, nz, lx, lz, dtype="f32", dt=1e-3, coeff=(0.0,),
                 param=(), nfield=None, ab_order=2, fd_order=2, alpha=0.51,
                 want_psi=False, want_phys=False, compensated=True, mean_split=True,
                 device=None):
        if problem not in PROBLEMS:
            raise ValueError(f"unknown problem {problem!r}")
        self.problem = problem
        # single precision: transform each channel's x mean apart from its
        # fluctuation (mlv_job_desc.mean_split)
        self.mean_split = bool(mean_split)
        self.nfield = FIELDS.get(problem, nfield if nfield is not None else len(coeff))
        if problem == "vorticity" and not 1 <= self.nfield <= 3:
            raise ValueError("vorticity jobs carry 1..3 fields")
        self.nx, self.nz, self.lx, self.lz = int(nx), int(nz), float(lx), float(lz)
        self.dtype = dtype
        self.double = dtype == "f64"
        if dtype not in ("f32", "f64"):
            raise ValueError("dtype must be 'f32' or 'f64'")
        self.real = torch.float64 if self.double else torch.float32
        self.cplx = torch.complex128 if self.double else torch.complex64
        self.dt = float(dt)
        # plain floats: numpy scalars would not survive the checkpoint's
        # literal metadata
        self.coeff = [float(c) for c in coeff] + [0.0] * (4 - len(coeff))
        self.param = [float(c) for c in param] + [0.0] * (4 - len(param))
        self.ab_order = int(ab_order)
        self.fd_order = int(fd_order)
        self.alpha = float(alpha)
        self.nn = (self.nx - 1) // 3
        self.nm = (self.nz - 1) // 3
        self.rows = 2 * self.nn + 1
        dev = device if device is not None else torch.device("cuda", torch.cuda.current_device())
        # device arrays are kz-major, (nm, rows): a kz column is contiguous,
        # which is how both x passes read and write the state
        shape = (self.nm, self.rows)
        nf = self.nfield
        self.buf = [[torch.zeros(shape, dtype=self.cplx, device=dev) for _ in range(nf)]
                    for _ in range(2)]
        self.hist = [torch.zeros((self.ab_order,) + shape, dtype=self.cplx, device=dev)
                     for _ in range(nf)]
        # single precision: Kahan residual of the state update (see the
        # x-forward kernel), value held = buf - carry; stored as bfloat16
        # (re, im) pairs: it only resolves a fraction of one FP32 ulp
        self.carry = ([torch.zeros(shape + (2,), dtype=torch.bfloat16, device=dev)
                       for _ in range(nf)]
                      if (compensated and not self.double) else None)
        nbytes = ctypes.c_size_t()
        _lib.call("mlv_job_work_bytes", PROBLEMS[problem], nf, self.nx, self.nz,
                  _lib.F64 if self.double else _lib.F32, ctypes.byref(nbytes))
        self.work = torch.empty(nbytes.value, dtype=torch.uint8, device=dev)
        self.psi = torch.zeros(shape, dtype=self.cplx, device=dev) if want_psi else None
        self.phys = ([torch.zeros((self.nx, self.nz), dtype=self.real, device=dev)
                      for _ in range(_phys_count(problem, nf))] if want_phys else None)
        self.curr = 0
        self.parity = 0
        self.steps = 0
        self.t = 0.0
        # the batch whose device job table holds this job's dt and t after a
        # run_device (None when the host values are current)
        self.time_owner = None

    # -- host access --------------------------------------------------------
    @property
    def spectral_shape(self):
        """Shape of the spectra the host sees (the reference layout)."""
        return (self.rows, self.nm)

    def hbm_bytes(self):
        n = sum(b.numel() * b.element_size() for row in self.buf for b in row)
        n += sum(h.numel() * h.element_size() for h in self.hist)
        if self.carry is not None:
            n += sum(c.numel() * c.element_size() for c in self.carry)
        n += self.work.numel()
        if self.psi is not None:
            n += self.psi.numel() * self.psi.element_size()
        if self.phys is not None:
            n += sum(p.numel() * p.element_size() for p in self.phys)
        return n

    def set_state(self, fields):
        """Load spectral fields (host or device arrays) into the current buffer."""
        if len(fields) != self.nfield:
            raise ValueError(f"expected {self.nfield} fields")
        for i, f in enumerate(fields):
            t = f if isinstance(f, torch.Tensor) else torch.as_tensor(np.asarray(f))
            if tuple(t.shape) != self.spectral_shape:
                raise ValueError(f"field {i} has shape {tuple(t.shape)}, expected "
                                 f"{self.spectral_shape}")
            t = t.to(self.buf[0][0].device, torch.complex128).transpose(0, 1)
            hi = t.to(self.cplx)
            self.buf[self.parity][i].copy_(hi)
            if self.carry is not None:
                r = torch.view_as_real(hi.to(torch.complex128) - t)
                self.carry[i].copy_(r.to(torch.bfloat16))

    def current(self):
        """Device tensors of the current spectral state, as (rows, nm) views
        of the kz-major buffers."""
        return [b.transpose(0, 1) for b in self.buf[self.parity]]

    def state(self):
        """Current spectral state as complex128 host arrays (including the
        single-precision carry when there is one: value = buf - carry)."""
        out = []
        for i, b in enumerate(self.buf[self.parity]):
            v = b.detach().cpu().numpy().astype(np.complex128)
            if self.carry is not None:
                r = self.carry[i].detach().to(torch.float64).cpu().numpy()
                v = v - (r[..., 0] + 1j * r[..., 1])
            out.append(np.ascontiguousarray(v.T))
        return out

    def psi_spectral(self):
        """Streamfunction of the last step (needs ``want_psi``), as a
        complex128 host array in the reference layout."""
        if self.psi is None:
            raise RuntimeError("job was created without psi output")
        return np.ascontiguousarray(self.psi.detach().cpu().numpy().astype(np.complex128).T)

    def physical(self):
        """Physical fields written by the last step run with write_phys."""
        if self.phys is None:
            raise RuntimeError("job was created without physical outputs")
        return [p.detach().cpu().numpy().astype(np.float64) for p in self.phys]

    def desc(self):
        d = _lib.JobDesc()
        d.problem = PROBLEMS[self.problem]
        d.nx, d.nz, d.lx, d.lz = self.nx, self.nz, self.lx, self.lz
        d.fd_order, d.ab_order, d.alpha, d.dt = self.fd_order, self.ab_order, self.alpha, self.dt
        for i in range(4):
            d.coeff[i] = self.coeff[i]
            d.param[i] = self.param[i]
        d.nfield = self.nfield
        for b in range(2):
            for f in range(self.nfield):
                d.state[b][f] = self.buf[b][f].data_ptr()
        for f in range(self.nfield):
            d.hist[f] = self.hist[f].data_ptr()
            if self.carry is not None:
                d.comp[f] = self.carry[f].data_ptr()
        d.work = self.work.data_ptr()
        d.psi = self.psi.data_ptr() if self.psi is not None else None
        if self.phys is not None:
            for i, p in enumerate(self.phys):
                d.phys[i] = p.data_ptr()
        d.curr, d.parity = self.curr, self.parity
        d.t, d.loops = self.t, self.steps
        d.mean_split = 1 if self.mean_split else 0
        return d


# Columns of the tracker records written by the on-device CFL check:
# kinetic energy as utility.py:42-59, enstrophy 0.5 mean(w^2), and
# scalar_flux = mean(f uz) of the job's first advected scalar (NaN without
# one), whose complement 1 - flux is double_diffusive_convection.py:20-23's
# Nusselt number.
TRACKER_COLUMNS = ("loop", "t", "dt", "kinetic_energy", "max_ux", "max_uz", "next_dt",
                   "enstrophy", "scalar_flux")


class DeviceRun:
    """On-device CFL ticker and tracker of a batch (the fully spectral
    ``Batch`` and the FD-z ``FdzBatch``); ``_C`` is the C entry-point prefix."""

    _C = "mlv_batch"

    def run_device(self, nsteps, cadence=10, cutoff=0.5, graph=False, stream=None):
        """Queue ``nsteps`` loops with the CFL ticker applied on the device
        (no host synchronisation; see mlv_batch_run and mlv_fdz_batch_run).
        Host-side dt and t of the jobs are stale until ``sync()``."""
        if nsteps <= 0:
            return
        if graph and (self.jobs[0].steps - self._steps0[0]) % 4:
            raise ValueError("graph replay needs the batch step count at a multiple of 4")
        _lib.call(self._C + "_run", self.handle, int(nsteps), int(cadence), float(cutoff),
                  1 if graph else 0, _lib.stream_handle(stream))
        for j in self.jobs:
            j.steps += nsteps
            j.curr = (j.curr + nsteps) % j.ab_order
            j.parity = (j.parity + nsteps) & 1
            # dt and t now live in this batch's device job table
            j.time_owner = self

    def sync(self):
        """Read every job's device-side dt and t; returns the indices of
        jobs whose CFL check found a breach."""
        breached = []
        for i, j in enumerate(self.jobs):
            dt, t = ctypes.c_double(), ctypes.c_double()
            loops, br = ctypes.c_longlong(), ctypes.c_int()
            _lib.call(self._C + "_job_time", self.handle, i, ctypes.byref(dt), ctypes.byref(t),
                      ctypes.byref(loops), ctypes.byref(br))
            j.dt, j.t = dt.value, t.value
            j.time_owner = None
            if br.value:
                breached.append(i)
        return breached

    def tracker(self, index, max_records=4096):
        """The job's tracker records (oldest first) as a dict of arrays
        named by TRACKER_COLUMNS; DDC jobs also get ``nusselt_number``."""
        nc = len(TRACKER_COLUMNS)
        buf = (ctypes.c_double * (nc * max_records))()
        n = ctypes.c_int()
        _lib.call(self._C + "_tracker", self.handle, int(index), buf, int(max_records),
                  ctypes.byref(n))
        a = np.frombuffer(buf, dtype=np.float64)[: nc * n.value].reshape(n.value, nc)
        out = {name: a[:, k].copy() for k, name in enumerate(TRACKER_COLUMNS)}
        if getattr(self.jobs[index], "problem", None) == "ddc":
            out["nusselt_number"] = 1.0 - out["scalar_flux"]
        return out

    def save_trackers(self, index, directory=".", cadence=100):
        """Write job ``index``'s tracker series as the reference's
        ScalarTracker files (ScalarTracker.py:15-19: ``t`` and ``values``
        arrays): kinetic_energy.npz, enstrophy.npz, and for DDC jobs
        nusselt_number.npz (double_diffusive_convection.py:80-88).  The
        reference's tracker ticker fires after loops 1, 1 + cadence, ...
        (Ticker.py:14-25, tracker_cadence = 100 by default); the device
        records those loops when cadence is a multiple of the CFL cadence
        of the run, and only those rows are written."""
        import os

        tr = self.tracker(index)
        keep = (tr["loop"].astype(np.int64) - 1) % int(cadence) == 0
        tr = {k: v[keep] for k, v in tr.items()}
        names = ["kinetic_energy", "enstrophy"]
        if "nusselt_number" in tr:
            names.append("nusselt_number")
        paths = []
        for name in names:
            p = os.path.join(directory, name + ".npz")
            np.savez(p, t=tr["t"], values=tr[name])
            paths.append(p)
        return paths


class Batch(DeviceRun):
    """A set of jobs advanced together; all jobs share one precision."""

    def __init__(self, jobs):
        if not jobs:
            raise ValueError("empty batch")
        dt = {j.dtype for j in jobs}
        if len(dt) != 1:
            raise ValueError("all jobs of a batch must share one precision")
        self.jobs = list(jobs)
        self.double = self.jobs[0].double
        arr = (_lib.JobDesc * len(self.jobs))(*[j.desc() for j in self.jobs])
        h = ctypes.c_void_p()
        _lib.call("mlv_batch_create", ctypes.byref(h), _lib.F64 if self.double else _lib.F32,
                  arr, len(self.jobs))
        self.handle = h
        self._steps0 = [j.steps for j in self.jobs]

    def rebuild(self, jobs):
        """Change the batch's membership in place (mlv_batch_rebuild): jobs
        not in ``jobs`` leave, new ones join, jobs that stay keep their
        tracker records; no memory is allocated while the new set fits the
        batch's tables.  Synchronises with the device."""
        jobs = list(jobs)
        if not jobs:
            raise ValueError("empty batch")
        if {j.dtype for j in jobs} != {self.jobs[0].dtype}:
            raise ValueError("all jobs of a batch must share one precision")
        if any(getattr(j, "time_owner", None) is self for j in self.jobs):
            self.sync()  # the device-side dt and t of run_device
        arr = (_lib.JobDesc * len(jobs))(*[j.desc() for j in jobs])
        _lib.call("mlv_batch_rebuild", self.handle, arr, len(jobs))
        self.jobs = jobs
        self._steps0 = [j.steps for j in self.jobs]

    def step(self, nsteps=1, write_phys=False, stream=None):
        """Queue ``nsteps`` steps of every job on the stream (asynchronous)."""
        if nsteps <= 0:
            return
        _lib.call("mlv_batch_step", self.handle, int(nsteps), 1 if write_phys else 0,
                  _lib.stream_handle(stream))
        for j in self.jobs:
            j.steps += nsteps
            j.curr = (j.curr + nsteps) % j.ab_order
            j.parity = (j.parity + nsteps) & 1
            j.t += nsteps * j.dt

    def graph_step(self, nsteps, stream=None):
        """Queue ``nsteps`` (a multiple of 4) steps as one HIP graph launch;
        the graph is captured on first use and replayed afterwards."""
        if nsteps % 4:
            raise ValueError("graph steps must be a multiple of 4")
        if (self.jobs[0].steps - self._steps0[0]) % 4:
            raise ValueError("graph replay needs the batch step count at a multiple of 4")
        _lib.call("mlv_batch_graph_step", self.handle, int(nsteps), _lib.stream_handle(stream))
        for j in self.jobs:
            j.steps += nsteps
            j.curr = (j.curr + nsteps) % j.ab_order
            j.parity = (j.parity + nsteps) & 1
            j.t += nsteps * j.dt

    def set_dt(self, index, dt, stream=None):
        job = self.jobs[index]
        job.dt = float(dt)
        _lib.call("mlv_batch_set_dt", self.handle, int(index), float(dt),
                  _lib.stream_handle(stream))

    # names of the kernel choices (mlv_batch_autotune)
    _SHAPE_A = ("512 threads", "256 threads", "512 threads, mixed ops")
    _PLAN_B = ("split-pass", "two-sweep")
    _SHAPE_C = ("narrow", "wide", "narrow, 512 threads")

    def _advance(self, n):
        for j in self.jobs:
            j.steps += n
            j.curr = (j.curr + n) % j.ab_order
            j.parity = (j.parity + n) & 1
            j.t += n * j.dt

    def autotune(self, steps=4, stream=None):
        """Pick each fused kernel's variant by timing ``steps`` real steps
        with each option (mlv_batch_autotune); the jobs advance by 2 * steps
        (3 * steps where the x-inverse has three shapes).  Returns the
        chosen plan (see ``plan``)."""
        chosen = (ctypes.c_int * 3)()
        before = _lib.lib().mlv_batch_steps(self.handle)
        _lib.call("mlv_batch_autotune", self.handle, int(steps), _lib.stream_handle(stream), chosen)
        self._advance(_lib.lib().mlv_batch_steps(self.handle) - before)
        return self.plan

    @property
    def plan(self):
        """The kernel choices: x-inverse block shape, z-pass plan, update
        tile shape, and which alternatives this job set admits."""
        p = (ctypes.c_int * 7)()
        _lib.call("mlv_batch_plan", self.handle, p)
        double = self.jobs[0].dtype == "f64"
        shape_a = ("256 threads", "128 threads", "256 threads, mixed ops") if double else self._SHAPE_A
        return {"x_inverse": shape_a[p[0]], "z_pass": self._PLAN_B[p[1]],
                "x_forward_update": self._SHAPE_C[p[2]],
                "x_inverse_small_block_available": bool(p[3]),
                "z_pass_split_available": bool(p[4]),
                "x_inverse_mixed_available": bool(p[5]),
                "x_forward_update_512_available": bool(p[6])}

    def set_plan(self, shape_xinv=0, plan_zpass=0, wide_update=0):
        """Set the kernel choices (x-inverse and update 0/1/2, z pass 0/1,
        as in ``plan``)."""
        _lib.call("mlv_batch_set_plan", self.handle, int(shape_xinv), int(plan_zpass),
                  int(wide_update))

    def close(self):
        if getattr(self, "handle", None) is not None and self.handle.value:
            if any(getattr(j, "time_owner", None) is self for j in self.jobs):
                self.sync()  # keep the device-side dt and t of run_device
            _lib.lib().mlv_batch_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def grid_points(self):
        return sum(j.nx * j.nz for j in self.jobs)


def cfl_dt(job, dt, cutoff=0.5):
    """The reference's CFL rule (Integrator.py:35-44) on a job's physical
    velocities (needs ``want_phys`` and a step run with write_phys)."""
    from . import hip

    nf = job.nfield
    ux, uz = job.phys[nf], job.phys[nf + 1]
    dx, dz = job.lx / job.nx, job.lz / job.nz
    with np.errstate(divide="ignore", invalid="ignore"):
        c = min(np.float64(dx) / np.float64(hip.max(ux)), np.float64(dz) / np.float64(hip.max(uz)))
    if dt > c or math.isnan(c):
        raise Exception("CFL condition breached")
    while dt > cutoff * c:
        dt = dt * 0.9
    return dt


def run_device(batch, nsteps, cfl_cadence=10, cutoff=0.5, chunk=None):
    """``run`` with the CFL ticker on the device: captured graphs of
    ``chunk`` loops (a multiple of 4 and of the cadence) replayed back to
    back, eager steps for what does not fit, one synchronisation at the
    end.  Raises as the reference does if any job's check found a breach
    (the device cannot stop mid-run, so the jobs' states after the breach
    are not meaningful)."""
    import math as _m

    unit = 4 * cfl_cadence // _m.gcd(4, cfl_cadence) if cfl_cadence else 4
    chunk = chunk or unit * max(1, 100 // unit)
    if chunk % unit:
        raise ValueError(f"chunk must be a multiple of {unit}")
    done = 0
    while done < nsteps:
        aligned = (batch.jobs[0].steps - batch._steps0[0]) % unit == 0
        if aligned and nsteps - done >= chunk:
            batch.run_device(chunk, cfl_cadence, cutoff, graph=True)
            done += chunk
        else:
            k = nsteps - done if aligned else min(nsteps - done,
                                                   unit - (batch.jobs[0].steps - batch._steps0[0]) % unit)
            batch.run_device(k, cfl_cadence, cutoff, graph=False)
            done += k
    if batch.sync():
        raise Exception("CFL condition breached")
    return done


def run(batch, nsteps, cfl_cadence=10, cutoff=0.5, on_cfl=None, cfl=None):
    """Advance every job ``nsteps`` loops with the reference's CFL ticker
    (Simulation.py:99-109, Ticker.py:14-25): a job's dt is checked after its
    loops 1, 1+cadence, ... from that loop's velocities and applies from the
    next loop.  Loops count from the start of the job's simulation
    (``job.steps``), so a job resumed from a checkpoint keeps the phase of
    its uninterrupted run, as the reference's restored ticker does.  Jobs
    need ``want_phys``; cfl_cadence=0 disables the check.
    ``cfl(job, dt, cutoff)`` is the dt rule (default: cfl_dt)."""
    cfl_dt_fn = cfl if cfl is not None else cfl_dt
    done = 0
    while done < nsteps:
        # loops before the next check loop of any job (a job's check loop
        # starts with its loop count a multiple of the cadence)
        k = min((-j.steps) % cfl_cadence for j in batch.jobs) if cfl_cadence else nsteps
        k = min(k, nsteps - done)
        if k > 0:
            batch.step(k)
            done += k
            continue
        batch.step(1, write_phys=True)
        done += 1
        for i, job in enumerate(batch.jobs):
            if (job.steps - 1) % cfl_cadence:
                continue
            new = cfl_dt_fn(job, job.dt, cutoff)
            if new != job.dt:
                batch.set_dt(i, new)
            if on_cfl is not None:
                on_cfl(i, job)
    return done
