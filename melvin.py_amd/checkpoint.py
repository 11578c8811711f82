"""Checkpoints of fused jobs: a pickle-free ``.npz`` with everything a
bit-exact resume needs, and a snapshot that does not stall the stream.

The reference dumps each variable's spectral data, its TimeDerivative ring
and ``curr_idx``, plus ``t``, ``dt`` and the loop counter
(Simulation.py:155-197, Variable.py:67-85, TimeDerivative.py:47-62).  A
fused job's file holds the same items -- the current spectral state and the
whole Adams-Bashforth ring in the reference layout (rows = kx), the ring
slot the next step writes, dt, t and the step count -- plus the job's
parameters and, for single-precision jobs, the rounding residual kept
beside the state (stored separately, so a resumed run is bit-identical).

``snapshot(job)`` takes the checkpoint asynchronously: a device-to-device
copy of the job's arrays is queued on the job's stream (a few microseconds
of HBM traffic), and the copy to pinned host memory runs on a side stream,
so the steps queued after the snapshot proceed while the data travels.
After ``run_device`` (CFL ticker on the device) the job's dt and t are read
back from the batch first, which synchronises with the device once.
"""
from __future__ import annotations

import numpy as np
import torch

from . import fdz, fused

FORMAT = "melvin_amd.fused-job/1"


def _arrays(job):
    """Device arrays of the job's current state, in the reference layout
    (views; fully spectral buffers are kz-major)."""
    cur = job.buf[job.parity]
    if isinstance(job, fused.Job):
        state = [b.transpose(0, 1) for b in cur]
        hist = [h.transpose(1, 2) for h in job.hist]
        # bfloat16 residual pairs, kept as their 16-bit patterns (numpy has
        # no bfloat16)
        carry = ([c.view(torch.int16).transpose(0, 1) for c in job.carry]
                 if job.carry is not None else None)
    else:
        state, hist, carry = list(cur), list(job.hist), job.carry
    return state, hist, carry


def _meta(job):
    owner = getattr(job, "time_owner", None)
    if owner is not None:
        # after run_device the CFL ticker may have changed dt on the device:
        # read the device-side dt and t (a host synchronisation) so that the
        # checkpoint resumes the run it was taken from
        owner.sync()
    m = {"format": FORMAT, "dtype": job.dtype, "nx": job.nx, "nz": job.nz, "lx": job.lx,
         "lz": job.lz, "ab_order": job.ab_order, "fd_order": job.fd_order, "alpha": job.alpha,
         "dt": job.dt, "t": job.t, "steps": job.steps, "curr": job.curr}
    if isinstance(job, fused.Job):
        m.update(kind="fused", problem=job.problem, nfield=job.nfield, coeff=list(job.coeff),
                 param=list(job.param), mean_split=job.mean_split)
    else:
        m.update(kind="fdz", Pr=job.Pr, Ra=job.Ra, scheme=job.scheme)
    return m


class Snapshot:
    """A checkpoint in flight; ``wait()`` returns it as a dict of host
    arrays (the content of the ``.npz``)."""

    def __init__(self, job, stream=None):
        dev_stream = stream if stream is not None else torch.cuda.current_stream()
        state, hist, carry = _arrays(job)
        self.meta = _meta(job)
        def fresh(a):  # a private contiguous copy, queued on the job's stream
            return torch.empty(a.shape, dtype=a.dtype, device=a.device).copy_(a)

        with torch.cuda.stream(dev_stream):
            staged = {f"state{i}": fresh(s) for i, s in enumerate(state)}
            staged.update({f"hist{i}": fresh(h) for i, h in enumerate(hist)})
            if carry is not None:
                staged.update({f"carry{i}": fresh(c) for i, c in enumerate(carry)})
            ready = torch.cuda.Event()
            ready.record(dev_stream)
        self._side = torch.cuda.Stream()
        self._host = {}
        with torch.cuda.stream(self._side):
            self._side.wait_event(ready)
            for k, v in staged.items():
                h = torch.empty(v.shape, dtype=v.dtype, pin_memory=True)
                h.copy_(v, non_blocking=True)
                self._host[k] = h
                v.record_stream(self._side)
            self._done = torch.cuda.Event()
            self._done.record(self._side)

    def wait(self):
        self._done.synchronize()
        out = {k: v.numpy() for k, v in self._host.items()}
        out["meta"] = np.array(repr(self.meta))
        return out


def snapshot(job, stream=None):
    """Start an asynchronous checkpoint of ``job`` (see module docstring)."""
    return Snapshot(job, stream)


def save(job, fname):
    """Write ``job``'s checkpoint to ``fname`` (.npz, no pickles)."""
    np.savez(fname, **snapshot(job).wait())


def _read(fname):
    import ast

    d = np.load(fname, allow_pickle=False)
    meta = ast.literal_eval(str(d["meta"]))
    if meta.get("format") != FORMAT:
        raise ValueError(f"{fname}: not a fused-job checkpoint")
    return d, meta


def load(fname, **kw):
    """A new job holding the checkpoint in ``fname``; ``kw`` passes job
    options that are not part of the state (want_psi, want_phys, ...).  Put
    it in a new batch to resume."""
    d, m = _read(fname)
    if m["kind"] == "fused":
        job = fused.Job(m["problem"], m["nx"], m["nz"], m["lx"], m["lz"], dtype=m["dtype"],
                        dt=m["dt"], coeff=m["coeff"], param=m["param"], nfield=m["nfield"],
                        ab_order=m["ab_order"], fd_order=m["fd_order"], alpha=m["alpha"],
                        compensated="carry0" in d.files, mean_split=m.get("mean_split", True),
                        **kw)
        nf = job.nfield
    else:
        job = fdz.FdzJob(m["nx"], m["nz"], m["lx"], m["lz"], m["Pr"], m["Ra"], dtype=m["dtype"],
                         dt=m["dt"], scheme=m["scheme"], ab_order=m["ab_order"],
                         fd_order=m["fd_order"], alpha=m["alpha"],
                         compensated="carry0" in d.files, **kw)
        nf = 2
    restore(job, d, m, nf)
    return job


def restore(job, d, m, nf):
    """Copy a checkpoint's arrays into ``job`` (same grid and options)."""
    state, hist, carry = _arrays(job)
    for i in range(nf):
        state[i].copy_(torch.as_tensor(d[f"state{i}"]))
        hist[i].copy_(torch.as_tensor(d[f"hist{i}"]))
        if carry is not None:
            carry[i].copy_(torch.as_tensor(d[f"carry{i}"]))
    job.curr = int(m["curr"])
    job.steps = int(m["steps"])
    job.t = float(m["t"])
    job.dt = float(m["dt"])
