"""Continuous-batching request loop over the fused batched step.

Requests (simulation jobs) arrive over time.  The server admits a waiting
request as soon as its HBM footprint fits the budget, keeps every admitted
job's spectral state resident on the GPU, and advances all resident jobs
together -- one fused step covers every job, whatever its size or problem
-- in chunks of steps.  Between chunks finished jobs leave (their memory is
released) and new arrivals join, so work of requests that arrive while
others run overlaps on the device.  Each job's arithmetic is independent of
what else is in the batch, so a result is bit-identical alone and in a mix.

Reported per request: admission wait and completion latency (completion
time minus arrival time); overall: aggregate grid-point-steps/s, the
p50 / p99 latency and admission wait, and the time spent changing the
batch's membership (in place with ``Batch.rebuild``, or by creating a new
batch each time with ``incremental=False``).
"""
from __future__ import annotations

import time
from collections import deque
from dataclasses import dataclass, field

import numpy as np
import torch

from . import fused, initial


@dataclass
class Request:
    rid: int
    arrival: float          # seconds after the start of the run
    problem: str            # "vortex" (vortex pair) or "mhd" (resistive tearing)
    n: int                  # grid size (n x n)
    steps: int
    seed: int = 0
    result: list = field(default=None, repr=False)
    admitted: float = None
    done: float = None

    @property
    def work(self):
        return self.n * self.n * self.steps

    @property
    def latency(self):
        return None if self.done is None else self.done - self.arrival


def poisson_requests(rate=20.0, duration=60.0, seed=0, sizes=(256, 512, 1024, 2048),
                     steps=(100, 5000), p_vortex=0.5):
    """Seeded Poisson arrivals: exponential inter-arrival times at ``rate``
    per second for ``duration`` seconds; each job is a vortex pair or a
    tearing run with equal odds, a grid drawn uniformly from ``sizes`` and a
    step count uniform in ``steps`` (inclusive)."""
    rng = np.random.default_rng(seed)
    t = 0.0
    out = []
    while True:
        t += rng.exponential(1.0 / rate)
        if t >= duration:
            break
        prob = "vortex" if rng.random() < p_vortex else "mhd"
        n = int(rng.choice(sizes))
        k = int(rng.integers(steps[0], steps[1] + 1))
        out.append(Request(len(out), t, prob, n, k, seed=int(rng.integers(1 << 31))))
    return out


def make_job(req, dtype="f32", nu=1e-4, eta=1e-4, cfl=0.1):
    """The job for a request: a vortex pair (vorticity problem, Re = 1/nu) or
    resistive tearing (omega and current j, viscosity nu and resistivity
    eta) on [0, 1]^2, AB2 semi-implicit, FD order 2, fixed dt = cfl / n
    (velocities and fields are O(1))."""
    dt = cfl / req.n
    double = dtype == "f64"
    if req.problem == "vortex":
        job = fused.Job("vorticity", req.n, req.n, 1.0, 1.0, dtype=dtype, dt=dt, coeff=[nu])
        job.set_state([initial.vortex_pair_device(req.n, seed=req.seed, double=double)])
    elif req.problem == "mhd":
        job = fused.Job("mhd", req.n, req.n, 1.0, 1.0, dtype=dtype, dt=dt, coeff=[nu, eta])
        job.set_state(list(initial.tearing_device(req.n, seed=req.seed, double=double)))
    else:
        raise ValueError(req.problem)
    return job


def footprint(req, dtype="f32"):
    """HBM bytes a request's job will hold (without allocating it)."""
    import ctypes

    from . import _lib

    n = req.n
    nn, nm = (n - 1) // 3, (n - 1) // 3
    nf = 1 if req.problem == "vortex" else 2
    c = 8 if dtype == "f32" else 16
    S = (2 * nn + 1) * nm * c
    work = ctypes.c_size_t()
    _lib.call("mlv_job_work_bytes", _lib.PROB_VORTICITY if nf == 1 else _lib.PROB_MHD, nf, n, n,
              _lib.F32 if dtype == "f32" else _lib.F64, ctypes.byref(work))
    carry = nf * S // 2 if dtype == "f32" else 0   # bfloat16 residual pairs
    return nf * (2 * S + 2 * S) + carry + work.value   # state x2, AB2 ring, residual, work


class Server:
    """Admits requests by HBM budget and advances resident jobs together.

    ``chunk`` bounds the steps run between admission points; ``keep_results``
    copies each finished job's spectral state to the host.
    """

    def __init__(self, hbm_budget=64 << 30, chunk=16, dtype="f32", keep_results=False,
                 job_factory=make_job, virtual_step_s=1e-3, incremental=True):
        self.incremental = bool(incremental)
        self.budget = int(hbm_budget)
        self.chunk = int(chunk)
        self.dtype = dtype
        self.keep = keep_results
        self.job_factory = job_factory
        self.virtual_step_s = virtual_step_s  # clock advance per step when not real time
        self.used = 0

    def run(self, requests, realtime=True, log=None):
        reqs = sorted(requests, key=lambda r: r.arrival)
        pending = deque(reqs)
        waiting = deque()
        resident = []          # (req, job, remaining, bytes)
        batch = None
        dirty = True
        t0 = time.perf_counter()
        clock = (lambda: time.perf_counter() - t0) if realtime else None
        sim_t = 0.0
        total_work = 0
        chunks = 0
        max_res = 0
        rebuild_s = 0.0
        rebuilds = 0
        while pending or waiting or resident:
            now = clock() if realtime else sim_t
            while pending and pending[0].arrival <= now:
                waiting.append(pending.popleft())
            while waiting:
                r = waiting[0]
                need = footprint(r, self.dtype)
                if need > self.budget:
                    raise ValueError(f"request {r.rid} needs {need} bytes, over the budget")
                if self.used + need > self.budget:
                    break
                waiting.popleft()
                job = self.job_factory(r, self.dtype)
                r.admitted = clock() if realtime else sim_t
                resident.append([r, job, r.steps, need])
                self.used += need
                dirty = True
            if not resident:
                if pending:
                    nxt = pending[0].arrival
                    if realtime:
                        time.sleep(max(0.0, nxt - clock()))
                    else:
                        sim_t = nxt
                continue
            if dirty:
                tb = time.perf_counter()
                members = [e[1] for e in resident]
                if batch is not None and self.incremental:
                    batch.rebuild(members)
                else:
                    if batch is not None:
                        batch.close()
                    batch = fused.Batch(members)
                rebuild_s += time.perf_counter() - tb
                rebuilds += 1
                dirty = False
            k = min(self.chunk, min(e[2] for e in resident))
            batch.step(k)
            chunks += 1
            max_res = max(max_res, len(resident))
            if not realtime:
                sim_t += k * self.virtual_step_s
            for e in resident:
                e[2] -= k
                total_work += e[0].n * e[0].n * k
            finished = [e for e in resident if e[2] == 0]
            if finished:
                torch.cuda.synchronize()
                tdone = clock() if realtime else sim_t
                for e in finished:
                    r, job = e[0], e[1]
                    r.done = tdone
                    if self.keep:
                        r.result = job.state()
                    self.used -= e[3]
                    resident.remove(e)
                del finished
                dirty = True
                if log is not None:
                    log(f"t={tdone:8.2f}s resident={len(resident)} waiting={len(waiting)} "
                        f"pending={len(pending)} hbm={self.used / 2**30:.1f}GiB")
        torch.cuda.synchronize()
        wall = (clock() if realtime else sim_t)
        if batch is not None:
            batch.close()
        lat = np.array([r.latency for r in reqs if r.latency is not None])
        wait = np.array([r.admitted - r.arrival for r in reqs if r.admitted is not None])
        return {
            "requests": len(reqs),
            "wall_s": wall,
            "grid_pt_steps": total_work,
            "grid_pt_steps_per_s": total_work / wall if wall > 0 else None,
            "p50_latency_s": float(np.percentile(lat, 50)) if len(lat) else None,
            "p99_latency_s": float(np.percentile(lat, 99)) if len(lat) else None,
            "mean_latency_s": float(lat.mean()) if len(lat) else None,
            "p50_admission_wait_s": float(np.percentile(wait, 50)) if len(wait) else None,
            "p99_admission_wait_s": float(np.percentile(wait, 99)) if len(wait) else None,
            "chunks": chunks,
            "max_resident": max_res,
            "membership_changes": rebuilds,
            "membership_change_s": rebuild_s,
            "membership_change_mode": "rebuild in place" if self.incremental else "new batch",
        }
