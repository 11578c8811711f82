"""Checkpoints of fused jobs: a run resumed from a saved file (or from a
snapshot taken while steps were queued behind it) is bit-identical to the
uninterrupted run, for both fused paths and both precisions."""
import numpy as np
import pytest

from oracle import melvin_oracle as O

pytestmark = pytest.mark.gpu


def _kh_job(mlv, dtype):
    p = O.make_params(dict(nx=64, nz=64, lx=2.0, lz=2.0, initial_dt=2e-3, integrator_order=4))
    w, c = O.ic_kelvin_helmholtz(p, seed=1, amp=1e-2)
    j = mlv.fused.Job("vorticity", 64, 64, 2.0, 2.0, dtype=dtype, dt=2e-3, coeff=[1e-3, 1e-3],
                      ab_order=4)
    j.set_state([O.to_spectral(p, w), O.to_spectral(p, c)])
    return j


def _rbc_job(mlv, dtype):
    p = O.make_params(dict(nx=32, nz=33, lx=2.0, lz=1.0, Pr=1.0, Ra=1e5,
                           discretisation=["spectral", "fdm"]))
    X, Z = O.grid(p)
    j = mlv.fdz.FdzJob(32, 33, 2.0, 1.0, 1.0, 1e5, dtype=dtype, dt=1e-5, scheme="pc")
    j.set_state([O.to_spectral(p, O.noise((32, 33), 1e-1, 2)),
                 O.to_spectral(p, 1 - Z + O.noise((32, 33), 1e-2, 3))])
    return j


def _batch(mlv, job):
    return (mlv.fused.Batch if isinstance(job, mlv.fused.Job) else mlv.fdz.FdzBatch)([job])


@pytest.mark.parametrize("kind", ["kh", "rbc"])
@pytest.mark.parametrize("dtype", ["f64", "f32"])
def test_resume_is_bit_exact(mlv, tmp_path, kind, dtype):
    make = _kh_job if kind == "kh" else _rbc_job
    job = make(mlv, dtype)
    b = _batch(mlv, job)
    b.step(7)
    f = tmp_path / "ck.npz"
    mlv.checkpoint.save(job, f)
    snap = mlv.checkpoint.snapshot(job)   # steps queued behind it must not leak in
    b.step(9)
    ref = job.state()
    ref_hist = [h.cpu().numpy() for h in job.hist]
    taken = snap.wait()
    b.close()
    for src in (f, None):
        if src is None:
            g = tmp_path / "snap.npz"
            np.savez(g, **taken)
            src = g
        job2 = mlv.checkpoint.load(src)
        assert job2.steps == 7 and job2.curr == 7 % job2.ab_order
        b2 = _batch(mlv, job2)
        b2.step(9)
        for x, y in zip(job2.state(), ref):
            np.testing.assert_array_equal(x, y)
        # the ring as stored (same layout for jobs of the same kind)
        for h2, h in zip(job2.hist, ref_hist):
            np.testing.assert_array_equal(h2.cpu().numpy(), h)
        b2.close()


@pytest.mark.parametrize("kind", ["kh", "rbc"])
def test_checkpoint_after_device_ticker_keeps_device_dt(mlv, tmp_path, kind):
    """A checkpoint taken right after run_device (the CFL ticker shrank dt on
    the device, and the host has not synchronised) stores the device-side dt
    and t, so the resumed run is bit-identical to the uninterrupted one."""
    make = _kh_job if kind == "kh" else _rbc_job
    job = make(mlv, "f64")
    dt0 = job.dt
    # a cutoff far below 1: the first check shrinks dt (dt *= 0.9 while
    # dt > cutoff * cfl) without a breach (dt < cfl)
    cut = 1e-3 if kind == "kh" else 1e-8
    b = _batch(mlv, job)
    b.run_device(11, cadence=10, cutoff=cut)
    f = tmp_path / "ck.npz"
    mlv.checkpoint.save(job, f)
    assert job.time_owner is None
    dt_saved = job.dt
    assert dt_saved < dt0
    b.run_device(9, cadence=10, cutoff=cut)
    assert not b.sync()
    ref = job.state()
    b.close()
    job2 = mlv.checkpoint.load(f)
    assert job2.dt == dt_saved and job2.steps == 11
    b2 = _batch(mlv, job2)
    b2.run_device(9, cadence=10, cutoff=cut)
    b2.sync()
    assert job2.dt == job.dt
    for x, y in zip(job2.state(), ref):
        np.testing.assert_array_equal(x, y)
    b2.close()


def test_numpy_scalar_parameters_round_trip(mlv, tmp_path):
    """Coefficients given as NumPy scalars are stored as plain floats, so the
    checkpoint's metadata reads back."""
    j = mlv.fused.Job("ddc", 32, 32, 10.0, 10.0, dtype="f64", dt=1e-3,
                      coeff=[np.float64(7.0), np.float32(1.0), np.float64(0.01)],
                      param=[np.float64(7.0), np.float64(1.5)])
    b = mlv.fused.Batch([j])
    b.step(2)
    f = tmp_path / "ddc.npz"
    mlv.checkpoint.save(j, f)
    b.close()
    j2 = mlv.checkpoint.load(f)
    assert j2.coeff[:3] == [7.0, 1.0, 0.01] and j2.param[:2] == [7.0, 1.5]
    assert all(type(c) is float for c in j2.coeff + j2.param)
