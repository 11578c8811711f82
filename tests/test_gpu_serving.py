"""The continuous-batching server: jobs that arrive while others run are
admitted by HBM budget and overlap on the device, and every result is
bit-identical to the same job run alone."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_results_alone_equal_results_in_mix(mlv):
    S = mlv.serving
    reqs = [S.Request(0, 0.000, "vortex", 64, 24, seed=1),
            S.Request(1, 0.003, "mhd", 32, 40, seed=2),
            S.Request(2, 0.004, "vortex", 32, 12, seed=3),
            S.Request(3, 0.010, "mhd", 64, 20, seed=4),
            S.Request(4, 0.011, "vortex", 128, 16, seed=5),
            S.Request(5, 0.030, "mhd", 64, 8, seed=6)]
    budget = 3 * max(S.footprint(r) for r in reqs)  # forces queueing
    srv = S.Server(hbm_budget=budget, chunk=4, keep_results=True)
    stats = srv.run(reqs, realtime=False)
    assert all(r.done is not None for r in reqs)
    assert stats["max_resident"] >= 2          # work overlapped on the device
    assert stats["grid_pt_steps"] == sum(r.work for r in reqs)
    for r in reqs:
        job = S.make_job(r)
        b = mlv.fused.Batch([job])
        b.step(r.steps)
        for x, y in zip(job.state(), r.result):
            np.testing.assert_array_equal(x, y)
        b.close()


def _jobs(mlv, dtype="f64"):
    S = mlv.serving
    reqs = [S.Request(0, 0.0, "vortex", 64, 0, seed=11), S.Request(1, 0.0, "mhd", 32, 0, seed=12),
            S.Request(2, 0.0, "vortex", 128, 0, seed=13), S.Request(3, 0.0, "mhd", 64, 0, seed=14)]
    return [S.make_job(r, dtype) for r in reqs]


@pytest.mark.parametrize("dtype", ["f64", "f32"])
def test_rebuild_changes_membership_exactly(mlv, dtype):
    """Jobs leave and join a batch in place (reordered, grown beyond the
    first tables, shrunk); every job equals the same steps run alone."""
    a, b, c, d = _jobs(mlv, dtype)
    batch = mlv.fused.Batch([a, b])
    batch.step(5)
    batch.rebuild([b, c, d, a])     # larger set: the tables grow
    batch.step(7)
    batch.rebuild([c])              # smaller set: the tables are reused
    batch.step(3)
    batch.close()
    ref = _jobs(mlv, dtype)
    for job, n in zip(ref, (12, 12, 10, 7)):
        bb = mlv.fused.Batch([job])
        bb.step(n)
        bb.close()
    for got, want in zip((a, b, c, d), ref):
        assert got.steps == want.steps
        for x, y in zip(got.state(), want.state()):
            np.testing.assert_array_equal(x, y)


def test_rebuild_keeps_device_clock_and_tracker(mlv):
    """With the CFL ticker on the device, a job that stays through a rebuild
    keeps its tracker records and its device-side dt and t: the same
    records, dt and t as an uninterrupted run."""
    a, b, c, _ = _jobs(mlv)
    for j in (a, b, c):
        j.dt *= 3           # below the CFL limit, above cutoff x limit: dt shrinks
    batch = mlv.fused.Batch([a, b])
    batch.run_device(23, cadence=10, cutoff=0.1)
    batch.rebuild([c, a])
    batch.run_device(22, cadence=10, cutoff=0.1)
    assert not batch.sync()
    got = batch.tracker(1)
    batch.close()
    ref = _jobs(mlv)[0]
    ref.dt *= 3
    dt0 = ref.dt
    rb = mlv.fused.Batch([ref])
    rb.run_device(45, cadence=10, cutoff=0.1)
    rb.sync()
    want = rb.tracker(0)
    rb.close()
    assert len(got["loop"]) == len(want["loop"]) == 5
    for k in want:
        np.testing.assert_array_equal(got[k], want[k])
    assert a.dt == ref.dt < dt0 and a.t == ref.t
    for x, y in zip(a.state(), ref.state()):
        np.testing.assert_array_equal(x, y)


def test_serving_modes_agree(mlv):
    """The server changes membership in place or by new batches: the same
    results either way."""
    S = mlv.serving
    out = {}
    for incremental in (True, False):
        reqs = [S.Request(0, 0.000, "vortex", 64, 24, seed=1),
                S.Request(1, 0.003, "mhd", 32, 40, seed=2),
                S.Request(2, 0.004, "vortex", 32, 12, seed=3),
                S.Request(3, 0.010, "mhd", 64, 20, seed=4)]
        budget = 2 * max(S.footprint(r) for r in reqs)
        srv = S.Server(hbm_budget=budget, chunk=4, keep_results=True, incremental=incremental)
        stats = srv.run(reqs, realtime=False)
        assert stats["membership_changes"] >= 3
        assert stats["p99_admission_wait_s"] >= stats["p50_admission_wait_s"] >= 0
        out[incremental] = [r.result for r in reqs]
    for x, y in zip(out[True], out[False]):
        for a, b in zip(x, y):
            np.testing.assert_array_equal(a, b)


@pytest.mark.parametrize("dtype", ["f64", "f32"])
def test_rebuild_between_size_specialised_and_generic_kernels(mlv, dtype):
    """A batch of 512^2 jobs runs the kernels compiled for that size alone;
    a 256^2 job joining switches the batch to the generic kernels, and its
    leaving switches it back.  Every job equals the same steps run alone (a
    512^2 job alone runs the specialised kernels throughout): both kernel
    forms compute the same bits."""
    S = mlv.serving

    def jobs():
        reqs = [S.Request(0, 0.0, "vortex", 512, 0, seed=21), S.Request(1, 0.0, "mhd", 512, 0, seed=22),
                S.Request(2, 0.0, "vortex", 256, 0, seed=23)]
        return [S.make_job(r, dtype) for r in reqs]

    a, b, c = jobs()
    batch = mlv.fused.Batch([a, b])
    print(f"[serving] specialised/generic rebuilds {dtype}", flush=True)
    batch.step(3)
    batch.rebuild([a, c, b])        # mixed sizes: generic kernels
    batch.step(4)
    batch.rebuild([b, a])           # one size again: specialised kernels
    batch.step(2)
    batch.close()
    ref = jobs()
    for job, n in zip(ref, (9, 9, 4)):
        bb = mlv.fused.Batch([job])
        bb.step(n)
        bb.close()
    for got, want in zip((a, b, c), ref):
        assert got.steps == want.steps
        for x, y in zip(got.state(), want.state()):
            np.testing.assert_array_equal(x, y)
