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
