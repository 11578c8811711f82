"""The CFL ticker and diagnostics on the device (no host synchronisation):
the same dt sequence, time and state as the host-driven ticker and the
reference's trajectory, tracker records equal to host-computed kinetic
energies at the checked loops, graph replays equal to eager runs, and a
breach reported as the reference raises it."""
import numpy as np
import pytest

from conftest import golden, rel
from oracle import melvin_oracle as O

pytestmark = pytest.mark.gpu


def _tgv(mlv, dtype="f64"):
    g, params, p = golden("tgv_cfl")
    j = mlv.fused.Job("vorticity", p.nx, p.nz, p.lx, p.lz, dtype=dtype, dt=p.initial_dt,
                      coeff=[0.05], ab_order=p.integrator_order, fd_order=p.spatial_derivative_order,
                      alpha=p.alpha, want_phys=True)
    j.set_state([g["w0_spec"]])
    return g, p, j


def test_device_ticker_matches_host_and_reference(mlv):
    g, p, host = _tgv(mlv)
    n = len(g["dts"])
    kes = {}

    def on_cfl(i, job):
        ux, uz = job.physical()[1:3]
        kes[job.steps] = O.kinetic_energy(p, ux, uz)

    bh = mlv.fused.Batch([host])
    mlv.fused.run(bh, n, on_cfl=on_cfl)
    for graph_chunk in (None, 20):
        _, _, dev = _tgv(mlv)
        bd = mlv.fused.Batch([dev])
        print(f"[device-cfl] tgv {n} loops, chunk {graph_chunk}", flush=True)
        if graph_chunk:
            mlv.fused.run_device(bd, n, chunk=graph_chunk)
        else:
            bd.run_device(n)
            assert not bd.sync()
        np.testing.assert_array_equal(dev.state()[0], host.state()[0])
        assert dev.dt == host.dt == float(g["final_dt"])
        assert dev.t == pytest.approx(float(g["t"]), rel=1e-15)
        assert rel(dev.state()[0], g["w_spec"]) < 1e-12
        tr = bd.tracker(0)
        loops = tr["loop"].astype(int)
        np.testing.assert_array_equal(loops, np.arange(1, n + 1, 10))
        for lp, ke in zip(loops, tr["kinetic_energy"]):
            assert ke == pytest.approx(kes[lp], rel=1e-12)
        bd.close()
    bh.close()


def test_device_ticker_in_a_mixed_batch(mlv):
    """Each job's dt follows its own velocities (sizes and precisions of a
    batch differ only in grid here)."""
    jobs, ref = [], []
    for n in (32, 64):
        p = O.make_params(dict(nx=n, nz=n, lx=2 * np.pi, lz=2 * np.pi, initial_dt=0.05))
        w = O.ic_taylor_green(p, seed=n, amp=1e-3)
        for dst in (jobs, ref):
            j = mlv.fused.Job("vorticity", n, n, p.lx, p.lz, dtype="f64", dt=0.05,
                              coeff=[1e-2], want_phys=True)
            j.set_state([w])
            dst.append(j)
    for j in ref:
        b = mlv.fused.Batch([j])
        mlv.fused.run(b, 40)
        b.close()
    b = mlv.fused.Batch(jobs)
    mlv.fused.run_device(b, 40, chunk=20)
    for j, r in zip(jobs, ref):
        assert j.dt == r.dt
        np.testing.assert_array_equal(j.state()[0], r.state()[0])
    b.close()


def test_device_ticker_reports_a_breach(mlv):
    """dt far above the CFL limit: the reference raises 'CFL condition
    breached' (Integrator.py:40-41); the device run flags it and the host
    raises after the run."""
    g, p, j = _tgv(mlv)
    j.dt = 50.0
    b = mlv.fused.Batch([j])
    with pytest.raises(Exception, match="CFL condition breached"):
        mlv.fused.run_device(b, 20, chunk=20)
    b.close()
