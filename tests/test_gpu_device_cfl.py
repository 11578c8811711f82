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
        w, ux, uz = job.physical()[0:3]
        kes[job.steps] = (O.kinetic_energy(p, ux, uz), 0.5 * np.mean(w * w))

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
        for lp, ke, ens in zip(loops, tr["kinetic_energy"], tr["enstrophy"]):
            assert ke == pytest.approx(kes[lp][0], rel=1e-12)
            assert ens == pytest.approx(kes[lp][1], rel=1e-12)
        assert np.isnan(tr["scalar_flux"]).all()  # one field: no scalar
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


def test_device_tracker_scalar_flux_and_files(mlv, tmp_path):
    """DDC: the scalar flux mean(T uz) and Nusselt number 1 - flux
    (double_diffusive_convection.py:20-23) of the device records equal the
    host values from the same loops' physical fields; save_trackers writes
    the reference's ScalarTracker files (t, values) at the tracker cadence."""
    g, params, p = golden("ddc")

    def mk():
        j = mlv.fused.Job("ddc", p.nx, p.nz, p.lx, p.lz, dtype="f64", dt=p.initial_dt,
                          coeff=[p.Pr, 1.0, p.tau], param=[p.Pr, p.R0],
                          ab_order=p.integrator_order, fd_order=p.spatial_derivative_order,
                          alpha=p.alpha, want_phys=True)
        j.set_state([g["w0_spec"], g["t0_spec"], g["s0_spec"]])
        return j

    want = {}

    def on_cfl(i, job):
        w, t, s, ux, uz = job.physical()[:5]
        want[job.steps] = (O.kinetic_energy(p, ux, uz), np.mean(t * uz), np.mean(np.abs(t * uz)),
                           job.t)

    host = mk()
    bh = mlv.fused.Batch([host])
    mlv.fused.run(bh, 41, on_cfl=on_cfl)
    dev = mk()
    bd = mlv.fused.Batch([dev])
    print("[device-cfl] ddc tracker 41 loops", flush=True)
    bd.run_device(41)
    assert not bd.sync()
    tr = bd.tracker(0)
    loops = tr["loop"].astype(int)
    np.testing.assert_array_equal(loops, [1, 11, 21, 31, 41])
    for k, lp in enumerate(loops):
        ke, flux, scale, t = want[lp]
        assert tr["kinetic_energy"][k] == pytest.approx(ke, rel=1e-12)
        assert tr["scalar_flux"][k] == pytest.approx(flux, abs=1e-12 * scale)
        assert tr["nusselt_number"][k] == 1.0 - tr["scalar_flux"][k]
        assert tr["t"][k] == pytest.approx(t, rel=1e-15)
    paths = bd.save_trackers(0, str(tmp_path), cadence=20)
    assert sorted(q.rsplit("/", 1)[1] for q in paths) == [
        "enstrophy.npz", "kinetic_energy.npz", "nusselt_number.npz"]
    d = np.load(str(tmp_path / "nusselt_number.npz"), allow_pickle=False)
    np.testing.assert_array_equal(d["t"], tr["t"][[0, 2, 4]])
    np.testing.assert_array_equal(d["values"], tr["nusselt_number"][[0, 2, 4]])
    bd.close()
    bh.close()


# ---------------------------------------------------------------------------
# FD-z batches (mlv_fdz_batch_run)
# ---------------------------------------------------------------------------
def _fdz_job(mlv, p, scheme, dtype="f64", dt=None, **kw):
    return mlv.fdz.FdzJob(p.nx, p.nz, p.lx, p.lz, p.Pr, p.Ra, dtype=dtype,
                          dt=p.initial_dt if dt is None else dt, scheme=scheme,
                          ab_order=p.integrator_order, fd_order=p.spatial_derivative_order,
                          alpha=p.alpha, **kw)


@pytest.mark.parametrize("name", ["rbc_o2", "rbc_o4"])
def test_fdz_device_ticker_vs_golden(mlv, name):
    """The reference's RBC trajectory (20 loops, its CFL ticker firing at
    loops 1 and 11) with the ticker on the device: state, dt and t equal to
    the golden values, tracker kinetic energy and mean(T uz) equal to the
    host values of the same loops."""
    g, params, p = golden(name)
    host = _fdz_job(mlv, p, "explicit", want_phys=True)
    host.set_state([g["w0_spec"], g["t0_spec"]])
    want = {}

    def on_cfl(i, job):
        ux, uz, w, t = job.physical()
        want[job.steps] = (O.kinetic_energy(p, ux, uz), np.mean(t * uz),
                           np.mean(np.abs(t * uz)), 0.5 * np.mean(w * w))

    bh = mlv.fdz.FdzBatch([host])
    mlv.fdz.run(bh, 20, on_cfl=on_cfl)
    dev = _fdz_job(mlv, p, "explicit")
    dev.set_state([g["w0_spec"], g["t0_spec"]])
    bd = mlv.fdz.FdzBatch([dev])
    print(f"[device-cfl] fdz {name} 20 loops (graph)", flush=True)
    mlv.fdz.run_device(bd, 20, chunk=20)
    w, t = dev.state()
    assert rel(w, g["w_spec"]) < 1e-12
    assert rel(t, g["t_spec"]) < 1e-12
    for a, b in zip(dev.state(), host.state()):
        np.testing.assert_array_equal(a, b)
    assert dev.dt == host.dt == pytest.approx(float(g["final_dt"]), rel=1e-15)
    assert dev.t == pytest.approx(host.t, rel=1e-14)
    tr = bd.tracker(0)
    np.testing.assert_array_equal(tr["loop"].astype(int), [1, 11])
    for k, lp in enumerate([1, 11]):
        ke, flux, scale, ens = want[lp]
        assert tr["kinetic_energy"][k] == pytest.approx(ke, rel=1e-12)
        assert tr["scalar_flux"][k] == pytest.approx(flux, abs=1e-12 * scale)
        assert tr["enstrophy"][k] == pytest.approx(ens, rel=1e-12)
    bd.close()
    bh.close()


@pytest.mark.parametrize("dtype", ["f64", "f32"])
def test_fdz_device_ticker_rebuilds_theta_operators(mlv, dtype):
    """A predictor-corrector run whose CFL rule shrinks dt: the
    device ticker (dt, theta bands and tables rebuilt on the device) gives
    the same states and dt sequence, bit for bit, as the host ticker that
    calls set_dt; eager and graph runs agree."""
    p = O.make_params(dict(nx=64, nz=65, lx=2.0, lz=1.0, Pr=1.0, Ra=1e4, alpha=0.51,
                           integrator_order=2, discretisation=["spectral", "fdm"]))
    X, Z = O.grid(p)
    t0 = O.to_spectral(p, 1 - Z + O.noise((p.nx, p.nz), 1e-2, 5))
    w0 = O.to_spectral(p, 400.0 * np.sin(np.pi * X) * np.sin(np.pi * Z))
    dts = []

    def on_cfl(i, job):
        dts.append(job.dt)

    host = _fdz_job(mlv, p, "pc", dtype=dtype, dt=2e-4, want_phys=True)
    host.set_state([w0, t0])
    bh = mlv.fdz.FdzBatch([host])
    mlv.fdz.run(bh, 60, on_cfl=on_cfl)
    assert bh.jobs[0].dt < 2e-4  # the rule fired
    for chunk in (None, 20):
        dev = _fdz_job(mlv, p, "pc", dtype=dtype, dt=2e-4)
        dev.set_state([w0, t0])
        bd = mlv.fdz.FdzBatch([dev])
        print(f"[device-cfl] fdz pc {dtype} chunk {chunk}", flush=True)
        if chunk:
            mlv.fdz.run_device(bd, 60, chunk=chunk)
        else:
            bd.run_device(60)
            assert not bd.sync()
        assert dev.dt == host.dt
        for a, b in zip(dev.state(), host.state()):
            np.testing.assert_array_equal(a, b)
        np.testing.assert_array_equal(bd.tracker(0)["next_dt"], dts)
        bd.close()
    bh.close()


def test_fdz_device_ticker_reports_a_breach(mlv):
    g, params, p = golden("rbc_o2")
    j = _fdz_job(mlv, p, "explicit", dt=1.0)
    j.set_state([g["w0_spec"], g["t0_spec"]])
    b = mlv.fdz.FdzBatch([j])
    with pytest.raises(Exception, match="CFL condition breached"):
        mlv.fdz.run_device(b, 20, chunk=20)
    b.close()


def _pc_case(mlv, dtype="f64"):
    p = O.make_params(dict(nx=64, nz=65, lx=2.0, lz=1.0, Pr=1.0, Ra=1e4, alpha=0.51,
                           integrator_order=2, discretisation=["spectral", "fdm"]))
    X, Z = O.grid(p)
    t0 = O.to_spectral(p, 1 - Z + O.noise((p.nx, p.nz), 1e-2, 5))
    w0 = O.to_spectral(p, 400.0 * np.sin(np.pi * X) * np.sin(np.pi * Z))
    j = _fdz_job(mlv, p, "pc", dtype=dtype, dt=2e-4, want_phys=True)
    j.set_state([w0, t0])
    return j


@pytest.mark.parametrize("kind", ["fused", "fdz"])
@pytest.mark.parametrize("ticker", ["host", "device"])
def test_resumed_run_keeps_the_ticker_phase(mlv, tmp_path, kind, ticker):
    """A run checkpointed after 15 loops and resumed for 18 more equals the
    uninterrupted 33-loop run, CFL ticker included (checks stay at loops 1,
    11, 21, 31 as the reference's restored Ticker keeps them); the resumed
    job batched with a fresh one equals both solo runs."""
    if kind == "fused":
        make = lambda: _tgv(mlv)[2]  # noqa: E731
        Batch, mod = mlv.fused.Batch, mlv.fused
    else:
        make = lambda: _pc_case(mlv)  # noqa: E731
        Batch, mod = mlv.fdz.FdzBatch, mlv.fdz
    go = (lambda b, n: mod.run(b, n)) if ticker == "host" else (lambda b, n: mod.run_device(b, n))
    ref = make()
    b = Batch([ref])
    go(b, 33)
    b.close()
    part = make()
    b = Batch([part])
    go(b, 15)
    f = str(tmp_path / "ck.npz")
    mlv.checkpoint.save(part, f)
    b.close()
    print(f"[device-cfl] resume {kind} {ticker}", flush=True)
    res = mlv.checkpoint.load(f, want_phys=True)
    assert res.steps == 15
    b = Batch([res])
    go(b, 18)
    b.close()
    assert res.dt == ref.dt
    for x, y in zip(res.state(), ref.state()):
        np.testing.assert_array_equal(x, y)
    # a resumed job and a fresh one in one batch: each as alone
    res2 = mlv.checkpoint.load(f, want_phys=True)
    fresh, solo = make(), make()
    b = Batch([solo])
    go(b, 18)
    b.close()
    b = Batch([fresh, res2])
    go(b, 18)
    b.close()
    for x, y in zip(res2.state(), ref.state()):
        np.testing.assert_array_equal(x, y)
    for x, y in zip(fresh.state(), solo.state()):
        np.testing.assert_array_equal(x, y)
    assert fresh.dt == solo.dt and res2.dt == ref.dt
