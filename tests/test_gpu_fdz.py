"""The fused Fourier-x / FD-z step (column and row HIP kernels over a job
table) against the reference's golden Rayleigh-Benard trajectories, the
oracle's theta-implicit and predictor-corrector schemes, the alone-versus-
in-a-mix identity, graph replay and the single-precision form."""
import numpy as np
import pytest

from conftest import golden, rel
from oracle import melvin_oracle as O

pytestmark = pytest.mark.gpu


def _job(mlv, p, scheme, dtype="f64", dt=None, **kw):
    return mlv.fdz.FdzJob(p.nx, p.nz, p.lx, p.lz, p.Pr, p.Ra, dtype=dtype,
                          dt=p.initial_dt if dt is None else dt, scheme=scheme,
                          ab_order=p.integrator_order, fd_order=p.spatial_derivative_order,
                          alpha=p.alpha, **kw)


@pytest.mark.parametrize("name", ["rbc_o2", "rbc_o4"])
def test_fdz_explicit_vs_golden(mlv, name):
    """20 loops of rayleigh_benard_convection.py's body with the reference's
    CFL ticker: omega, T, psi and the last loop's physical velocities."""
    g, params, p = golden(name)
    job = _job(mlv, p, "explicit", want_psi=True, want_phys=True)
    job.set_state([g["w0_spec"], g["t0_spec"]])
    b = mlv.fdz.FdzBatch([job])
    print(f"[fdz] {name} nx={p.nx} nz={p.nz}", flush=True)
    mlv.fdz.run(b, 19)
    b.step(1, write_phys=True)
    w, t = job.state()
    assert rel(w, g["w_spec"]) < 1e-12
    assert rel(t, g["t_spec"]) < 1e-12
    assert rel(job.step_psi(), g["psi_spec"]) < 1e-12
    ux, uz, _, _ = job.physical()
    assert rel(ux, g["ux_phys"]) < 1e-12
    assert rel(uz, g["uz_phys"]) < 1e-12
    assert job.dt == pytest.approx(float(g["final_dt"]), rel=1e-15)
    b.close()


def _rbc_start(p, seed=3):
    X, Z = O.grid(p)
    t = 1 - Z + O.noise((p.nx, p.nz), 1e-2, seed)
    w = O.noise((p.nx, p.nz), 1e-1, seed + 1)
    return O.to_spectral(p, w), O.to_spectral(p, t)


def _oracle_run(p, step, w0, t0, dt, n):
    st = O.new_state(p, ("w", "tmp"), dt=dt)
    st["w"], st["tmp"] = w0.copy(), t0.copy()
    for _ in range(n):
        step(p, st)
    return st


@pytest.mark.parametrize("scheme,order,ab,nz", [
    ("semi-implicit", 2, 2, 33), ("semi-implicit", 4, 4, 33), ("pc", 2, 2, 33), ("pc", 4, 4, 33),
    ("pc", 2, 2, 257), ("semi-implicit", 2, 2, 513), ("pc", 4, 2, 700)])
def test_fdz_theta_schemes_vs_oracle(mlv, scheme, order, ab, nz):
    """theta-implicit diffusion and the AB/AM predictor-corrector (no
    reference counterpart) against the oracle's restatement, FP64.  The z
    sizes cover every row-kernel variant family (one element per lane, and
    5, 9 and 11 elements per lane with partly empty lanes)."""
    p = O.make_params(dict(nx=64, nz=nz, lx=2.0, lz=1.0, Pr=0.7, Ra=2e5, alpha=0.51,
                           integrator_order=ab, spatial_derivative_order=order,
                           discretisation=["spectral", "fdm"]))
    w0, t0 = _rbc_start(p)
    dt, n = 2e-5, 24
    step = O.step_rbc_semi_implicit if scheme == "semi-implicit" else O.step_rbc_pc
    st = _oracle_run(p, step, w0, t0, dt, n)
    job = _job(mlv, p, scheme, dt=dt, want_psi=True)
    job.set_state([w0, t0])
    b = mlv.fdz.FdzBatch([job])
    print(f"[fdz] {scheme} order={order} ab={ab} nz={nz}", flush=True)
    b.step(n)
    w, t = job.state()
    assert rel(w, st["w"]) < 1e-12
    assert rel(t, st["tmp"]) < 1e-12
    assert rel(job.step_psi(), st["psi"]) < 1e-12
    b.close()


def _mixed_jobs(mlv, dtype="f64"):
    specs = [(64, 33, "pc", 2, 2), (32, 17, "explicit", 2, 2), (128, 65, "semi-implicit", 4, 4),
             (64, 100, "pc", 4, 2), (16, 9, "explicit", 4, 4), (32, 300, "pc", 2, 2),
             (32, 513, "pc", 2, 2), (16, 700, "semi-implicit", 2, 4), (256, 200, "pc", 2, 2)]
    jobs = []
    for k, (nx, nz, scheme, order, ab) in enumerate(specs):
        p = O.make_params(dict(nx=nx, nz=nz, lx=2.0, lz=1.0, Pr=1.0, Ra=1e5, alpha=0.51,
                               integrator_order=ab, spatial_derivative_order=order,
                               discretisation=["spectral", "fdm"]))
        w0, t0 = _rbc_start(p, seed=10 + k)
        j = _job(mlv, p, scheme, dtype=dtype, dt=2e-6, want_psi=True)
        j.set_state([w0, t0])
        jobs.append(j)
    return jobs


def test_fdz_mixed_batch_equals_alone(mlv):
    """Every job's result is bit-identical alone and in a mixed batch
    (sizes, schemes, orders and row-kernel variants)."""
    alone = []
    for j in _mixed_jobs(mlv):
        b = mlv.fdz.FdzBatch([j])
        b.step(12)
        alone.append(j.state() + [j.step_psi()])
        b.close()
    jobs = _mixed_jobs(mlv)
    b = mlv.fdz.FdzBatch(jobs)
    b.step(12)
    for j, ref in zip(jobs, alone):
        got = j.state() + [j.step_psi()]
        for a, r in zip(got, ref):
            np.testing.assert_array_equal(a, r)
    b.close()


def test_fdz_graph_replay_equals_eager(mlv):
    """Captured steps replayed several times give the eager results."""
    eager = _mixed_jobs(mlv)
    be = mlv.fdz.FdzBatch(eager)
    graph = _mixed_jobs(mlv)
    bg = mlv.fdz.FdzBatch(graph)
    for rep in range(3):
        be.step(8)
        bg.graph_step(8)
        for a, g in zip(eager, graph):
            for x, y in zip(a.state(), g.state()):
                np.testing.assert_array_equal(x, y)
    be.close()
    bg.close()


def test_fdz_set_dt_rebuilds_the_implicit_operators(mlv):
    """A dt change mid-run equals the oracle run with the same dt change."""
    p = O.make_params(dict(nx=32, nz=25, lx=2.0, lz=1.0, Pr=1.0, Ra=1e5, alpha=0.51,
                           integrator_order=2, discretisation=["spectral", "fdm"]))
    w0, t0 = _rbc_start(p)
    st = O.new_state(p, ("w", "tmp"), dt=2e-5)
    st["w"], st["tmp"] = w0.copy(), t0.copy()
    for k in range(10):
        if k == 5:
            st["dt"] = 1.3e-5
        O.step_rbc_pc(p, st)
    job = _job(mlv, p, "pc", dt=2e-5)
    job.set_state([w0, t0])
    b = mlv.fdz.FdzBatch([job])
    b.step(5)
    b.set_dt(0, 1.3e-5)
    b.step(5)
    w, t = job.state()
    assert rel(w, st["w"]) < 1e-12
    assert rel(t, st["tmp"]) < 1e-12
    b.close()


def test_fdz_single_precision_pc(mlv):
    """FP32 transforms and state (FP64 solves) against the FP64 run of the
    same predictor-corrector case: relative L2 error of omega and T."""
    p = O.make_params(dict(nx=256, nz=129, lx=2.0, lz=1.0, Pr=1.0, Ra=1e6, alpha=0.51,
                           integrator_order=2, discretisation=["spectral", "fdm"]))
    w0, t0 = _rbc_start(p)
    out = {}
    for dtype in ("f64", "f32"):
        job = _job(mlv, p, "pc", dtype=dtype, dt=1e-5)
        job.set_state([w0, t0])
        b = mlv.fdz.FdzBatch([job])
        b.step(200)
        out[dtype] = job.state()
        b.close()
    ew = rel(out["f32"][0], out["f64"][0])
    et = rel(out["f32"][1], out["f64"][1])
    print(f"[fdz] f32 vs f64 after 200 steps: omega {ew:.2e} T {et:.2e}", flush=True)
    assert ew < 1e-5 and et < 1e-5


def test_fdz_prime_after_set_state(mlv):
    """Replacing a job's state between steps and priming the batch gives the
    run that started from that state."""
    p = O.make_params(dict(nx=32, nz=17, lx=2.0, lz=1.0, Pr=1.0, Ra=1e5, alpha=0.51,
                           integrator_order=2, discretisation=["spectral", "fdm"]))
    w0, t0 = _rbc_start(p)
    w1, t1 = _rbc_start(p, seed=9)
    job = _job(mlv, p, "semi-implicit", dt=1e-5)
    job.set_state([w0, t0])
    b = mlv.fdz.FdzBatch([job])
    b.step(4)
    job.set_state([w1, t1])
    job.hist[0].zero_()
    job.hist[1].zero_()
    b.prime()
    b.step(6)
    st = _oracle_run(p, O.step_rbc_semi_implicit, w1, t1, 1e-5, 6)
    w, t = job.state()
    assert rel(w, st["w"]) < 1e-12
    assert rel(t, st["tmp"]) < 1e-12
    b.close()
