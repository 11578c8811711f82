"""The fused batched step (three HIP kernels per stage over a job table)
against the reference's golden trajectories and the FP64 oracle, the
alone-versus-in-a-mix identity, and the precision form."""
import numpy as np
import pytest

from conftest import golden, rel
from oracle import melvin_oracle as O

pytestmark = pytest.mark.gpu


def _job(mlv, problem, params, dtype="f64", **kw):
    p = O.make_params(params)
    return mlv.fused.Job(problem, p.nx, p.nz, p.lx, p.lz, dtype=dtype, dt=p.initial_dt,
                         ab_order=p.integrator_order, fd_order=p.spatial_derivative_order,
                         alpha=p.alpha, **kw)


def _golden_jobs(mlv, dtype="f64"):
    """(job, check) pairs for every fully spectral golden trajectory."""
    out = []
    g, params, _ = golden("tgv_cfl")
    j = _job(mlv, "vorticity", params, dtype, coeff=[0.05], want_psi=True, want_phys=True)
    j.set_state([g["w0_spec"]])
    out.append((j, len(g["dts"]), lambda j, g=g: [(j.state()[0], g["w_spec"]),
                                                  (j.psi_spectral(), g["psi_spec"])]))
    g, params, _ = golden("kh_ab4")
    j = _job(mlv, "vorticity", params, dtype, coeff=[1e-3, 1e-3], want_phys=True)
    j.set_state([g["w0_spec"], g["c0_spec"]])
    out.append((j, len(g["dts"]), lambda j, g=g: [(j.state()[0], g["w_spec"]),
                                                  (j.state()[1], g["c_spec"])]))
    g, params, _ = golden("vort_o4")
    j = _job(mlv, "vorticity", params, dtype, coeff=[2e-3], want_phys=True)
    j.set_state([g["w0_spec"]])
    out.append((j, len(g["dts"]), lambda j, g=g: [(j.state()[0], g["w_spec"])]))
    g, params, p = golden("ddc")
    j = _job(mlv, "ddc", params, dtype, coeff=[p.Pr, 1.0, p.tau], param=[p.Pr, p.R0],
             want_phys=True)
    j.set_state([g["w0_spec"], g["t0_spec"], g["s0_spec"]])
    out.append((j, 15, lambda j, g=g: [(j.state()[0], g["w_spec"]), (j.state()[1], g["t_spec"]),
                                       (j.state()[2], g["s_spec"])]))
    g, params, p = golden("mhd")
    j = _job(mlv, "mhd", params, dtype, coeff=[1 / p.Re, 1 / p.S], want_phys=True)
    j.set_state([g["w0_spec"], g["j0_spec"]])
    out.append((j, 15, lambda j, g=g: [(j.state()[0], g["w_spec"]), (j.state()[1], g["j_spec"])]))
    return out


def test_fused_single_jobs_vs_golden(mlv):
    for job, n, check in _golden_jobs(mlv):
        b = mlv.fused.Batch([job])
        mlv.fused.run(b, n)
        for got, want in check(job):
            assert rel(got, want) < 1e-12, job.problem
        b.close()


def test_fused_cfl_matches_reference_dt(mlv):
    g, params, _ = golden("tgv_cfl")
    j = _job(mlv, "vorticity", params, coeff=[0.05], want_phys=True)
    j.set_state([g["w0_spec"]])
    b = mlv.fused.Batch([j])
    mlv.fused.run(b, len(g["dts"]))
    assert j.dt == float(g["final_dt"])
    assert j.t == pytest.approx(float(g["t"]), rel=1e-15)


def test_mixed_batch_equals_alone(mlv):
    """Every job's result is identical alone and inside a mixed batch."""
    alone = []
    for job, n, _ in _golden_jobs(mlv):
        b = mlv.fused.Batch([job])
        b.step(12)
        alone.append(job.state())
        b.close()
    jobs = [j for j, _, _ in _golden_jobs(mlv)]
    b = mlv.fused.Batch(jobs)
    b.step(12)
    for j, ref in zip(jobs, alone):
        for x, y in zip(j.state(), ref):
            np.testing.assert_array_equal(x, y)


def test_fused_physical_outputs(mlv):
    g, params, p = golden("kh_ab4")
    j = _job(mlv, "vorticity", params, coeff=[1e-3, 1e-3], want_phys=True, want_psi=True)
    j.set_state([g["w0_spec"], g["c0_spec"]])
    b = mlv.fused.Batch([j])
    b.step(1, write_phys=True)
    w, c, ux, uz = j.physical()
    psi, oux, ouz, _, _ = O.velocity_from_vorticity(p, g["w0_spec"])
    # uz here is mostly high-kx noise, the least favourable case for rounding
    for got, want in ((w, O.to_physical(p, g["w0_spec"])), (c, O.to_physical(p, g["c0_spec"])),
                      (ux, oux), (uz, ouz), (j.psi_spectral(), psi)):
        assert rel(got, want) < 1e-12


@pytest.mark.parametrize("nx,nz", [(64, 32), (128, 128), (256, 64), (2048, 16), (16, 1024)])
def test_fused_sizes_vs_oracle(mlv, nx, nz):
    """Non-square grids of many sizes, 3 steps, AB2, FD order 2."""
    params = dict(nx=nx, nz=nz, lx=2.0, lz=1.5, initial_dt=1e-4)
    p = O.make_params(params)
    rng = np.random.default_rng(nx + 3 * nz)
    w0 = O.to_spectral(p, rng.standard_normal((nx, nz)))
    j = _job(mlv, "vorticity", params, coeff=[1e-3])
    j.set_state([w0])
    b = mlv.fused.Batch([j])
    b.step(3)
    st = O.new_state(p, ("w",))
    st["w"] = w0.copy()
    for _ in range(3):
        O.step_vorticity(p, st, 1e-3)
    assert rel(j.state()[0], st["w"]) < 1e-12


def test_config1_taylor_green_256_fp64(mlv):
    """Taylor-Green 256^2 on [0, 2pi]^2, nu = 1e-2, AB2 semi-implicit, dt = 1e-3,
    1000 steps: <= 1e-12 relative L2 in omega and psi against the oracle, and
    the (1,1) mode decays as exp(-2 nu t) (the theta-scheme error at this dt
    is ~1e-9)."""
    params = dict(nx=256, nz=256, lx=2 * np.pi, lz=2 * np.pi, initial_dt=1e-3)
    p = O.make_params(params)
    w0 = O.ic_taylor_green(p, seed=0, amp=1e-6, kmax=8)
    j = _job(mlv, "vorticity", params, coeff=[1e-2], want_psi=True, want_phys=True)
    j.set_state([w0])
    b = mlv.fused.Batch([j])
    mlv.fused.run(b, 1000)
    st = O.new_state(p, ("w",))
    st["w"] = w0.copy()
    O.run(p, st, lambda p, s: O.step_vorticity(p, s, 1e-2), 1000)
    assert j.dt == st["dt"] == 1e-3
    assert rel(j.state()[0], st["w"]) <= 1e-12
    assert rel(j.psi_spectral(), st["psi"]) <= 1e-12
    amp = j.state()[0][1, 1] / w0[1, 1]
    assert abs(amp - np.exp(-2 * 1e-2 * 1.0)) < 1e-6


def test_precision_form_fp32_close_to_fp64(mlv):
    g, params, _ = golden("kh_ab4")
    res = {}
    for dt in ("f64", "f32"):
        j = _job(mlv, "vorticity", params, dt, coeff=[1e-3, 1e-3])
        j.set_state([g["w0_spec"], g["c0_spec"]])
        b = mlv.fused.Batch([j])
        b.step(50)
        res[dt] = j.state()
    assert rel(res["f32"][0], res["f64"][0]) < 1e-5
    assert rel(res["f32"][1], res["f64"][1]) < 1e-5


def test_graph_replays_match_eager_steps(mlv):
    """Captured 4-step graphs replayed three times, with eager work, syncs and
    a dt change between replays; every replay equals eager stepping."""
    import torch

    g, params, _ = golden("kh_ab4")
    res = {}
    for mode in ("eager", "graph"):
        jobs = []
        for dt in ("f32",):
            j = _job(mlv, "vorticity", params, dt, coeff=[1e-3, 1e-3])
            j.set_state([g["w0_spec"], g["c0_spec"]])
            jobs.append(j)
        b = mlv.fused.Batch(jobs)
        out = []
        for rep in range(3):
            if mode == "eager":
                b.step(4)
            else:
                b.graph_step(4)
            torch.cuda.synchronize()
            out.append(jobs[0].state())
            _ = torch.ones(1000, device="cuda").sum().item()  # unrelated eager work
            if rep == 1:
                b.set_dt(0, jobs[0].dt * 0.5)
        res[mode] = out
        b.close()
    for a, b_ in zip(res["eager"], res["graph"]):
        for x, y in zip(a, b_):
            np.testing.assert_array_equal(x, y)


@pytest.mark.parametrize("nx,nz", [(2048, 32), (32, 2048)])
@pytest.mark.parametrize("problem", ["vort1", "vort2", "ddc", "mhd"])
def test_every_program_at_2048_vs_oracle(mlv, problem, nx, nz):
    """Each z-pass program at N = 2048 along x and along z (16 points per
    thread where the program fits two LDS slots, 32 otherwise; fewer
    columns or rows than wavefronts per block), 3 steps FP64 against the
    oracle, and bit-identical when batched with a small job."""
    params = dict(nx=nx, nz=nz, lx=2.0, lz=1.5, initial_dt=1e-5, Pr=7.0, R0=1.5, tau=0.01,
                  Re=1e3, S=1e3)
    p = O.make_params(params)
    rng = np.random.default_rng(nx + nz)
    f = [O.to_spectral(p, rng.standard_normal((nx, nz))) for _ in range(3)]
    spec = {
        "vort1": (dict(problem="vorticity", coeff=[1e-3]), ("w",),
                  lambda p, st: O.step_vorticity(p, st, 1e-3)),
        "vort2": (dict(problem="vorticity", coeff=[1e-3, 1e-3]), ("w", "c"),
                  lambda p, st: O.step_vorticity(p, st, 1e-3, ("c",))),
        "ddc": (dict(problem="ddc", coeff=[p.Pr, 1.0, p.tau], param=[p.Pr, p.R0]),
                ("w", "tmp", "xi"), O.step_ddc),
        "mhd": (dict(problem="mhd", coeff=[1 / p.Re, 1 / p.S]), ("w", "j"), O.step_mhd),
    }
    kw, names, body = spec[problem]
    init = f[:len(names)]

    def make():
        k = dict(kw)
        j = _job(mlv, k.pop("problem"), params, **k)
        j.set_state(init)
        return j

    print(f"[fused] {problem} {nx}x{nz}", flush=True)
    job = make()
    b = mlv.fused.Batch([job])
    b.step(3)
    st = O.new_state(p, names, dt=1e-5)
    for k, v in zip(names, init):
        st[k] = v.copy()
    for _ in range(3):
        body(p, st)
    for got, name in zip(job.state(), names):
        assert rel(got, st[name]) < 1e-12, name
    b.close()
    ps = dict(nx=32, nz=32, lx=1.0, lz=1.0, initial_dt=1e-4)
    small = _job(mlv, "vorticity", ps, coeff=[1e-3])
    small.set_state([O.to_spectral(O.make_params(ps), rng.standard_normal((32, 32)))])
    mixed = make()
    b = mlv.fused.Batch([small, mixed])
    b.step(3)
    for x, y in zip(mixed.state(), job.state()):
        np.testing.assert_array_equal(x, y)
    b.close()


@pytest.mark.parametrize("dtype", ["f64", "f32"])
def test_autotune_steps_are_real_steps(mlv, dtype):
    """mlv_batch_autotune times real steps with both block shapes of the
    x-inverse kernel and both tile shapes of the update kernel (the golden
    grids are below the z pass's split-pass sizes, so its plan stays):
    afterwards every job equals the same number of plain steps bit for bit
    (the shapes compute identical values), and a batch of the same job
    shapes created later starts with the chosen plan."""
    tuned = [j for j, _, _ in _golden_jobs(mlv, dtype)]
    plain = [j for j, _, _ in _golden_jobs(mlv, dtype)]
    b = mlv.fused.Batch(tuned)
    print(f"[fused] autotune {dtype}", flush=True)
    choice = b.autotune(steps=3)
    assert not choice["z_pass_split_available"] and choice["z_pass"] == "split-pass"
    assert choice["x_forward_update"] in ("wide", "narrow")
    assert choice == b.plan
    b.step(4)
    b.close()
    again = mlv.fused.Batch([j for j, _, _ in _golden_jobs(mlv, dtype)])
    assert again.plan == choice
    again.close()
    b = mlv.fused.Batch(plain)
    b.step(10)
    b.close()
    for t, p in zip(tuned, plain):
        assert t.steps == p.steps == 10
        for x, y in zip(t.state(), p.state()):
            np.testing.assert_array_equal(x, y)


def test_fp32_mean_split_keeps_equilibrium_accuracy(mlv):
    """Resistive tearing (an O(10) current sheet with 1e-4 vorticity noise,
    whose Lorentz force cancels exactly): with the x-mean split the FP32 run
    stays within 1e-5 of FP64 in omega and j; transforming the channels
    whole, FP32 rounding of the sheet's products swamps omega.  A job without
    the split gives the same bits alone and batched beside a split job."""
    S = mlv.serving
    req = S.Request(0, 0.0, "mhd", 128, 120, seed=3)

    def run(dtype, split, partner=None):
        job = S.make_job(req, dtype)
        job.mean_split = split
        jobs = [job] + ([partner] if partner is not None else [])
        b = mlv.fused.Batch(jobs)
        b.step(req.steps)
        b.close()
        return job.state()

    ref = run("f64", False)
    on = run("f32", True)
    off = run("f32", False)
    e_on = [rel(a, b) for a, b in zip(on, ref)]
    e_off = [rel(a, b) for a, b in zip(off, ref)]
    print(f"[fused] tearing 128^2 FP32 vs FP64 (omega, j): split {e_on}, whole {e_off}", flush=True)
    assert e_on[0] < 1e-5 and e_on[1] < 1e-6
    assert e_off[0] > 100 * e_on[0]
    partner = S.make_job(S.Request(1, 0.0, "vortex", 64, 10, seed=4), "f32")
    assert partner.mean_split
    mixed = run("f32", False, partner)
    for x, y in zip(mixed, off):
        np.testing.assert_array_equal(x, y)
