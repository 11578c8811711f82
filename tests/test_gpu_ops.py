"""Operator-level parity of the HIP mirror classes (called through the C ABI)
with the reference's golden vectors and the FP64 oracle."""
import numpy as np
import pytest

from conftest import golden, rel
from oracle import melvin_oracle as O

pytestmark = pytest.mark.gpu

OPS = ["ops_fs_o2", "ops_fs_o4", "ops_fdz_o2", "ops_fdz_o4"]


def _setup(mlv, params, precision="double"):
    q = dict(params)
    q["precision"] = precision
    p = mlv.Parameters(q, validate=False)
    af = mlv.ArrayFactory(p, mlv.hip)
    st = mlv.SpectralTransformer(p, mlv.hip, af)
    sd = mlv.SpatialDifferentiator(p, mlv.hip, af)
    fs = p.is_fully_spectral()
    B = mlv.BasisFunctions
    bfs = [B.COMPLEX_EXP, B.COMPLEX_EXP if fs else B.FDM]
    return p, af, st, sd, bfs


def _h(x):
    return x.detach().cpu().numpy()


@pytest.mark.parametrize("name", OPS)
def test_transforms_vs_golden(mlv, name):
    g, params, _ = golden(name)
    p, af, st, sd, bfs = _setup(mlv, params)
    assert rel(_h(st.to_spectral(g["phys"], basis_functions=bfs)), g["to_spectral"]) < 1e-14
    assert rel(_h(st.to_physical(g["spec"], basis_functions=bfs)), g["to_physical"]) < 1e-14


@pytest.mark.parametrize("name", OPS)
def test_transforms_fp32_vs_golden(mlv, name):
    g, params, _ = golden(name)
    p, af, st, sd, bfs = _setup(mlv, params, "single")
    assert rel(_h(st.to_spectral(g["phys"], basis_functions=bfs)), g["to_spectral"]) < 2e-6
    assert rel(_h(st.to_physical(g["spec"], basis_functions=bfs)), g["to_physical"]) < 2e-6


@pytest.mark.parametrize("nx,nz", [(8, 8), (16, 64), (64, 16), (128, 256), (256, 256),
                                   (512, 128), (1024, 1024), (2048, 64), (4096, 32)])
def test_transform_sizes_vs_numpy(mlv, nx, nz):
    params = dict(nx=nx, nz=nz, lx=1.3, lz=0.7, final_time=1.0)
    po = O.make_params(params)
    rng = np.random.default_rng(nx * 7 + nz)
    phys = rng.standard_normal((nx, nz))
    spec = rng.standard_normal(po.spectral_shape) + 1j * rng.standard_normal(po.spectral_shape)
    for prec, tol in (("double", 1e-14), ("single", 3e-6)):
        p, af, st, sd, bfs = _setup(mlv, params, prec)
        assert rel(_h(st.to_spectral(phys)), O.to_spectral(po, phys)) < tol, prec
        assert rel(_h(st.to_physical(spec)), O.to_physical(po, spec)) < tol, prec


@pytest.mark.parametrize("nx,nz", [(32, 9), (64, 513), (1024, 33), (256, 257)])
def test_fdz_transform_sizes_vs_numpy(mlv, nx, nz):
    params = dict(nx=nx, nz=nz, lx=2.0, lz=1.0, final_time=1.0,
                  discretisation=["spectral", "fdm"])
    po = O.make_params(params)
    rng = np.random.default_rng(nx + nz)
    phys = rng.standard_normal((nx, nz))
    spec = rng.standard_normal(po.spectral_shape) + 1j * rng.standard_normal(po.spectral_shape)
    p, af, st, sd, bfs = _setup(mlv, params)
    assert rel(_h(st.to_spectral(phys, basis_functions=bfs)), O.to_spectral(po, phys)) < 1e-14
    assert rel(_h(st.to_physical(spec, basis_functions=bfs)), O.to_physical(po, spec)) < 1e-14


@pytest.mark.parametrize("name", OPS)
def test_derivatives_and_poisson_vs_golden(mlv, name):
    g, params, po = golden(name)
    p, af, st, sd, bfs = _setup(mlv, params)
    assert rel(_h(sd.pddx(g["phys"])), g["pddx"]) <= 1e-15
    assert rel(_h(sd.pddz(g["phys"])), g["pddz"]) <= 1e-15
    assert rel(_h(sd.pd2dz2(g["phys"])), g["pd2dz2"]) <= 1e-15
    assert rel(_h(sd.sddx(g["spec"], bfs[0])), g["sddx"]) <= 1e-15
    assert rel(_h(sd.sd2dx2(g["spec"], bfs[0])), g["sd2dx2"]) <= 1e-15
    assert rel(_h(sd.sd2dz2(g["spec"], bfs[1])), g["sd2dz2"]) <= 1e-15
    if p.is_fully_spectral():
        assert rel(_h(sd.sddz(g["spec"], bfs[1])), g["sddz"]) <= 1e-15
        assert rel(_h(sd.calc_lap(bfs)), g["lap"]) <= 1e-15
    var = mlv.Variable(p, mlv.hip, sd=sd, st=st, array_factory=af, basis_functions=bfs)
    var.sets(g["spec"])
    assert rel(_h(var.snabla2()), g["snabla2"]) <= 1e-15
    assert rel(_h(var.vec_dot_nabla(g["ux"], g["uz"])), g["vec_dot_nabla"]) < 1e-14
    ls = mlv.LaplacianSolver(p, mlv.hip, bfs, spatial_diff=sd, array_factory=af)
    tol = 1e-15 if p.is_fully_spectral() else 1e-13
    assert rel(_h(ls.solve(g["spec"])), g["poisson"]) <= tol


@pytest.mark.parametrize("name", OPS)
def test_integrators_vs_golden(mlv, name):
    g, params, _ = golden(name)
    for order in (2, 4):
        for scheme in ("semi-implicit", "explicit"):
            key = f"int_{scheme}_{order}"
            if key + "_var" not in g.files:
                continue
            q = dict(params)
            q.update(integrator_order=order, integrator=scheme, initial_dt=0.013)
            p, af, st, sd, bfs = _setup(mlv, q)
            integ = mlv.Integrator(p, mlv.hip)
            td = mlv.TimeDerivative(p, mlv.hip)
            td.load(g[key + "_hist_in"])
            td.set_curr_idx(order - 1)
            v = mlv.Variable(p, mlv.hip, sd=sd, st=st, array_factory=af, basis_functions=bfs)
            v.sets(g["spec"])
            integ.integrate(v, td, mlv.hip.asarray(g[key + "_lin"]))
            assert rel(_h(v.gets()), g[key + "_var"]) <= 1e-15, key
            assert rel(_h(td.get_all()), g[key + "_hist_out"]) <= 1e-15, key
            assert td.get_curr_idx() == int(g[key + "_curr_out"])
            # predictor / corrector alone
            td.load(g[key + "_hist_in"])
            td.set_curr_idx(order - 1)
            h = O.History(order, p.spectral_shape)
            h.data[:] = g[key + "_hist_in"]
            h.curr = order - 1
            assert rel(_h(integ.predictor(td)), O.adams_bashforth(h, 0.013)) <= 1e-15
            assert rel(_h(integ.corrector(td)), O.adams_moulton(h, 0.013)) <= 1e-15


def test_reductions(mlv):
    rng = np.random.default_rng(0)
    for dt in (np.float64, np.float32):
        a = rng.standard_normal((257, 129)).astype(dt)
        b = rng.standard_normal((257, 129)).astype(dt)
        ta, tb = mlv.hip.asarray(a), mlv.hip.asarray(b)
        a64, b64 = a.astype(np.float64), b.astype(np.float64)
        assert mlv.hip.max(ta) == float(np.max(a64))
        assert mlv.hip.sum(ta) == pytest.approx(float(np.sum(a64)), rel=1e-13, abs=1e-12)
        assert mlv.hip.mean(ta) == pytest.approx(float(np.mean(a64)), rel=1e-12, abs=1e-14)
        assert mlv.hip.sum_sq2(ta, tb) == pytest.approx(float(np.sum(a64 ** 2 + b64 ** 2)), rel=1e-13)
        assert mlv.hip.dot(ta, tb) == pytest.approx(float(np.sum(a64 * b64)), rel=1e-12, abs=1e-12)
    a[3, 4] = np.nan
    assert np.isnan(mlv.hip.max(mlv.hip.asarray(a)))
    neg = mlv.hip.asarray(-np.ones((10, 10)))
    assert mlv.hip.max(neg) == -1.0  # signed max, as the reference's xp.max


def test_cfl_set_dt(mlv):
    params = dict(nx=16, nz=16, lx=1.0, lz=1.0, final_time=1.0, initial_dt=0.03)
    p, af, st, sd, bfs = _setup(mlv, params)
    integ = mlv.Integrator(p, mlv.hip)
    ux = mlv.Variable(p, mlv.hip, sd=sd, st=st, array_factory=af, basis_functions=bfs)
    uz = mlv.Variable(p, mlv.hip, sd=sd, st=st, array_factory=af, basis_functions=bfs)
    ux.setp(np.full((16, 16), 2.0))
    uz.setp(np.full((16, 16), 1.0))
    integ.set_dt(ux, uz)
    assert integ.get_dt() == pytest.approx(0.03 * 0.9 ** 7)
    integ.override_dt(0.04)
    with pytest.raises(Exception, match="CFL"):
        integ.set_dt(ux, uz)
    ux.setp(np.full((16, 16), np.nan))
    integ.override_dt(0.001)
    with pytest.raises(Exception, match="CFL"):
        integ.set_dt(ux, uz)
