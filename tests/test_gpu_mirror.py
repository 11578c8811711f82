"""The reference's example loop bodies, run unchanged through the MI355X
mirror classes (every operator a HIP kernel), against golden trajectories
written by the reference itself."""
import numpy as np
import pytest

import _bodies
from conftest import golden, rel

pytestmark = pytest.mark.gpu


def _h(x):
    return x.detach().cpu().numpy()


@pytest.mark.parametrize("name,coeff,scalar", [("tgv_cfl", 0.05, False),
                                               ("kh_ab4", 1e-3, True),
                                               ("vort_o4", 2e-3, False)])
def test_vorticity_bodies(mlv, name, coeff, scalar):
    g, params, _ = golden(name)
    out = _bodies.vorticity(mlv, mlv.hip, params, g["w0_spec"], len(g["dts"]), coeff,
                            g["c0_spec"] if scalar else None)
    assert rel(_h(out["w"]), g["w_spec"]) < 1e-12
    assert rel(_h(out["psi"]), g["psi_spec"]) < 1e-12
    assert rel(_h(out["ux"]), g["ux_phys"]) < 1e-12
    assert rel(_h(out["uz"]), g["uz_phys"]) < 1e-12
    if scalar:
        assert rel(_h(out["c"]), g["c_spec"]) < 1e-12
    # CFL ticker and time accumulation identical to the reference
    np.testing.assert_array_equal(np.array(out["dts"]), g["dts"])
    assert out["t"] == float(g["t"])
    assert out["ke"] == pytest.approx(float(g["ke"]), rel=1e-12)


def test_ddc_body(mlv):
    g, params, _ = golden("ddc")
    out = _bodies.ddc(mlv, mlv.hip, params, g["w0_spec"], g["t0_spec"], g["s0_spec"], 15)
    for k, gk in (("w", "w_spec"), ("t", "t_spec"), ("s", "s_spec"), ("psi", "psi_spec")):
        assert rel(_h(out[k]), g[gk]) < 1e-12, k


def test_mhd_body(mlv):
    g, params, _ = golden("mhd")
    out = _bodies.mhd(mlv, mlv.hip, params, g["w0_spec"], g["j0_spec"], 15)
    for k, gk in (("w", "w_spec"), ("j", "j_spec"), ("psi", "psi_spec"), ("phi", "phi_spec")):
        assert rel(_h(out[k]), g[gk]) < 1e-12, k


@pytest.mark.parametrize("name", ["rbc_o2", "rbc_o4"])
def test_rbc_body(mlv, name):
    g, params, _ = golden(name)
    out = _bodies.rbc(mlv, mlv.hip, params, g["w0_spec"], g["t0_spec"], 20)
    for k, gk in (("w", "w_spec"), ("t", "t_spec"), ("psi", "psi_spec"),
                  ("ux", "ux_phys"), ("uz", "uz_phys")):
        assert rel(_h(out[k]), g[gk]) < 1e-12, k


def test_dump_and_resume_bit_exact(mlv, tmp_path):
    """Simulation.dump / load restore state, history, tickers and clock, and
    the resumed run continues bit-for-bit."""
    g, params, _ = golden("kh_ab4")
    B = mlv.BasisFunctions
    bfs = [B.COMPLEX_EXP] * 2

    def build():
        p = mlv.Parameters(params)
        sim = mlv.Simulation(p, mlv.hip)
        w = sim.make_variable("w", bfs)
        dw = sim.make_derivative("dw")
        psi, ux, uz = (sim.make_variable(n, bfs) for n in ("psi", "ux", "uz"))
        sim.init_laplacian_solver(bfs)
        sim.config_cfl(ux, uz)
        sim.config_dump([w], [dw])
        return sim, w, dw, psi, ux, uz

    def steps(sim, w, dw, psi, ux, uz, n):
        from melvin_amd import utility as ut

        for _ in range(n):
            ut.calc_velocity_from_vorticity(w, psi, ux, uz, sim.get_laplacian_solver())
            dw[:] = -w.vec_dot_nabla(ux.getp(), uz.getp())
            sim._integrator.integrate(w, dw, 1e-3 * w.lap())
            sim._loop_counter += 1
            sim._t += sim._integrator._dt

    a = build()
    a[1].load(g["w0_spec"])
    steps(*a, 7)
    fname = a[0].dump(fname=str(tmp_path / "d.npz"))
    steps(*a, 5)
    b = build()
    b[0].load(fname)
    steps(*b, 5)
    np.testing.assert_array_equal(_h(a[1].gets()), _h(b[1].gets()))
    np.testing.assert_array_equal(_h(a[2].get_all()), _h(b[2].get_all()))
    assert a[0]._t == b[0]._t and a[0]._loop_counter == b[0]._loop_counter


@pytest.mark.parametrize("disc", [["spectral", "spectral"], ["spectral", "fdm"]])
def test_init_var_with_noise_matches_reference_rule(mlv, disc):
    """utility.init_var_with_noise (reference utility.py:31-39): epsilon *
    (2U - 1) from default_rng(seed) loaded as physical data, then the
    truncating forward transform -- the oracle's noise and to_spectral on
    the same seed, for both discretisations."""
    from oracle import melvin_oracle as O

    nz = 33 if disc[1] == "fdm" else 32
    params = dict(nx=64, nz=nz, lx=2.0, lz=1.0, final_time=1.0, discretisation=disc)
    p = mlv.Parameters(params)
    sim = mlv.Simulation(p, mlv.hip)
    bfs = [mlv.BasisFunctions.COMPLEX_EXP,
           mlv.BasisFunctions.FDM if disc[1] == "fdm" else mlv.BasisFunctions.COMPLEX_EXP]
    v = sim.make_variable("v", bfs)
    for seed in (0, 11):
        mlv.utility.init_var_with_noise(v, 1e-3, seed=seed)
        op = O.make_params(params)
        want = O.to_spectral(op, O.noise((64, nz), 1e-3, seed))
        assert rel(_h(v.gets()), want) <= 1e-14
    # seed 0 is the default, as in the reference
    mlv.utility.init_var_with_noise(v, 1e-3)
    assert rel(_h(v.gets()), O.to_spectral(O.make_params(params),
                                           O.noise((64, nz), 1e-3, 0))) <= 1e-14


@pytest.mark.parametrize("precision", ["double", "single"])
@pytest.mark.parametrize("disc", [["spectral", "spectral"], ["spectral", "fdm"]])
def test_load_regrids_spectral_data(mlv, disc, precision):
    """Variable.load and TimeDerivative.load carry spectral data of another
    grid size onto the simulation's grid (reference Variable.py:76-85,
    140-151, TimeDerivative.py:56-62): the oracle's regrid_spectral, bit for
    bit, from coarser and from finer grids; data of the right shape loads
    unchanged; a different FD grid is refused."""
    from oracle import melvin_oracle as O

    fdm = disc[1] == "fdm"
    nz = 33 if fdm else 64
    params = dict(nx=64, nz=nz, lx=2.0, lz=1.0, final_time=1.0, discretisation=disc,
                  precision=precision, integrator_order=4)
    p = mlv.Parameters(params)
    sim = mlv.Simulation(p, mlv.hip)
    bfs = [mlv.BasisFunctions.COMPLEX_EXP,
           mlv.BasisFunctions.FDM if fdm else mlv.BasisFunctions.COMPLEX_EXP]
    v = sim.make_variable("v", bfs)
    d = sim.make_derivative("dv")
    op = O.make_params(params)
    cplx = np.complex128 if precision == "double" else np.complex64
    rng = np.random.default_rng(7)
    for nx_src, nz_src in ((48, 48), (96, 128), (64, 64)):
        src_p = O.make_params(dict(nx=nx_src, nz=nz if fdm else nz_src, lx=2.0, lz=1.0,
                                   discretisation=disc))
        shape = src_p.spectral_shape
        src = rng.standard_normal(shape) + 1j * rng.standard_normal(shape)
        print(f"[mirror] load {disc} {precision} from {shape}", flush=True)
        v.load(src)
        want = O.regrid_spectral(op, src).astype(cplx)
        np.testing.assert_array_equal(_h(v.gets()), want)
        levels = [src * (k + 1) for k in range(4)]
        d.load(levels)
        for k in range(4):
            np.testing.assert_array_equal(_h(d.get_all()[k]),
                                          O.regrid_spectral(op, levels[k]).astype(cplx))
    if fdm:
        with pytest.raises(ValueError):
            v.load(np.zeros((10, nz + 2), dtype=complex))
