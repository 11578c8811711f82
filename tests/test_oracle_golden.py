"""The oracle against golden vectors written by the reference itself
(tests/golden/make_golden.py): single operators and multi-step loop bodies
(AB zero start, sequential update order, CFL ticker, FD-z wall writes)."""
import numpy as np
import pytest

from conftest import golden, rel
from oracle import melvin_oracle as O

OPS = ["ops_fs_o2", "ops_fs_o4", "ops_fdz_o2", "ops_fdz_o4"]


@pytest.mark.parametrize("name", OPS)
def test_single_operators(name):
    g, _, p = golden(name)
    got = {
        "to_spectral": O.to_spectral(p, g["phys"]),
        "to_physical": O.to_physical(p, g["spec"]),
        "pddx": O.pddx(p, g["phys"]),
        "pddz": O.pddz(p, g["phys"]),
        "pd2dz2": O.pd2dz2(p, g["phys"]),
        "sddx": O.sddx(p, g["spec"]),
        "sd2dx2": O.sd2dx2(p, g["spec"]),
        "sd2dz2": O.sd2dz2(p, g["spec"]),
        "vec_dot_nabla": O.vec_dot_nabla(p, g["spec"], g["ux"], g["uz"]),
        "snabla2": O.snabla2(p, g["spec"]),
        "poisson": O.poisson_solve(p, g["spec"]),
    }
    if p.fully_spectral:
        got["sddz"] = O.sddz(p, g["spec"])
        got["lap"] = O.calc_lap(p)
    for key, val in got.items():
        tol = 1e-13 if key == "poisson" and not p.fully_spectral else 1e-15
        assert rel(val, g[key]) <= tol, key


@pytest.mark.parametrize("name", OPS)
def test_integrators(name):
    g, _, p = golden(name)
    for order in (2, 4):
        for scheme in ("semi-implicit", "explicit"):
            key = f"int_{scheme}_{order}"
            if key + "_var" not in g.files:
                continue
            h = O.History(order, p.spectral_shape)
            h.data[:] = g[key + "_hist_in"]
            h.curr = order - 1
            if scheme == "semi-implicit":
                v = O.integrate_semi_implicit(g["spec"].copy(), h, g[key + "_lin"], 0.013, 0.51)
            else:
                v = O.integrate_explicit(g["spec"].copy(), h, g[key + "_lin"], 0.013)
            assert rel(v, g[key + "_var"]) <= 1e-15, key
            assert rel(h.data, g[key + "_hist_out"]) <= 1e-15, key
            assert h.curr == int(g[key + "_curr_out"])


@pytest.mark.parametrize("name,coeff,scalar", [("tgv_cfl", 0.05, False),
                                               ("kh_ab4", 1e-3, True),
                                               ("vort_o4", 2e-3, False)])
def test_vorticity_trajectories(name, coeff, scalar):
    g, _, p = golden(name)
    fields = ("w", "c") if scalar else ("w",)
    st = O.new_state(p, fields)
    st["w"] = g["w0_spec"].copy()
    if scalar:
        st["c"] = g["c0_spec"].copy()
    t, dts = O.run(p, st, lambda p, s: O.step_vorticity(p, s, coeff, fields[1:]), len(g["dts"]))
    assert rel(st["w"], g["w_spec"]) <= 1e-14
    assert rel(st["psi"], g["psi_spec"]) <= 1e-14
    if scalar:
        assert rel(st["c"], g["c_spec"]) <= 1e-14
    np.testing.assert_array_equal(np.array(dts), g["dts"])
    assert st["dt"] == float(g["final_dt"]) and t == float(g["t"])


def test_cfl_ticker_fired():
    """The tgv_cfl fixture starts above the CFL cut-off: dt must shrink at loop 1."""
    g, _, _ = golden("tgv_cfl")
    assert g["dts"][0] == pytest.approx(0.12) and g["dts"][1] < 0.12


def test_ddc_trajectory():
    g, _, p = golden("ddc")
    st = O.new_state(p, ("w", "tmp", "xi"))
    st["w"], st["tmp"], st["xi"] = g["w0_spec"].copy(), g["t0_spec"].copy(), g["s0_spec"].copy()
    O.run(p, st, O.step_ddc, 15)
    for k, gk in (("w", "w_spec"), ("tmp", "t_spec"), ("xi", "s_spec"), ("psi", "psi_spec")):
        assert rel(st[k], g[gk]) <= 1e-14, k


def test_mhd_trajectory():
    g, _, p = golden("mhd")
    st = O.new_state(p, ("w", "j"))
    st["w"], st["j"] = g["w0_spec"].copy(), g["j0_spec"].copy()
    O.run(p, st, O.step_mhd, 15)
    for k, gk in (("w", "w_spec"), ("j", "j_spec"), ("psi", "psi_spec"), ("phi", "phi_spec")):
        assert rel(st[k], g[gk]) <= 1e-14, k


@pytest.mark.parametrize("name", ["rbc_o2", "rbc_o4"])
def test_rbc_trajectory(name):
    g, _, p = golden(name)
    st = O.new_state(p, ("w", "tmp"))
    st["w"], st["tmp"] = g["w0_spec"].copy(), g["t0_spec"].copy()
    O.run(p, st, O.step_rbc, 20)
    # Thomas here vs SuperLU in the reference: agreement to rounding
    assert rel(st["w"], g["w_spec"]) <= 1e-12
    assert rel(st["tmp"], g["t_spec"]) <= 1e-12
    assert rel(st["psi"], g["psi_spec"]) <= 1e-12
    lo, di, up = O.fdm_laplacian_bands(p)
    for n in (0, 1):
        A = np.diag(di[n]) + np.diag(lo[n, 1:], -1) + np.diag(up[n, :-1], 1)
        np.testing.assert_array_equal(A, g[f"laps{n}"])
