#!/usr/bin/env python3
"""Write golden input/output vectors from the reference Melvin.py.

Run in the build container only (the reference is not present elsewhere):

    python tests/golden/make_golden.py /root/reference

It imports the reference package with the NumPy backend, replays the
example loop bodies (examples/*.py) through the reference's own
``Simulation.end_loop`` bookkeeping for a fixed number of loops on small
grids, and saves inputs plus outputs as ``tests/golden/*.npz`` (data only,
no code).  The fixtures pin the composite behaviour that closed forms leave
open: the Adams-Bashforth zero start, the order of the sequential field
updates, the CFL ticker timing, the FD-z boundary writes and the exact
FD-z Laplacian rows.
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))


def _import_reference(ref_root):
    sys.path.insert(0, ref_root)
    import melvin  # noqa: E402
    from melvin import utility  # noqa: E402
    return melvin, utility


def _grid(p, endpoint_z=False):
    x = np.linspace(0, p.lx, p.nx, endpoint=False)
    z = np.linspace(0, p.lz, p.nz, endpoint=endpoint_z)
    return np.meshgrid(x, z, indexing="ij")


def _noise(shape, eps, seed):
    rng = np.random.default_rng(seed)
    return eps * (2 * rng.random(shape) - 1.0)


def case_vorticity(melvin, ut, name, params, w0_phys, nsteps, scalar0=None,
                   coeff=None):
    """taylor_green_vortex.py:85-95 / kelvin_helmholtz (ink enabled)."""
    p = melvin.Parameters(params)
    sim = melvin.Simulation(p, np)
    bfs = [melvin.BasisFunctions.COMPLEX_EXP] * 2
    w = sim.make_variable("w", bfs)
    dw = sim.make_derivative("dw")
    psi = sim.make_variable("psi", bfs)
    ux = sim.make_variable("ux", bfs)
    uz = sim.make_variable("uz", bfs)
    ink = dink = None
    if scalar0 is not None:
        ink = sim.make_variable("ink", bfs)
        dink = sim.make_derivative("dink")
    sim.init_laplacian_solver(bfs)
    sim.config_cfl(ux, uz)
    w.load(w0_phys.copy(), is_physical=True)
    if ink is not None:
        ink.load(scalar0.copy(), is_physical=True)
    out = {"w0_spec": w.gets().copy()}
    if ink is not None:
        out["c0_spec"] = ink.gets().copy()
    dts = []
    for _ in range(nsteps):
        ut.calc_velocity_from_vorticity(w, psi, ux, uz,
                                        sim.get_laplacian_solver())
        lin_op = coeff * w.lap()
        dw[:] = -w.vec_dot_nabla(ux.getp(), uz.getp())
        sim._integrator.integrate(w, dw, lin_op)
        if ink is not None:
            lin_op = coeff * ink.lap()
            dink[:] = -ink.vec_dot_nabla(ux.getp(), uz.getp())
            sim._integrator.integrate(ink, dink, lin_op)
        dts.append(sim._integrator._dt)
        sim.end_loop()
    out.update(
        w_spec=w.gets().copy(), psi_spec=psi.gets().copy(),
        ux_phys=ux.getp().copy(), uz_phys=uz.getp().copy(),
        w_phys=w.getp().copy(), dw_hist=dw.get_all().copy(),
        dw_curr=dw.get_curr_idx(), dts=np.array(dts), t=sim._t,
        final_dt=sim._integrator._dt, loops=sim._loop_counter,
        ke=ut.calc_kinetic_energy(ux, uz, np, p))
    if ink is not None:
        out["c_spec"] = ink.gets().copy()
    out["params"] = np.array(repr(params))
    np.savez(os.path.join(HERE, name + ".npz"), **out)
    print(name, "steps", nsteps, "final dt", sim._integrator._dt)


def case_ddc(melvin, ut, name, params, nsteps):
    """double_diffusive_convection.py:178-126 loop body."""
    p = melvin.Parameters(params)
    sim = melvin.Simulation(p, np)
    bfs = [melvin.BasisFunctions.COMPLEX_EXP] * 2
    w = sim.make_variable("w", bfs)
    tmp = sim.make_variable("tmp", bfs)
    xi = sim.make_variable("xi", bfs)
    dw, dtmp, dxi = (sim.make_derivative(n) for n in ("dw", "dtmp", "dxi"))
    psi = sim.make_variable("psi", bfs)
    ux = sim.make_variable("ux", bfs)
    uz = sim.make_variable("uz", bfs)
    sim.init_laplacian_solver(psi._basis_functions)
    sim.config_cfl(ux, uz)
    # Distinct seeds so the three fields differ (the example reuses seed 0).
    ut.init_var_with_noise(w, 1e-2, seed=1)
    ut.init_var_with_noise(tmp, 1e-2, seed=2)
    ut.init_var_with_noise(xi, 1e-2, seed=3)
    out = {"w0_spec": w.gets().copy(), "t0_spec": tmp.gets().copy(),
           "s0_spec": xi.gets().copy()}
    for _ in range(nsteps):
        ut.calc_velocity_from_vorticity(w, psi, ux, uz,
                                        sim.get_laplacian_solver())
        lin_op = p.Pr * w.lap()
        dw[:] = (-w.vec_dot_nabla(ux.getp(), uz.getp()) + p.Pr * xi.sddx()
                 - p.Pr * tmp.sddx())
        sim._integrator.integrate(w, dw, lin_op)
        lin_op = tmp.lap()
        dtmp[:] = -tmp.vec_dot_nabla(ux.getp(), uz.getp()) - uz[:]
        sim._integrator.integrate(tmp, dtmp, lin_op)
        lin_op = p.tau * xi.lap()
        dxi[:] = -xi.vec_dot_nabla(ux.getp(), uz.getp()) - uz[:] / p.R0
        sim._integrator.integrate(xi, dxi, lin_op)
        tmp[:, 0] = 0.0
        xi[:, 0] = 0.0
        sim.end_loop()
    out.update(w_spec=w.gets().copy(), t_spec=tmp.gets().copy(),
               s_spec=xi.gets().copy(), psi_spec=psi.gets().copy(),
               ux_phys=ux.getp().copy(), uz_phys=uz.getp().copy(),
               t=sim._t, final_dt=sim._integrator._dt,
               params=np.array(repr(params)))
    np.savez(os.path.join(HERE, name + ".npz"), **out)
    print(name, "steps", nsteps, "final dt", sim._integrator._dt)


def case_mhd(melvin, ut, name, params, nsteps):
    """resistive_tearing_instability.py:268-292 loop body."""
    p = melvin.Parameters(params)
    sim = melvin.Simulation(p, np)
    bfs = [melvin.BasisFunctions.COMPLEX_EXP] * 2
    w = sim.make_variable("w", bfs)
    j = sim.make_variable("j", bfs)
    dw, dj = sim.make_derivative("dw"), sim.make_derivative("dj")
    psi = sim.make_variable("psi", bfs)
    ux = sim.make_variable("ux", bfs)
    uz = sim.make_variable("uz", bfs)
    phi = sim.make_variable("phi", bfs)
    bx = sim.make_variable("bx", bfs)
    bz = sim.make_variable("bz", bfs)
    sim.init_laplacian_solver(psi._basis_functions)
    sim.config_cfl(ux, uz)
    X, Z = _grid(p)
    j0 = -np.power(1.0 / np.cosh((Z - 0.5) / 0.05), 2) / 0.05
    j0 = j0 + 0.01 * (2 * np.random.default_rng(0).random(X.shape) - 1.0)
    w0 = 0.05 * np.sin(2 * np.pi * X / p.lx) * np.cos(2 * np.pi * Z / p.lz)
    j.load(j0, is_physical=True)
    w.load(w0, is_physical=True)
    out = {"w0_spec": w.gets().copy(), "j0_spec": j.gets().copy()}
    for _ in range(nsteps):
        ut.calc_velocity_from_vorticity(w, psi, ux, uz,
                                        sim.get_laplacian_solver())
        ut.calc_velocity_from_vorticity(j, phi, bx, bz,
                                        sim.get_laplacian_solver())
        lin_op = 1.0 / p.Re * w.lap()
        dw[:] = -w.vec_dot_nabla(ux.getp(), uz.getp()) + j.vec_dot_nabla(
            bx.getp(), bz.getp())
        sim._integrator.integrate(w, dw, lin_op)
        lin_op = 1.0 / p.S * j.lap()
        dj[:] = -j.vec_dot_nabla(ux.getp(), uz.getp()) + w.vec_dot_nabla(
            bx.getp(), bz.getp())
        sim._integrator.integrate(j, dj, lin_op)
        sim.end_loop()
    out.update(w_spec=w.gets().copy(), j_spec=j.gets().copy(),
               psi_spec=psi.gets().copy(), phi_spec=phi.gets().copy(),
               t=sim._t, final_dt=sim._integrator._dt,
               params=np.array(repr(params)))
    np.savez(os.path.join(HERE, name + ".npz"), **out)
    print(name, "steps", nsteps, "final dt", sim._integrator._dt)


def case_rbc(melvin, ut, name, params, nsteps):
    """rayleigh_benard_convection.py:95-145 loop body (explicit, FD-z)."""
    p = melvin.Parameters(params)
    sim = melvin.Simulation(p, np)
    bfs = [melvin.BasisFunctions.COMPLEX_EXP, melvin.BasisFunctions.FDM]
    w = sim.make_variable("w", bfs)
    tmp = sim.make_variable("tmp", bfs)
    dw, dtmp = sim.make_derivative("dw"), sim.make_derivative("dtmp")
    psi = sim.make_variable("psi", bfs)
    ux = sim.make_variable("ux", bfs)
    uz = sim.make_variable("uz", bfs)
    sim.init_laplacian_solver(psi._basis_functions)
    sim.config_cfl(ux, uz)
    X, Z = _grid(p)
    tmp.load(1 - Z + 1e-2 * np.sin(np.pi * X / p.lx), is_physical=True)
    ut.init_var_with_noise(w, 1e-2)
    out = {"w0_spec": w.gets().copy(), "t0_spec": tmp.gets().copy(),
           "laps0": sim.get_laplacian_solver().laps[0].toarray(),
           "laps1": sim.get_laplacian_solver().laps[1].toarray()}
    o4 = p.spatial_derivative_order == 4
    for _ in range(nsteps):
        ut.calc_velocity_from_vorticity(w, psi, ux, uz,
                                        sim.get_laplacian_solver())
        diffusion_term = p.Pr * w.snabla2()
        dw[:] = (-w.vec_dot_nabla(ux.getp(), uz.getp())
                 - p.Pr * p.Ra * tmp.sddx())
        sim._integrator.integrate(w, dw, diffusion_term)
        diffusion_term = tmp.snabla2()
        dtmp[:] = -tmp.vec_dot_nabla(ux.getp(), uz.getp())
        sim._integrator.integrate(tmp, dtmp, diffusion_term)
        if not o4:
            w[1:, 0] = 0.0
            w[1:, -1] = 0.0
            psi[1:, 0] = 0.0
            psi[1:, -1] = 0.0
            tmp[0, 0] = 1.0
            tmp[0, -1] = 0.0
            tmp[1:, 0] = 0.0
            tmp[1:, -1] = 0.0
        else:
            w[1:, :2] = 0.0
            w[1:, -2:] = 0.0
            psi[1:, :2] = 0.0
            psi[1:, -2:] = 0.0
            tmp[0, :2] = 1.0
            tmp[0, -2:] = 0.0
            tmp[1:, :2] = 0.0
            tmp[1:, -2:] = 0.0
        psi[0, :] = 0.0
        w[0, :] = 0.0
        sim.end_loop()
    out.update(w_spec=w.gets().copy(), t_spec=tmp.gets().copy(),
               psi_spec=psi.gets().copy(), ux_phys=ux.getp().copy(),
               uz_phys=uz.getp().copy(), t=sim._t,
               final_dt=sim._integrator._dt, params=np.array(repr(params)))
    np.savez(os.path.join(HERE, name + ".npz"), **out)
    print(name, "steps", nsteps, "final dt", sim._integrator._dt)


def case_ops(melvin, name, params, seed):
    """Single-operator vectors: transforms, FD stencils, spectral
    derivatives, Laplacian symbol/solve and the integrators."""
    p = melvin.Parameters(params, validate=False)
    af = melvin.ArrayFactory(p, np)
    st = melvin.SpectralTransformer(p, np, af)
    sd = melvin.SpatialDifferentiator(p, np, af)
    rng = np.random.default_rng(seed)
    fs = p.is_fully_spectral()
    bfs = [melvin.BasisFunctions.COMPLEX_EXP,
           melvin.BasisFunctions.COMPLEX_EXP if fs else
           melvin.BasisFunctions.FDM]
    phys = rng.standard_normal((p.nx, p.nz))
    ux = rng.standard_normal((p.nx, p.nz))
    uz = rng.standard_normal((p.nx, p.nz))
    spec = (rng.standard_normal(p.spectral_shape)
            + 1j * rng.standard_normal(p.spectral_shape))
    out = {"phys": phys, "spec": spec, "ux": ux, "uz": uz}
    out["to_spectral"] = st.to_spectral(phys.copy(), basis_functions=bfs)
    out["to_physical"] = st.to_physical(spec.copy(), basis_functions=bfs)
    out["pddx"] = sd.pddx(phys)
    out["pddz"] = sd.pddz(phys)
    out["pd2dz2"] = sd.pd2dz2(phys)
    out["sddx"] = sd.sddx(spec, bfs[0])
    out["sd2dx2"] = sd.sd2dx2(spec, bfs[0])
    out["sd2dz2"] = sd.sd2dz2(spec, bfs[1])
    if fs:
        out["sddz"] = sd.sddz(spec, bfs[1])
        out["lap"] = sd.calc_lap(bfs)
    var = melvin.Variable(p, np, sd=sd, st=st, array_factory=af,
                          basis_functions=bfs)
    var.sets(spec.copy())
    out["vec_dot_nabla"] = var.vec_dot_nabla(ux, uz)
    var.sets(spec.copy())
    out["snabla2"] = var.snabla2()
    ls = melvin.LaplacianSolver(p, np, bfs, spatial_diff=sd,
                                array_factory=af)
    out["poisson"] = ls.solve(spec.copy())
    # Integrators: history of random RHS values.
    for order in (2, 4):
        for scheme in ("semi-implicit", "explicit"):
            if scheme == "semi-implicit" and not fs:
                continue
            q = dict(params)
            q.update(integrator_order=order, integrator=scheme,
                     initial_dt=0.013)
            pq = melvin.Parameters(q, validate=False)
            integ = melvin.Integrator(pq, np)
            td = melvin.TimeDerivative(pq, np)
            hist = (rng.standard_normal((order,) + p.spectral_shape)
                    + 1j * rng.standard_normal((order,) + p.spectral_shape))
            td._data[:] = hist
            td.set_curr_idx(order - 1)
            v = melvin.Variable(pq, np, sd=sd, st=st, array_factory=af,
                                basis_functions=bfs)
            v.sets(spec.copy())
            lin = (0.7 * sd.calc_lap(bfs) if fs else
                   rng.standard_normal(p.spectral_shape)
                   + 1j * rng.standard_normal(p.spectral_shape))
            integ.integrate(v, td, lin)
            key = f"int_{scheme}_{order}"
            out[key + "_hist_in"] = hist
            out[key + "_lin"] = lin
            out[key + "_var"] = v.gets().copy()
            out[key + "_hist_out"] = td.get_all().copy()
            out[key + "_curr_out"] = td.get_curr_idx()
    out["params"] = np.array(repr(params))
    np.savez(os.path.join(HERE, name + ".npz"), **out)
    print(name, "ops written")


def main():
    ref = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
    melvin, ut = _import_reference(ref)

    # Single-operator vectors (non-square grids, unequal lengths).
    case_ops(melvin, "ops_fs_o2", dict(nx=32, nz=64, lx=2.0, lz=1.5,
                                       final_time=1.0), seed=11)
    case_ops(melvin, "ops_fs_o4", dict(nx=64, nz=32, lx=1.0, lz=3.0,
                                       final_time=1.0,
                                       spatial_derivative_order=4), seed=12)
    case_ops(melvin, "ops_fdz_o2", dict(nx=32, nz=17, lx=2.0, lz=1.0,
                                        final_time=1.0,
                                        discretisation=["spectral", "fdm"]),
             seed=13)
    case_ops(melvin, "ops_fdz_o4", dict(nx=64, nz=21, lx=1.5, lz=1.0,
                                        final_time=1.0,
                                        spatial_derivative_order=4,
                                        discretisation=["spectral", "fdm"]),
             seed=14)

    # Taylor-Green-like, with a large dt so the CFL ticker shrinks it.
    P = dict(nx=32, nz=32, lx=2 * np.pi, lz=2 * np.pi, final_time=10.0,
             initial_dt=0.12, integrator_order=2, cfl_cutoff=0.5)
    p = melvin.Parameters(P)
    X, Z = _grid(p)
    w0 = -2 * np.cos(X) * np.cos(Z) + _noise(X.shape, 0.3, 5)
    case_vorticity(melvin, ut, "tgv_cfl", P, w0, 25, coeff=0.05)

    # Kelvin-Helmholtz-like omega + passive scalar, AB4, non-square.
    P = dict(nx=64, nz=32, lx=2.0, lz=1.0, final_time=10.0,
             initial_dt=2e-3, integrator_order=4)
    p = melvin.Parameters(P)
    X, Z = _grid(p)
    u = np.tanh((Z - 0.25) / 0.08) - np.tanh((Z - 0.75) / 0.08)
    w0 = (1 / np.cosh((Z - 0.25) / 0.08) ** 2
          - 1 / np.cosh((Z - 0.75) / 0.08) ** 2) / 0.08
    w0 = w0 + _noise(X.shape, 1e-2, 7)
    case_vorticity(melvin, ut, "kh_ab4", P, w0, 20, scalar0=u.copy(),
                   coeff=1e-3)

    # 4th-order FD advection, fully spectral.
    P = dict(nx=32, nz=64, lx=1.0, lz=2.0, final_time=10.0,
             initial_dt=5e-3, spatial_derivative_order=4)
    p = melvin.Parameters(P)
    X, Z = _grid(p)
    w0 = np.sin(2 * np.pi * X) * np.cos(np.pi * Z) + _noise(X.shape, 0.2, 9)
    case_vorticity(melvin, ut, "vort_o4", P, w0, 15, coeff=2e-3)

    case_ddc(melvin, ut, "ddc", dict(nx=32, nz=32, lx=6.0, lz=4.0,
                                     final_time=10.0, initial_dt=1e-2,
                                     Pr=7.0, R0=1.1, tau=1.0 / 3.0), 15)
    case_mhd(melvin, ut, "mhd", dict(nx=32, nz=32, lx=16.0 / 9, lz=1.0,
                                     final_time=10.0, initial_dt=5e-4,
                                     Re=1e3, S=1e3), 15)
    case_rbc(melvin, ut, "rbc_o2", dict(nx=32, nz=17, lx=2.0, lz=1.0,
                                        final_time=10.0, initial_dt=2e-5,
                                        Pr=1.0, Ra=1e5,
                                        integrator="explicit",
                                        discretisation=["spectral", "fdm"]),
             20)
    case_rbc(melvin, ut, "rbc_o4", dict(nx=32, nz=21, lx=2.0, lz=1.0,
                                        final_time=10.0, initial_dt=1e-5,
                                        Pr=1.0, Ra=1e5,
                                        integrator="explicit",
                                        spatial_derivative_order=4,
                                        discretisation=["spectral", "fdm"]),
             20)


if __name__ == "__main__":
    main()
