"""The example loop bodies written against a Melvin-style package ``m``
(the reference's API) with array module ``xp``, so the same code drives the
MI355X mirror classes in the GPU tests.  Mirrors examples/*.py; the
initial conditions are those of tests/golden/make_golden.py."""
import numpy as np


def _grid(p):
    x = np.linspace(0, p.lx, p.nx, endpoint=False)
    z = np.linspace(0, p.lz, p.nz, endpoint=False)
    return np.meshgrid(x, z, indexing="ij")


def vorticity(m, xp, params, w0_spec, nsteps, coeff, c0_spec=None):
    """taylor_green_vortex.py:85-95 (+ the KH ink block)."""
    p = m.Parameters(params)
    sim = m.Simulation(p, xp)
    bfs = [m.BasisFunctions.COMPLEX_EXP] * 2
    w = sim.make_variable("w", bfs)
    dw = sim.make_derivative("dw")
    psi = sim.make_variable("psi", bfs)
    ux = sim.make_variable("ux", bfs)
    uz = sim.make_variable("uz", bfs)
    ink = dink = None
    if c0_spec is not None:
        ink = sim.make_variable("ink", bfs)
        dink = sim.make_derivative("dink")
        ink.load(c0_spec)
    sim.init_laplacian_solver(bfs)
    sim.config_cfl(ux, uz)
    w.load(w0_spec)
    dts = []
    from importlib import import_module

    ut = import_module(m.__name__ + ".utility")
    for _ in range(nsteps):
        ut.calc_velocity_from_vorticity(w, psi, ux, uz, sim.get_laplacian_solver())
        lin_op = coeff * w.lap()
        dw[:] = -w.vec_dot_nabla(ux.getp(), uz.getp())
        sim._integrator.integrate(w, dw, lin_op)
        if ink is not None:
            lin_op = coeff * ink.lap()
            dink[:] = -ink.vec_dot_nabla(ux.getp(), uz.getp())
            sim._integrator.integrate(ink, dink, lin_op)
        dts.append(sim._integrator._dt)
        sim.end_loop()
    out = dict(w=w.gets(), psi=psi.gets(), ux=ux.getp(), uz=uz.getp(), dts=dts, t=sim._t,
               sim=sim, ke=ut.calc_kinetic_energy(ux, uz, xp, p))
    if ink is not None:
        out["c"] = ink.gets()
    return out


def ddc(m, xp, params, w0, t0, s0, nsteps):
    """double_diffusive_convection.py:252-279."""
    from importlib import import_module

    ut = import_module(m.__name__ + ".utility")
    p = m.Parameters(params)
    sim = m.Simulation(p, xp)
    bfs = [m.BasisFunctions.COMPLEX_EXP] * 2
    w, tmp, xi = (sim.make_variable(n, bfs) for n in ("w", "tmp", "xi"))
    dw, dtmp, dxi = (sim.make_derivative(n) for n in ("dw", "dtmp", "dxi"))
    psi = sim.make_variable("psi", bfs)
    ux = sim.make_variable("ux", bfs)
    uz = sim.make_variable("uz", bfs)
    sim.init_laplacian_solver(bfs)
    sim.config_cfl(ux, uz)
    w.load(w0)
    tmp.load(t0)
    xi.load(s0)
    for _ in range(nsteps):
        ut.calc_velocity_from_vorticity(w, psi, ux, uz, sim.get_laplacian_solver())
        lin_op = p.Pr * w.lap()
        dw[:] = -w.vec_dot_nabla(ux.getp(), uz.getp()) + p.Pr * xi.sddx() - p.Pr * tmp.sddx()
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
    return dict(w=w.gets(), t=tmp.gets(), s=xi.gets(), psi=psi.gets())


def mhd(m, xp, params, w0, j0, nsteps):
    """resistive_tearing_instability.py:268-292."""
    from importlib import import_module

    ut = import_module(m.__name__ + ".utility")
    p = m.Parameters(params)
    sim = m.Simulation(p, xp)
    bfs = [m.BasisFunctions.COMPLEX_EXP] * 2
    w = sim.make_variable("w", bfs)
    j = sim.make_variable("j", bfs)
    dw, dj = sim.make_derivative("dw"), sim.make_derivative("dj")
    psi, ux, uz = (sim.make_variable(n, bfs) for n in ("psi", "ux", "uz"))
    phi, bx, bz = (sim.make_variable(n, bfs) for n in ("phi", "bx", "bz"))
    sim.init_laplacian_solver(bfs)
    sim.config_cfl(ux, uz)
    w.load(w0)
    j.load(j0)
    for _ in range(nsteps):
        ut.calc_velocity_from_vorticity(w, psi, ux, uz, sim.get_laplacian_solver())
        ut.calc_velocity_from_vorticity(j, phi, bx, bz, sim.get_laplacian_solver())
        lin_op = 1.0 / p.Re * w.lap()
        dw[:] = -w.vec_dot_nabla(ux.getp(), uz.getp()) + j.vec_dot_nabla(bx.getp(), bz.getp())
        sim._integrator.integrate(w, dw, lin_op)
        lin_op = 1.0 / p.S * j.lap()
        dj[:] = -j.vec_dot_nabla(ux.getp(), uz.getp()) + w.vec_dot_nabla(bx.getp(), bz.getp())
        sim._integrator.integrate(j, dj, lin_op)
        sim.end_loop()
    return dict(w=w.gets(), j=j.gets(), psi=psi.gets(), phi=phi.gets())


def rbc(m, xp, params, w0, t0, nsteps):
    """rayleigh_benard_convection.py:95-145 (explicit, Fourier-x / FD-z)."""
    from importlib import import_module

    ut = import_module(m.__name__ + ".utility")
    p = m.Parameters(params)
    sim = m.Simulation(p, xp)
    bfs = [m.BasisFunctions.COMPLEX_EXP, m.BasisFunctions.FDM]
    w = sim.make_variable("w", bfs)
    tmp = sim.make_variable("tmp", bfs)
    dw, dtmp = sim.make_derivative("dw"), sim.make_derivative("dtmp")
    psi, ux, uz = (sim.make_variable(n, bfs) for n in ("psi", "ux", "uz"))
    sim.init_laplacian_solver(bfs)
    sim.config_cfl(ux, uz)
    w.load(w0)
    tmp.load(t0)
    o4 = p.spatial_derivative_order == 4
    for _ in range(nsteps):
        ut.calc_velocity_from_vorticity(w, psi, ux, uz, sim.get_laplacian_solver())
        diffusion_term = p.Pr * w.snabla2()
        dw[:] = -w.vec_dot_nabla(ux.getp(), uz.getp()) - p.Pr * p.Ra * tmp.sddx()
        sim._integrator.integrate(w, dw, diffusion_term)
        diffusion_term = tmp.snabla2()
        dtmp[:] = -tmp.vec_dot_nabla(ux.getp(), uz.getp())
        sim._integrator.integrate(tmp, dtmp, diffusion_term)
        k = 2 if o4 else 1
        w[1:, :k] = 0.0
        w[1:, -k:] = 0.0
        psi[1:, :k] = 0.0
        psi[1:, -k:] = 0.0
        tmp[0, :k] = 1.0
        tmp[0, -k:] = 0.0
        tmp[1:, :k] = 0.0
        tmp[1:, -k:] = 0.0
        psi[0, :] = 0.0
        w[0, :] = 0.0
        sim.end_loop()
    return dict(w=w.gets(), t=tmp.gets(), psi=psi.gets(), ux=ux.getp(), uz=uz.getp())
