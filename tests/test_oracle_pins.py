"""Pins of the FP64 oracle that need no reference import: the reference's own
known-answer tests restated (test/*.py), closed forms, brute-force and dense
solves.  Each is chosen so a plausible slip (sign, index, transposed axis,
missing normalisation, wrong band) fails it."""
import numpy as np
import pytest

from oracle import melvin_oracle as O


@pytest.fixture
def p256():
    # test/conftest.py:12-21
    return O.make_params(dict(nx=256, nz=256, lx=1.0, lz=1.0))


@pytest.fixture
def p256_fdm():
    # test/conftest.py:24-34
    return O.make_params(dict(nx=256, nz=256, lx=1.0, lz=1.0,
                              discretisation=["spectral", "fdm"]))


# ---------------------------------------------------------------------------
# The reference's known-answer tests
# ---------------------------------------------------------------------------
def test_kat_transform_periodic(p256):
    """SpectralTransformer_test.py:27-61: layout, sign and normalisation."""
    X, Z = O.grid(p256)
    phys = np.cos(2 * np.pi * X) + 2.0 * np.sin(2 * 2 * np.pi * Z)
    spec = O.to_spectral(p256, phys)
    want = np.zeros(p256.spectral_shape, dtype=complex)
    want[1, 0] = 0.5
    want[-1, 0] = 0.5
    want[0, 2] = 2.0j / -2.0
    np.testing.assert_array_almost_equal(spec, want)
    np.testing.assert_array_almost_equal(O.to_physical(p256, spec), phys)


def test_kat_transform_fdm(p256_fdm):
    """SpectralTransformer_test.py:313-348."""
    p = p256_fdm
    x = np.linspace(0, 1.0, p.nx, endpoint=False)
    z = np.linspace(0, 1.0, p.nz)
    X, Z = np.meshgrid(x, z, indexing="ij")
    phys = 3.0 + np.cos(2 * np.pi * X) + 2.0 * np.cos(2 * 2 * np.pi * X)
    spec = O.to_spectral(p, phys)
    want = np.zeros(p.spectral_shape, dtype=complex)
    want[0, :] = 3.0
    want[1, :] = 0.5
    want[2, :] = 1.0
    np.testing.assert_array_almost_equal(spec, want)
    np.testing.assert_array_almost_equal(O.to_physical(p, spec), phys)


def test_kat_spatial_derivatives(p256):
    """Variable_test.py:16-46."""
    X, Z = O.grid(p256)
    f = np.cos(2 * np.pi * X) * np.cos(2 * np.pi * Z)
    x = X[:, 0]
    z = Z[0, :]
    assert O.pddx(p256, f)[:, 0] == pytest.approx(-2 * np.pi * np.sin(2 * np.pi * x), rel=1e-3)
    assert O.pddz(p256, f)[0, :] == pytest.approx(-2 * np.pi * np.sin(2 * np.pi * z), rel=1e-3)


def test_kat_fdm_nabla2(p256_fdm):
    """Variable_test.py:49-73."""
    p = p256_fdm
    z = np.linspace(0, p.lz, p.nz, endpoint=False)
    var = np.zeros(p.spectral_shape, dtype=complex)
    var[1] = 0.5 * z ** 2
    got = O.snabla2(p, var)
    want = np.zeros_like(got)
    want[1] = -((2 * np.pi) ** 2) * var[1] + 1
    np.testing.assert_array_almost_equal(got[1, 1:-1], want[1, 1:-1])


def test_kat_spectral_nabla2(p256):
    """Variable_test.py:76-100."""
    got = O.snabla2(p256, np.ones(p256.spectral_shape, dtype=complex))
    n, m = O.mode_numbers(p256)
    want = -((2 * np.pi / p256.lx * n) ** 2) - (2 * np.pi / p256.lz * m) ** 2
    np.testing.assert_array_almost_equal(got, want)


def test_kat_fdm_laplacian_solver(p256_fdm):
    """LaplacianSolver_test.py:37-59 (with the kx^2 term present)."""
    p = p256_fdm
    lo, di, up = O.fdm_laplacian_bands(p)
    x = np.ones(p.spectral_shape, dtype=complex)
    rhs = di * x
    rhs[:, 1:] += lo[:, 1:] * x[:, :-1]
    rhs[:, :-1] += up[:, :-1] * x[:, 1:]
    np.testing.assert_array_almost_equal(O.poisson_solve(p, rhs), x)


def test_kat_array_shapes(p256, p256_fdm):
    """ArrayFactory_test.py:7-24."""
    assert p256.spectral_shape == (2 * p256.nn + 1, p256.nm)
    assert p256_fdm.spectral_shape == (p256_fdm.nn, p256_fdm.nz)


# ---------------------------------------------------------------------------
# Closed forms and brute force
# ---------------------------------------------------------------------------
def _dft2(a):
    nx, nz = a.shape
    x = np.arange(nx)
    z = np.arange(nz)
    fx = np.exp(-2j * np.pi * np.outer(x, x) / nx)
    fz = np.exp(-2j * np.pi * np.outer(z, z) / nz)
    return fx @ a @ fz


def test_to_spectral_matches_brute_force_dft():
    """Asymmetric grid: catches a transposed axis or a wrong kept band."""
    p = O.make_params(dict(nx=16, nz=8, lx=2.0, lz=3.0))
    rng = np.random.default_rng(3)
    phys = rng.standard_normal((16, 8))
    full = _dft2(phys) / (16 * 8)
    want = np.zeros(p.spectral_shape, dtype=complex)
    for r in range(p.spectral_shape[0]):
        kx = r if r <= p.nn else r - p.spectral_shape[0]
        for c in range(p.nm):
            want[r, c] = full[kx % 16, c]
    np.testing.assert_allclose(O.to_spectral(p, phys), want, atol=1e-14)


def test_to_physical_single_mode_sums():
    """A lone mode (kx, kz) and its real-field image: ignoring the imaginary
    part of the kz = 0 bin, as numpy's c2r does."""
    p = O.make_params(dict(nx=16, nz=16, lx=1.0, lz=1.0))
    X, Z = O.grid(p)
    spec = np.zeros(p.spectral_shape, dtype=complex)
    spec[2, 3] = 0.25 - 0.5j
    got = O.to_physical(p, spec)
    want = 2 * np.real((0.25 - 0.5j) * np.exp(2j * np.pi * (2 * X + 3 * Z)))
    np.testing.assert_allclose(got, want, atol=1e-13)
    spec = np.zeros(p.spectral_shape, dtype=complex)
    spec[0, 0] = 1.0 + 7.0j  # imaginary DC is dropped
    np.testing.assert_allclose(O.to_physical(p, spec), np.ones((16, 16)), atol=1e-14)


@pytest.mark.parametrize("order", [2, 4])
def test_fd_modified_wavenumber(order):
    """Periodic central differences of exp(i k x) are i*sin-type multiples."""
    p = O.make_params(dict(nx=32, nz=16, lx=3.0, lz=1.0, spatial_derivative_order=order))
    X, _ = O.grid(p)
    k = 5
    th = 2 * np.pi * k / p.nx
    f = np.exp(1j * 2 * np.pi * k * X / p.lx)
    if order == 2:
        sym = 1j * np.sin(th) / p.dx
    else:
        sym = 1j * (8 * np.sin(th) - np.sin(2 * th)) / (6 * p.dx)
    np.testing.assert_allclose(O.pddx(p, f), sym * f, atol=1e-12)


def test_fd_z_boundaries_stay_zero():
    p = O.make_params(dict(nx=8, nz=9, lx=1.0, lz=1.0, discretisation=["spectral", "fdm"]))
    f = np.random.default_rng(1).standard_normal((8, 9))
    d = O.pddz(p, f)
    assert np.all(d[:, 0] == 0) and np.all(d[:, -1] == 0)
    np.testing.assert_allclose(d[:, 3], (f[:, 4] - f[:, 2]) / (2 * p.dz))
    d2 = O.pd2dz2(p, f)
    assert np.all(d2[:, 0] == 0) and np.all(d2[:, -1] == 0)


def test_taylor_green_flux_divergence_vanishes():
    """On dx = dz the discrete flux form is zero for Taylor-Green, so TGV is an
    exact discrete steady state of the advection operator."""
    p = O.make_params(dict(nx=32, nz=32, lx=2 * np.pi, lz=2 * np.pi))
    w = O.to_spectral(p, 2 * np.sin(O.grid(p)[0]) * np.sin(O.grid(p)[1]))
    _, ux, uz, _, _ = O.velocity_from_vorticity(p, w)
    adv = O.vec_dot_nabla(p, w, ux, uz)
    assert np.max(np.abs(adv)) < 1e-13


def test_velocity_closed_form():
    """psi = w / k^2; ux = -dpsi/dz, uz = dpsi/dx for omega = sin(2x) cos(z)."""
    p = O.make_params(dict(nx=32, nz=32, lx=2 * np.pi, lz=2 * np.pi))
    X, Z = O.grid(p)
    w = O.to_spectral(p, np.sin(2 * X) * np.cos(Z))
    psi, ux, uz, _, _ = O.velocity_from_vorticity(p, w)
    np.testing.assert_allclose(O.to_physical(p, psi), np.sin(2 * X) * np.cos(Z) / 5, atol=1e-13)
    np.testing.assert_allclose(ux, np.sin(2 * X) * np.sin(Z) / 5, atol=1e-13)
    np.testing.assert_allclose(uz, 2 * np.cos(2 * X) * np.cos(Z) / 5, atol=1e-13)
    # mean mode: psi[0,0] = -w[0,0]
    w2 = w.copy()
    w2[0, 0] = 0.3
    assert O.poisson_solve(p, -w2)[0, 0] == pytest.approx(-0.3)


def test_theta_scheme_decay_factor():
    p = O.make_params(dict(nx=16, nz=16, lx=2 * np.pi, lz=2 * np.pi))
    h = O.History(2, p.spectral_shape)
    v = np.zeros(p.spectral_shape, dtype=complex)
    v[1, 1] = 1.0
    lin = -0.3 * np.ones(p.spectral_shape)
    out = O.integrate_semi_implicit(v, h, lin, 0.1, 0.51)
    assert out[1, 1] == pytest.approx((1 + 0.49 * 0.1 * -0.3) / (1 + 0.51 * 0.1 * 0.3))


@pytest.mark.parametrize("q", [2, 4])
def test_adams_bashforth_polynomial_exactness(q):
    """AB_q integrates polynomials of degree < q in t exactly."""
    dt = 0.1
    for deg in range(q):
        h = O.History(q, (1, 1))
        ts = [0.0 - (q - 1 - i) * dt for i in range(q)]  # oldest ... newest
        for i, t in enumerate(ts):
            h.curr = i
            h.set(np.array([[t ** deg]], dtype=complex))
        h.curr = q - 1
        inc = O.adams_bashforth(h, dt)[0, 0].real
        exact = ((dt) ** (deg + 1) - 0.0) / (deg + 1)
        assert inc == pytest.approx(exact, rel=1e-12, abs=1e-15)


def test_thomas_matches_dense_solve():
    p = O.make_params(dict(nx=16, nz=11, lx=1.7, lz=1.0, discretisation=["spectral", "fdm"]))
    lo, di, up = O.fdm_laplacian_bands(p)
    rng = np.random.default_rng(5)
    rhs = rng.standard_normal(p.spectral_shape) + 1j * rng.standard_normal(p.spectral_shape)
    got = O.poisson_solve(p, rhs)
    for n in range(p.nn):
        A = np.diag(di[n]) + np.diag(lo[n, 1:], -1) + np.diag(up[n, :-1], 1)
        np.testing.assert_allclose(got[n], np.linalg.solve(A, rhs[n]), rtol=1e-12, atol=1e-14)
    # wall rows of the operator: psi_wall = rhs_wall
    np.testing.assert_allclose(got[:, 0], rhs[:, 0])
    np.testing.assert_allclose(got[:, -1], rhs[:, -1])


def test_cfl_rule():
    p = O.make_params(dict(nx=16, nz=16, lx=1.0, lz=1.0, cfl_cutoff=0.5))
    ux = np.full((16, 16), 2.0)
    uz = np.full((16, 16), 1.0)
    # cfl_dt = min(dx/2, dz/1) = 1/32
    assert O.cfl_update_dt(p, 0.01, ux, uz) == pytest.approx(0.01)
    assert O.cfl_update_dt(p, 0.03, ux, uz) == pytest.approx(0.03 * 0.9 ** 7)
    with pytest.raises(Exception, match="CFL"):
        O.cfl_update_dt(p, 0.04, ux, uz)
    with pytest.raises(Exception, match="CFL"):
        O.cfl_update_dt(p, 0.01, ux * np.nan, uz)


def test_taylor_green_decay_semi_implicit():
    """A pure TGV mode decays by the theta-scheme factor each step."""
    p = O.make_params(dict(nx=32, nz=32, lx=2 * np.pi, lz=2 * np.pi))
    st = O.new_state(p, ("w",), dt=1e-2)
    st["w"] = O.ic_taylor_green(p, amp=0.0)
    w0 = st["w"].copy()
    nu = 0.05
    for _ in range(20):
        O.step_vorticity(p, st, nu)
    L = -2 * nu
    g = (1 + 0.49 * 1e-2 * L) / (1 - 0.51 * 1e-2 * L)
    np.testing.assert_allclose(st["w"], w0 * g ** 20, atol=1e-13)
    assert g ** 20 == pytest.approx(np.exp(-2 * nu * 0.2), rel=1e-6)


def test_fdz_semi_implicit_solves_its_system():
    """The theta-implicit FD-z step (an extension; the reference has none)
    solves (I - a dt c L) v_new = (I + (1-a) dt c L) v + AB exactly, mode by
    mode, and for a = 0 it is the forward-Euler/AB update of the same L."""
    p = O.make_params(dict(nx=16, nz=9, lx=2.0, lz=1.0, discretisation=["spectral", "fdm"]))
    rng = np.random.default_rng(7)
    v = rng.standard_normal(p.spectral_shape) + 1j * rng.standard_normal(p.spectral_shape)
    hv = rng.standard_normal((2,) + p.spectral_shape) + 0j
    for alpha in (0.51, 0.0):
        h = O.History(2, p.spectral_shape)
        h.data[:] = hv
        h.curr = 1
        ab = O.adams_bashforth(h, 1e-3)
        out = O.integrate_fdz_semi_implicit(p, v.copy(), h, 0.7, 1e-3, alpha)
        for n in range(p.nn):
            L = np.zeros((p.nz, p.nz))
            for j in range(1, p.nz - 1):
                L[j, j - 1] = L[j, j + 1] = 1 / p.dz ** 2
                L[j, j] = -2 / p.dz ** 2 - (2 * np.pi * n / p.lx) ** 2
            A = np.eye(p.nz) - alpha * 1e-3 * 0.7 * L
            b = v[n] + (1 - alpha) * 1e-3 * 0.7 * (L @ v[n]) + ab[n]
            np.testing.assert_allclose(out[n], np.linalg.solve(A, b), rtol=1e-12, atol=1e-14)


def _rbc_run(step, dt, tend, alpha=0.5, Ra=1e4):
    """Rayleigh-Benard on 16 x 9 from a smooth start, with the ring filled
    with the initial RHS (so the zero-history start of the reference loop
    does not cost an order)."""
    p = O.make_params(dict(nx=16, nz=9, lx=2.0, lz=1.0, discretisation=["spectral", "fdm"],
                           Pr=1.0, Ra=Ra, alpha=alpha, integrator_order=2))
    X, Z = np.meshgrid(np.linspace(0, 2, 16, endpoint=False),
                       np.linspace(0, 1, 9, endpoint=False), indexing="ij")
    st = O.new_state(p, ("w", "tmp"), dt=dt)
    st["w"] = O.to_spectral(p, 50 * np.sin(np.pi * X) * np.sin(np.pi * Z))
    st["tmp"] = O.to_spectral(p, 1 - Z + 0.1 * np.cos(np.pi * X) * np.sin(np.pi * Z))
    O._rbc_walls(p, {"w": st["w"], "tmp": st["tmp"], "psi": np.zeros_like(st["w"])})
    rw, rt, *_ = O._rbc_rhs(p, st["w"], st["tmp"])
    st["dw"].data[:] = rw
    st["dtmp"].data[:] = rt
    for _ in range(int(round(tend / dt))):
        step(p, st)
    return st["w"]


def test_fdz_pc_and_theta_schemes_converge_at_second_order():
    """The FD-z theta scheme (a = 1/2) and the AB2/AM2 predictor-corrector
    are second order in time, and the corrector cuts the error of the
    AB2-only scheme by far more than 5x (a corrector that repeated the AB
    increment, or read the wrong ring slots, fails one of these).  Both
    schemes are extensions: parity unpinned against the reference."""
    tend = 4e-3
    e_si = None
    for step in (O.step_rbc_semi_implicit, O.step_rbc_pc):
        ref = _rbc_run(step, tend / 256, tend)
        errs = [np.linalg.norm(_rbc_run(step, tend / k, tend) - ref) / np.linalg.norm(ref)
                for k in (16, 32, 64)]
        for a, b in zip(errs, errs[1:]):
            assert 3.3 < a / b < 4.8, (step.__name__, errs)
        if step is O.step_rbc_semi_implicit:
            e_si = errs[0]
        else:
            assert errs[0] < e_si / 5, (errs[0], e_si)


# ---------------------------------------------------------------------------
# Cosine and sine bases
# ---------------------------------------------------------------------------
import _basis_kats  # noqa: E402


@pytest.mark.parametrize("name", sorted(_basis_kats.KATS))
def test_kat_transform_bases(p256, name):
    """SpectralTransformer_test.py:64-310: the seven cosine / sine cases,
    forward to the listed coefficients and back to the field."""
    bases, phys, want = _basis_kats.case(name)
    spec = O.to_spectral_basis(p256, phys, bases)
    np.testing.assert_array_almost_equal(spec, want)
    np.testing.assert_array_almost_equal(O.to_physical_basis(p256, spec, bases), phys)


def test_cosine_sine_forward_match_dct_dst_type1():
    """A cosine axis is scipy's DCT-I / (n-1) (k = 0 halved), a sine axis
    (field zero at both ends) scipy's DST-I of the interior / (n-1); the
    other axis periodic and constant."""
    import scipy.fft as sf

    p = O.make_params(dict(nx=64, nz=16, lx=1.0, lz=1.0))
    rng = np.random.default_rng(4)
    f = rng.standard_normal(p.nx)
    phys = np.repeat(f[:, None], p.nz, axis=1)
    got = O.to_spectral_basis(p, phys, (O.COSINE, O.COMPLEX_EXP))[:, 0]
    want = sf.dct(f, type=1) / (p.nx - 1)
    want[0] *= 0.5
    np.testing.assert_allclose(got[: p.nn + 1].real, want[: p.nn + 1], atol=1e-13)
    np.testing.assert_allclose(got[: p.nn + 1].imag, 0, atol=1e-13)
    g = f.copy()
    g[0] = g[-1] = 0.0
    phys = np.repeat(g[:, None], p.nz, axis=1)
    got = O.to_spectral_basis(p, phys, (O.SINE, O.COMPLEX_EXP))[:, 0]
    want = sf.dst(g[1:-1], type=1) / (p.nx - 1)
    np.testing.assert_allclose(got[1: p.nn + 1].real, want[: p.nn], atol=1e-13)
    np.testing.assert_allclose(got[1: p.nn + 1].imag, 0, atol=1e-13)
    assert abs(got[0]) < 1e-13


def test_basis_derivatives_closed_forms(p256):
    """d/dx of a cosine series is the sine series with the factor -pi n
    (and d/dz of a sine series the cosine series with pi m); second
    derivatives keep the basis with -(pi n)^2: each checked in physical
    space against the analytic derivative."""
    x = np.linspace(0, 1.0, p256.nx)
    z = np.linspace(0, 1.0, p256.nz)
    X, Z = np.meshgrid(x, z, indexing="ij")
    f = np.cos(np.pi * X) + 2.0 * np.cos(3 * np.pi * X)
    s = O.to_spectral_basis(p256, f, (O.COSINE, O.COSINE))
    d = O.spectral_op_basis(p256, "ddx", s, (O.COSINE, O.COSINE))
    want = -np.pi * np.sin(np.pi * X) - 6.0 * np.pi * np.sin(3 * np.pi * X)
    np.testing.assert_allclose(O.to_physical_basis(p256, d, (O.SINE, O.COSINE)), want,
                               atol=1e-10)
    g = np.sin(np.pi * Z) + 0.5 * np.sin(2 * np.pi * Z)
    b = (O.COMPLEX_EXP, O.SINE)
    s = O.to_spectral_basis(p256, g, b)
    d = O.spectral_op_basis(p256, "ddz", s, b)
    want = np.pi * np.cos(np.pi * Z) + np.pi * np.cos(2 * np.pi * Z)
    np.testing.assert_allclose(O.to_physical_basis(p256, d, (O.COMPLEX_EXP, O.COSINE)), want,
                               atol=1e-10)
    d2 = O.spectral_op_basis(p256, "d2dz2", s, b)
    want = -np.pi ** 2 * np.sin(np.pi * Z) - 2.0 * np.pi ** 2 * np.sin(2 * np.pi * Z)
    np.testing.assert_allclose(O.to_physical_basis(p256, d2, b), want, atol=1e-9)


def test_basis_poisson_inverts_the_laplacian(p256):
    """sin(pi x) sin(2 pi z) solves lap(psi) = -(pi^2 + 4 pi^2) psi, and
    poisson(lap * f) = f off mode (0, 0), for every basis pair."""
    x = np.linspace(0, 1.0, p256.nx)
    X, Z = np.meshgrid(x, x, indexing="ij")
    f = np.sin(np.pi * X) * np.sin(2 * np.pi * Z)
    b = (O.SINE, O.SINE)
    rhs = O.to_spectral_basis(p256, -5 * np.pi ** 2 * f, b)
    psi = O.spectral_op_basis(p256, "poisson", rhs, b)
    np.testing.assert_allclose(O.to_physical_basis(p256, psi, b), f, atol=1e-12)
    rng = np.random.default_rng(1)
    s = rng.standard_normal(p256.spectral_shape) + 1j * rng.standard_normal(p256.spectral_shape)
    s[0, 0] = 0
    for bx in (O.COMPLEX_EXP, O.COSINE, O.SINE):
        for bz in (O.COMPLEX_EXP, O.COSINE, O.SINE):
            lap = O.spectral_op_basis(p256, "lap", None, (bx, bz))
            back = O.spectral_op_basis(p256, "poisson", lap * s, (bx, bz))
            np.testing.assert_allclose(back, s, rtol=1e-13, atol=1e-13)
