"""The reference's own test suite (test/*.py), each test restated against the
mirror classes with the HIP backend as ``xp``: the same calls, arguments
and assertions, on the 256 x 256 unit box of test/conftest.py:12-34.  The
seven cosine/sine transform tests of SpectralTransformer_test.py:64-310 are
in test_gpu_bases.py."""
import numpy as np
import pytest
from numpy.testing import assert_array_almost_equal
from pytest import approx

pytestmark = pytest.mark.gpu


def _h(x):
    return x.detach().cpu().numpy() if hasattr(x, "detach") else np.asarray(x)


@pytest.fixture
def parameters(mlv):
    # test/conftest.py:12-21
    return mlv.Parameters({"nx": 4 ** 4, "nz": 2 ** 8, "lx": 1.0, "lz": 1.0, "final_time": 1.0},
                          validate=False)


@pytest.fixture
def fdm_parameters(mlv):
    # test/conftest.py:24-34
    return mlv.Parameters({"nx": 4 ** 4, "nz": 2 ** 8, "lx": 1.0, "lz": 1.0, "final_time": 1.0,
                           "discretisation": ["spectral", "fdm"]}, validate=False)


def test_fully_spectral_array_production(mlv, parameters):
    """ArrayFactory_test.py:7-14."""
    af = mlv.ArrayFactory(parameters, mlv.hip)
    assert tuple(af.make_spectral().shape) == (2 * parameters.nn + 1, parameters.nm)
    assert tuple(af.make_physical().shape) == (parameters.nx, parameters.nz)


def test_fdm_array_production(mlv, fdm_parameters):
    """ArrayFactory_test.py:17-24."""
    af = mlv.ArrayFactory(fdm_parameters, mlv.hip)
    assert tuple(af.make_spectral().shape) == (fdm_parameters.nn, fdm_parameters.nz)
    assert tuple(af.make_physical().shape) == (fdm_parameters.nx, fdm_parameters.nz)


def test_transform_periodic(mlv, parameters):
    """SpectralTransformer_test.py:27-61."""
    p = parameters
    af = mlv.ArrayFactory(p, mlv.hip)
    st = mlv.SpectralTransformer(p, mlv.hip, af)
    spectral, physical = af.make_spectral(), af.make_physical()
    x = np.linspace(0, 1.0, p.nx, endpoint=False)
    z = np.linspace(0, 1.0, p.nz, endpoint=False)
    X, Z = np.meshgrid(x, z, indexing="ij")
    true_physical = np.cos(2 * np.pi * X) + 2.0 * np.sin(2 * 2 * np.pi * Z)
    ce = [mlv.BasisFunctions.COMPLEX_EXP, mlv.BasisFunctions.COMPLEX_EXP]
    st.to_spectral(mlv.hip.asarray(true_physical), spectral, basis_functions=ce)
    true_spectral = np.zeros(tuple(spectral.shape), dtype=complex)
    true_spectral[1, 0] = 1.0 / 2.0
    true_spectral[-1, 0] = 1.0 / 2.0
    true_spectral[0, 2] = 2.0j / -2.0
    assert_array_almost_equal(_h(spectral), true_spectral)
    st.to_physical(spectral, physical, basis_functions=ce)
    assert_array_almost_equal(_h(physical), true_physical)


def test_transform_periodic_x_fdm_z(mlv, fdm_parameters):
    """SpectralTransformer_test.py:313-348."""
    p = fdm_parameters
    af = mlv.ArrayFactory(p, mlv.hip)
    st = mlv.SpectralTransformer(p, mlv.hip, array_factory=af)
    spectral, physical = af.make_spectral(), af.make_physical()
    x = np.linspace(0, 1.0, p.nx, endpoint=False)
    z = np.linspace(0, 1.0, p.nz)
    X, Z = np.meshgrid(x, z, indexing="ij")
    true_physical = 3.0 + np.cos(2 * np.pi * X) + 2.0 * np.cos(2 * 2 * np.pi * X)
    bf = [mlv.BasisFunctions.COMPLEX_EXP, mlv.BasisFunctions.FDM]
    st.to_spectral(mlv.hip.asarray(true_physical), spectral, basis_functions=bf)
    true_spectral = np.zeros(tuple(spectral.shape), dtype=complex)
    true_spectral[0, :] = 3.0
    true_spectral[1, :] = 1.0 / 2
    true_spectral[2, :] = 2.0 / 2
    assert_array_almost_equal(_h(spectral), true_spectral)
    st.to_physical(spectral, physical, basis_functions=bf)
    assert_array_almost_equal(_h(physical), true_physical)


def test_spatial_derivatives(mlv, parameters):
    """Variable_test.py:16-46."""
    p = parameters
    af = mlv.ArrayFactory(p, mlv.hip)
    st = mlv.SpectralTransformer(p, mlv.hip, af)
    sd = mlv.SpatialDifferentiator(p, mlv.hip, af)
    x = np.linspace(0, p.lx, p.nx, endpoint=False)
    z = np.linspace(0, p.lz, p.nz, endpoint=False)
    X, Z = np.meshgrid(x, z, indexing="ij")
    var = mlv.Variable(p, mlv.hip, st=st, sd=sd, array_factory=af,
                       basis_functions=[mlv.BasisFunctions.COMPLEX_EXP] * 2)
    var.setp(np.cos(2 * np.pi * X) * np.cos(2 * np.pi * Z))
    dvardx = _h(var.pddx())
    assert dvardx[:, 0] == approx(-2 * np.pi * np.sin(2 * np.pi * x), rel=1e-3)
    dvardz = _h(var.pddz())
    assert dvardz[0, :] == approx(-2 * np.pi * np.sin(2 * np.pi * z), rel=1e-3)


def test_fdm_nabla2(mlv, fdm_parameters):
    """Variable_test.py:49-73."""
    p = fdm_parameters
    z = np.linspace(0, p.lz, p.nz, endpoint=False)
    af = mlv.ArrayFactory(p, mlv.hip)
    sd = mlv.SpatialDifferentiator(p, mlv.hip, array_factory=af)
    var = mlv.Variable(p, mlv.hip, sd=sd, array_factory=af,
                       basis_functions=[mlv.BasisFunctions.COMPLEX_EXP] * 2)
    var[1] = 0.5 * z ** 2
    nabla2 = _h(var.snabla2())
    true_nabla2 = np.zeros_like(nabla2)
    true_nabla2[1] = -((2 * np.pi) ** 2) * _h(var[1]) + 1
    assert_array_almost_equal(nabla2[1, 1:-1], true_nabla2[1, 1:-1])


def test_spectral_nabla2(mlv, parameters):
    """Variable_test.py:76-100."""
    p = parameters
    af = mlv.ArrayFactory(p, mlv.hip)
    sd = mlv.SpatialDifferentiator(p, mlv.hip, array_factory=af)
    var = mlv.Variable(p, mlv.hip, sd=sd, array_factory=af,
                       basis_functions=[mlv.BasisFunctions.COMPLEX_EXP] * 2)
    var[:] = 1
    nabla2 = _h(var.snabla2())
    n, m = (_h(a) for a in af.make_mode_number_matrices())
    true_nabla2 = -((2 * np.pi / p.lx * n) ** 2) - (2 * np.pi / p.lz * m) ** 2
    assert_array_almost_equal(nabla2, true_nabla2)


def test_fdm_laplacian_solver(mlv, fdm_parameters):
    """LaplacianSolver_test.py:37-59, including its call pattern: the
    Variable itself passed where the basis functions go."""
    p = fdm_parameters
    af = mlv.ArrayFactory(p, mlv.hip)
    psi = mlv.Variable(p, mlv.hip, array_factory=af,
                       basis_functions=[mlv.BasisFunctions.COMPLEX_EXP] * 2)
    laplacian_solver = mlv.LaplacianSolver(p, mlv.hip, psi, array_factory=af)
    true_soln = af.make_spectral()
    rhs = af.make_spectral()
    true_soln[:, :] = 1.0
    for n in range(p.nn):
        rhs[n, :] = laplacian_solver.laps[n] @ true_soln[n, :]
    soln = laplacian_solver.solve(rhs)
    assert_array_almost_equal(_h(soln), _h(true_soln))
