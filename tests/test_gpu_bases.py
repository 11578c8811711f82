"""Cosine and sine bases on the GPU: the reference's known-answer tests
through the mirror SpectralTransformer, every basis pair of the transforms
and spectral operators against the oracle, the Laplacian solver, the
reference's in-place doubling of cosine k = 0 terms and the single-precision
plan."""
import numpy as np
import pytest
import torch

import _basis_kats
from conftest import rel
from oracle import melvin_oracle as O

pytestmark = pytest.mark.gpu

PAIRS = [(bx, bz) for bx in (0, 1, 2) for bz in (0, 1, 2) if (bx, bz) != (0, 0)]


def _h(x):
    return x.detach().cpu().numpy()


def _params(mlv, nx, nz, lx=1.0, lz=1.0, dtype=None):
    d = dict(nx=nx, nz=nz, lx=lx, lz=lz, final_time=1.0)
    if dtype:
        d["precision"] = dtype
    return mlv.Parameters(d, validate=False), O.make_params(dict(nx=nx, nz=nz, lx=lx, lz=lz))


def _bases(mlv, pair):
    return [mlv.BasisFunctions(b) for b in pair]


@pytest.mark.parametrize("name", sorted(_basis_kats.KATS))
def test_kat_transforms(mlv, name):
    """SpectralTransformer_test.py:64-310 through the mirror class, with the
    reference's sequence: to_spectral, then to_physical of the same array."""
    params, _ = _params(mlv, 256, 256)
    st = mlv.SpectralTransformer(params, mlv.hip)
    af = mlv.ArrayFactory(params, mlv.hip)
    spectral, physical = af.make_spectral(), af.make_physical()
    bases, phys, want = _basis_kats.case(name)
    bf = _bases(mlv, bases)
    print(f"[bases] kat {name}", flush=True)
    st.to_spectral(torch.as_tensor(phys, device=spectral.device), spectral, basis_functions=bf)
    np.testing.assert_array_almost_equal(_h(spectral), want)
    st.to_physical(spectral, physical, basis_functions=bf)
    np.testing.assert_array_almost_equal(_h(physical), phys)


@pytest.mark.parametrize("pair", PAIRS)
@pytest.mark.parametrize("nx,nz", [(64, 128), (256, 32)])
def test_transforms_vs_oracle(mlv, pair, nx, nz):
    """Random physical fields forward and random spectra inverse, FP64,
    against the oracle's mirror-and-FFT restatement."""
    params, p = _params(mlv, nx, nz, lx=2.0, lz=0.7)
    st = mlv.SpectralTransformer(params, mlv.hip)
    rng = np.random.default_rng(nx + 7 * nz + 3 * pair[0] + pair[1])
    phys = rng.standard_normal((nx, nz))
    spec = rng.standard_normal(p.spectral_shape) + 1j * rng.standard_normal(p.spectral_shape)
    bf = _bases(mlv, pair)
    print(f"[bases] transforms {pair} {nx}x{nz}", flush=True)
    got = _h(st.to_spectral(torch.as_tensor(phys, device="cuda"), basis_functions=bf))
    assert rel(got, O.to_spectral_basis(p, phys, pair)) < 1e-12
    got = _h(st.to_physical(torch.as_tensor(spec, device="cuda"), basis_functions=bf))
    assert rel(got, O.to_physical_basis(p, spec, pair)) < 1e-12


def test_to_physical_doubles_cosine_k0_in_place(mlv):
    """SpectralTransformer.py:117-120: the caller's spectral array has its
    kx = 0 row (cosine x) and kz = 0 column (cosine z) doubled."""
    params, p = _params(mlv, 64, 64)
    st = mlv.SpectralTransformer(params, mlv.hip)
    rng = np.random.default_rng(5)
    spec = rng.standard_normal(p.spectral_shape) + 1j * rng.standard_normal(p.spectral_shape)
    for pair in [(1, 1), (1, 2), (2, 1), (2, 2)]:
        t = torch.as_tensor(spec, device="cuda").clone()
        phys = st.to_physical(t, basis_functions=_bases(mlv, pair))
        want = spec.copy()
        if pair[0] == 1:
            want[0] *= 2
        if pair[1] == 1:
            want[:, 0] *= 2
        np.testing.assert_array_equal(_h(t), want)
        assert rel(_h(phys), O.to_physical_basis(p, spec, pair)) < 1e-12


@pytest.mark.parametrize("pair", PAIRS)
def test_spectral_ops_vs_oracle(mlv, pair):
    """sddx / sddz / sd2dx2 / sd2dz2 / snabla2 / calc_lap with bases and the
    fully spectral Laplacian solve, against the oracle."""
    params, p = _params(mlv, 64, 32, lx=1.5, lz=0.8)
    sd = mlv.SpatialDifferentiator(params, mlv.hip)
    rng = np.random.default_rng(11 + 3 * pair[0] + pair[1])
    f = rng.standard_normal(p.spectral_shape) + 1j * rng.standard_normal(p.spectral_shape)
    ft = torch.as_tensor(f, device="cuda")
    bx, bz = _bases(mlv, pair)
    print(f"[bases] ops {pair}", flush=True)
    got = {"ddx": sd.sddx(ft, bx), "ddz": sd.sddz(ft, bz), "d2dx2": sd.sd2dx2(ft, bx),
           "d2dz2": sd.sd2dz2(ft, bz), "snabla2": sd.snabla2(ft, [bx, bz])}
    for kind, g in got.items():
        assert rel(_h(g), O.spectral_op_basis(p, kind, f, pair)) < 1e-14, kind
    lap = _h(sd.calc_lap([bx, bz]))
    np.testing.assert_allclose(lap, O.spectral_op_basis(p, "lap", None, pair), rtol=1e-15)
    ls = mlv.LaplacianSolver(params, mlv.hip, [bx, bz], spatial_diff=sd)
    psi = _h(ls.solve(ft))
    assert rel(psi, O.spectral_op_basis(p, "poisson", f, pair)) < 1e-14


def test_single_precision_plan(mlv):
    """FP32 plan: the dense transforms and operators within 1e-5 of the
    FP64 oracle (relative L2)."""
    params, p = _params(mlv, 128, 64, dtype="single")
    st = mlv.SpectralTransformer(params, mlv.hip)
    rng = np.random.default_rng(2)
    phys = rng.standard_normal((128, 64))
    for pair in [(1, 2), (2, 1), (1, 0)]:
        bf = _bases(mlv, pair)
        got = _h(st.to_spectral(torch.as_tensor(phys, dtype=torch.float32, device="cuda"),
                                basis_functions=bf))
        assert rel(got, O.to_spectral_basis(p, phys.astype(np.float32), pair)) < 1e-5
        spec = O.to_spectral_basis(p, phys, pair)
        got = _h(st.to_physical(torch.as_tensor(spec, dtype=torch.complex64, device="cuda"),
                                basis_functions=bf))
        assert rel(got, O.to_physical_basis(p, spec.astype(np.complex64), pair)) < 1e-5


def test_basis_errors(mlv):
    """Cosine / sine need a fully spectral plan; unknown bases are argument
    errors; the mirror rejects FD-z bases other than (COMPLEX_EXP, FDM)."""
    from ctypes import c_void_p

    L = mlv._lib
    lib = L.lib()
    params, _ = _params(mlv, 32, 32)
    pl = mlv._plans.plan_for(params)
    buf = torch.zeros(4096, dtype=torch.complex128, device="cuda")
    ptr = c_void_p(buf.data_ptr())
    assert lib.mlv_to_spectral_basis(pl.handle, 3, 0, ptr, ptr, None) == L.MLV_ERR_ARG
    assert lib.mlv_spec_op_basis(pl.handle, 9, 1, 1, ptr, ptr, None) == L.MLV_ERR_ARG
    fdz = mlv.Parameters(dict(nx=32, nz=33, lx=1.0, lz=1.0, final_time=1.0,
                              discretisation=["spectral", "fdm"]), validate=False)
    pf = mlv._plans.plan_for(fdz)
    assert lib.mlv_to_physical_basis(pf.handle, 1, 0, ptr, ptr, None) == L.MLV_ERR_UNSUPPORTED
    assert b"fully spectral" in lib.mlv_last_error()
    st = mlv.SpectralTransformer(fdz, mlv.hip)
    with pytest.raises(NotImplementedError):
        st.to_spectral(torch.zeros((32, 33), dtype=torch.float64, device="cuda"),
                       basis_functions=[mlv.BasisFunctions.COSINE, mlv.BasisFunctions.FDM])
    torch.cuda.synchronize()
