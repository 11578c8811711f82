"""Pins of the oracle's spectral re-grid on load (regrid_spectral, the
restatement of Variable.py:76-85,140-151 and TimeDerivative.py:56-62):
identity at the same shape, exact spectral interpolation onto a grid twice
as fine (samples at the coarse points agree), down-grid of an up-grid is
the identity, a closed-form field carried onto a finer grid, truncation of
modes the coarser grid cannot hold, and the Fourier-x / FD-z layout.  A
re-grid that maps a kx row across the sign of kx, drops imaginary parts or
rescales coefficients fails the interpolation and closed-form pins."""
import numpy as np
import pytest

from oracle import melvin_oracle as O


def _p(n, m=None, lx=1.0, lz=2.0, fdm=False):
    d = dict(nx=n, nz=m or n, lx=lx, lz=lz)
    if fdm:
        d["discretisation"] = ["spectral", "fdm"]
    return O.make_params(d)


def _random_field(p, seed):
    rng = np.random.default_rng(seed)
    return rng.standard_normal(p.physical_shape)


def test_same_shape_is_identity():
    p = _p(48)
    spec = O.to_spectral(p, _random_field(p, 0))
    spec += 1j * np.random.default_rng(1).standard_normal(spec.shape) * 1e-3
    np.testing.assert_array_equal(O.regrid_spectral(p, spec), spec)


def test_upgrid_interpolates_exactly():
    """A band-limited field carried onto a grid twice as fine: the fine
    field's samples at the coarse grid points are the coarse field."""
    pc, pf = _p(48), _p(96)
    phys = _random_field(pc, 2)
    spec_c = O.to_spectral(pc, phys)
    fine = O.to_physical(pf, O.regrid_spectral(pf, spec_c))
    np.testing.assert_allclose(fine[::2, ::2], O.to_physical(pc, spec_c), atol=1e-13)
    # 48 -> 64 points (nn 15 -> 21, fewer than the 31 rows of the input):
    # the grids share every 3rd / 4th point
    pm = _p(64)
    mid = O.to_physical(pm, O.regrid_spectral(pm, spec_c))
    np.testing.assert_allclose(mid[::4, ::4], O.to_physical(pc, spec_c)[::3, ::3], atol=1e-13)


def test_downgrid_of_upgrid_is_identity():
    pc, pf = _p(48, 36), _p(96, 80)
    spec_c = O.to_spectral(pc, _random_field(pc, 3))
    back = O.regrid_spectral(pc, O.regrid_spectral(pf, spec_c))
    np.testing.assert_array_equal(back, spec_c)


def test_closed_form_onto_finer_grid():
    pc, pf = _p(48), _p(128)

    def field(p):
        X, Z = O.grid(p)
        return (np.cos(2 * np.pi * 3 * X / p.lx) * np.cos(2 * np.pi * 2 * Z / p.lz)
                + np.sin(2 * np.pi * 5 * X / p.lx - 2 * np.pi * 4 * Z / p.lz))

    got = O.to_physical(pf, O.regrid_spectral(pf, O.to_spectral(pc, field(pc))))
    np.testing.assert_allclose(got, field(pf), atol=1e-13)


def test_downgrid_truncates_modes_beyond_the_coarse_grid():
    pf, pc = _p(96), _p(48)   # nn 31 -> 15, nm 31 -> 15
    X, Z = O.grid(pf)
    low = np.cos(2 * np.pi * 4 * X / pf.lx) * np.sin(2 * np.pi * 3 * Z / pf.lz)
    high = np.cos(2 * np.pi * 20 * X / pf.lx) + np.sin(2 * np.pi * 25 * Z / pf.lz)
    got = O.regrid_spectral(pc, O.to_spectral(pf, low + high))
    Xc, Zc = O.grid(pc)
    want = O.to_spectral(pc, np.cos(2 * np.pi * 4 * Xc / pc.lx) * np.sin(2 * np.pi * 3 * Zc / pc.lz))
    np.testing.assert_allclose(got, want, atol=1e-15)


def test_fdz_layout_regrids_along_x_only():
    pc, pf = _p(48, 33, fdm=True), _p(96, 33, fdm=True)
    spec_c = O.to_spectral(pc, _random_field(pc, 4))
    fine = O.to_physical(pf, O.regrid_spectral(pf, spec_c))
    np.testing.assert_allclose(fine[::2], O.to_physical(pc, spec_c), atol=1e-13)
    np.testing.assert_array_equal(O.regrid_spectral(pc, O.regrid_spectral(pf, spec_c)), spec_c)
    with pytest.raises(ValueError):
        O.regrid_spectral(_p(48, 65, fdm=True), spec_c)
