"""Spectral and finite-difference derivatives on the GPU (mirror of
melvin/SpatialDifferentiator.py).

Spectral multiplies by i*k and -k^2 (50-74); physical central differences
of order 2 or 4 with the periodic wrap only in spectral directions
(76-185).  In Fourier-x / FD-z, sddz and sd2dz2 are the FD z-derivatives
of the (complex) spectral rows (35-40).
"""
from __future__ import annotations

import torch

from . import _lib, hip
from ._arrays import as_input, finish, output
from ._plans import plan_for, ptr
from .ArrayFactory import ArrayFactory
from .BasisFunctions import BasisFunctions, gen_diff2_factors, gen_diff_factors


class SpatialDifferentiator:
    def __init__(self, params, xp=hip, array_factory=None):
        self._params = params
        self._xp = hip
        if params.discretisation[0] == "fdm":
            raise NotImplementedError("Finite difference not implemented in x direction")
        if params.spatial_derivative_order not in (2, 4):
            raise ValueError("spatial_derivative_order must be 2 or 4")
        self._x_periodic = params.discretisation[0] == "spectral"
        self._z_periodic = params.discretisation[1] == "spectral"
        af = array_factory or ArrayFactory(params)
        self._n, self._m = af.make_mode_number_matrices()
        self._plan = plan_for(params)
        self._diffx_factors = gen_diff_factors(params.lx)
        self._diffz_factors = gen_diff_factors(params.lz)
        self._diff2x_factors = gen_diff2_factors(params.lx)
        self._diff2z_factors = gen_diff2_factors(params.lz)
        if params.discretisation[1] == "fdm":
            self.sddz = lambda var, bs=None: self.pddz(var)
            self.sd2dz2 = lambda var, bs=None: self.pd2dz2(var)

    # -- spectral ---------------------------------------------------------
    def _spec(self, kind, var, bx, bz=BasisFunctions.COMPLEX_EXP):
        """One spectral multiply; bx / bz are the bases whose factors the
        operator uses (BasisFunctions.py:26-59)."""
        pl = self._plan
        src = as_input(var, pl.cplx, (pl.rows, pl.cols))
        out = torch.empty_like(src)
        uses_x = kind in (_lib.SPEC_DDX, _lib.SPEC_D2DX2)
        uses_z = kind in (_lib.SPEC_DDZ, _lib.SPEC_D2DZ2)
        if (uses_x and bx is BasisFunctions.FDM) or (uses_z and bz is BasisFunctions.FDM):
            out.zero_()  # the FDM basis has a zero derivative factor
            return out
        if int(bx) in (1, 2) or int(bz) in (1, 2):
            _lib.call("mlv_spec_op_basis", pl.handle, kind, int(bx), int(bz), ptr(src),
                      ptr(out), _lib.stream_handle())
        else:
            _lib.call("mlv_spec_op", pl.handle, kind, ptr(src), ptr(out),
                      _lib.stream_handle())
        return out

    def sddx(self, var, basis_fn=BasisFunctions.COMPLEX_EXP):
        return self._spec(_lib.SPEC_DDX, var, basis_fn)

    def sddz(self, var, basis_fn=BasisFunctions.COMPLEX_EXP):
        return self._spec(_lib.SPEC_DDZ, var, BasisFunctions.COMPLEX_EXP, basis_fn)

    def sd2dx2(self, var, basis_fn=BasisFunctions.COMPLEX_EXP):
        return self._spec(_lib.SPEC_D2DX2, var, basis_fn)

    def sd2dz2(self, var, basis_fn=BasisFunctions.COMPLEX_EXP):
        return self._spec(_lib.SPEC_D2DZ2, var, BasisFunctions.COMPLEX_EXP, basis_fn)

    def snabla2(self, var, basis_fns=(BasisFunctions.COMPLEX_EXP,
                                      BasisFunctions.COMPLEX_EXP)):
        """sd2dx2 + sd2dz2 in one pass (Variable.py:111-113)."""
        return self._spec(_lib.SPEC_SNABLA2, var, basis_fns[0], basis_fns[1])

    def calc_lap(self, basis_fns=(BasisFunctions.COMPLEX_EXP,
                                  BasisFunctions.COMPLEX_EXP)):
        """float64 symbol -(kx^2) - (kz^2) (SpatialDifferentiator.py:70-74)."""
        pl = self._plan
        out = torch.empty((pl.rows, pl.cols), dtype=torch.float64, device=hip.device())
        b0, b1 = (int(b) for b in basis_fns)
        if b0 in (1, 2) or b1 in (1, 2):
            _lib.call("mlv_spec_op_basis", pl.handle, _lib.SPEC_LAP, b0, b1, ctypes_null(),
                      ptr(out), _lib.stream_handle())
        else:
            _lib.call("mlv_spec_op", pl.handle, _lib.SPEC_LAP, ctypes_null(), ptr(out),
                      _lib.stream_handle())
        return out

    # -- physical finite differences -------------------------------------
    def _fd(self, kind, var, out):
        pl = self._plan
        t = hip.asarray(var)
        dtype = pl.cplx if t.is_complex() else pl.real
        src = as_input(t, dtype)
        rows, cols = src.shape
        dst, ret = output(out, dtype, (rows, cols))
        _lib.call("mlv_fd_op", pl.handle, kind, 1 if src.is_complex() else 0, rows,
                  cols, ptr(src), ptr(dst), _lib.stream_handle())
        return finish(dst, ret)

    def pddx(self, var, out=None):
        return self._fd(_lib.FD_DDX, var, out)

    def pddz(self, var, out=None):
        return self._fd(_lib.FD_DDZ, var, out)

    def pd2dz2(self, var, out=None):
        return self._fd(_lib.FD_D2DZ2, var, out)


def ctypes_null():
    import ctypes

    return ctypes.c_void_p(0)
