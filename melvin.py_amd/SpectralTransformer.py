"""Spectral <-> physical transforms on the GPU (mirror of
melvin/SpectralTransformer.py).

Fully spectral: the compact spectrum (2nn+1, nm) is zero-padded, scaled by
nx*nz and inverse transformed (SpectralTransformer.py:90-150), and the
forward direction divides the rfft2 by nx*nz and truncates (152-199).  The
padding and truncation are never materialised: libmelvin_hip runs one x
pass and one z pass of its own Stockham FFT.  Fourier-x / FD-z uses the
x-only transform pair (33-88).  Cosine and sine bases (108-196) -- the
reference's transform of the signal mirrored about its end point -- run as
dense per-axis transforms on the GPU (``mlv_to_spectral_basis``).
"""
from __future__ import annotations

from . import _lib, hip
from ._arrays import as_input, finish, output
from ._plans import plan_for, ptr
from .ArrayFactory import ArrayFactory
from .BasisFunctions import BasisFunctions

_CE = BasisFunctions.COMPLEX_EXP
_FDM = BasisFunctions.FDM


_SPECTRAL = (BasisFunctions.COMPLEX_EXP, BasisFunctions.COSINE, BasisFunctions.SINE)


def check_basis(params, basis_functions):
    b0, b1 = basis_functions
    if params.is_fully_spectral():
        if b0 not in _SPECTRAL or b1 not in _SPECTRAL:
            raise Exception("fully spectral transforms need spectral bases")
    elif b1 is not _FDM:
        raise Exception("One basis function must be FDM")
    elif b0 is not _CE:
        raise NotImplementedError("the Fourier-x / FD-z transform is along x only")


class SpectralTransformer:
    def __init__(self, params, xp=hip, array_factory=None):
        self._p = params
        self._xp = hip
        self._array_factory = array_factory or ArrayFactory(params)
        self._plan = plan_for(params)

    def to_physical(self, in_arr, out=None,
                    basis_functions=(BasisFunctions.COMPLEX_EXP,
                                     BasisFunctions.COMPLEX_EXP)):
        if not self._p.is_fully_spectral() and tuple(basis_functions) == (_CE, _CE):
            basis_functions = (_CE, _FDM)
        check_basis(self._p, basis_functions)
        pl = self._plan
        src = as_input(in_arr, pl.cplx, (pl.rows, pl.cols))
        dst, ret = output(out, pl.real, (pl.nx, pl.nz))
        b0, b1 = (int(b) for b in basis_functions)
        if self._p.is_fully_spectral() and (b0, b1) != (0, 0):
            _lib.call("mlv_to_physical_basis", pl.handle, b0, b1, ptr(src), ptr(dst),
                      _lib.stream_handle())
            # The reference doubles the k = 0 cosine term of its *input* in
            # place before transforming (SpectralTransformer.py:117-120); the
            # transform here folds the doubling into its matrices and the
            # caller's array sees the same side effect.
            if basis_functions[0] is BasisFunctions.COSINE:
                in_arr[0] *= 2
            if basis_functions[1] is BasisFunctions.COSINE:
                in_arr[:, 0] *= 2
        else:
            _lib.call("mlv_to_physical", pl.handle, ptr(src), ptr(dst), _lib.stream_handle())
        return finish(dst, ret)

    def to_spectral(self, in_arr, out=None,
                    basis_functions=(BasisFunctions.COMPLEX_EXP,
                                     BasisFunctions.COMPLEX_EXP)):
        if not self._p.is_fully_spectral() and tuple(basis_functions) == (_CE, _CE):
            basis_functions = (_CE, _FDM)
        check_basis(self._p, basis_functions)
        pl = self._plan
        src = as_input(in_arr, pl.real, (pl.nx, pl.nz))
        dst, ret = output(out, pl.cplx, (pl.rows, pl.cols))
        b0, b1 = (int(b) for b in basis_functions)
        if self._p.is_fully_spectral() and (b0, b1) != (0, 0):
            _lib.call("mlv_to_spectral_basis", pl.handle, b0, b1, ptr(src), ptr(dst),
                      _lib.stream_handle())
        else:
            _lib.call("mlv_to_spectral", pl.handle, ptr(src), ptr(dst), _lib.stream_handle())
        return finish(dst, ret)
