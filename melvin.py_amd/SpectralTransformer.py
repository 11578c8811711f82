"""Spectral <-> physical transforms on the GPU (mirror of
melvin/SpectralTransformer.py).

Fully spectral: the compact spectrum (2nn+1, nm) is zero-padded, scaled by
nx*nz and inverse transformed (SpectralTransformer.py:90-150), and the
forward direction divides the rfft2 by nx*nz and truncates (152-199).  The
padding and truncation are never materialised: libmelvin_hip runs one x
pass and one z pass of its own Stockham FFT.  Fourier-x / FD-z uses the
x-only transform pair (33-88).  Cosine and sine bases are not implemented.
"""
from __future__ import annotations

from . import _lib, hip
from ._arrays import as_input, finish, output
from ._plans import plan_for, ptr
from .ArrayFactory import ArrayFactory
from .BasisFunctions import BasisFunctions

_CE = BasisFunctions.COMPLEX_EXP
_FDM = BasisFunctions.FDM


def check_basis(params, basis_functions):
    b0, b1 = basis_functions
    if b0 in (BasisFunctions.COSINE, BasisFunctions.SINE) or b1 in (
            BasisFunctions.COSINE, BasisFunctions.SINE):
        raise NotImplementedError("cosine/sine bases are not implemented")
    if params.is_fully_spectral():
        if b0 is not _CE or b1 is not _CE:
            raise Exception("fully spectral transforms need COMPLEX_EXP bases")
    elif b1 is not _FDM:
        raise Exception("One basis function must be FDM")


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
        _lib.call("mlv_to_spectral", pl.handle, ptr(src), ptr(dst), _lib.stream_handle())
        return finish(dst, ret)
