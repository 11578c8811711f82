"""Array allocation for the MI355X backend (mirror of melvin/ArrayFactory.py).

Spectral arrays are (2nn+1, nm) complex (fully spectral) or (nn, nz)
(Fourier-x / FD-z); physical arrays are (nx, nz) real; z is contiguous
(ArrayFactory.py:8-62).
"""
from __future__ import annotations

from . import hip


class ArrayFactory:
    def __init__(self, params, xp=hip):
        self._p = params
        self._xp = hip

    def make_mode_number_matrices(self):
        """Integer mode numbers n, m on the spectral grid, 'ij' indexing
        (ArrayFactory.py:8-26)."""
        p = self._p
        xp = self._xp
        if p.is_fully_spectral():
            n = xp.concatenate((xp.arange(0, p.nn + 1), xp.arange(-p.nn, 0)))
            m = xp.arange(0, p.nm)
        elif p.discretisation[0] == "fdm":
            n = xp.arange(0, p.nx)
            m = xp.arange(0, p.nm)
        else:
            n = xp.arange(0, p.nn)
            m = xp.arange(0, p.nz)
        return xp.meshgrid(n, m, indexing="ij")

    def make_spectral(self, ni=None, nj=None):
        ni = self._p.spectral_shape[0] if ni is None else ni
        nj = self._p.spectral_shape[1] if nj is None else nj
        return self._xp.zeros((ni, nj), dtype=self._p.complex)

    def make_physical(self, nx=None, nz=None):
        nx = self._p.nx if nx is None else nx
        nz = self._p.nz if nz is None else nz
        return self._xp.zeros((nx, nz), dtype=self._p.float)
