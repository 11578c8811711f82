"""Cache of libmelvin_hip plans, one per (grid, discretisation, precision,
FD order, device)."""
from __future__ import annotations

import ctypes

import numpy as np

from . import _lib


class Plan:
    """Owns one ``mlv_plan``."""

    def __init__(self, nx, nz, lx, lz, fully_spectral, double, fd_order):
        import torch

        self.device = torch.cuda.current_device()
        self.nx, self.nz = int(nx), int(nz)
        self.fully = bool(fully_spectral)
        self.dtype = _lib.F64 if double else _lib.F32
        self.real = torch.float64 if double else torch.float32
        self.cplx = torch.complex128 if double else torch.complex64
        h = ctypes.c_void_p()
        _lib.call("mlv_plan_create", ctypes.byref(h), self.nx, self.nz,
                  float(lx), float(lz),
                  _lib.FULLY_SPECTRAL if self.fully else _lib.FOURIER_X_FD_Z,
                  self.dtype, int(fd_order))
        self.handle = h
        vals = [ctypes.c_int() for _ in range(4)]
        _lib.call("mlv_plan_info", h, *[ctypes.byref(v) for v in vals])
        self.nn, self.nm, self.rows, self.cols = (v.value for v in vals)

    def __del__(self):
        h = getattr(self, "handle", None)
        if h is not None and h.value:
            try:
                _lib.lib().mlv_plan_destroy(h)
            except Exception:
                pass


_CACHE: dict = {}


def plan_for(params, fd_order=None):
    import torch

    fully = params.is_fully_spectral()
    if not fully and params.discretisation[0] != "spectral":
        raise NotImplementedError("Finite difference not implemented in x direction")
    double = np.dtype(params.float) == np.float64
    order = params.spatial_derivative_order if fd_order is None else fd_order
    key = (params.nx, params.nz, float(params.lx), float(params.lz), fully,
           double, int(order), torch.cuda.current_device())
    p = _CACHE.get(key)
    if p is None:
        p = Plan(params.nx, params.nz, params.lx, params.lz, fully, double,
                 order)
        _CACHE[key] = p
    return p


def ptr(t):
    """Device pointer of a contiguous tensor (raises otherwise)."""
    if not t.is_cuda:
        raise TypeError("expected a GPU tensor")
    if not t.is_contiguous():
        raise ValueError("expected a contiguous tensor")
    return ctypes.c_void_p(t.data_ptr())
