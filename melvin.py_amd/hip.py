"""The ``xp`` array module of the MI355X backend.

Pass ``melvin_amd.hip`` wherever the reference takes ``xp`` (numpy or
cupy).  Arrays are torch tensors on the current GPU; torch only allocates
and moves memory.  The reductions the reference's drivers call through
``xp`` (``max`` for CFL, Integrator.py:38-39; ``sum`` for the kinetic
energy, utility.py:58; ``mean`` for the Nusselt number,
double_diffusive_convection.py:22) run as FP64-accumulating HIP kernels.
"""
from __future__ import annotations

import ctypes

import numpy as np
import torch

from . import _lib


float64 = np.float64
float32 = np.float32
complex128 = np.complex128
complex64 = np.complex64

_TORCH = {
    np.dtype(np.float64): torch.float64,
    np.dtype(np.float32): torch.float32,
    np.dtype(np.complex128): torch.complex128,
    np.dtype(np.complex64): torch.complex64,
    np.dtype(np.int64): torch.int64,
    np.dtype(np.int32): torch.int32,
    np.dtype(np.bool_): torch.bool,
}


def device():
    if not torch.cuda.is_available():
        raise _lib.MelvinHipError("melvin_amd.hip needs a GPU (no CPU fallback)")
    return torch.device("cuda", torch.cuda.current_device())


def torch_dtype(dtype):
    if isinstance(dtype, torch.dtype):
        return dtype
    if dtype is float:
        dtype = np.float64
    elif dtype is complex:
        dtype = np.complex128
    elif dtype is int:
        dtype = np.int64
    return _TORCH[np.dtype(dtype)]


def zeros(shape, dtype=float):
    return torch.zeros(shape, dtype=torch_dtype(dtype), device=device())


def zeros_like(a, dtype=None):
    return torch.zeros_like(a, dtype=None if dtype is None else torch_dtype(dtype))


def ones(shape, dtype=float):
    return torch.ones(shape, dtype=torch_dtype(dtype), device=device())


def arange(*args, dtype=None):
    t = torch.arange(*args, device=device())
    return t if dtype is None else t.to(torch_dtype(dtype))


def concatenate(arrays, axis=0):
    return torch.cat([asarray(a) for a in arrays], dim=axis)


def meshgrid(*arrays, indexing="xy"):
    return list(torch.meshgrid(*[asarray(a) for a in arrays], indexing=indexing))


def asarray(a, dtype=None):
    if isinstance(a, torch.Tensor):
        t = a if a.device.type == "cuda" else a.to(device())
        return t if dtype is None else t.to(torch_dtype(dtype))
    arr = np.asarray(a, dtype=dtype)
    return torch.as_tensor(arr, device=device())


array = asarray


def asnumpy(a):
    if isinstance(a, torch.Tensor):
        return a.detach().cpu().numpy()
    return np.asarray(a)


def _reduce(op, a, b=None, scale=1.0):
    a = asarray(a).contiguous()
    if a.dtype == torch.float64:
        dt = _lib.F64
    elif a.dtype == torch.float32:
        dt = _lib.F32
    else:
        raise TypeError(f"reduction over {a.dtype} is not supported")
    if b is not None:
        b = asarray(b).contiguous()
        if b.dtype != a.dtype or b.shape != a.shape:
            raise ValueError("reduction operands differ in dtype or shape")
    out = torch.empty(1, dtype=torch.float64, device=a.device)
    _lib.call("mlv_reduce", None, op, dt, ctypes.c_void_p(a.data_ptr()),
              ctypes.c_void_p(0 if b is None else b.data_ptr()), a.numel(),
              float(scale), ctypes.c_void_p(out.data_ptr()), _lib.stream_handle())
    return out


def max(a):  # noqa: A001
    """Signed maximum (NaN-propagating), FP64, as a Python float."""
    return float(_reduce(_lib.RED_MAX, a).item())


def sum(a):  # noqa: A001
    """Compensated FP64 sum as a Python float."""
    return float(_reduce(_lib.RED_SUM, a).item())


def mean(a):
    n = asarray(a).numel()
    return float(_reduce(_lib.RED_SUM, a, scale=1.0 / n).item())


def sum_sq2(a, b):
    """sum(a**2 + b**2) in FP64."""
    return float(_reduce(_lib.RED_SUMSQ2, a, b).item())


def dot(a, b):
    """sum(a*b) in FP64."""
    return float(_reduce(_lib.RED_DOT, a, b).item())


def save(fname, a):
    np.save(fname, asnumpy(a))


def savez(fname, **arrays):
    np.savez(fname, **{k: asnumpy(v) for k, v in arrays.items()})


def load(fname):
    return np.load(fname, allow_pickle=False)


class _FFT:
    """``xp.fft`` of this backend: numpy.fft's real transforms on device
    arrays, computed by the library's Stockham kernels (mlv_rfft /
    mlv_irfft) -- the four calls the reference's transformer makes through
    its array module (SpectralTransformer.py:58 irfft, 85 rfft, 132 irfft2,
    191 rfft2).  Arrays are 1-D or 2-D, float32/64 (complex64/128 for the
    inverses), transformed lengths powers of two in [8, 4096]; ``n`` / ``s``
    other than the natural lengths and ``norm`` other than "backward" raise
    NotImplementedError (there is no fallback)."""

    _AXES = {0: _lib.FFT_AXIS0, 1: _lib.FFT_AXIS1}

    @staticmethod
    def _real_dtype(t):
        if t.dtype in (torch.float64, torch.complex128):
            return _lib.F64, torch.float64, torch.complex128
        if t.dtype in (torch.float32, torch.complex64):
            return _lib.F32, torch.float32, torch.complex64
        raise TypeError(f"fft of {t.dtype} is not supported")

    @staticmethod
    def _check(norm, extra):
        if norm not in (None, "backward"):
            raise NotImplementedError("only norm='backward' (numpy's default)")
        if extra is not None:
            raise NotImplementedError(f"unsupported argument {extra!r}")

    def _run(self, name, axes, nx, nz, src, out, work_shape, cdt):
        dt = _lib.F64 if cdt == torch.complex128 else _lib.F32
        work = torch.empty(work_shape, dtype=cdt, device=src.device) if work_shape else None
        _lib.call(name, dt, axes, int(nx), int(nz), ctypes.c_void_p(src.data_ptr()),
                  ctypes.c_void_p(out.data_ptr()),
                  ctypes.c_void_p(0 if work is None else work.data_ptr()), _lib.stream_handle())
        return out

    def _as2d(self, a, axis):
        t = asarray(a)
        if t.ndim == 1:
            if axis not in (0, -1):
                raise ValueError("axis out of range")
            return t.reshape(1, -1).contiguous(), 1, True
        if t.ndim != 2:
            raise NotImplementedError("1-D and 2-D arrays only")
        ax = axis % 2
        return t.contiguous(), ax, False

    def rfft(self, a, n=None, axis=-1, norm=None):
        self._check(norm, None if n is None or n == asarray(a).shape[axis] else n)
        t, ax, flat = self._as2d(a, axis)
        dt, rdt, cdt = self._real_dtype(t)
        t = t.to(rdt)
        nx, nz = t.shape
        shape = (nx // 2 + 1, nz) if ax == 0 else (nx, nz // 2 + 1)
        out = torch.empty(shape, dtype=cdt, device=t.device)
        self._run("mlv_rfft", self._AXES[ax], nx, nz, t, out, None, cdt)
        return out.reshape(-1) if flat else out

    def irfft(self, a, n=None, axis=-1, norm=None):
        t, ax, flat = self._as2d(a, axis)
        m = t.shape[ax]
        self._check(norm, None if n is None or n == 2 * (m - 1) else n)
        dt, rdt, cdt = self._real_dtype(t)
        t = t.to(cdt)
        nx, nz = (2 * (m - 1), t.shape[1]) if ax == 0 else (t.shape[0], 2 * (m - 1))
        out = torch.empty((nx, nz), dtype=rdt, device=t.device)
        self._run("mlv_irfft", self._AXES[ax], nx, nz, t, out, None, cdt)
        return out.reshape(-1) if flat else out

    def rfft2(self, a, s=None, axes=(-2, -1), norm=None):
        t = asarray(a)
        self._check(norm, None if s is None or tuple(s) == tuple(t.shape[-2:]) else s)
        if t.ndim != 2 or tuple(ax % 2 for ax in axes) != (0, 1):
            raise NotImplementedError("2-D arrays over axes (-2, -1) only")
        dt, rdt, cdt = self._real_dtype(t)
        t = t.to(rdt).contiguous()
        nx, nz = t.shape
        out = torch.empty((nx, nz // 2 + 1), dtype=cdt, device=t.device)
        return self._run("mlv_rfft", _lib.FFT_AXES01, nx, nz, t, out, (nx, nz // 2 + 1), cdt)

    def irfft2(self, a, s=None, axes=(-2, -1), norm=None):
        t = asarray(a)
        if t.ndim != 2 or tuple(ax % 2 for ax in axes) != (0, 1):
            raise NotImplementedError("2-D arrays over axes (-2, -1) only")
        nx, nz = t.shape[0], 2 * (t.shape[1] - 1)
        self._check(norm, None if s is None or tuple(s) == (nx, nz) else s)
        dt, rdt, cdt = self._real_dtype(t)
        t = t.to(cdt).contiguous()
        out = torch.empty((nx, nz), dtype=rdt, device=t.device)
        return self._run("mlv_irfft", _lib.FFT_AXES01, nx, nz, t, out, tuple(t.shape), cdt)


fft = _FFT()
