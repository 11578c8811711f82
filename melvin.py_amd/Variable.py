"""A field held in spectral and physical form (mirror of
melvin/Variable.py:6-151).  ``vec_dot_nabla`` is the reference's advection
operator: physical-space FD flux divergence of u*f, then a forward
transform (Variable.py:119-128)."""
from __future__ import annotations

import numpy as np

from . import _lib, hip
from ._arrays import as_input
from ._plans import plan_for, ptr


class Variable:
    def __init__(self, params, xp=hip, sd=None, st=None, dt=None,
                 array_factory=None, dump_name=None, basis_functions=None):
        self._params = params
        self._xp = hip
        self._st = st
        self._sd = sd
        self._dt = dt
        self._array_factory = array_factory
        self._dump_name = dump_name
        if basis_functions is None:
            raise Exception("Basis functions must be specified.")
        self._basis_functions = basis_functions
        self._sdata = array_factory.make_spectral()
        self._pdata = array_factory.make_physical()

    def __setitem__(self, index, value):
        if not isinstance(value, (int, float, complex)):
            value = hip.asarray(value)
        self._sdata[index] = value

    def __getitem__(self, index):
        return self._sdata[index]

    def sets(self, data):
        self._sdata[:, :] = hip.asarray(data)[:, :]

    def setp(self, data):
        self._pdata[:, :] = hip.asarray(data)[:, :]

    def getp(self):
        return self._pdata

    def gets(self):
        return self._sdata

    def to_physical(self):
        self._st.to_physical(self._sdata, self._pdata, self._basis_functions)

    def to_spectral(self):
        self._st.to_spectral(self._pdata, self._sdata, self._basis_functions)

    def load(self, data, is_physical=False):
        if isinstance(data, str):
            data = np.load(data, allow_pickle=False)
        data = hip.asarray(data)
        if is_physical:
            self.setp(data)
            self.to_spectral()
        else:
            if tuple(data.shape) != tuple(self._sdata.shape):
                data = regrid(self._params, data)
            self.sets(data)

    def pddx(self):
        return self._sd.pddx(self.getp())

    def pddz(self):
        return self._sd.pddz(self.getp())

    def sddx(self):
        return self._sd.sddx(self.gets(), self._basis_functions[0])

    def sddz(self):
        return self._sd.sddz(self.gets(), self._basis_functions[1])

    def sd2dx2(self):
        return self._sd.sd2dx2(self.gets(), self._basis_functions[0])

    def sd2dz2(self):
        return self._sd.sd2dz2(self.gets(), self._basis_functions[1])

    def snabla2(self):
        """Spectral Laplacian, one pass (Variable.py:111-113)."""
        return self._sd.snabla2(self.gets(), self._basis_functions)

    def lap(self):
        return self._sd.calc_lap(self._basis_functions)

    def vec_dot_nabla(self, ux, uz, out=None, convert_to_physical=True):
        pl = plan_for(self._params)
        if out is None:
            out = hip.zeros_like(self._pdata)
        if convert_to_physical:
            self.to_physical()
        a = as_input(ux, pl.real, (pl.nx, pl.nz))
        b = as_input(uz, pl.real, (pl.nx, pl.nz))
        f = as_input(self._pdata, pl.real, (pl.nx, pl.nz))
        dst = out if (out.is_contiguous() and out.dtype == pl.real) else hip.zeros_like(f)
        _lib.call("mlv_flux_divergence", pl.handle, ptr(a), ptr(b), ptr(f), ptr(dst),
                  _lib.stream_handle())
        if dst is not out:
            out.copy_(dst)
        return self._st.to_spectral(out, basis_functions=self._basis_functions)

    def save(self, dump_counter):
        np.save(self._dump_name + f"{dump_counter:04d}.npy", hip.asnumpy(self._pdata))

    def on_host(self):
        return hip.asnumpy(self.gets())

    def get_name(self):
        return self._dump_name


def scale_variable(var, outsize, xp=hip):
    """Fully spectral coefficients ``(2 nn' + 1, nm')`` on a grid of
    ``outsize = (nn, nm)``: shape ``(2 nn + 1, nm)``, complex, on the device
    (Variable.py:140-151).  The coefficients are amplitudes, so each mode
    present on both grids keeps its value -- rows ``|kx| <= min(nn', nn)``,
    each to the row of the same kx, columns ``kz < min(nm', nm)`` -- and the
    other modes are zero.  Unlike the reference, whose output array is real
    (it keeps only real parts) and whose positive-row bound counts the
    input's rows of both signs of kx (mapping negative-kx rows onto positive
    ones when nn' < nn), the values stay complex and every kx maps to
    itself."""
    src = hip.asarray(var)
    if src.ndim != 2 or src.shape[0] % 2 != 1:
        raise ValueError(f"fully spectral data has shape (2nn+1, nm), got {tuple(src.shape)}")
    nn, nm = int(outsize[0]), int(outsize[1])
    dtype = src.dtype if src.is_complex() else hip.complex128
    out = hip.zeros((2 * nn + 1, nm), dtype=dtype)
    k = min((src.shape[0] - 1) // 2, nn)
    m = min(src.shape[1], nm)
    out[: k + 1, :m] = src[: k + 1, :m]
    if k > 0:
        out[-k:, :m] = src[-k:, :m]
    return out


def regrid(params, data):
    """Spectral data of another grid size on ``params``' grid, in its
    layout (the load-time re-grid of Variable.py:76-85 and
    TimeDerivative.py:56-62): fully spectral by ``scale_variable``;
    Fourier-x / FD-z ``(nn', nz)`` keeps rows ``kx < min(nn', nn)`` and
    FD-x / Fourier-z ``(nx, nm')`` columns ``kz < min(nm', nm)``.  The
    finite-difference axis must already match (another FD grid needs
    interpolation, not a re-grid)."""
    src = hip.asarray(data)
    shape = tuple(params.spectral_shape)
    if params.is_fully_spectral():
        return scale_variable(src, (params.nn, params.nm)).to(hip.torch_dtype(params.complex))
    if src.ndim != 2:
        raise ValueError(f"spectral data must be 2-D, got {tuple(src.shape)}")
    out = hip.zeros(shape, dtype=params.complex)
    if params.discretisation[1] == "fdm":
        if src.shape[1] != shape[1]:
            raise ValueError(f"FD z grid has nz = {src.shape[1]}, expected {shape[1]}")
        k = min(src.shape[0], shape[0])
        out[:k] = src[:k]
    else:
        if src.shape[0] != shape[0]:
            raise ValueError(f"FD x grid has nx = {src.shape[0]}, expected {shape[0]}")
        m = min(src.shape[1], shape[1])
        out[:, :m] = src[:, :m]
    return out
