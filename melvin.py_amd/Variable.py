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
                raise ValueError(f"spectral data has shape {tuple(data.shape)}, "
                                 f"expected {tuple(self._sdata.shape)}")
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
