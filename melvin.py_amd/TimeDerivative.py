"""Ring of right-hand sides for multistep integration (mirror of
melvin/TimeDerivative.py:4-65): zero-initialised, ``get(i)`` indexes
relative to the current slot with Python's negative wrap, ``advance``
moves the current slot."""
from __future__ import annotations

from . import hip
from .Variable import regrid


class TimeDerivative:
    def __init__(self, params, xp=hip, dump_name="", array_factory=None):
        self._params = params
        self._xp = hip
        self._curr_idx = 0
        self._data = hip.zeros((params.integrator_order, params.spectral_shape[0],
                                params.spectral_shape[1]), dtype=params.complex)
        self._dump_name = dump_name

    def __setitem__(self, index, value):
        if not isinstance(value, (int, float, complex)):
            value = hip.asarray(value)
        self._data[self._curr_idx][index] = value

    def __getitem__(self, index):
        return self._data[self._curr_idx][index]

    def get_curr_idx(self):
        return self._curr_idx

    def set_curr_idx(self, curr_idx):
        self._curr_idx = int(curr_idx)

    def advance(self):
        self._curr_idx = (self._curr_idx + 1) % self._params.integrator_order

    def get(self, idx=0):
        return self._data[self._curr_idx + idx]

    def get_all(self):
        return self._data

    def set(self, data, idx=0):
        self._data[self._curr_idx + idx] = hip.asarray(data)

    def load(self, data):
        """Load the history, one level per integrator order
        (TimeDerivative.py:56-62): ``data[i]`` for i < integrator_order, each
        re-gridded onto this grid when its shape differs (Variable.regrid)."""
        order = self._params.integrator_order
        if len(data) < order:
            raise ValueError(f"history has {len(data)} levels, expected {order}")
        shape = tuple(self._data.shape[1:])
        for i in range(order):
            level = hip.asarray(data[i])
            if tuple(level.shape) != shape:
                level = regrid(self._params, level)
            self._data[i].copy_(level.to(self._data.dtype))

    def get_name(self):
        return self._dump_name
