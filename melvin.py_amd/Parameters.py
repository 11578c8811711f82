"""Run configuration (mirror of melvin/Parameters.py).

Same defaults (Parameters.py:9-30), required keys (7), validation (46-55),
derived sizes and dtypes (63-92) and JSON save (94-96) as the reference;
``load_json`` adds the loader the reference's README promises but lacks.
"""
from __future__ import annotations

import json
import sys

import numpy as np


class Parameters:
    required_params = ["nx", "nz", "lx", "lz", "final_time"]

    integrator_order = 2
    integrator = "semi-implicit"
    spatial_derivative_order = 2
    alpha = 0.51
    cfl_cutoff = 0.5
    cfl_cadence = 10
    tracker_cadence = 100
    save_cadence = 100
    load_from = None

    discretisation = ["spectral", "spectral"]
    precision = "double"

    nx = None
    nz = None
    lx = None
    lz = None
    final_time = None

    def __init__(self, params, validate=True):
        if validate and not self.is_valid(params):
            sys.exit(-1)
        self.load_from_dict(params)
        self._original_params = params

    @classmethod
    def load_json(cls, fname, validate=True):
        with open(fname) as fp:
            return cls(json.load(fp), validate=validate)

    def load_from_dict(self, params):
        for key, value in params.items():
            setattr(self, key, value)
        self.set_derived_params(params)

    def is_valid(self, params):
        ok = True
        for key in self.required_params:
            if key not in params:
                print(key, "missing from input parameters.")
                ok = False
        # Evaluated before the dict is applied, as in the reference
        # (Parameters.py:52-54), so it reads the class defaults.
        if "fdm" in self.discretisation and self.integrator == "explicit":
            print("FDM and implicit method currently not supported.")
            ok = False
        return int(ok)

    def is_fully_spectral(self):
        return (self.discretisation[0] == "spectral"
                and self.discretisation[1] == "spectral")

    def set_derived_params(self, params):
        if "dump_cadence" not in params:
            self.dump_cadence = 0.1 * self.final_time
        if self.discretisation[0] == "spectral":
            self.nn = (self.nx - 1) // 3
        if self.discretisation[1] == "spectral":
            self.nm = (self.nz - 1) // 3
        if self.is_fully_spectral():
            self.spectral_shape = (2 * self.nn + 1, self.nm)
        elif self.discretisation[0] == "fdm":
            self.spectral_shape = (self.nx, self.nm)
        elif self.discretisation[1] == "fdm":
            self.spectral_shape = (self.nn, self.nz)
        self.physical_shape = (self.nx, self.nz)
        self.dx = self.lx / self.nx
        self.dz = self.lz / self.nz
        if self.precision == "double":
            self.complex, self.float = np.complex128, np.float64
        elif self.precision == "single":
            self.complex, self.float = np.complex64, np.float32
        if "initial_dt" not in params:
            self.initial_dt = 0.2 * min(self.dx, self.dz)

    def save(self, fname="params.json"):
        with open(fname, "w") as fp:
            json.dump(self._original_params, fp)
