"""Helpers of the reference's utility module (melvin/utility.py:9-79) on the
GPU: the velocity from vorticity, kinetic energy, noise initial
conditions."""
from __future__ import annotations

import ctypes

import numpy as np
from numpy.random import default_rng

from . import _lib, hip
from ._plans import plan_for, ptr
from .BasisFunctions import BasisFunctions


def sech(x):
    return 1.0 / np.cosh(x)


def init_var_with_noise(var, epsilon, seed=0):
    """Uniform noise epsilon*(2U-1) from default_rng(seed), loaded as physical
    data (utility.py:31-39)."""
    rng = default_rng(seed)
    shape = tuple(var.getp().shape)
    data = np.zeros(shape)
    data += epsilon * (2 * rng.random(shape) - 1.0)
    var.load(data, is_physical=True)


def calc_kinetic_energy(ux, uz, xp=hip, params=None):
    """0.5 * sum(ux^2 + uz^2) / (nx*nz), FP64 accumulation (utility.py:42-59)."""
    p = params
    return 0.5 * hip.sum_sq2(uz.getp(), ux.getp()) / (p.nx * p.nz)


def calc_velocity_from_vorticity(vorticity, streamfunction, ux, uz,
                                 laplacian_solver=None):
    """psi = solve(-w); ux = -d(psi)/dz, uz = d(psi)/dx in physical space
    (utility.py:62-79).  The spectral part is one HIP kernel (or the FD-z
    tridiagonal solve); the physical fields come from the transforms."""
    params = vorticity._params
    pl = plan_for(params)
    w = vorticity.gets()
    psi = streamfunction._sdata
    null = ctypes.c_void_p(0)
    if not (w.is_contiguous() and psi.is_contiguous()):
        raise ValueError("vorticity and streamfunction must be contiguous")
    fdm_z = streamfunction._basis_functions[1] is BasisFunctions.FDM
    _lib.call("mlv_velocity", pl.handle, ptr(w), ptr(psi),
              null if fdm_z else ptr(ux._sdata), ptr(uz._sdata),
              _lib.stream_handle())
    if fdm_z:
        streamfunction.to_physical()
        ux.setp(-streamfunction.pddz())
    else:
        ux.to_physical()
    uz.to_physical()
