"""Time integration on the GPU (mirror of melvin/Integrator.py).

``integrate`` is one fused HIP kernel: it combines the history ring with
Adams-Bashforth 2 or 4 (Integrator.py:5-18) and applies the semi-implicit
theta-scheme for the diffusion (58-63) or the explicit update that first
adds the diffusion into the stored RHS (53-56); the ring then advances.
``corrector`` (Adams-Moulton, 20-33) is provided as in the reference.
CFL control (35-44) reduces the physical velocities on the GPU in FP64.
"""
from __future__ import annotations

import numpy as np
import torch

from . import _lib, hip
from ._arrays import as_input
from ._plans import plan_for, ptr


class Integrator:
    def __init__(self, params, xp=hip):
        self._p = params
        self._dt = params.initial_dt
        self._dx = params.dx
        self._dz = params.dz
        self._cfl_cutoff = params.cfl_cutoff
        self._xp = hip
        self._order = params.integrator_order
        if self._order not in (2, 4):
            raise ValueError("integrator_order must be 2 or 4")
        self._plan = plan_for(params)
        if params.integrator == "semi-implicit":
            if params.is_fully_spectral():
                self.integrate = self._semi_implicit_spectral
                self._alpha = params.alpha
        elif params.integrator == "explicit":
            self.integrate = self._explicit

    def _combine(self, dvar, moulton):
        pl = self._plan
        out = torch.empty((pl.rows, pl.cols), dtype=pl.cplx, device=hip.device())
        data = dvar.get_all()
        _lib.call("mlv_combine", pl.handle, moulton, self._order, float(self._dt),
                  ptr(data), dvar.get_curr_idx() % self._order, ptr(out),
                  _lib.stream_handle())
        return out

    def predictor(self, dvar):
        return self._combine(dvar, 0)

    def corrector(self, dvar):
        return self._combine(dvar, 1)

    def set_dt(self, ux, uz):
        """Sets dt based on the CFL limit (signed maxima, as the reference)."""
        with np.errstate(divide="ignore", invalid="ignore"):
            cfl_dt = min(np.float64(self._dx) / np.float64(hip.max(ux.getp())),
                         np.float64(self._dz) / np.float64(hip.max(uz.getp())))
        if self._dt > cfl_dt or np.isnan(cfl_dt):
            raise Exception("CFL condition breached")
        while self._dt > self._cfl_cutoff * cfl_dt:
            self._dt = self._dt * 0.9

    def override_dt(self, dt):
        self._dt = dt

    def get_dt(self):
        return self._dt

    def _update(self, var, dvar, lin, scheme):
        pl = self._plan
        data = dvar.get_all()
        state = var.gets()
        if not state.is_contiguous() or not data.is_contiguous():
            raise ValueError("state and history must be contiguous")
        _lib.call("mlv_update", pl.handle, scheme, self._order,
                  float(getattr(self, "_alpha", self._p.alpha)), float(self._dt),
                  ptr(lin), ptr(state), ptr(data), dvar.get_curr_idx() % self._order,
                  _lib.stream_handle())
        dvar.advance()

    def _explicit(self, var, dvar, diffusion_term):
        pl = self._plan
        dif = as_input(diffusion_term, pl.cplx, (pl.rows, pl.cols))
        self._update(var, dvar, dif, _lib.SCHEME_EXPLICIT)

    def _semi_implicit_spectral(self, var, dvar, lin_op):
        pl = self._plan
        t = hip.asarray(lin_op)
        if t.is_complex():
            raise TypeError("the semi-implicit linear operator must be real")
        lin = t.to(torch.float64).expand(pl.rows, pl.cols).contiguous()
        self._update(var, dvar, lin, _lib.SCHEME_SEMI_IMPLICIT)
