"""Loop bookkeeping (mirror of melvin/Simulation.py:18-221): owns the shared
operators, builds variables, runs the tickers after each loop (CFL every
``cfl_cadence`` loops, trackers, saves, dumps).

``dump``/``load`` write and read a self-describing ``.npz`` without pickled
objects: spectral variables, every derivative ring with its own current
slot, the ticker counters, dt, t and the loop counter, plus the parameter
dict as JSON.  (The reference's dump pickles tickers and Parameters and its
load path drops the imaginary part, Simulation.py:155-197, Variable.py:
130-151; this format restores a run bit-exactly.)
"""
from __future__ import annotations

import json

import numpy as np

from . import hip
from .ArrayFactory import ArrayFactory
from .DataTransferer import DataTransferer
from .Integrator import Integrator
from .LaplacianSolver import LaplacianSolver
from .ScalarTracker import ScalarTracker
from .SpatialDifferentiator import SpatialDifferentiator
from .SpectralTransformer import SpectralTransformer
from .Ticker import Ticker
from .TimeDerivative import TimeDerivative
from .Timer import Timer
from .Variable import Variable


class Simulation:
    def __init__(self, params, xp=hip):
        self._params = params
        self._xp = hip
        self._data_trans = DataTransferer(hip)
        self._integrator = Integrator(params, hip)
        self._array_factory = ArrayFactory(params, hip)
        self._spatial_diff = SpatialDifferentiator(params, hip, self._array_factory)
        self._spectral_trans = SpectralTransformer(params, hip, self._array_factory)
        self._t = 0
        self._loop_counter = 0
        self._dump_vars = []
        self._dump_dvars = []
        self._dump_idx = 0
        self._save_vars = []
        self._tickers = []
        self._trackers = []
        self._timer = Timer()
        self._wallclock_remaining = 0.0
        self._wallclock_ticker = Ticker(100, self.calc_time_remaining, is_loop_counter=True)
        self.register_ticker(self._wallclock_ticker)

    def make_variable(self, name, basis_fns):
        return Variable(self._params, hip, sd=self._spatial_diff, st=self._spectral_trans,
                        dt=self._data_trans, array_factory=self._array_factory,
                        dump_name=name, basis_functions=basis_fns)

    def make_derivative(self, name):
        return TimeDerivative(self._params, hip, dump_name=name)

    def init_laplacian_solver(self, basis_fns):
        self._laplacian_solver = LaplacianSolver(self._params, hip, basis_fns,
                                                 spatial_diff=self._spatial_diff,
                                                 array_factory=self._array_factory)

    def config_dump(self, variables, derivatives):
        self._dump_vars = variables
        self._dump_dvars = derivatives
        self._dump_ticker = Ticker(self._params.dump_cadence, self.dump,
                                   dump_name="dump_ticker", is_loop_counter=False)
        self.register_ticker(self._dump_ticker)

    def config_save(self, variables):
        self._save_vars = variables
        self._save_ticker = Ticker(self._params.save_cadence, self.save,
                                   dump_name="save_ticker", is_loop_counter=False)
        self.register_ticker(self._save_ticker)

    def config_cfl(self, ux, uz):
        self._ux = ux
        self._uz = uz
        self._cfl_ticker = Ticker(self._params.cfl_cadence, self.set_dt,
                                  dump_name="cfl_ticker", is_loop_counter=True)
        self.register_ticker(self._cfl_ticker)

    def config_scalar_trackers(self, trackers):
        self._tracker_fns = [trackers[name] for name in trackers]
        self._trackers = [ScalarTracker(self._params, hip, name) for name in trackers]
        self._tracker_ticker = Ticker(self._params.tracker_cadence, self.track_scalars,
                                      dump_name="tracker_ticker", is_loop_counter=True)
        self.register_ticker(self._tracker_ticker)
        self._save_vars += self._trackers

    def track_scalars(self, ticker):
        for tracker, fn in zip(self._trackers, self._tracker_fns):
            tracker.append(self._t, fn())

    def set_dt(self, ticker):
        self._integrator.set_dt(self._ux, self._uz)

    def save(self, ticker):
        self.print_info()
        for var in self._save_vars:
            var.save(ticker.times_fired)

    def print_info(self):
        hours = int(self._wallclock_remaining / 3600)
        minutes = int((self._wallclock_remaining / 3600.0 - hours) * 60)
        print(f"{self._t / self._params.final_time * 100:.2f}% complete",
              f"t = {self._t:.2e}", f"dt = {self._integrator._dt:.2e}",
              f"Remaining: {hours} hr, {minutes} min")

    def form_dumpname(self, index):
        return f"dump{index:04d}.npz"

    def dump(self, ticker=None, fname=None):
        if fname is None:
            fname = self.form_dumpname(ticker.times_fired if ticker else self._dump_idx)
        data = {"var__" + v.get_name(): hip.asnumpy(v.gets()) for v in self._dump_vars}
        for d in self._dump_dvars:
            data["dvar__" + d.get_name()] = hip.asnumpy(d.get_all())
            data["curr__" + d.get_name()] = np.array(d.get_curr_idx())
        for tk in self._tickers:
            data["ticker__" + tk.get_name()] = np.array(
                [tk._counter, tk.times_fired], dtype=np.float64)
        data.update(dt=np.array(self._integrator._dt), t=np.array(self._t),
                    loop_counter=np.array(self._loop_counter),
                    params=np.array(json.dumps(self._params._original_params,
                                               default=str)))
        np.savez(fname, **data)
        return fname

    def load(self, index_or_fname):
        fname = (index_or_fname if isinstance(index_or_fname, str)
                 else self.form_dumpname(index_or_fname))
        with np.load(fname, allow_pickle=False) as z:
            for v in self._dump_vars:
                v.load(z["var__" + v.get_name()])
            for d in self._dump_dvars:
                d.load(z["dvar__" + d.get_name()])
                d.set_curr_idx(int(z["curr__" + d.get_name()]))
            for tk in self._tickers:
                key = "ticker__" + tk.get_name()
                if key in z:
                    counter, fired = z[key]
                    tk.restore({"counter": counter.item(), "times_fired": int(fired)})
            self._integrator._dt = float(z["dt"])
            self._t = float(z["t"])
            self._loop_counter = int(z["loop_counter"])

    def is_running(self):
        return self._t < self._params.final_time

    def register_ticker(self, ticker):
        self._tickers.append(ticker)

    def end_loop(self):
        self._loop_counter += 1
        self._t += self._integrator._dt
        for ticker in self._tickers:
            ticker.tick(self._t, self._loop_counter)

    def calc_time_remaining(self, ticker):
        self._timer.split()
        per_step = self._timer.diff / ticker._cadence
        self._wallclock_remaining = (per_step * (self._params.final_time - self._t)
                                     / self._integrator._dt)

    def get_laplacian_solver(self):
        return self._laplacian_solver
