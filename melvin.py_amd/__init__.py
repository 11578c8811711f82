"""melvin_amd: MI355X-native backend of Melvin.py's vorticity-streamfunction
timestep.

The classes keep the reference package's names and signatures (the
reference's melvin/__init__.py:1-15), so a loop body written for
``melvin`` runs unchanged with ``xp = melvin_amd.hip``: every operator is a
hand-written HIP kernel in libmelvin_hip.so reached through ctypes, and
arrays are torch tensors on the GPU.  ``melvin_amd.fused`` adds whole-step
kernels over batches of simulations and ``melvin_amd.serving`` a
continuous-batching request loop.

The directory is ``melvin.py_amd``; it is imported under the package name
``melvin_amd`` (see ``__graft_entry__.load_package``).
"""
from . import hip
from . import checkpoint, fdz, fused, initial, serving, utility
from .ArrayFactory import ArrayFactory
from .BasisFunctions import BasisFunctions
from .DataTransferer import DataTransferer
from .Integrator import Integrator
from .LaplacianSolver import LaplacianSolver
from .Parameters import Parameters
from .ScalarTracker import ScalarTracker
from .SpatialDifferentiator import SpatialDifferentiator
from .SpectralTransformer import SpectralTransformer
from .Ticker import Ticker
from .TimeDerivative import TimeDerivative
from .Timer import Timer
from .Variable import Variable
from .Simulation import Simulation

__all__ = [
    "hip", "fused", "initial", "serving", "utility", "ArrayFactory", "BasisFunctions", "DataTransferer", "Integrator",
    "LaplacianSolver", "Parameters", "ScalarTracker", "SpatialDifferentiator",
    "SpectralTransformer", "Ticker", "TimeDerivative", "Timer", "Variable",
    "Simulation",
]
