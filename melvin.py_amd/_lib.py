"""ctypes binding of libmelvin_hip.so (the C ABI in include/melvin_hip.h).

The library is built in-tree by ``__graft_entry__.build()`` (or ``make`` in
``csrc/``).  There is no fallback: if the library or a GPU is missing every
call raises.
"""
from __future__ import annotations

import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libmelvin_hip.so")

MLV_OK = 0
MLV_ERR_ARG = 1
MLV_ERR_HIP = 2
MLV_ERR_UNSUPPORTED = 3

FULLY_SPECTRAL = 0
FOURIER_X_FD_Z = 1
F32 = 0
F64 = 1
FFT_AXIS0, FFT_AXIS1, FFT_AXES01 = 0, 1, 2

SPEC_DDX, SPEC_DDZ, SPEC_D2DX2, SPEC_D2DZ2, SPEC_LAP, SPEC_SNABLA2, SPEC_POISSON = range(7)
FD_DDX, FD_DDZ, FD_D2DZ2 = range(3)
SCHEME_SEMI_IMPLICIT, SCHEME_EXPLICIT = 0, 1
RED_MAX, RED_SUM, RED_SUMSQ2, RED_DOT = range(4)
PROB_VORTICITY, PROB_VORT_SCALAR, PROB_DDC, PROB_MHD = range(4)
FDZ_EXPLICIT, FDZ_SEMI_IMPLICIT, FDZ_PC = range(3)

# Every symbol include/melvin_hip.h declares (checked by the tests).
EXPORTS = (
    "mlv_last_error", "mlv_abi_version", "mlv_plan_create", "mlv_plan_destroy",
    "mlv_plan_info", "mlv_to_physical", "mlv_to_spectral", "mlv_spec_op",
    "mlv_fd_op", "mlv_flux_divergence", "mlv_poisson", "mlv_update",
    "mlv_reduce", "mlv_reduce_noalloc", "mlv_combine", "mlv_velocity", "mlv_job_work_bytes", "mlv_batch_create", "mlv_batch_rebuild",
    "mlv_batch_destroy", "mlv_batch_step", "mlv_batch_graph_step", "mlv_batch_set_dt",
    "mlv_batch_steps", "mlv_batch_profile", "mlv_batch_profile_read",
    "mlv_batch_run", "mlv_batch_job_time", "mlv_batch_tracker", "mlv_batch_autotune",
    "mlv_batch_plan", "mlv_batch_set_plan", "mlv_rfft", "mlv_irfft",
    "mlv_to_spectral_basis", "mlv_to_physical_basis", "mlv_spec_op_basis",
    "mlv_fdz_work_bytes", "mlv_fdz_batch_create", "mlv_fdz_batch_destroy",
    "mlv_fdz_batch_prime", "mlv_fdz_batch_step", "mlv_fdz_batch_graph_step",
    "mlv_fdz_batch_set_dt", "mlv_fdz_batch_steps", "mlv_fdz_batch_profile",
    "mlv_fdz_batch_profile_read", "mlv_fdz_batch_run", "mlv_fdz_batch_job_time",
    "mlv_fdz_batch_tracker",
)


class MelvinHipError(RuntimeError):
    """A libmelvin_hip call returned a non-zero status."""


class JobDesc(ctypes.Structure):
    """Mirror of ``mlv_job_desc``."""

    _fields_ = [
        ("problem", ctypes.c_int),
        ("nx", ctypes.c_int),
        ("nz", ctypes.c_int),
        ("lx", ctypes.c_double),
        ("lz", ctypes.c_double),
        ("fd_order", ctypes.c_int),
        ("ab_order", ctypes.c_int),
        ("alpha", ctypes.c_double),
        ("dt", ctypes.c_double),
        ("coeff", ctypes.c_double * 4),
        ("param", ctypes.c_double * 4),
        ("nfield", ctypes.c_int),
        ("state", (ctypes.c_void_p * 4) * 2),
        ("hist", ctypes.c_void_p * 4),
        ("comp", ctypes.c_void_p * 4),
        ("work", ctypes.c_void_p),
        ("psi", ctypes.c_void_p),
        ("phys", ctypes.c_void_p * 8),
        ("curr", ctypes.c_int),
        ("parity", ctypes.c_int),
        ("t", ctypes.c_double),
        ("loops", ctypes.c_longlong),
        ("mean_split", ctypes.c_int),
    ]


class FdzJobDesc(ctypes.Structure):
    """Mirror of ``mlv_fdz_job_desc``."""

    _fields_ = [
        ("nx", ctypes.c_int),
        ("nz", ctypes.c_int),
        ("lx", ctypes.c_double),
        ("lz", ctypes.c_double),
        ("fd_order", ctypes.c_int),
        ("ab_order", ctypes.c_int),
        ("scheme", ctypes.c_int),
        ("alpha", ctypes.c_double),
        ("dt", ctypes.c_double),
        ("Pr", ctypes.c_double),
        ("Ra", ctypes.c_double),
        ("state", (ctypes.c_void_p * 2) * 2),
        ("hist", ctypes.c_void_p * 2),
        ("comp", ctypes.c_void_p * 2),
        ("psi", ctypes.c_void_p * 2),
        ("work", ctypes.c_void_p),
        ("phys", ctypes.c_void_p * 4),
        ("curr", ctypes.c_int),
        ("parity", ctypes.c_int),
        ("t", ctypes.c_double),
        ("loops", ctypes.c_longlong),
    ]


_lib = None


def _declare(lib):
    vp, i, d, sz = ctypes.c_void_p, ctypes.c_int, ctypes.c_double, ctypes.c_size_t
    pp = ctypes.POINTER(ctypes.c_void_p)
    ip = ctypes.POINTER(ctypes.c_int)
    sig = {
        "mlv_last_error": ([], ctypes.c_char_p),
        "mlv_abi_version": ([], i),
        "mlv_plan_create": ([pp, i, i, d, d, i, i, i], i),
        "mlv_plan_destroy": ([vp], i),
        "mlv_plan_info": ([vp, ip, ip, ip, ip], i),
        "mlv_to_physical": ([vp, vp, vp, vp], i),
        "mlv_to_spectral": ([vp, vp, vp, vp], i),
        "mlv_spec_op": ([vp, i, vp, vp, vp], i),
        "mlv_fd_op": ([vp, i, i, i, i, vp, vp, vp], i),
        "mlv_flux_divergence": ([vp, vp, vp, vp, vp, vp], i),
        "mlv_poisson": ([vp, vp, vp, vp], i),
        "mlv_update": ([vp, i, i, d, d, vp, vp, vp, i, vp], i),
        "mlv_reduce": ([vp, i, i, vp, vp, sz, d, vp, vp], i),
        "mlv_reduce_noalloc": ([i, i, vp, vp, sz, d, vp, vp], i),
        "mlv_combine": ([vp, i, i, d, vp, i, vp, vp], i),
        "mlv_velocity": ([vp, vp, vp, vp, vp, vp], i),
        "mlv_job_work_bytes": ([i, i, i, i, i, ctypes.POINTER(sz)], i),
        "mlv_batch_create": ([pp, i, ctypes.POINTER(JobDesc), i], i),
        "mlv_batch_rebuild": ([vp, ctypes.POINTER(JobDesc), i], i),
        "mlv_batch_destroy": ([vp], i),
        "mlv_batch_step": ([vp, i, i, vp], i),
        "mlv_batch_graph_step": ([vp, i, vp], i),
        "mlv_batch_set_dt": ([vp, i, d, vp], i),
        "mlv_batch_steps": ([vp], ctypes.c_longlong),
        "mlv_batch_profile": ([vp, i], i),
        "mlv_batch_profile_read": ([vp, ctypes.POINTER(d), ctypes.POINTER(ctypes.c_longlong)], i),
        "mlv_to_spectral_basis": ([vp, i, i, vp, vp, vp], i),
        "mlv_to_physical_basis": ([vp, i, i, vp, vp, vp], i),
        "mlv_spec_op_basis": ([vp, i, i, i, vp, vp, vp], i),
        "mlv_batch_run": ([vp, i, i, d, i, vp], i),
        "mlv_batch_autotune": ([vp, i, vp, ip], i),
        "mlv_batch_plan": ([vp, ip], i),
        "mlv_rfft": ([i, i, i, i, vp, vp, vp, vp], i),
        "mlv_irfft": ([i, i, i, i, vp, vp, vp, vp], i),
        "mlv_batch_set_plan": ([vp, i, i, i], i),
        "mlv_batch_job_time": ([vp, i, ctypes.POINTER(d), ctypes.POINTER(d),
                                ctypes.POINTER(ctypes.c_longlong), ip], i),
        "mlv_batch_tracker": ([vp, i, ctypes.POINTER(d), i, ip], i),
        "mlv_fdz_work_bytes": ([i, i, i, ctypes.POINTER(sz)], i),
        "mlv_fdz_batch_create": ([pp, i, ctypes.POINTER(FdzJobDesc), i], i),
        "mlv_fdz_batch_destroy": ([vp], i),
        "mlv_fdz_batch_prime": ([vp, vp], i),
        "mlv_fdz_batch_step": ([vp, i, i, vp], i),
        "mlv_fdz_batch_graph_step": ([vp, i, vp], i),
        "mlv_fdz_batch_set_dt": ([vp, i, d, vp], i),
        "mlv_fdz_batch_steps": ([vp], ctypes.c_longlong),
        "mlv_fdz_batch_profile": ([vp, i], i),
        "mlv_fdz_batch_profile_read": ([vp, ctypes.POINTER(d),
                                        ctypes.POINTER(ctypes.c_longlong)], i),
        "mlv_fdz_batch_run": ([vp, i, i, d, i, vp], i),
        "mlv_fdz_batch_job_time": ([vp, i, ctypes.POINTER(d), ctypes.POINTER(d),
                                    ctypes.POINTER(ctypes.c_longlong), ip], i),
        "mlv_fdz_batch_tracker": ([vp, i, ctypes.POINTER(d), i, ip], i),
    }
    for name, (args, res) in sig.items():
        fn = getattr(lib, name)
        fn.argtypes = args
        fn.restype = res


def lib():
    """The loaded library; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise MelvinHipError(
                f"{LIB_PATH} is missing: build it with __graft_entry__.build() "
                "(there is no CPU fallback)")
        handle = ctypes.CDLL(LIB_PATH)
        _declare(handle)
        _lib = handle
    return _lib


def check(status, what=""):
    if status != MLV_OK:
        msg = lib().mlv_last_error().decode(errors="replace")
        if status == MLV_ERR_ARG:
            raise ValueError(f"{what}: {msg}")
        if status == MLV_ERR_UNSUPPORTED:
            raise NotImplementedError(f"{what}: {msg}")
        raise MelvinHipError(f"{what}: {msg} (status {status})")


def call(name, *args):
    check(getattr(lib(), name)(*args), name)


def stream_handle(stream=None):
    """HIP stream handle of ``stream`` or of torch's current stream."""
    import torch

    s = stream if stream is not None else torch.cuda.current_stream()
    return ctypes.c_void_p(s.cuda_stream)
