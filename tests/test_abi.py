"""The C ABI without a GPU: the library loads, exports every symbol the
header declares, the ctypes mirror of mlv_job_desc matches the C layout,
and argument validation fails cleanly before touching the device."""
import ctypes
import os
import re
import subprocess

import pytest

from conftest import ROOT

HEADER = os.path.join(ROOT, "include", "melvin_hip.h")


def _lib():
    import __graft_entry__

    m = __graft_entry__.load_package()
    from melvin_amd import _lib as L  # noqa: F401

    return m._lib


def _declared():
    text = open(HEADER).read()
    return sorted(set(re.findall(r"^\w[\w\s\*]*?\b(mlv_\w+)\s*\(", text, re.M)))


def test_library_exports_every_declared_symbol():
    L = _lib()
    lib = L.lib()
    declared = _declared()
    assert len(declared) >= 20
    for name in declared:
        assert hasattr(lib, name), name
    assert set(declared) == set(L.EXPORTS)
    assert lib.mlv_abi_version() == 2


def test_job_desc_layout_matches_c(tmp_path):
    L = _lib()
    src = tmp_path / "layout.c"
    src.write_text(
        '#include <stdio.h>\n#include <stddef.h>\n#include "melvin_hip.h"\n'
        "int main(void){printf(\"%zu %zu %zu %zu %zu\\n\", sizeof(mlv_job_desc),"
        " offsetof(mlv_job_desc, dt), offsetof(mlv_job_desc, state),"
        " offsetof(mlv_job_desc, phys), offsetof(mlv_job_desc, parity));return 0;}\n")
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-I", os.path.dirname(HEADER), str(src), "-o", str(exe)], check=True)
    vals = [int(v) for v in subprocess.check_output([str(exe)]).split()]
    J = L.JobDesc
    assert vals == [ctypes.sizeof(J), J.dt.offset, J.state.offset, J.phys.offset,
                    J.parity.offset]


def test_argument_errors_are_reported():
    L = _lib()
    lib = L.lib()
    assert lib.mlv_plan_create(None, 32, 32, 1.0, 1.0, 0, 1, 2) == L.MLV_ERR_ARG
    assert b"NULL" in lib.mlv_last_error()
    h = ctypes.c_void_p()
    assert lib.mlv_plan_create(ctypes.byref(h), 48, 32, 1.0, 1.0, 0, 1, 2) == L.MLV_ERR_UNSUPPORTED
    assert b"power of two" in lib.mlv_last_error()
    assert lib.mlv_plan_create(ctypes.byref(h), 32, 32, 1.0, 1.0, 0, 1, 3) == L.MLV_ERR_ARG
    assert lib.mlv_plan_create(ctypes.byref(h), 32, 32, -1.0, 1.0, 0, 1, 2) == L.MLV_ERR_ARG
    with pytest.raises(NotImplementedError):
        L.call("mlv_plan_create", ctypes.byref(h), 32, 33, 1.0, 1.0, 0, 1, 2)
    assert lib.mlv_to_physical(None, None, None, None) == L.MLV_ERR_ARG
    assert lib.mlv_to_spectral_basis(None, 1, 1, None, None, None) == L.MLV_ERR_ARG
    assert lib.mlv_to_physical_basis(None, 1, 1, None, None, None) == L.MLV_ERR_ARG
    assert lib.mlv_spec_op_basis(None, 0, 1, 1, None, None, None) == L.MLV_ERR_ARG
    assert lib.mlv_batch_step(None, 1, 0, None) == L.MLV_ERR_ARG
    assert lib.mlv_batch_steps(None) == -1
    assert lib.mlv_batch_plan(None, None) == L.MLV_ERR_ARG
    assert lib.mlv_batch_set_plan(None, 0, 0, 0) == L.MLV_ERR_ARG
    # xp.fft entry points: argument checks come before any device call
    buf = ctypes.c_void_p(1)
    assert lib.mlv_rfft(0, 2, 32, 32, None, buf, buf, None) == L.MLV_ERR_ARG
    assert lib.mlv_rfft(0, 3, 32, 32, buf, buf, buf, None) == L.MLV_ERR_ARG
    assert lib.mlv_rfft(2, 2, 32, 32, buf, buf, buf, None) == L.MLV_ERR_ARG
    assert lib.mlv_rfft(0, 0, 48, 33, buf, buf, None, None) == L.MLV_ERR_UNSUPPORTED
    assert b"nx must be a power of two" in lib.mlv_last_error()
    assert lib.mlv_irfft(1, 1, 33, 2, buf, buf, None, None) == L.MLV_ERR_UNSUPPORTED
    assert lib.mlv_irfft(1, 2, 16, 8192, buf, buf, buf, None) == L.MLV_ERR_UNSUPPORTED
    n = ctypes.c_size_t()
    assert lib.mlv_job_work_bytes(9, 1, 32, 32, 0, ctypes.byref(n)) == L.MLV_ERR_ARG
    assert lib.mlv_job_work_bytes(L.PROB_VORTICITY, 1, 32, 32, 0, ctypes.byref(n)) == 0
    # TGV channels: w, ux, uz + Px, Pz -> 5 channels of nx x (nm rounded up
    # to whole groups of 8 kz) complex64, then the x-mean profiles of six
    # channels (3 * nz complex64)
    assert n.value == 5 * 32 * 16 * 8 + 3 * 32 * 8


def test_fdz_job_desc_layout_matches_c(tmp_path):
    L = _lib()
    src = tmp_path / "layout_fdz.c"
    src.write_text(
        '#include <stdio.h>\n#include <stddef.h>\n#include "melvin_hip.h"\n'
        "int main(void){printf(\"%zu %zu %zu %zu %zu %zu\\n\", sizeof(mlv_fdz_job_desc),"
        " offsetof(mlv_fdz_job_desc, alpha), offsetof(mlv_fdz_job_desc, state),"
        " offsetof(mlv_fdz_job_desc, work), offsetof(mlv_fdz_job_desc, phys),"
        " offsetof(mlv_fdz_job_desc, parity));return 0;}\n")
    exe = tmp_path / "layout_fdz"
    subprocess.run(["gcc", "-I", os.path.dirname(HEADER), str(src), "-o", str(exe)], check=True)
    vals = [int(v) for v in subprocess.check_output([str(exe)]).split()]
    J = L.FdzJobDesc
    assert vals == [ctypes.sizeof(J), J.alpha.offset, J.state.offset, J.work.offset,
                    J.phys.offset, J.parity.offset]


def test_fdz_argument_errors_are_reported():
    """Descriptors are validated on the host before any device call."""
    L = _lib()
    lib = L.lib()
    n = ctypes.c_size_t()
    assert lib.mlv_fdz_work_bytes(32, 17, 0, ctypes.byref(n)) == 0
    # Qx/Qz of two fields, velocity of three states, predicted state
    assert n.value == 12 * 10 * 17 * 8
    assert lib.mlv_fdz_work_bytes(32, 3, 0, ctypes.byref(n)) == L.MLV_ERR_ARG
    d = L.FdzJobDesc()
    d.nx, d.nz, d.lx, d.lz = 32, 17, 2.0, 1.0
    d.fd_order, d.ab_order, d.scheme = 2, 2, L.FDZ_PC
    d.alpha, d.dt, d.Pr, d.Ra = 0.51, 1e-4, 1.0, 1e5
    dummy = ctypes.c_void_p(16)
    for b in range(2):
        for f in range(2):
            d.state[b][f] = dummy
    d.hist[0] = d.hist[1] = dummy
    d.work = dummy
    h = ctypes.c_void_p()
    for field, bad, word in (("nx", 48, b"power of two"), ("nx", 4096, b"power of two"),
                             ("nz", 769, b"nz"),
                             ("scheme", 7, b"scheme"), ("fd_order", 3, b"fd_order"),
                             ("ab_order", 3, b"ab_order"), ("dt", 0.0, b"positive")):
        good = getattr(d, field)
        setattr(d, field, bad)
        assert lib.mlv_fdz_batch_create(ctypes.byref(h), 0, ctypes.byref(d), 1) == L.MLV_ERR_ARG
        assert word in lib.mlv_last_error(), field
        setattr(d, field, good)
    d.work = None
    assert lib.mlv_fdz_batch_create(ctypes.byref(h), 0, ctypes.byref(d), 1) == L.MLV_ERR_ARG
    assert lib.mlv_fdz_batch_step(None, 1, 0, None) == L.MLV_ERR_ARG
    assert lib.mlv_fdz_batch_steps(None) == -1
