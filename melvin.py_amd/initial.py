"""Initial conditions of the benchmark configurations, built on the host in
FP64 (seeded NumPy PCG64, as utility.init_var_with_noise uses,
utility.py:31-39) and returned as compact spectra ready for ``Job.set_state``.

The transforms here are the reference's (rfft2 / (nx*nz), then truncation,
SpectralTransformer.py:152-199) done with numpy.fft on the host, once per
job: initial conditions are host data in the reference too (Variable.load,
Variable.py:67-80).
"""
from __future__ import annotations

import numpy as np


def _sizes(nx, nz):
    nn, nm = (nx - 1) // 3, (nz - 1) // 3
    return nn, nm


def spectral(phys):
    """Compact spectrum of a real (nx, nz) field."""
    nx, nz = phys.shape
    nn, nm = _sizes(nx, nz)
    f = np.fft.rfft2(phys) / (nx * nz)
    out = np.zeros((2 * nn + 1, nm), dtype=np.complex128)
    out[: nn + 1] = f[: nn + 1, :nm]
    out[-nn:] = f[-nn:, :nm]
    return out


def grid(nx, nz, lx, lz):
    x = np.linspace(0, lx, nx, endpoint=False)
    z = np.linspace(0, lz, nz, endpoint=False)
    return np.meshgrid(x, z, indexing="ij")


def noise(shape, amp, seed):
    rng = np.random.default_rng(seed)
    return amp * (2 * rng.random(shape) - 1.0)


def taylor_green(n, seed=0, amp=1e-6, kmax=8):
    """omega = 2 sin x sin z on [0, 2pi]^2 plus a seeded Gaussian perturbation
    on the modes 0 < |k| <= kmax (Hermitian on the kz = 0 column)."""
    X, Z = grid(n, n, 2 * np.pi, 2 * np.pi)
    w = spectral(2 * np.sin(X) * np.sin(Z))
    nn, nm = _sizes(n, n)
    rows = 2 * nn + 1
    kx = np.concatenate((np.arange(nn + 1), np.arange(-nn, 0)))[:, None]
    kz = np.arange(nm)[None, :]
    sel = (kx ** 2 + kz ** 2 <= kmax ** 2) & ~((kx == 0) & (kz == 0))
    rng = np.random.default_rng(seed)
    v = rng.standard_normal((rows, nm)) + 1j * rng.standard_normal((rows, nm))
    pert = np.where(sel, amp * v, 0)
    col = pert[:, 0].copy()
    for r in range(1, nn + 1):
        col[-r] = np.conj(col[r])
    col[0] = col[0].real
    pert[:, 0] = col
    return w + pert


def kelvin_helmholtz(n, seed=0, amp=1e-3, delta=0.05, length=2.0):
    """Double shear layer on [0, L]^2: u = tanh((z-L/4)/d) - tanh((z-3L/4)/d),
    omega = du/dz (omega = -lap psi, ux = -dpsi/dz), a passive scalar equal
    to u, and seeded uniform omega noise.  Returns (omega, scalar) spectra."""
    X, Z = grid(n, n, length, length)
    z1, z2 = 0.25 * length, 0.75 * length
    u = np.tanh((Z - z1) / delta) - np.tanh((Z - z2) / delta)
    w = (1 / np.cosh((Z - z1) / delta) ** 2 - 1 / np.cosh((Z - z2) / delta) ** 2) / delta
    w = w + noise((n, n), amp, seed)
    return spectral(w), spectral(u)


def _periodic_r2(X, Z, x0, z0, lx, lz):
    dx = (X - x0 + 0.5 * lx) % lx - 0.5 * lx
    dz = (Z - z0 + 0.5 * lz) % lz - 0.5 * lz
    return dx * dx + dz * dz


def vortex_pair(n, seed=0, radius=0.1, separation=0.25, length=1.0):
    """Two Lamb-Oseen vortices of circulation +1 and -1 and radius ``radius``
    (the profile of vortex_pairs.py:24-28) at seeded positions, on [0, L]^2
    with the nearest periodic image."""
    rng = np.random.default_rng(seed)
    X, Z = grid(n, n, length, length)
    cx, cz = rng.uniform(0.3, 0.7, size=2) * length
    th = rng.uniform(0, 2 * np.pi)
    ox, oz = 0.5 * separation * np.cos(th), 0.5 * separation * np.sin(th)
    w = np.zeros((n, n))
    for sgn, (x0, z0) in ((1.0, (cx + ox, cz + oz)), (-1.0, (cx - ox, cz - oz))):
        w += sgn / (np.pi * radius ** 2) * np.exp(-_periodic_r2(X, Z, x0, z0, length, length)
                                                  / radius ** 2)
    return spectral(w)


def tearing(n, seed=0, width=0.1, amp=1e-4, length=1.0):
    """Resistive tearing: a periodic double Harris sheet
    Bx = tanh((z-L/4)/a) - tanh((z-3L/4)/a) (summed over periodic images)
    with current j = dBx/dz (the reference evolves j,
    resistive_tearing_instability.py:36-60), omega = 0 plus seeded noise of
    amplitude ``amp``.  Returns (omega, j) spectra."""
    X, Z = grid(n, n, length, length)
    w = noise((n, n), amp, seed)
    return spectral(w), spectral(_sheet_current(Z, width, length, np))


# Periodic images each side of the domain summed into the double sheet:
# sheets of width a at L/4 and 3L/4 leave sech^2(L/(4a)) at the domain edge
# (2.7e-2 of the peak at a = L/10), which a single image would turn into a
# jump of j there; with five images the remainder is below sech^2(4.5 L/a).
_IMAGES = 2


def _sheet_current(Z, width, length, mod):
    """j = dBx/dz of the periodic double Harris sheet, Bx = sum over images
    of tanh((z - L/4 - kL)/a) - tanh((z - 3L/4 - kL)/a) (``mod`` is numpy
    or torch)."""
    j = 0.0 * Z
    for k in range(-_IMAGES, _IMAGES + 1):
        z1, z2 = (0.25 + k) * length, (0.75 + k) * length
        j = j + (1 / mod.cosh((Z - z1) / width) ** 2 - 1 / mod.cosh((Z - z2) / width) ** 2) / width
    return j


def ddc_noise(n, length, seed=0, amp=1e-3):
    """Double-diffusive start: omega = 0, independent seeded T and S noise."""
    w = np.zeros((n, n))
    t = noise((n, n), amp, 2 * seed + 1)
    s = noise((n, n), amp, 2 * seed + 2)
    return spectral(w), spectral(t), spectral(s)


# ---------------------------------------------------------------------------
# Device-side construction, for admission paths that must not stall on the
# host: the physical field is formed with torch element-wise operations on
# the GPU and brought to spectral form by the HIP forward transform.
# ---------------------------------------------------------------------------
def _device_grid(n, length, dtype):
    import torch

    x = torch.arange(n, device="cuda", dtype=dtype) * (length / n)
    return torch.meshgrid(x, x, indexing="ij")


def _to_spectral_device(phys, length, double):
    import torch

    from . import _lib
    from ._plans import Plan, ptr

    n = phys.shape[0]
    key = (n, length, double)
    plan = _DEVICE_PLANS.get(key)
    if plan is None:
        plan = _DEVICE_PLANS[key] = Plan(n, n, length, length, True, double, 2)
    out = torch.empty((plan.rows, plan.cols), dtype=plan.cplx, device="cuda")
    src = phys.to(plan.real).contiguous()
    _lib.call("mlv_to_spectral", plan.handle, ptr(src), ptr(out), _lib.stream_handle())
    return out


_DEVICE_PLANS: dict = {}


def vortex_pair_device(n, seed=0, radius=0.1, separation=0.25, length=1.0, double=False):
    """``vortex_pair`` formed on the GPU (same positions for a given seed)."""
    import torch

    rng = np.random.default_rng(seed)
    cx, cz = rng.uniform(0.3, 0.7, size=2) * length
    th = rng.uniform(0, 2 * np.pi)
    ox, oz = 0.5 * separation * np.cos(th), 0.5 * separation * np.sin(th)
    X, Z = _device_grid(n, length, torch.float64)
    w = torch.zeros((n, n), dtype=torch.float64, device="cuda")
    for sgn, (x0, z0) in ((1.0, (cx + ox, cz + oz)), (-1.0, (cx - ox, cz - oz))):
        dx = torch.remainder(X - x0 + 0.5 * length, length) - 0.5 * length
        dz = torch.remainder(Z - z0 + 0.5 * length, length) - 0.5 * length
        w += sgn / (np.pi * radius ** 2) * torch.exp(-(dx * dx + dz * dz) / radius ** 2)
    return _to_spectral_device(w, length, double)


def tearing_device(n, seed=0, width=0.1, amp=1e-4, length=1.0, double=False):
    """``tearing`` formed on the GPU; the noise comes from torch's seeded
    generator on the device (deterministic per seed, not NumPy's stream)."""
    import torch

    X, Z = _device_grid(n, length, torch.float64)
    j = _sheet_current(Z, width, length, torch)
    g = torch.Generator(device="cuda")
    g.manual_seed(int(seed))
    w = amp * (2 * torch.rand((n, n), generator=g, device="cuda", dtype=torch.float64) - 1)
    return _to_spectral_device(w, length, double), _to_spectral_device(j, length, double)


def _spectral_fdz(phys):
    """Fourier-x / FD-z spectrum of a real (nx, nz) field: rfft along x / nx,
    rows kx < nn (SpectralTransformer.py:63-88)."""
    nx = phys.shape[0]
    nn = (nx - 1) // 3
    return (np.fft.rfft(phys, axis=0) / nx)[:nn].astype(np.complex128)


def rbc_conduction(nx, nz, lx=2.0, lz=1.0, seed=0, amp=1e-3):
    """Rayleigh-Benard start: omega = 0 and the conductive profile
    T = 1 - z/lz plus seeded uniform noise of amplitude ``amp`` on the
    z grid of rayleigh_benard_convection.py:22-25 (endpoint=False).
    Returns (omega, T) Fourier-x / FD-z spectra."""
    x = np.linspace(0, lx, nx, endpoint=False)
    z = np.linspace(0, lz, nz, endpoint=False)
    _, Z = np.meshgrid(x, z, indexing="ij")
    t = 1.0 - Z / lz + noise((nx, nz), amp, seed)
    return _spectral_fdz(np.zeros((nx, nz))), _spectral_fdz(t)
