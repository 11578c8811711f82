"""FP64 NumPy restatement of Melvin.py's vorticity-streamfunction timestep.

TEST INFRASTRUCTURE ONLY.  This module is the checker for the HIP product
path: only ``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline``
leg of ``bench.py`` may import it.  The product package never calls it and
has no CPU fallback.

Everything here is plain FP64 NumPy (``numpy.fft`` = pocketfft, the same
library the reference calls with ``xp = numpy``).  Each function cites the
reference file:line whose behaviour it restates.  Arrays follow the
reference layouts:

* fully spectral: spectral ``(2nn+1, nm)`` complex, row ``r`` <-> ``kx = r``
  for ``r <= nn`` else ``r - (2nn+1)``; column ``c`` <-> ``kz = c``.
* Fourier-x / FD-z: spectral ``(nn, nz)`` complex, row = ``kx`` in
  ``[0, nn)``, column = z grid point.
* physical ``(nx, nz)`` real, z contiguous.

Pins (see tests/test_oracle_*.py): the reference's own known-answer tests
restated, closed forms (single modes, modified wavenumbers, theta-scheme
decay, polynomial exactness of Adams-Bashforth), dense / brute-force
solves on small grids, and golden vectors written by the reference itself
(tests/golden/make_golden.py).  A function with none of these says
"parity unpinned" in its docstring.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

COMPLEX_EXP, COSINE, SINE, FDM = 0, 1, 2, 3


# --------------------------------------------------------------------------
# Parameters (melvin/Parameters.py:9-30 defaults, 63-92 derived values)
# --------------------------------------------------------------------------
@dataclass
class Params:
    nx: int
    nz: int
    lx: float
    lz: float
    final_time: float = 1.0
    integrator_order: int = 2
    integrator: str = "semi-implicit"
    spatial_derivative_order: int = 2
    alpha: float = 0.51
    cfl_cutoff: float = 0.5
    cfl_cadence: int = 10
    tracker_cadence: int = 100
    discretisation: tuple = ("spectral", "spectral")
    initial_dt: float | None = None
    extra: dict = field(default_factory=dict)

    def __post_init__(self):
        # Parameters.py:67-92
        self.discretisation = tuple(self.discretisation)
        if self.discretisation[0] == "spectral":
            self.nn = (self.nx - 1) // 3
        if self.discretisation[1] == "spectral":
            self.nm = (self.nz - 1) // 3
        if self.fully_spectral:
            self.spectral_shape = (2 * self.nn + 1, self.nm)
        elif self.discretisation[1] == "fdm":
            self.spectral_shape = (self.nn, self.nz)
        else:
            raise NotImplementedError("finite differences in x")
        self.physical_shape = (self.nx, self.nz)
        self.dx = self.lx / self.nx
        self.dz = self.lz / self.nz
        if self.initial_dt is None:
            self.initial_dt = 0.2 * min(self.dx, self.dz)

    @property
    def fully_spectral(self):
        return self.discretisation == ("spectral", "spectral")

    def __getattr__(self, key):
        extra = self.__dict__.get("extra", {})
        if key in extra:
            return extra[key]
        raise AttributeError(key)


def make_params(d: dict) -> Params:
    known = {f for f in Params.__dataclass_fields__ if f != "extra"}
    kw = {k: v for k, v in d.items() if k in known}
    extra = {k: v for k, v in d.items() if k not in known}
    return Params(**kw, extra=extra)


# --------------------------------------------------------------------------
# Mode numbers and derivative factors
# (ArrayFactory.py:8-26, BasisFunctions.py:26-59)
# --------------------------------------------------------------------------
def mode_numbers(p: Params):
    """Integer (n, m) meshgrid, 'ij' indexing (ArrayFactory.py:8-26)."""
    if p.fully_spectral:
        n = np.concatenate((np.arange(0, p.nn + 1), np.arange(-p.nn, 0)))
        m = np.arange(0, p.nm)
    else:
        n = np.arange(0, p.nn)
        m = np.arange(0, p.nz)
    return np.meshgrid(n, m, indexing="ij")


def kx_factor(p):
    """COMPLEX_EXP first-derivative factor 2*pi*i/lx (BasisFunctions.py:26-37,49-52)."""
    return 1j * 2 * np.pi / p.lx


def kz_factor(p):
    return 1j * 2 * np.pi / p.lz


def k2x_factor(p):
    """Second-derivative factor -(2*pi)^2/lx^2 (BasisFunctions.py:40-43,55-59)."""
    return -np.abs(1j * 2 * np.pi) ** 2 / p.lx ** 2


def k2z_factor(p):
    return -np.abs(1j * 2 * np.pi) ** 2 / p.lz ** 2


# --------------------------------------------------------------------------
# Spectral transforms (SpectralTransformer.py)
# --------------------------------------------------------------------------
def to_physical(p: Params, spec: np.ndarray) -> np.ndarray:
    """Spectral -> physical.

    Fully spectral (SpectralTransformer.py:90-150 with 21-24): zero-pad the
    compact spectrum into the rfft2 half-spectrum ``(nx, nz//2+1)``, scale
    by ``nx*nz`` and ``irfft2``.  Fourier-x/FD-z (33-61 with 26-31): pad
    rows ``[0, nn)`` into ``(nx//2+1, nz)``, scale by ``nx``, ``irfft``
    along x.
    """
    if p.fully_spectral:
        up = np.zeros((p.nx, p.nz // 2 + 1), dtype=np.complex128)
        up[: p.nn + 1, : p.nm] = spec[: p.nn + 1, : p.nm]
        up[-p.nn:, : p.nm] = spec[-p.nn:, : p.nm]
        up *= p.nx * p.nz
        return np.fft.irfft2(up)
    up = np.zeros((p.nx // 2 + 1, p.nz), dtype=np.complex128)
    up[: p.nn] = spec[: p.nn]
    return np.fft.irfft(up * p.nx, axis=0)


def to_spectral(p: Params, phys: np.ndarray) -> np.ndarray:
    """Physical -> spectral: rfft2 / (nx*nz) then truncate
    (SpectralTransformer.py:152-199); 1-D: rfft along x / nx, keep rows
    ``[0, nn)`` (63-88)."""
    out = np.zeros(p.spectral_shape, dtype=np.complex128)
    if p.fully_spectral:
        f = np.fft.rfft2(phys) / (p.nx * p.nz)
        out[: p.nn + 1, : p.nm] = f[: p.nn + 1, : p.nm]
        out[-p.nn:, : p.nm] = f[-p.nn:, : p.nm]
        return out
    f = np.fft.rfft(phys, axis=0) / p.nx
    out[: p.nn] = f[: p.nn]
    return out


def regrid_spectral(p: Params, spec: np.ndarray) -> np.ndarray:
    """Spectral coefficients of another grid size carried onto p's grid:
    the re-grid ``Variable.load`` and ``TimeDerivative.load`` apply to
    spectral arrays (Variable.py:76-85, ``scale_variable`` at
    Variable.py:140-151, TimeDerivative.py:56-62).  The coefficients are
    amplitudes (the forward transform divides by the point count), so each
    mode present on both grids keeps its value and the others are zero:

    * fully spectral ``(2 nn' + 1, nm')`` -> ``(2 nn + 1, nm)``: modes
      ``|kx| <= min(nn', nn)``, ``kz < min(nm', nm)``, each row mapped to
      the row of the same kx (positive rows from the top, negative from
      the bottom);
    * Fourier-x / FD-z ``(nn', nz)`` -> ``(nn, nz)``: rows ``kx <
      min(nn', nn)``; the z grid must match (a different FD grid needs
      interpolation, not a re-grid).

    Two behaviours of the reference are not restated: its output array is
    real (``xp.zeros`` defaults to float64), so the assignment keeps only
    the real parts; and it bounds the positive rows by ``min(2 nn' + 1, nn)``
    -- the input's row count over both signs of kx -- which, when nn' < nn,
    copies negative-kx rows onto positive ones.  The restatement keeps the
    complex values and maps every kx to itself.
    """
    spec = np.asarray(spec)
    out = np.zeros(p.spectral_shape, dtype=np.complex128)
    if p.fully_spectral:
        if spec.ndim != 2 or spec.shape[0] % 2 != 1:
            raise ValueError(f"fully spectral data has shape (2nn+1, nm), got {spec.shape}")
        k = min((spec.shape[0] - 1) // 2, p.nn)
        m = min(spec.shape[1], p.nm)
        out[: k + 1, :m] = spec[: k + 1, :m]
        if k > 0:
            out[-k:, :m] = spec[-k:, :m]
        return out
    if spec.ndim != 2 or spec.shape[1] != p.nz:
        raise ValueError(f"Fourier-x / FD-z data needs nz = {p.nz} columns, got {spec.shape}")
    k = min(spec.shape[0], p.nn)
    out[:k] = spec[:k]
    return out


# --------------------------------------------------------------------------
# Cosine and sine bases (SpectralTransformer.py:90-199, BasisFunctions.py)
# --------------------------------------------------------------------------
def _basis_period(b, n):
    """Length of the periodic signal a basis transforms along an axis of n
    samples: n for COMPLEX_EXP, the 2(n-1) of the mirror image otherwise."""
    return n if b == COMPLEX_EXP else 2 * (n - 1)


def _basis_norm(b, n):
    """Normalisation of a basis's coefficients (SpectralTransformer.py:104-115,
    167-180): n, n-1 (cosine) or -i(n-1) (sine)."""
    return {COMPLEX_EXP: n, COSINE: n - 1, SINE: -1j * (n - 1)}[b]


def _mirror(a, b, axis):
    """Extend ``a`` along ``axis`` to its basis period: samples 0..n-2, then
    n-1..1 reflected (negated for the sine basis)."""
    if b == COMPLEX_EXP:
        return a
    n = a.shape[axis]
    head = np.take(a, np.arange(0, n - 1), axis=axis)
    tail = np.take(a, np.arange(n - 1, 0, -1), axis=axis)
    return np.concatenate((head, -tail if b == SINE else tail), axis=axis)


def to_spectral_basis(p: Params, phys: np.ndarray, bases) -> np.ndarray:
    """Fully spectral forward transform with a basis per axis
    (SpectralTransformer.py:152-199): the signal mirrored to its basis
    period on each cosine / sine axis, ``rfft2``, divided by the two
    normalisations, the k = 0 cosine coefficients halved, then truncated as
    ``to_spectral``."""
    bx, bz = bases
    ext = _mirror(_mirror(np.asarray(phys, dtype=np.float64), bx, 0), bz, 1)
    f = np.fft.rfft2(ext) / (_basis_norm(bx, p.nx) * _basis_norm(bz, p.nz))
    if bx == COSINE:
        f[0, :] *= 0.5
    if bz == COSINE:
        f[:, 0] *= 0.5
    out = np.zeros(p.spectral_shape, dtype=np.complex128)
    out[: p.nn + 1] = f[: p.nn + 1, : p.nm]
    out[-p.nn:] = f[-p.nn:, : p.nm]
    return out


def to_physical_basis(p: Params, spec: np.ndarray, bases) -> np.ndarray:
    """Inverse of ``to_spectral_basis`` (SpectralTransformer.py:90-150):
    k = 0 cosine coefficients doubled, the compact spectrum placed in the
    half-spectrum of the basis periods, scaled by both normalisations,
    ``irfft2`` and cropped to the first nx x nz samples.  The input is not
    modified (the reference doubles its caller's k = 0 cosine rows and
    columns in place, SpectralTransformer.py:117-120)."""
    bx, bz = bases
    lx_, lz_ = _basis_period(bx, p.nx), _basis_period(bz, p.nz)
    s = np.array(spec, dtype=np.complex128)
    if bx == COSINE:
        s[0, :] *= 2
    if bz == COSINE:
        s[:, 0] *= 2
    up = np.zeros((lx_, lz_ // 2 + 1), dtype=np.complex128)
    up[: p.nn + 1, : p.nm] = s[: p.nn + 1]
    up[-p.nn:, : p.nm] = s[-p.nn:]
    up *= _basis_norm(bx, p.nx) * _basis_norm(bz, p.nz)
    return np.fft.irfft2(up, s=(lx_, lz_))[: p.nx, : p.nz]


def basis_diff_factor(b, length):
    """First-derivative factor of a basis over a domain length
    (BasisFunctions.py:26-37, 49-52): 2 pi i, pi (sine), -pi (cosine), 0
    (FDM), over the length."""
    w = {COMPLEX_EXP: 2j * np.pi, SINE: np.pi, COSINE: -np.pi}.get(b, 0.0)
    return w / length


def basis_diff2_factor(b, length):
    """Second-derivative factor -|w|^2 / length^2 (BasisFunctions.py:40-43,
    55-59)."""
    return -np.abs(basis_diff_factor(b, 1.0)) ** 2 / length ** 2


def spectral_op_basis(p: Params, kind: str, f, bases):
    """The spectral operators with bases (SpatialDifferentiator.py:50-74,
    Variable.py:111-113, LaplacianSolver.py:58-68): ``kind`` is one of
    ddx, ddz, d2dx2, d2dz2, snabla2, lap (``f`` ignored) and poisson (f
    divided by the Laplacian symbol, with 1 at mode (0, 0))."""
    bx, bz = bases
    n, m = mode_numbers(p)
    d2x = basis_diff2_factor(bx, p.lx) * n ** 2
    d2z = basis_diff2_factor(bz, p.lz) * m ** 2
    if kind == "ddx":
        return basis_diff_factor(bx, p.lx) * n * f
    if kind == "ddz":
        return basis_diff_factor(bz, p.lz) * m * f
    if kind == "d2dx2":
        return d2x * f
    if kind == "d2dz2":
        return d2z * f
    if kind == "snabla2":
        return d2x * f + d2z * f
    lap = d2x + d2z
    if kind == "lap":
        return lap
    if kind == "poisson":
        lap = lap.copy()
        lap[0, 0] = 1.0
        return f / lap
    raise ValueError(kind)


# --------------------------------------------------------------------------
# Spectral derivatives (SpatialDifferentiator.py:50-74)
# --------------------------------------------------------------------------
def sddx(p, f):
    n, _ = mode_numbers(p)
    return kx_factor(p) * n * f


def sddz(p, f):
    """Spectral d/dz; in FD-z the reference routes sddz to pddz
    (SpatialDifferentiator.py:35-38)."""
    if not p.fully_spectral:
        return pddz(p, f)
    _, m = mode_numbers(p)
    return kz_factor(p) * m * f


def sd2dx2(p, f):
    n, _ = mode_numbers(p)
    return k2x_factor(p) * n ** 2 * f


def sd2dz2(p, f):
    """In FD-z this is the physical-space FD second derivative applied to
    complex spectral rows (SpatialDifferentiator.py:36-40, 106-128)."""
    if not p.fully_spectral:
        return pd2dz2(p, f)
    _, m = mode_numbers(p)
    return k2z_factor(p) * m ** 2 * f


def calc_lap(p):
    """Real Laplacian symbol (SpatialDifferentiator.py:70-74)."""
    n, m = mode_numbers(p)
    return k2x_factor(p) * n ** 2 + k2z_factor(p) * m ** 2


def snabla2(p, f):
    """Variable.snabla2 (Variable.py:111-113)."""
    return sd2dx2(p, f) + sd2dz2(p, f)


# --------------------------------------------------------------------------
# Physical-space central finite differences (SpatialDifferentiator.py:76-185)
# --------------------------------------------------------------------------
def _x_periodic(p):
    return p.discretisation[0] == "spectral"


def _z_periodic(p):
    return p.discretisation[1] == "spectral"


def pddx(p, f):
    """d/dx by 2nd (76-89) or 4th (130-157) order central differences;
    periodic wrap when x is spectral, else the end rows stay zero."""
    out = np.zeros(f.shape, dtype=np.result_type(f, np.float64))
    dx = p.dx
    if p.spatial_derivative_order == 2:
        out[1:-1] = (f[2:] - f[:-2]) / (2 * dx)
        if _x_periodic(p):
            out[0] = (f[1] - f[-1]) / (2 * dx)
            out[-1] = (f[0] - f[-2]) / (2 * dx)
    else:
        out[2:-2] = (-0.25 * f[4:] + 2 * f[3:-1] - 2 * f[1:-3]
                     + 0.25 * f[:-4]) / (3 * dx)
        if _x_periodic(p):
            fp = np.concatenate((f[-2:], f, f[:2]), axis=0)
            for i in (0, 1, f.shape[0] - 2, f.shape[0] - 1):
                j = i + 2
                out[i] = (-0.25 * fp[j + 2] + 2 * fp[j + 1] - 2 * fp[j - 1]
                          + 0.25 * fp[j - 2]) / (3 * dx)
    return out


def pddz(p, f):
    """d/dz, as pddx along axis 1 (SpatialDifferentiator.py:91-104,159-185)."""
    out = np.zeros(f.shape, dtype=np.result_type(f, np.float64))
    dz = p.dz
    if p.spatial_derivative_order == 2:
        out[:, 1:-1] = (f[:, 2:] - f[:, :-2]) / (2 * dz)
        if _z_periodic(p):
            out[:, 0] = (f[:, 1] - f[:, -1]) / (2 * dz)
            out[:, -1] = (f[:, 0] - f[:, -2]) / (2 * dz)
    else:
        out[:, 2:-2] = (-0.25 * f[:, 4:] + 2 * f[:, 3:-1] - 2 * f[:, 1:-3]
                        + 0.25 * f[:, :-4]) / (3 * dz)
        if _z_periodic(p):
            fp = np.concatenate((f[:, -2:], f, f[:, :2]), axis=1)
            for i in (0, 1, f.shape[1] - 2, f.shape[1] - 1):
                j = i + 2
                out[:, i] = (-0.25 * fp[:, j + 2] + 2 * fp[:, j + 1]
                             - 2 * fp[:, j - 1] + 0.25 * fp[:, j - 2]) / (3 * dz)
    return out


def pd2dz2(p, f):
    """d2/dz2 central differences on interior points only; the end columns
    stay zero (SpatialDifferentiator.py:106-128)."""
    out = np.zeros_like(f)
    dz = p.dz
    if p.spatial_derivative_order == 2:
        out[:, 1:-1] = (f[:, 2:] - 2 * f[:, 1:-1] + f[:, :-2]) / dz ** 2
    else:
        out[:, 2:-2] = (-1.0 / 12 * f[:, 4:] + 4.0 / 3 * f[:, 3:-1]
                        - 5.0 / 2 * f[:, 2:-2] + 4.0 / 3 * f[:, 1:-3]
                        - 1.0 / 12 * f[:, :-4]) / dz ** 2
    return out


# --------------------------------------------------------------------------
# Advection (Variable.py:119-128)
# --------------------------------------------------------------------------
def flux_divergence(p, ux, uz, fphys):
    """pddx(ux*f) + pddz(uz*f) in physical space (Variable.py:125-127)."""
    return pddx(p, ux * fphys) + pddz(p, uz * fphys)


def vec_dot_nabla(p, fspec, ux, uz):
    """Transform f to physical, FD flux divergence, transform back
    (Variable.py:119-128)."""
    return to_spectral(p, flux_divergence(p, ux, uz, to_physical(p, fspec)))


# --------------------------------------------------------------------------
# Poisson / Laplacian solve (LaplacianSolver.py)
# --------------------------------------------------------------------------
def fdm_laplacian_bands(p, ddx_abs=None):
    """Per-mode tridiagonal bands of the FD-z Laplacian with identity
    Dirichlet rows (LaplacianSolver.py:21-52).  Returns (lower, diag,
    upper), each (nn, nz); lower[:,0] and upper[:,-1] are unused."""
    if ddx_abs is None:
        ddx_abs = 2 * np.pi / p.lx
    nn, nz, dz = p.nn, p.nz, p.dz
    n = np.arange(nn)[:, None]
    diag = -((n * ddx_abs) ** 2 + 2.0 / dz ** 2) * np.ones((nn, nz))
    lower = np.full((nn, nz), 1.0 / dz ** 2)
    upper = np.full((nn, nz), 1.0 / dz ** 2)
    diag[:, 0] = 1.0
    upper[:, 0] = 0.0
    diag[:, -1] = 1.0
    lower[:, -1] = 0.0
    return lower, diag, upper


def thomas(lower, diag, upper, rhs):
    """Row-batched Thomas algorithm: solves sum of the tridiagonal bands
    times x = rhs along axis 1 for every row."""
    nz = rhs.shape[1]
    c = np.zeros(rhs.shape, dtype=np.float64)
    d = np.zeros(rhs.shape, dtype=np.result_type(rhs, np.complex128))
    c[:, 0] = upper[:, 0] / diag[:, 0]
    d[:, 0] = rhs[:, 0] / diag[:, 0]
    for j in range(1, nz):
        den = diag[:, j] - lower[:, j] * c[:, j - 1]
        c[:, j] = upper[:, j] / den
        d[:, j] = (rhs[:, j] - lower[:, j] * d[:, j - 1]) / den
    x = np.zeros_like(d)
    x[:, -1] = d[:, -1]
    for j in range(nz - 2, -1, -1):
        x[:, j] = d[:, j] - c[:, j] * x[:, j + 1]
    return x


def poisson_solve(p, rhs):
    """Solve lap(psi) = rhs.  Fully spectral: ``rhs / lap`` with
    ``lap[0,0]`` taken as 1 (LaplacianSolver.py:58-68).  FD-z: per-mode
    tridiagonal solve with identity wall rows (LaplacianSolver.py:70-78)."""
    if p.fully_spectral:
        lap = calc_lap(p).astype(np.float64)
        lap[0, 0] = 1.0
        return rhs / lap
    lo, di, up = fdm_laplacian_bands(p)
    return thomas(lo, di, up, rhs)


# --------------------------------------------------------------------------
# Time integration (Integrator.py, TimeDerivative.py)
# --------------------------------------------------------------------------
AB_COEFFS = {2: (3.0, -1.0), 4: (55.0, -59.0, 37.0, -9.0)}
AB_DENOM = {2: 2.0, 4: 24.0}
AM_COEFFS = {2: (1.0, 1.0), 4: (9.0, 19.0, -5.0, 1.0)}


class History:
    """Zero-initialised ring of RHS values (TimeDerivative.py:9-45)."""

    def __init__(self, q, shape):
        self.q = q
        self.data = np.zeros((q,) + tuple(shape), dtype=np.complex128)
        self.curr = 0

    def get(self, idx=0):
        return self.data[(self.curr + idx) % self.q]

    def set(self, value):
        self.data[self.curr] = value

    def advance(self):
        self.curr = (self.curr + 1) % self.q


def adams_bashforth(hist: History, dt):
    """AB2 / AB4 increment (Integrator.py:5-18)."""
    q = hist.q
    if q == 2:
        return dt / 2 * (3 * hist.get() - hist.get(-1))
    return dt / 24 * (55 * hist.get() - 59 * hist.get(-1)
                      + 37 * hist.get(-2) - 9 * hist.get(-3))


def adams_moulton(hist: History, dt):
    """AM2 / AM4 combinations (Integrator.py:20-33); defined but never called
    by the reference's examples.  Parity unpinned beyond polynomial
    exactness."""
    if hist.q == 2:
        return dt / 2 * (hist.get() + hist.get(-1))
    return dt / 24 * (9 * hist.get() + 19 * hist.get(-1)
                      - 5 * hist.get(-2) + hist.get(-3))


def integrate_semi_implicit(var, hist: History, lin_op, dt, alpha):
    """theta-scheme diffusion + AB explicit part, then advance
    (Integrator.py:58-63)."""
    rhs = (1 + (1 - alpha) * dt * lin_op) * var + adams_bashforth(hist, dt)
    out = rhs / (1 - alpha * dt * lin_op)
    hist.advance()
    return out


def integrate_explicit(var, hist: History, diffusion, dt):
    """Diffusion added into the stored RHS, AB update, advance
    (Integrator.py:53-56)."""
    hist.data[hist.curr] += diffusion
    out = var + adams_bashforth(hist, dt)
    hist.advance()
    return out


def cfl_update_dt(p, dt, ux, uz):
    """CFL control (Integrator.py:35-44): signed maxima, raise on breach or
    NaN, else shrink dt by 0.9 until dt <= cutoff*cfl_dt."""
    cfl_dt = min(p.dx / np.max(ux), p.dz / np.max(uz))
    if dt > cfl_dt or np.isnan(cfl_dt):
        raise Exception("CFL condition breached")
    while dt > p.cfl_cutoff * cfl_dt:
        dt = dt * 0.9
    return dt


# --------------------------------------------------------------------------
# Velocity and diagnostics (utility.py:42-79)
# --------------------------------------------------------------------------
def velocity_from_vorticity(p, w_spec):
    """psi = solve(-w); ux, uz physical (utility.py:62-79).  Returns
    (psi_spec, ux_phys, uz_phys, uz_spec, psi_phys).  In FD-z ux comes from
    the FD z-derivative of physical psi, and psi_phys is that physical psi;
    fully spectral returns psi_phys None."""
    psi = poisson_solve(p, -w_spec)
    psi_phys = None
    if not p.fully_spectral:
        psi_phys = to_physical(p, psi)
        ux = -pddz(p, psi_phys)
    else:
        ux = to_physical(p, -sddz(p, psi))
    uz_spec = sddx(p, psi)
    uz = to_physical(p, uz_spec)
    return psi, ux, uz, uz_spec, psi_phys


def kinetic_energy(p, ux, uz):
    """0.5 * sum(ux^2 + uz^2) / (nx*nz) (utility.py:42-59)."""
    return 0.5 * np.sum(uz ** 2 + ux ** 2) / (p.nx * p.nz)


def nusselt(tphys, uz):
    """1 - mean(T*uz) (double_diffusive_convection.py:20-23)."""
    return 1.0 - np.mean(tphys * uz)


# --------------------------------------------------------------------------
# Example loop bodies (one step each).  ``st`` is a dict of spectral
# fields plus histories; every function returns the physical velocities
# used in the step (for the CFL / tracker logic of Simulation.end_loop).
# --------------------------------------------------------------------------
def step_vorticity(p, st, coeff, scalars=()):
    """Taylor-Green / vortex-pair / Kelvin-Helmholtz body
    (taylor_green_vortex.py:87-93, vortex_pairs.py:122-134,
    kelvin_helmholtz_instability.py:117-129 with the ink block enabled):
    every passively advected field in ``scalars`` uses the same velocity
    and the same diffusion coefficient."""
    psi, ux, uz, uz_spec, _ = velocity_from_vorticity(p, st["w"])
    st["psi"] = psi
    for name in ("w",) + tuple(scalars):
        lin = coeff * calc_lap(p)
        st["d" + name].set(-vec_dot_nabla(p, st[name], ux, uz))
        st[name] = integrate_semi_implicit(st[name], st["d" + name], lin,
                                           st["dt"], p.alpha)
    return ux, uz


def step_ddc(p, st):
    """Double-diffusive convection body (double_diffusive_convection.py:
    102-124): omega, T, S; every RHS uses the state at the start of the step
    (omega is updated first, and dT, dS only read uz spectral from before the
    update); then the kz=0 column of T and S is zeroed."""
    Pr, R0, tau = p.Pr, p.R0, p.tau
    psi, ux, uz, uz_spec, _ = velocity_from_vorticity(p, st["w"])
    st["psi"] = psi
    lap = calc_lap(p)
    T0, S0 = st["tmp"].copy(), st["xi"].copy()
    st["dw"].set(-vec_dot_nabla(p, st["w"], ux, uz) + Pr * sddx(p, S0)
                 - Pr * sddx(p, T0))
    st["w"] = integrate_semi_implicit(st["w"], st["dw"], Pr * lap, st["dt"],
                                      p.alpha)
    st["dtmp"].set(-vec_dot_nabla(p, T0, ux, uz) - uz_spec)
    st["tmp"] = integrate_semi_implicit(T0, st["dtmp"], lap, st["dt"],
                                        p.alpha)
    st["dxi"].set(-vec_dot_nabla(p, S0, ux, uz) - uz_spec / R0)
    st["xi"] = integrate_semi_implicit(S0, st["dxi"], tau * lap, st["dt"],
                                       p.alpha)
    st["tmp"][:, 0] = 0.0
    st["xi"][:, 0] = 0.0
    return ux, uz


def step_mhd(p, st):
    """Resistive-tearing MHD body (resistive_tearing_instability.py:
    127-146): flux function from j by the same Poisson solve; dw from the
    old w and j; dj uses the *updated* w in its w.vec_dot_nabla(b) term."""
    psi, ux, uz, _, _ = velocity_from_vorticity(p, st["w"])
    phi, bx, bz, _, _ = velocity_from_vorticity(p, st["j"])
    st["psi"], st["phi"] = psi, phi
    lap = calc_lap(p)
    st["dw"].set(-vec_dot_nabla(p, st["w"], ux, uz)
                 + vec_dot_nabla(p, st["j"], bx, bz))
    st["w"] = integrate_semi_implicit(st["w"], st["dw"], lap / p.Re,
                                      st["dt"], p.alpha)
    st["dj"].set(-vec_dot_nabla(p, st["j"], ux, uz)
                 + vec_dot_nabla(p, st["w"], bx, bz))
    st["j"] = integrate_semi_implicit(st["j"], st["dj"], lap / p.S,
                                      st["dt"], p.alpha)
    return ux, uz


def step_rbc(p, st):
    """Rayleigh-Benard body, Fourier-x / FD-z, explicit
    (rayleigh_benard_convection.py:97-142): AB with the diffusion stored
    into the history, then the spectral Dirichlet writes."""
    Pr, Ra = p.Pr, p.Ra
    psi, ux, uz, uz_spec, psi_phys = velocity_from_vorticity(p, st["w"])
    st["psi"] = psi
    T0 = st["tmp"].copy()
    diff = Pr * snabla2(p, st["w"])
    st["dw"].set(-vec_dot_nabla(p, st["w"], ux, uz) - Pr * Ra * sddx(p, T0))
    st["w"] = integrate_explicit(st["w"], st["dw"], diff, st["dt"])
    diff = snabla2(p, T0)
    st["dtmp"].set(-vec_dot_nabla(p, T0, ux, uz))
    st["tmp"] = integrate_explicit(T0, st["dtmp"], diff, st["dt"])
    w, t, ps = st["w"], st["tmp"], st["psi"]
    if p.spatial_derivative_order == 2:
        w[1:, 0] = 0.0
        w[1:, -1] = 0.0
        ps[1:, 0] = 0.0
        ps[1:, -1] = 0.0
        t[0, 0] = 1.0
        t[0, -1] = 0.0
        t[1:, 0] = 0.0
        t[1:, -1] = 0.0
    else:
        w[1:, :2] = 0.0
        w[1:, -2:] = 0.0
        ps[1:, :2] = 0.0
        ps[1:, -2:] = 0.0
        t[0, :2] = 1.0
        t[0, -2:] = 0.0
        t[1:, :2] = 0.0
        t[1:, -2:] = 0.0
    ps[0, :] = 0.0
    w[0, :] = 0.0
    return ux, uz


def fdz_implicit_bands(p, coeff, dt, alpha):
    """Bands of I - alpha*dt*coeff*L_n for every x mode n, where L_n is the
    FD-z Laplacian -kx^2 + D2z on interior points; the wall rows are
    identity.  Extension of this build: the reference has no semi-implicit
    FD-z scheme (Integrator.py:78-83).  Parity unpinned against the
    reference; pinned by a dense solve and the explicit limit in tests."""
    nn, nz, dz = p.nn, p.nz, p.dz
    n = np.arange(nn)[:, None]
    k2 = (n * 2 * np.pi / p.lx) ** 2
    a = alpha * dt * coeff
    diag = (1.0 + a * (k2 + 2.0 / dz ** 2)) * np.ones((nn, nz))
    lower = np.full((nn, nz), -a / dz ** 2)
    upper = np.full((nn, nz), -a / dz ** 2)
    diag[:, 0] = diag[:, -1] = 1.0
    upper[:, 0] = lower[:, -1] = 0.0
    return lower, diag, upper


def fdz_laplacian_apply(p, f):
    """L_n f = -kx^2 f + D2z f on interior points, zero on the walls (the
    operator of snabla2 in FD-z, SpatialDifferentiator.py:36-40, order 2)."""
    n = np.arange(p.nn)[:, None]
    out = -((n * 2 * np.pi / p.lx) ** 2) * f
    out[:, 0] = 0.0
    out[:, -1] = 0.0
    out[:, 1:-1] += (f[:, 2:] - 2 * f[:, 1:-1] + f[:, :-2]) / p.dz ** 2
    return out


def integrate_fdz_semi_implicit(p, var, hist: History, coeff, dt, alpha):
    """theta-scheme diffusion for FD-z (order-2 z differences): solve
    (I - a dt c L) v_new = (I + (1-a) dt c L) v + AB(hist) mode by mode,
    identity wall rows, then advance the ring."""
    rhs = var + (1 - alpha) * dt * coeff * fdz_laplacian_apply(p, var) + adams_bashforth(hist, dt)
    lo, di, up = fdz_implicit_bands(p, coeff, dt, alpha)
    out = thomas(lo, di, up, rhs)
    hist.advance()
    return out


def _rbc_walls(p, st):
    w, t, ps = st["w"], st["tmp"], st["psi"]
    k = 1 if p.spatial_derivative_order == 2 else 2
    w[1:, :k] = 0.0
    w[1:, -k:] = 0.0
    ps[1:, :k] = 0.0
    ps[1:, -k:] = 0.0
    t[0, :k] = 1.0
    t[0, -k:] = 0.0
    t[1:, :k] = 0.0
    t[1:, -k:] = 0.0
    ps[0, :] = 0.0
    w[0, :] = 0.0


def step_rbc_semi_implicit(p, st):
    """Rayleigh-Benard with theta-implicit diffusion (this build's extension
    for stiff FD-z diffusion): the advection and buoyancy terms of
    rayleigh_benard_convection.py:101-110 in the AB ring, diffusion by
    integrate_fdz_semi_implicit, then the same wall writes (113-142)."""
    Pr, Ra = p.Pr, p.Ra
    psi, ux, uz, uz_spec, psi_phys = velocity_from_vorticity(p, st["w"])
    st["psi"] = psi
    T0 = st["tmp"].copy()
    st["dw"].set(-vec_dot_nabla(p, st["w"], ux, uz) - Pr * Ra * sddx(p, T0))
    st["w"] = integrate_fdz_semi_implicit(p, st["w"], st["dw"], Pr, st["dt"], p.alpha)
    st["dtmp"].set(-vec_dot_nabla(p, T0, ux, uz))
    st["tmp"] = integrate_fdz_semi_implicit(p, T0, st["dtmp"], 1.0, st["dt"], p.alpha)
    _rbc_walls(p, st)
    return ux, uz


def _rbc_rhs(p, w, t):
    """Explicit RHS of the Rayleigh-Benard equations at (w, t): advection and
    buoyancy (rayleigh_benard_convection.py:101-110).  Returns the RHS pair,
    the streamfunction and the physical velocities."""
    psi, ux, uz, _, _ = velocity_from_vorticity(p, w)
    rw = -vec_dot_nabla(p, w, ux, uz) - p.Pr * p.Ra * sddx(p, t)
    rt = -vec_dot_nabla(p, t, ux, uz)
    return rw, rt, psi, ux, uz


def _fdz_theta_solve(p, var, incr, coeff, dt, alpha):
    """theta-scheme FD-z diffusion with a given explicit increment ``incr``:
    solve (I - a dt c L) v_new = (I + (1-a) dt c L) var + incr."""
    rhs = var + (1 - alpha) * dt * coeff * fdz_laplacian_apply(p, var) + incr
    lo, di, up = fdz_implicit_bands(p, coeff, dt, alpha)
    return thomas(lo, di, up, rhs)


def step_rbc_pc(p, st):
    """Rayleigh-Benard with an Adams-Bashforth predictor and Adams-Moulton
    corrector (PECE) for the explicit terms and theta-implicit FD-z diffusion
    in both stages.  Extension of this build: the reference defines the
    AB/AM formulas (Integrator.py:5-33) but never calls the corrector.

    P: N_n = RHS(v_n) into the ring slot curr; v* = theta-solve with the AB
       increment over slots curr, curr-1, ...; wall writes on v*.
    E: N* = RHS(v*) into slot curr+1 (the oldest slot, no longer needed).
    C: v_{n+1} = theta-solve from v_n with the AM increment over slots
       curr+1, curr, ... ; wall writes; the ring ends advanced once.
    """
    Pr = p.Pr
    dt, a = st["dt"], p.alpha
    w0, t0 = st["w"].copy(), st["tmp"].copy()
    rw, rt, psi, ux, uz = _rbc_rhs(p, w0, t0)
    st["psi"] = psi
    hw, ht = st["dw"], st["dtmp"]
    hw.set(rw)
    ht.set(rt)
    ws = _fdz_theta_solve(p, w0, adams_bashforth(hw, dt), Pr, dt, a)
    ts = _fdz_theta_solve(p, t0, adams_bashforth(ht, dt), 1.0, dt, a)
    star = {"w": ws, "tmp": ts, "psi": np.zeros_like(ws)}
    _rbc_walls(p, star)
    rw2, rt2, _, _, _ = _rbc_rhs(p, star["w"], star["tmp"])
    hw.advance()
    ht.advance()
    hw.set(rw2)
    ht.set(rt2)
    st["w"] = _fdz_theta_solve(p, w0, adams_moulton(hw, dt), Pr, dt, a)
    st["tmp"] = _fdz_theta_solve(p, t0, adams_moulton(ht, dt), 1.0, dt, a)
    _rbc_walls(p, st)
    return ux, uz


PROBLEMS = {
    "vorticity": (("w",), step_vorticity),
    "ddc": (("w", "tmp", "xi"), step_ddc),
    "mhd": (("w", "j"), step_mhd),
    "rbc": (("w", "tmp"), step_rbc),
    "rbc_si": (("w", "tmp"), step_rbc_semi_implicit),
    "rbc_pc": (("w", "tmp"), step_rbc_pc),
}


def new_state(p, fields, dt=None):
    st = {"dt": p.initial_dt if dt is None else dt}
    for f in fields:
        st[f] = np.zeros(p.spectral_shape, dtype=np.complex128)
        st["d" + f] = History(p.integrator_order, p.spectral_shape)
    return st


def run(p, st, body, nsteps, cfl=True, on_step=None):
    """Drive ``body`` for ``nsteps`` loops with the Simulation.end_loop
    bookkeeping (Simulation.py:99-109,199-209; Ticker.py:14-25): t += dt
    after each step, the CFL ticker fires at loops 1, 1+cadence, ... and
    may shrink dt for the following steps.  Returns (t, loops, dt_trace)."""
    t = 0.0
    counter = 0
    dts = []
    for loop in range(1, nsteps + 1):
        ux, uz = body(p, st)
        dts.append(st["dt"])
        t += st["dt"]
        if cfl and counter < loop:
            st["dt"] = cfl_update_dt(p, st["dt"], ux, uz)
            counter += p.cfl_cadence
        if on_step is not None:
            on_step(loop, st, ux, uz)
    st["t"] = t
    return t, dts


# --------------------------------------------------------------------------
# Initial conditions
# --------------------------------------------------------------------------
def grid(p, z_endpoint=False):
    """Cell-origin grid as in every example (e.g. taylor_green_vortex.py:
    23-25): linspace(0, L, N, endpoint=False), 'ij' meshgrid."""
    x = np.linspace(0, p.lx, p.nx, endpoint=False)
    z = np.linspace(0, p.lz, p.nz, endpoint=z_endpoint)
    return np.meshgrid(x, z, indexing="ij")


def noise(shape, epsilon, seed=0):
    """epsilon*(2U-1) from default_rng(seed) (utility.py:31-39)."""
    rng = np.random.default_rng(seed)
    return epsilon * (2 * rng.random(shape) - 1.0)


def ic_taylor_green(p, seed=0, amp=1e-6, kmax=8):
    """Taylor-Green 2 sin x sin z plus a seeded Gaussian perturbation of
    amplitude ``amp`` on the modes 0 < |k| <= kmax, Hermitian on the kz=0
    column (so the physical field is real).  Returns spectral omega."""
    X, Z = grid(p)
    w = to_spectral(p, 2 * np.sin(X) * np.sin(Z))
    rng = np.random.default_rng(seed)
    pert = np.zeros(p.spectral_shape, dtype=np.complex128)
    n, m = mode_numbers(p)
    sel = (n ** 2 + m ** 2 <= kmax ** 2) & ~((n == 0) & (m == 0))
    vals = rng.standard_normal(p.spectral_shape) + 1j * rng.standard_normal(
        p.spectral_shape)
    pert[sel] = amp * vals[sel]
    # kz = 0 column: make pert[-n,0] = conj(pert[n,0]).
    col = pert[:, 0].copy()
    for r in range(1, p.nn + 1):
        col[-r] = np.conj(col[r])
    col[0] = col[0].real
    pert[:, 0] = col
    return w + pert


def ic_kelvin_helmholtz(p, seed=0, amp=1e-3, delta=0.05):
    """Double shear layer u = tanh((z-0.5)/d) - tanh((z-1.5)/d); omega =
    du/dz in the reference's convention (omega = -lap psi, ux = -dpsi/dz);
    a passive scalar equal to u; uniform omega noise of amplitude ``amp``.
    Returns (omega_phys, scalar_phys)."""
    X, Z = grid(p)
    u = np.tanh((Z - 0.5) / delta) - np.tanh((Z - 1.5) / delta)
    w = (1.0 / np.cosh((Z - 0.5) / delta) ** 2
         - 1.0 / np.cosh((Z - 1.5) / delta) ** 2) / delta
    w = w + noise((p.nx, p.nz), amp, seed)
    return w, u.copy()
