"""Initial conditions of the benchmark workloads, generated on the host in
FP64 with numpy (inputs of the path, not part of it).

Follows the reference's IC generators where one exists:
  * vortex_ring, plane 'xy' with scal = 1, no disturbance (src/ic.f90:352-406)
    and the Pade core get_rr (src/ic.f90:888-901);
  * fermi, the Thomas-Fermi profile (src/ic.f90:273-305);
  * vortex_line along z without disturbance (src/ic.f90:309-348, get_r and
    get_theta 728-884) with amp(r) = 1 - exp(-0.7 r^1.15) (313-324).
The random-phase field is the workload's own construction (an inverse FFT of
a Gaussian shell spectrum with uniform random phases); the reference's
random_phase IC uses FFTW2 and is not reproduced.
"""
from __future__ import annotations

import os

import numpy as np

from .params import RunParams


def _grid(p: RunParams):
    x, y, z = p.axes()
    Z, Y, X = np.meshgrid(z, y, x, indexing="ij")
    return X, Y, Z


def get_rr(r: np.ndarray) -> np.ndarray:
    """R in psi = R(r) r e^{i theta}: sqrt((0.3437 + 0.0286 r^2)/(1 + 0.3333 r^2 + 0.0286 r^4))."""
    r2 = r * r
    return np.sqrt((0.3437 + 0.0286 * r2) / (1.0 + 0.3333 * r2 + 0.0286 * r2 * r2))


def vortex_ring_xy(p: RunParams, x0: float, y0: float, z0: float, r0: float, direction: float) -> np.ndarray:
    """Vortex ring of radius r0 in the plane z = z0 centred at (x0, y0)."""
    X, Y, Z = _grid(p)
    s = np.sqrt((X - x0) ** 2 + (Y - y0) ** 2)
    pp = Z - z0
    d1 = np.sqrt(pp ** 2 + (s + r0) ** 2)
    d2 = np.sqrt(pp ** 2 + (s - r0) ** 2)
    return (get_rr(d1) * (pp + direction * 1j * (s + r0))) * (get_rr(d2) * (pp - direction * 1j * (s - r0)))


def fermi(p: RunParams) -> np.ndarray:
    """Thomas-Fermi profile sqrt(max(0, (mu - V)/g)), V = (omx^2 x^2 + omy^2 y^2 + omz^2 z^2)/2."""
    X, Y, Z = _grid(p)
    V = 0.5 * ((p.omx ** 2) * X ** 2 + (p.omy ** 2) * Y ** 2 + (p.omz ** 2) * Z ** 2)
    f = (p.mu - V) / p.g
    return np.where(f >= 0.0, np.sqrt(np.maximum(f, 0.0)), 0.0).astype(np.complex128)


def vortex_line_z(p: RunParams, x0: float = 0.0, y0: float = 0.0, sgn: float = 1.0) -> np.ndarray:
    """Straight singly quantised vortex line along z through (x0, y0)."""
    X, Y, _ = _grid(p)
    r = np.sqrt((X - x0) ** 2 + (Y - y0) ** 2)
    theta = sgn * np.arctan2(Y - y0, X - x0)
    return (1.0 - np.exp(-0.7 * r ** 1.15)) * np.exp(1j * theta)


def complex_noise(shape, amp: float, seed: int) -> np.ndarray:
    rng = np.random.default_rng(seed)
    return amp * (rng.standard_normal(shape) + 1j * rng.standard_normal(shape))


def random_phase(p: RunParams, k0: float = 1.0, sigma: float = 0.2, seed: int = 0) -> np.ndarray:
    """Inverse 3D FFT of amplitudes exp(-(|k|-k0)^2/(2 sigma^2)) with uniform
    random phases, k = 2 pi m/(n dx), rescaled to mean |psi|^2 = 1."""
    rng = np.random.default_rng(seed)
    kx = 2 * np.pi * np.fft.fftfreq(p.nx, d=p.dx)
    ky = 2 * np.pi * np.fft.fftfreq(p.ny, d=p.dy)
    kz = 2 * np.pi * np.fft.fftfreq(p.nz, d=p.dz)
    KZ, KY, KX = np.meshgrid(kz, ky, kx, indexing="ij")
    k = np.sqrt(KX ** 2 + KY ** 2 + KZ ** 2)
    amp = np.exp(-((k - k0) ** 2) / (2 * sigma ** 2))
    phase = rng.uniform(0.0, 2 * np.pi, size=amp.shape)
    spec = amp * np.exp(1j * phase)
    del k, amp, phase, KX, KY, KZ
    try:
        import scipy.fft as sfft

        psi = sfft.ifftn(spec, workers=min(16, os.cpu_count() or 1))
    except ImportError:
        psi = np.fft.ifftn(spec)
    del spec
    psi *= 1.0 / np.sqrt(np.mean(np.abs(psi) ** 2))
    return np.ascontiguousarray(psi.astype(np.complex128))


# ---------------------------------------------------------------------------
# The benchmark workloads (grid, equation and IC of each).

def config_ring64(seed: int = 0):
    """64^3 periodic, dx = 0.5, 2nd order, eqn 1, RK4 dt = 0.01 x 200 steps.
    psi = 1 x Pade vortex ring of radius 8 in the x-y plane at the centre,
    plus complex Gaussian noise of amplitude 1e-4."""
    from .params import xr_for_spacing

    n = 64
    xr = xr_for_spacing(n, 0.5)
    p = RunParams(nx=n, ny=n, nz=n, xr=xr, yr=xr, zr=xr, tau=0.01, end_time=199.5 * 0.01, scheme="rk4",
                  eqn_to_solve=1, bcs=1, order=2)
    psi = vortex_ring_xy(p, 0.0, 0.0, 0.0, 8.0, 1.0) + complex_noise(p.shape, 1e-4, seed)
    return p, np.ascontiguousarray(psi)


def config_trap256(storage: str = "c128", n: int = 256):
    """256^3 reflective, dx = 0.2, 4th order, eqn 4 (g = 1, mu = 10,
    omega = 0.05), RK4 dt = 0.002 x 2000.  IC: Thomas-Fermi x vortex line."""
    from .params import xr_for_spacing

    xr = xr_for_spacing(n, 0.2)
    p = RunParams(nx=n, ny=n, nz=n, xr=xr, yr=xr, zr=xr, tau=0.002, end_time=1999.5 * 0.002, scheme="rk4",
                  eqn_to_solve=4, bcs=2, order=4, g=1.0, mu=10.0, omx=0.05, omy=0.05, omz=0.05, storage=storage)
    psi = fermi(p) * vortex_line_z(p)
    return p, np.ascontiguousarray(psi)


def config_turb(n: int = 512, storage: str = "c128", seed: int = 0, scheme: str = "rk45", tau: float = 0.01):
    """Random-phase turbulence: n^3 periodic, dx = 0.3, 4th order, eqn 1."""
    from .params import xr_for_spacing

    xr = xr_for_spacing(n, 0.3)
    p = RunParams(nx=n, ny=n, nz=n, xr=xr, yr=xr, zr=xr, tau=tau, end_time=20.0, scheme=scheme, eqn_to_solve=1,
                  bcs=1, order=4, storage=storage)
    return p, random_phase(p, seed=seed)


def config_two_rings(n: int = 384, dx: float = 0.5, storage: str = "c128"):
    """Two coaxial x-y rings of radius 6 and 10 at z = -10 and z = +10,
    oriented to move towards each other; n^3 periodic, 4th order, RKF45."""
    from .params import xr_for_spacing

    xr = xr_for_spacing(n, dx)
    p = RunParams(nx=n, ny=n, nz=n, xr=xr, yr=xr, zr=xr, tau=0.01, end_time=40.0, scheme="rk45", eqn_to_solve=1,
                  bcs=1, order=4, storage=storage)
    psi = vortex_ring_xy(p, 0.0, 0.0, -10.0, 6.0, 1.0) * vortex_ring_xy(p, 0.0, 0.0, 10.0, 10.0, -1.0)
    return p, np.ascontiguousarray(psi)
