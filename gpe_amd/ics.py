"""Initial conditions, generated on the host in FP64 with numpy (inputs of
the path, not part of it).

The reference's IC generators (src/ic.f90), restated with the same
parameters and operation order; the reference composes them in ic.in as
`init_cond = <expression>` and these compose the same way as numpy arrays
(z, y, x):
  * vortex_ring, planes 'xy' / 'xz' / 'yz' with the planar (amp, mm) and
    helical (r1, kk) disturbances (src/ic.f90:352-509), core get_rr
    (888-901);
  * vortex_line along 'x' / 'y' / 'z' with the sinusoidal displacement
    (a1, a2, ll), sign and phase-only imprint (327-348, get_r 730-795,
    get_theta 819-884, amp 309-323);
  * fermi, the Thomas-Fermi profile (273-305), and Vtrap (708-726);
  * pade_pulse_ring 'ring' / 'pulse' (513-582), vortex_pair (586-631);
  * the masks sphere, sphere2 and wall (635-704);
  * random_phase_ref, the reference's random-phase state (190-269): modes
    inside the cutoff kc2 with amplitude comp_amp and random phases, then the
    normalised forward transform of `fft` (905-1001).
The benchmark workloads' random-phase field (random_phase) is their own
construction: an inverse FFT of a Gaussian shell spectrum with uniform
random phases.
"""
from __future__ import annotations

import dataclasses
import os

import numpy as np

from .params import RunParams

PI = 3.1415926535897932384626433832795  # src/parameters.f90:122


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


@dataclasses.dataclass(frozen=True)
class RingParam:
    """ring_param (parameters.in of the reference): plane of the ring, centre,
    radius r0, planar disturbance amp cos(mm theta), helical disturbance of
    radius r1 and wavenumber kk, propagation direction dir (+1 / -1)."""
    plane: str = "xy"
    x0: float = 0.0
    y0: float = 0.0
    z0: float = 0.0
    r0: float = 8.0
    amp: float = 0.0
    mm: int = 0
    r1: float = 0.0
    kk: int = 0
    dir: float = 1.0


@dataclasses.dataclass(frozen=True)
class LineParam:
    """line_param (parameters.in of the reference): a point on the line, the
    displacement amplitudes a1 (sin) and a2 (cos) of wavelength ll, the sign
    of the winding, the direction of the line and whether only the phase is
    imprinted (no core)."""
    x0: float = 0.0
    y0: float = 0.0
    z0: float = 0.0
    amp1: float = 0.0
    amp2: float = 0.0
    ll: float = 0.0
    sgn: float = 1.0
    dir: str = "z"
    imprint_phase: bool = False


def vtrap(p: RunParams) -> np.ndarray:
    """Vtrap (src/ic.f90:708-726): (omx^2 x^2 + omy^2 y^2 + omz^2 z^2) / 2."""
    X, Y, Z = _grid(p)
    return 0.5 * ((p.omx ** 2) * X ** 2 + (p.omy ** 2) * Y ** 2 + (p.omz ** 2) * Z ** 2)


def vortex_ring(p: RunParams, vr: RingParam, scal: float = 1.0) -> np.ndarray:
    """vortex_ring (src/ic.f90:352-509).  With (a, b, c) the in-plane axes
    and n the normal axis of vr.plane: theta = atan2(b - b0, a - a0),
    s = |(a, b) - (a0, b0)| - r1 sin(kk theta), pp = n - r1 cos(kk theta),
    dist = amp cos(mm theta); the field is the product of the two Pade cores
    at (pp - n0, s +- r0 - dist), each times its winding factor."""
    X, Y, Z = _grid(p)
    if vr.plane == "xy":
        a, b, n, a0, b0, n0 = X, Y, Z, vr.x0, vr.y0, vr.z0
    elif vr.plane == "xz":
        a, b, n, a0, b0, n0 = X, Z, Y, vr.x0, vr.z0, vr.y0
    elif vr.plane == "yz":
        a, b, n, a0, b0, n0 = Y, Z, X, vr.y0, vr.z0, vr.x0
    else:
        raise ValueError("ERROR: Unrecognised vortex ring plane.")
    theta = np.arctan2(b - b0, a - a0)
    s = np.sqrt((a - a0) ** 2 + (b - b0) ** 2) - vr.r1 * np.sin(float(vr.kk) * theta)
    dist = vr.amp * np.cos(float(vr.mm) * theta)
    pp = n - vr.r1 * np.cos(float(vr.kk) * theta)
    q = scal * (pp - n0)
    u1 = scal * (s + vr.r0 - dist)
    u2 = scal * (s - vr.r0 - dist)
    rr1 = get_rr(np.sqrt(q ** 2 + u1 ** 2))
    rr2 = get_rr(np.sqrt(q ** 2 + u2 ** 2))
    return rr1 * (q + vr.dir * 1j * u1) * rr2 * (q - vr.dir * 1j * u2)


def _line_safe(vl: LineParam):
    """The zero-wavelength guard of get_r / get_theta (src/ic.f90:741-755)."""
    eps = np.finfo(np.float64).eps
    safe1 = safe2 = 0.0
    if abs(vl.ll) < eps:
        if abs(vl.amp1) < eps:
            safe1 = 1.0
        else:
            raise ValueError("ERROR: vortex line disturbance wavelength is zero, but amplitude is non-zero.")
        if abs(vl.amp2) < eps:
            safe2 = 1.0
        else:
            raise ValueError("ERROR: vortex line disturbance wavelength is zero, but amplitude is non-zero.")
    return safe1, safe2


def _line_offsets(p: RunParams, vl: LineParam):
    """The displaced transverse coordinates (u, v) of get_r and get_theta:
    for a line along 'x', u = y - y0 - a1 sin(2 pi (x - x0)/ll) and
    v = z - z0 - a2 cos(...); 'y' uses (z, x), 'z' uses (x, y)."""
    X, Y, Z = _grid(p)
    safe1, safe2 = _line_safe(vl)
    if vl.dir == "x":
        t, t0, u, u0, v, v0 = X, vl.x0, Y, vl.y0, Z, vl.z0
    elif vl.dir == "y":
        t, t0, u, u0, v, v0 = Y, vl.y0, Z, vl.z0, X, vl.x0
    elif vl.dir == "z":
        t, t0, u, u0, v, v0 = Z, vl.z0, X, vl.x0, Y, vl.y0
    else:
        raise ValueError(f"vortex line direction {vl.dir!r} is not one of 'x', 'y', 'z'")
    du = u - u0 - vl.amp1 * np.sin(2.0 * PI * (t - t0) / (vl.ll + safe1))
    dv = v - v0 - vl.amp2 * np.cos(2.0 * PI * (t - t0) / (vl.ll + safe2))
    return du, dv


def line_amp(r: np.ndarray) -> np.ndarray:
    """amp (src/ic.f90:309-323): 1 - exp(-0.7 r^1.15)."""
    return 1.0 - np.exp(-0.7 * r ** 1.15)


def vortex_line(p: RunParams, vl: LineParam, scal: float = 1.0) -> np.ndarray:
    """vortex_line (src/ic.f90:327-348): exp(i theta), times amp(r) unless
    only the phase is imprinted; theta = sgn atan2(scal dv, scal du) and
    r = sqrt(scal du^2 + scal dv^2) (get_theta, get_r)."""
    du, dv = _line_offsets(p, vl)
    theta = vl.sgn * np.arctan2(scal * dv, scal * du)
    phase = np.exp(1j * theta)
    if vl.imprint_phase:
        return phase
    r = np.sqrt(scal * du ** 2 + scal * dv ** 2)
    return line_amp(r) * phase


_PADE = {
    # a(1..9), U, m (src/ic.f90:558-580)
    "ring": ((-1.1, 0.0170524, 0.0289452, -0.953, -0.0049767, -0.0594346, 0.04, 0.21, 0.00612), 0.6, 8.97),
    "pulse": ((-0.79792, 0.00388059, 0.00882276, -0.7981, -0.012783, -0.0574092, 0.0399, 0.199, 0.0058), 0.63, 8.37),
}


def pade_pulse_ring(p: RunParams, kind: str, x0: float, y0: float, z0: float, r0: float) -> np.ndarray:
    """pade_pulse_ring (src/ic.f90:513-582): the Pade approximation of a
    moving vortex ring ('ring', radius r0 about the x axis through (y0, z0))
    or rarefaction pulse ('pulse'), propagating along x."""
    a, U, m = _PADE[kind]
    pw = 7.0 / 4.0
    X, Y, Z = _grid(p)
    s = np.sqrt((Y - y0) ** 2 + (Z - z0) ** 2)
    one_2u = 1.0 - 2.0 * U ** 2
    x2 = (X - x0) ** 2
    s2 = (s - r0) ** 2
    q = x2 + one_2u * s2
    denom = (1.0 + a[7] * x2 + a[6] * s2 + a[8] * q ** 2) ** pw
    ma = m * (a[8] ** pw)
    uu = 1.0 + (a[0] + a[2] * x2 + a[1] * s2 + ma * U * (2.0 * x2 - one_2u * s2)) * q / denom
    vv = (X - x0) * (a[3] + a[5] * x2 + a[4] * s2 - ma * q ** 2) / denom
    return uu + 1j * vv


def vortex_pair(p: RunParams) -> np.ndarray:
    """vortex_pair (src/ic.f90:586-631): the Pade approximation of a 2D
    vortex pair moving along x, uniform in z."""
    a0, a1, a2 = -1.14026, -0.150112, -0.0294564
    b0, b1, b2 = -0.830953, -0.108296, -0.073641
    c0, c1, c2, c3, c4 = 0.35022, 0.03032, 0.15905, 0.04123, 0.01402
    x, y, _ = p.axes()
    X, Y = x[None, :], y[:, None]
    x2, x4, y2 = X ** 2, X ** 4, Y ** 2
    denom = 1.0 + c0 * x2 + c1 * x4 + c2 * y2 + c3 * x2 * y2 + c4 * Y ** 4
    uu = 1.0 + (a0 + a1 * x2 + a2 * y2) / denom
    vv = X * (b0 + b1 * x2 + b2 * y2) / denom
    return np.ascontiguousarray(np.broadcast_to((uu + 1j * vv)[None], p.shape))


def sphere(p: RunParams) -> np.ndarray:
    """sphere (src/ic.f90:635-657): (1 + tanh(|r| - 10 - 2)) / 2."""
    X, Y, Z = _grid(p)
    return (0.5 * (1.0 + np.tanh(np.sqrt(X ** 2 + Y ** 2 + Z ** 2) - 10.0 - 2.0))).astype(np.complex128)


def sphere2(p: RunParams) -> np.ndarray:
    """sphere2 (src/ic.f90:661-679): 100 exp(-|r|^2 / 10)."""
    X, Y, Z = _grid(p)
    return (100.0 * np.exp(-(X ** 2 + Y ** 2 + Z ** 2) / 10.0)).astype(np.complex128)


def wall(p: RunParams) -> np.ndarray:
    """wall (src/ic.f90:683-704): 0 for z <= -30, 1 above."""
    _, _, z = p.axes()
    w = np.where(z <= -30.0, 0.0, 1.0).astype(np.complex128)
    return np.ascontiguousarray(np.broadcast_to(w[:, None, None], p.shape))


def kc_amp(p: RunParams, nv: float, enerv: float):
    """get_kc_amp (src/ic.f90:256-269): the cutoff kc2 (squared wavenumber
    in grid units) and the mode amplitude comp_amp of the random-phase state
    with mean density nv and energy enerv per particle."""
    ev = ((0.5 * p.nx / PI) ** 2) * enerv
    comp_amp = np.sqrt(((0.11095279993106999 * nv ** 2.5) * (p.nx * p.ny * p.nz)) / (ev ** 1.5))
    kc2 = (1.666666666666666 * ev) / nv
    return float(kc2), float(comp_amp)


def _mode_sq(n: int) -> np.ndarray:
    """ii2 of random_phase (src/ic.f90:219-236): i^2 for i <= n/2, else (n - i)^2."""
    i = np.arange(n)
    return np.where(i <= (n - 1) // 2 + 1, i, n - i) ** 2


def random_phase_ref(p: RunParams, nv: float = 0.75, enerv: float = 0.75, seed: int = 0,
                     phases: np.ndarray | None = None) -> np.ndarray:
    """random_phase (src/ic.f90:190-252): a(k) = comp_amp exp(2 pi i phi) for
    the modes with kx^2 + ky^2 + kz^2 <= kc2 (grid units), 0 elsewhere, and
    psi = forward DFT of a / sqrt(nx ny nz), as `fft(a, psi, 'forward')`
    (FFTW_FORWARD, src/ic.f90:905-1001).  The reference draws phi from the
    Fortran generator seeded from the clock, so no two reference runs share
    phases; here phi comes from `phases` (one value per mode inside the
    cutoff, in the reference's loop order: z slowest, x fastest) or from a
    seeded numpy generator.  The reference unpacks FFTW's transposed output
    with the y extent as the z stride, which is the transform only when
    ny == nz; this restatement is the transform itself."""
    kc2, comp_amp = kc_amp(p, nv, enerv)
    m2 = _mode_sq(p.nz)[:, None, None] + _mode_sq(p.ny)[None, :, None] + _mode_sq(p.nx)[None, None, :]
    inside = m2 <= kc2
    nin = int(np.count_nonzero(inside))
    if phases is None:
        phases = np.random.default_rng(seed).random(nin)
    phases = np.asarray(phases, dtype=np.float64)
    if phases.shape != (nin,):
        raise ValueError(f"random_phase_ref: {nin} phases needed, {phases.size} given")
    a = np.zeros(p.shape, dtype=np.complex128)
    a[inside] = comp_amp * np.exp(2.0 * PI * 1j * phases)
    psi = np.fft.fftn(a) / np.sqrt(float(p.nx * p.ny * p.nz))
    return np.ascontiguousarray(psi)


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
