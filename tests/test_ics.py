"""Initial conditions (gpe_amd/ics.py) against the reference's own ics
compiled with each case's expression as its ic.in
(oracle/ref/ic_cases.txt -> tests/golden/ref_ic_*.npz, 16 x 12 x 10 grid),
and against closed forms the generators must satisfy.

Tolerance for the fixtures: 1e-14 times max(1, |psi|) per cell (measured:
7e-16 at most).  numpy's sin, cos, atan2, exp and pow may differ from the
libm calls of the Fortran build by a few ulp, and the generators have no
cancellation that would amplify that.
"""
import glob
import json
import os

import numpy as np
import pytest

from gpe_amd import RunParams, ics

HERE = os.path.dirname(os.path.abspath(__file__))
FIXTURES = sorted(glob.glob(os.path.join(HERE, "golden", "ref_ic_*.npz")))

# The Python composition of each case of oracle/ref/ic_cases.txt.
CASES = {
    "ring_xy": lambda p, s: ics.vortex_ring(p, ics.RingParam("xy", 0.5, -0.4, 0.3, 3.0, 0.3, 3, 0.4, 2, 1.0), s),
    "ring_xz": lambda p, s: ics.vortex_ring(p, ics.RingParam("xz", -0.3, 0.2, 0.6, 2.5, 0.2, 4, 0.3, 3, -1.0), s),
    "ring_yz": lambda p, s: ics.vortex_ring(p, ics.RingParam("yz", 0.4, 0.1, -0.2, 3.5, 0.0, 5, 0.0, 10, -1.0), s),
    "line_x": lambda p, s: ics.vortex_line(p, ics.LineParam(0.3, -0.5, 0.2, 0.4, 0.3, 6.0, 1.0, "x", False), s),
    "line_y_phase": lambda p, s: ics.vortex_line(p, ics.LineParam(-0.2, 0.1, 0.4, 0.5, 0.2, 5.0, -1.0, "y", True), s),
    "line_z": lambda p, s: ics.vortex_line(p, ics.LineParam(0.25, -0.35, 0.0, 0.0, 0.0, 0.0, 1.0, "z", False), s),
    "trap_line": lambda p, s: ics.fermi(p) * ics.vortex_line(p, ics.LineParam(-1.5, 0, 0, 0, 0, 0, 1.0, "z", True), s),
    "pade_ring": lambda p, s: ics.pade_pulse_ring(p, "ring", 0.5, -0.2, 0.3, 3.0),
    "pade_pulse": lambda p, s: ics.pade_pulse_ring(p, "pulse", -0.4, 0.3, -0.1, 0.0),
    "pair_sphere": lambda p, s: ics.vortex_pair(p) * ics.sphere(p),
    "wall_sphere2": lambda p, s: ics.wall(p) + ics.sphere2(p),
}


def ids(paths):
    return [os.path.basename(x)[7:-4] for x in paths]


def test_every_case_has_a_fixture():
    assert sorted(CASES) == sorted(ids(FIXTURES))


@pytest.mark.parametrize("path", FIXTURES, ids=ids(FIXTURES))
def test_generator_matches_reference_build(path):
    z = np.load(path, allow_pickle=False)
    p = RunParams(**json.loads(str(z["params"])))
    name = os.path.basename(path)[7:-4]
    got = CASES[name](p, float(z["scal"]))
    ref = z["out"]
    assert got.shape == ref.shape
    err = np.abs(got - ref) / np.maximum(1.0, np.abs(ref))
    print(f"{name}: max scaled difference {err.max():.2e}")
    assert err.max() <= 1e-14


def test_pade_core_limits():
    """R(r) r -> 1 far from the core and ~ sqrt(0.3437) r near it."""
    r = np.array([1e-6, 1e4])
    rr = ics.get_rr(r) * r
    assert abs(rr[0] / 1e-6 - np.sqrt(0.3437)) < 1e-9
    assert abs(rr[1] - 1.0) < 1e-6


def winding(field_on_loop):
    ph = np.angle(field_on_loop)
    d = np.diff(np.concatenate([ph, ph[:1]]))
    return int(np.round(((d + np.pi) % (2 * np.pi) - np.pi).sum() / (2 * np.pi)))


@pytest.mark.parametrize("plane", ["xy", "xz", "yz"])
@pytest.mark.parametrize("direction", [1.0, -1.0])
def test_ring_phase_winds_once_about_each_core(plane, direction):
    """A loop around the core cross-section on either side of the ring picks
    up a phase of 2 pi, with opposite signs at the two sides (a closed ring),
    and the sign flips with dir."""
    n = 97
    p = RunParams(nx=n, ny=n, nz=n, xr=6.0, yr=6.0, zr=6.0)
    vr = ics.RingParam(plane, 0.0, 0.0, 0.0, 3.0, dir=direction)
    psi = ics.vortex_ring(p, vr)
    c = n // 2
    step = 2 * 6.0 / (n - 1)
    i_core = c + int(round(3.0 / step))  # the core crosses the first in-plane axis at +r0
    i_opp = c - int(round(3.0 / step))
    th = np.linspace(0, 2 * np.pi, 64, endpoint=False)
    rad = 6  # cells
    a_off = np.round(rad * np.cos(th)).astype(int)
    n_off = np.round(rad * np.sin(th)).astype(int)
    # (z, y, x) index of (in-plane axis a, normal axis n), the second in-plane axis at its centre
    axes = {"xy": (2, 0, 1), "xz": (2, 1, 0), "yz": (1, 2, 0)}[plane]

    def loop(i0):
        idx = [None, None, None]
        idx[axes[0]] = i0 + a_off
        idx[axes[1]] = c + n_off
        idx[axes[2]] = np.full_like(a_off, c)
        return psi[tuple(idx)]

    w1, w2 = winding(loop(i_core)), winding(loop(i_opp))
    assert abs(w1) == 1 and w2 == -w1
    vr2 = ics.RingParam(plane, 0.0, 0.0, 0.0, 3.0, dir=-direction)
    psi = ics.vortex_ring(p, vr2)
    assert winding(loop(i_core)) == -w1


def test_line_phase_imprint_and_core():
    p = RunParams(nx=33, ny=33, nz=9, xr=4.0, yr=4.0, zr=1.0)
    vl = ics.LineParam(0.0, 0.0, 0.0, 0.0, 0.0, 0.0, -1.0, "z", True)
    psi = ics.vortex_line(p, vl)
    assert np.allclose(np.abs(psi), 1.0, atol=1e-15)
    x, y, _ = p.axes()
    assert np.allclose(np.angle(psi[0]), -np.arctan2(y[:, None], x[None, :]), atol=1e-14)
    psi = ics.vortex_line(p, ics.LineParam(dir="z"))
    # amplitude 1 - exp(-0.7 r^1.15): 0 on the axis
    assert abs(psi[0, 16, 16]) == 0.0
    with pytest.raises(ValueError):
        ics.vortex_line(p, ics.LineParam(amp1=0.5, ll=0.0))


def test_displaced_line_follows_its_sinusoid():
    """The core of a line along x with displacement a1 sin(2 pi x / ll) in y
    sits where |psi| is smallest in each x-plane."""
    p = RunParams(nx=16, ny=161, nz=161, xr=4.0, yr=4.0, zr=4.0)
    vl = ics.LineParam(0.0, 0.0, 0.0, 1.0, 0.0, 8.0, 1.0, "x", False)
    psi = ics.vortex_line(p, vl)
    x, y, z = p.axes()
    for i in (0, 3, 7, 11):
        plane = np.abs(psi[:, :, i])
        kz, jy = np.unravel_index(np.argmin(plane), plane.shape)
        assert abs(y[jy] - np.sin(2 * np.pi * x[i] / 8.0)) <= 0.05 + 1e-12
        assert abs(z[kz]) <= 0.05 + 1e-12


def test_pade_and_pair_tend_to_one_far_away():
    p = RunParams(nx=9, ny=9, nz=9, xr=400.0, yr=400.0, zr=400.0)
    for psi in (ics.pade_pulse_ring(p, "ring", 0, 0, 0, 3.0), ics.pade_pulse_ring(p, "pulse", 0, 0, 0, 0.0)):
        assert abs(psi[0, 0, 0] - 1.0) < 1e-2
    psi = ics.vortex_pair(p)
    assert abs(psi[0, 0, 0] - 1.0) < 1e-2


def test_random_phase_constants_and_spectrum():
    """get_kc_amp's constant is 3 / (4 pi (5/3)^1.5): the mean density of
    comp_amp^2 over a sphere of radius sqrt(kc2) modes is then nv, and the
    mean squared mode number over that sphere, (3/5) kc2, is ev/nv.  The
    state's transform holds comp_amp at every mode inside the cutoff and
    nothing outside."""
    assert abs(0.11095279993106999 - 3.0 / (4.0 * np.pi * (5.0 / 3.0) ** 1.5)) < 1e-16
    n = 64
    p = RunParams(nx=n, ny=n, nz=n, xr=10.0, yr=10.0, zr=10.0)
    nv, enerv = 0.8, 0.6
    kc2, comp_amp = ics.kc_amp(p, nv, enerv)
    psi = ics.random_phase_ref(p, nv, enerv, seed=4)
    spec = np.fft.ifftn(psi) * np.sqrt(float(n ** 3))
    m = np.minimum(np.arange(n), n - np.arange(n)) ** 2
    m2 = m[:, None, None] + m[None, :, None] + m[None, None, :]
    inside = m2 <= kc2
    assert np.allclose(np.abs(spec[inside]), comp_amp, rtol=1e-12)
    assert np.abs(spec[~inside]).max() < 1e-12 * comp_amp
    # lattice count vs the sphere's volume: a few per cent at kc2 ~ 50
    dens = float(np.mean(np.abs(psi) ** 2))
    assert abs(dens / nv - 1.0) < 0.05
    ev = ((0.5 * n / np.pi) ** 2) * enerv
    assert abs(m2[inside].mean() / (ev / nv) - 1.0) < 0.05
    # phases in the reference's loop order: z slowest, x fastest
    ph = np.random.default_rng(1).random(int(inside.sum()))
    psi2 = ics.random_phase_ref(p, nv, enerv, phases=ph)
    spec2 = np.fft.ifftn(psi2) * np.sqrt(float(n ** 3))
    k, j, i = np.nonzero(inside)
    first = (k[0], j[0], i[0])
    assert first == (0, 0, 0)
    assert abs(np.angle(spec2[first]) - np.angle(np.exp(2j * np.pi * ph[0]))) < 1e-12
