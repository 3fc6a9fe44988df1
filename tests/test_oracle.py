"""Pins of the CPU oracle (oracle/gpe_oracle.c) against what the mathematics
fixes, since the reference ships no tests, fixtures or golden vectors:

* index maps against the boundary-condition definitions (periodic wrap,
  even reflection about the end nodes);
* FD operators against their plane-wave symbols (periodic) and cosine
  eigenmodes (reflective, including the boundary rows);
* get_rhs against pointwise closed forms for plane waves per equation;
* euler/rk2/rk4 against the scalar ODE the PDE reduces to for a plane wave,
  integrated in 50-digit mpmath;
* Cash-Karp against its order of accuracy (5th-order solution, 4th-order
  embedded estimate) on a linear problem, and rkqs against the controller
  rules of src/solve.f90:173-229 driven by the oracle's own rkck;
* the norm quadrature against exact integrals of low-degree polynomials,
  energy and mass against plane-wave closed forms.
"""
import math

import mpmath as mp
import numpy as np
import pytest

from gpe_amd.params import RunParams

mp.mp.dps = 50


def grid_params(nx=10, ny=8, nz=12, order=2, bcs=1, eqn=1, **kw):
    return RunParams(nx=nx, ny=ny, nz=nz, xr=2.3, yr=1.7, zr=2.9, order=order, bcs=bcs, eqn_to_solve=eqn, **kw)


def plane_wave(p, m, amp=1.0 + 0j):
    """A e^{i k.x} with k = 2 pi m / (n d) (periodic on n d)."""
    x, y, z = p.axes()
    kx, ky, kz = (2 * np.pi * m[0] / (p.nx * p.dx), 2 * np.pi * m[1] / (p.ny * p.dy), 2 * np.pi * m[2] / (p.nz * p.dz))
    Z, Y, X = np.meshgrid(z, y, x, indexing="ij")
    return amp * np.exp(1j * (kx * X + ky * Y + kz * Z)), (kx * p.dx, ky * p.dy, kz * p.dz)


def sym1(th, d, order):
    return 1j * math.sin(th) / d if order == 2 else 1j * (8 * math.sin(th) - math.sin(2 * th)) / (6 * d)


def sym2(th, d, order):
    if order == 2:
        return (2 * math.cos(th) - 2) / d ** 2
    return (-2 * math.cos(2 * th) + 32 * math.cos(th) - 30) / (12 * d ** 2)


def rel(a, b):
    return np.linalg.norm((a - b).ravel()) / max(np.linalg.norm(np.ravel(b)), 1e-300)


# ---------------------------------------------------------------------------
# index maps (src/derivs.f90:276-398)

def ref_map(i, N, bcs):
    if bcs == 1:
        return i % N
    P = 2 * (N - 1)
    i %= P
    return P - i if i >= N else i


@pytest.mark.parametrize("order", [2, 4])
@pytest.mark.parametrize("bcs", [1, 2])
def test_index_maps_match_bc_definition(orc, order, bcs):
    N = 9
    n = N - 1
    for i in range(N):
        mi, pl = orc.index_map(bcs, order, i, n, True)
        for k in range(len(mi)):
            assert mi[k] == ref_map(i - k - 1, N, bcs)
            assert pl[k] == ref_map(i + k + 1, N, bcs)
        # y/z maps: periodic indices are left to the ghost planes, reflective ones map
        mi, pl = orc.index_map(bcs, order, i, n, False)
        for k in range(len(mi)):
            if bcs == 1:
                assert mi[k] == i - k - 1 and pl[k] == i + k + 1
            else:
                assert mi[k] == ref_map(i - k - 1, N, bcs) and pl[k] == ref_map(i + k + 1, N, bcs)


# ---------------------------------------------------------------------------
# FD operators

@pytest.mark.parametrize("order", [2, 4])
def test_periodic_plane_wave_symbols(orc, order):
    p = grid_params(order=order, bcs=1)
    f, th = plane_wave(p, (2, 1, 3))
    d = (p.dx, p.dy, p.dz)
    for op in range(3):
        assert rel(orc.deriv(p, op, f), sym1(th[op], d[op], order) * f) < 1e-13
        assert rel(orc.deriv(p, op + 3, f), sym2(th[op], d[op], order) * f) < 1e-13
    lam = sum(sym2(th[a], d[a], order) for a in range(3))
    assert rel(orc.deriv(p, 6, f), lam * f) < 1e-13


@pytest.mark.parametrize("order", [2, 4])
def test_reflective_cosine_eigenmodes(orc, order):
    # cos(pi m i/(N-1)) is even about both end nodes, so the reflected stencil
    # reproduces the interior symbol on every row, boundary rows included.
    p = grid_params(nx=11, ny=9, nz=13, order=order, bcs=2)
    d = (p.dx, p.dy, p.dz)
    ns = (p.nx, p.ny, p.nz)
    for axis in range(3):
        m = 3
        th = math.pi * m / (ns[axis] - 1)
        idx = np.arange(ns[axis])
        c = np.cos(th * idx)
        s = np.sin(th * idx)
        shape = [1, 1, 1]
        shape[2 - axis] = ns[axis]
        f = np.broadcast_to(c.reshape(shape), p.shape).astype(np.complex128) * (1 + 0.5j)
        g = np.broadcast_to(s.reshape(shape), p.shape).astype(np.complex128) * (1 + 0.5j)
        assert rel(orc.deriv(p, axis + 3, f), sym2(th, d[axis], order) * f) < 1e-12
        # first derivative of cos(theta i) is -(symbol/i) sin(theta i), exactly, including at the ends
        assert rel(orc.deriv(p, axis, f), (sym1(th, d[axis], order) / 1j) * -g) < 1e-12


@pytest.mark.parametrize("order", [2, 4])
def test_polynomial_exactness_interior(orc, order):
    p = grid_params(nx=14, ny=14, nz=14, order=order, bcs=2)
    x, y, z = p.axes()
    Z, Y, X = np.meshgrid(z, y, x, indexing="ij")
    deg = 2 if order == 2 else 4
    f = (X ** deg + 0.5 * Y ** 3 + Z ** 2 + 0j)
    h = order // 2
    inner = (slice(h, -h),) * 3
    lap = orc.deriv(p, 6, f)
    exact = deg * (deg - 1) * X ** (deg - 2) + 3.0 * Y + 2.0
    assert np.max(np.abs(lap[inner] - exact[inner])) < 1e-9
    dx = orc.deriv(p, 0, f)
    assert np.max(np.abs(dx[inner] - deg * X[inner] ** (deg - 1))) < (1e-9 if order == 4 else 1e-9)


# ---------------------------------------------------------------------------
# get_rhs closed forms (src/solve.f90:414-437)

def test_rhs_plane_wave_closed_forms(orc):
    for order in (2, 4):
        amp = 0.8 - 0.3j
        # eqn 1: 0.5 i (lam + 1 - |A|^2) psi
        p = grid_params(order=order, eqn=1)
        f, th = plane_wave(p, (1, 2, 1), amp)
        lam = sum(sym2(th[a], (p.dx, p.dy, p.dz)[a], order) for a in range(3))
        assert rel(orc.rhs(p, f), 0.5j * (lam + 1 - abs(amp) ** 2) * f) < 1e-13
        # eqn 2: i (lam - |A|^2) psi
        p2 = p.replace(eqn_to_solve=2)
        assert rel(orc.rhs(p2, f), 1j * (lam - abs(amp) ** 2) * f) < 1e-13
        # eqn 1 with Urhs: + U * i k-symbol psi
        p3 = p.replace(Urhs=0.7)
        s1 = sym1(th[0], p.dx, order)
        assert rel(orc.rhs(p3, f), (0.5j * (lam + 1 - abs(amp) ** 2) + 0.7 * s1) * f) < 1e-13


def test_rhs_trap_constant_field(orc):
    # eqn 4 on a constant field (laplacian 0 with reflective BCs):
    # f = -i (V + g|c|^2 - mu) c pointwise
    p = grid_params(order=4, bcs=2, eqn=4, g=1.3, mu=2.1, omx=0.3, omy=0.2, omz=0.1)
    c = 0.6 + 0.2j
    f = np.full(p.shape, c, dtype=np.complex128)
    x, y, z = p.axes()
    Z, Y, X = np.meshgrid(z, y, x, indexing="ij")
    V = 0.5 * (0.3 ** 2 * X ** 2 + 0.2 ** 2 * Y ** 2 + 0.1 ** 2 * Z ** 2)
    assert rel(orc.rhs(p, f), -1j * (V + 1.3 * abs(c) ** 2 - 2.1) * c) < 1e-13
    # eqn 1 with the V extension: 0.5 i (1 - |c|^2 - V) c
    p1 = p.replace(eqn_to_solve=1, vscale=1.0)
    assert rel(orc.rhs(p1, f), 0.5j * (1 - abs(c) ** 2 - V) * c) < 1e-13
    # eqn 3: eqn 1 times sphere() = 0.5 (1 + tanh(r - 12))
    # (in a box reaching r ~ 12, where the mask is O(1))
    p3 = p.replace(eqn_to_solve=3, xr=14.0, yr=13.0, zr=15.0)
    x, y, z = p3.axes()
    Z, Y, X = np.meshgrid(z, y, x, indexing="ij")
    m = 0.5 * (1 + np.tanh(np.sqrt(X ** 2 + Y ** 2 + Z ** 2) - 12.0))
    assert rel(orc.rhs(p3, f), 0.5j * (1 - abs(c) ** 2) * c * m) < 1e-13


def test_rhs_dissipation_term(orc):
    # diss = amp * sum_axis (1 + tanh(s - a) tanh(s + a)), a = coordinate of index n-3
    p = grid_params(order=2, bcs=2, eqn=1, diss_amp=0.05)
    c = 0.9 + 0.1j
    f = np.full(p.shape, c, dtype=np.complex128)
    x, y, z = p.axes()
    Z, Y, X = np.meshgrid(z, y, x, indexing="ij")
    a, b, cc = x[-3], y[-3], z[-3]
    diss = 0.05 * ((1 + np.tanh(X - a) * np.tanh(X + a)) + (1 + np.tanh(Y - b) * np.tanh(Y + b)) +
                   (1 + np.tanh(Z - cc) * np.tanh(Z + cc)))
    assert rel(orc.rhs(p, f), 0.5 * (1j + diss) * (1 - abs(c) ** 2) * c) < 1e-13


# ---------------------------------------------------------------------------
# steppers: plane waves keep |psi| uniform, so the PDE step is the scalar
# step of y' = F(y) = 0.5 i (lam + 1 - |y|^2) y (eqn 1), computed in mpmath.

def scalar_F(lam):
    return lambda y: 0.5j * (lam + 1 - abs(y) ** 2) * y


def mp_step(scheme, F, y, dt):
    y = mp.mpc(y)
    dt = mp.mpc(dt)
    if scheme == "euler":
        return y + dt * F(y)
    if scheme == "rk2":
        k1 = F(y)
        return y + dt * F(y + dt / 2 * k1)
    k1 = F(y)
    k2 = F(y + dt / 2 * k1)
    k3 = F(y + dt / 2 * k2)
    k4 = F(y + dt * k3)
    return y + dt * (k1 / 6 + k2 / 3 + k3 / 3 + k4 / 6)


@pytest.mark.parametrize("scheme", ["euler", "rk2", "rk4"])
@pytest.mark.parametrize("order", [2, 4])
def test_steppers_scalar_reduction(orc, scheme, order):
    p = grid_params(order=order)
    amp = 0.7 + 0.4j
    f, th = plane_wave(p, (1, 1, 2), amp)
    lam = sum(sym2(th[a], (p.dx, p.dy, p.dz)[a], order) for a in range(3))
    dt = 0.05 + 0j
    out, t = orc.step(p, scheme, f, dt, (0.0, 0.0))
    y = complex(mp_step(scheme, scalar_F(mp.mpf(lam)), amp, dt))
    assert rel(out, y * (f / amp)) < 1e-14
    assert t == (0.05, 0.0)
    # imaginary time: im_t advances by |aimag(dt)|
    out, t = orc.step(p, scheme, f, -0.05j, (1.0, 2.0))
    assert t == (1.0, 2.05)
    y = complex(mp_step(scheme, scalar_F(mp.mpf(lam)), amp, -0.05j))
    assert rel(out, y * (f / amp)) < 1e-14


def test_rk4_convergence_order(orc):
    p = grid_params(order=2)
    amp = 0.9 + 0.2j
    f, th = plane_wave(p, (1, 0, 1), amp)
    lam = sum(sym2(th[a], (p.dx, p.dy, p.dz)[a], 2) for a in range(3))
    F = scalar_F(mp.mpf(lam))
    errs = []
    for dt in (0.2, 0.1):
        out, _ = orc.step(p, "rk4", f, complex(dt), (0, 0))
        # exact solution of the scalar ODE: |y| constant -> y(t) = y0 exp(0.5 i (lam + 1 - |y0|^2) t)
        exact = complex(mp.mpc(amp) * mp.exp(0.5j * (lam + 1 - abs(amp) ** 2) * dt))
        errs.append(abs(out.ravel()[0] / (f.ravel()[0] / amp) - exact))
    assert 2 ** 4.6 < errs[0] / errs[1] < 2 ** 5.4  # local error O(dt^5)
    del F


def test_cash_karp_order(orc):
    # linear problem (eqn 4, g = 0, V = 0): y' = z y with z = i (lam/2 + mu)
    p = grid_params(order=2, eqn=4, g=0.0, mu=1.5)
    amp = 1.0 + 0j
    f, th = plane_wave(p, (1, 1, 1), amp)
    lam = sum(sym2(th[a], (p.dx, p.dy, p.dz)[a], 2) for a in range(3))
    zr = 1j * (0.5 * lam + 1.5)
    e5, e4 = [], []
    for dt in (0.4 / abs(zr), 0.2 / abs(zr)):
        new, err = orc.rkck(p, f, complex(dt))
        R = new.ravel()[0] / f.ravel()[0]
        E = err.ravel()[0] / f.ravel()[0]
        ex = complex(mp.exp(mp.mpc(zr) * dt))
        e5.append(abs(R - ex))
        e4.append(abs((R - E) - ex))
    assert 2 ** 5.5 < e5[0] / e5[1] < 2 ** 6.5   # 5th-order solution: local error O(dt^6)
    assert 2 ** 4.5 < e4[0] / e4[1] < 2 ** 5.5   # embedded 4th order: O(dt^5)


def test_rkqs_controller_rules(orc):
    # drive the reference controller rules (src/solve.f90:173-229) with the
    # oracle's own rkck and compare with orc_rkqs, including rejections
    p = grid_params(order=2, eqn=1)
    x, y, z = p.axes()
    rng = np.random.default_rng(3)
    psi = (1.0 + 0.3 * rng.standard_normal(p.shape) + 0.3j * rng.standard_normal(p.shape))
    errcon = (5.0 / 0.9) ** (1.0 / -0.2)
    assert abs(errcon - 1.8895680e-4) < 1e-10
    dt = 0.5 + 0j   # large: forces rejections
    t = (0.0, 0.0)
    cur = psi.copy()
    for _ in range(3):
        new, dtn, tn, errmax, nrej, nrhs = orc.rkqs(p, cur, dt, t)
        # restatement of the rules
        f0 = orc.rhs(p, cur)
        scal = np.abs(cur) + dt * np.abs(f0) + 1e-30
        d = dt
        rej = 0
        while True:
            cand, err = orc.rkck(p, cur, d)
            em = np.max(np.abs(err) / np.abs(scal)) / 1e-6
            if em <= 1.0:
                break
            rej += 1
            dtemp = 0.9 * d * em ** -0.25
            d = complex(max(abs(dtemp), 0.1 * abs(d)), 0.0)
        dnext = 0.9 * d * em ** -0.2 if em > errcon else 5.0 * d
        assert nrej == rej
        assert abs(errmax - em) <= 1e-12 * em
        assert abs(dtn - dnext) <= 1e-14 * abs(dnext)
        assert tn[0] == t[0] + d.real
        assert nrhs == 1 + 6 * (rej + 1)
        assert rel(new, cand) < 1e-15
        cur, dt, t = new, dtn, tn
    assert t[0] > 0


# ---------------------------------------------------------------------------
# reductions (src/variables.f90:537-611, 698-709, 713-823)

def test_norm_quadrature_exact_for_cubics(orc):
    p = grid_params(nx=11, ny=9, nz=13)
    one = np.ones(p.shape, dtype=np.complex128)
    assert abs(orc.norm(p, one) - 8 * p.xr * p.yr * p.zr) < 1e-12 * 8 * p.xr * p.yr * p.zr
    x, y, z = p.axes()
    Z, Y, X = np.meshgrid(z, y, x, indexing="ij")
    # |psi|^2 = x^2 (1 + y) with psi real: int = (2 xr^3/3)(2 yr)(2 zr)
    psi = (X * np.sqrt(1.0 + Y / 10.0)).astype(np.complex128)
    exact = (2 * p.xr ** 3 / 3) * (2 * p.yr) * (2 * p.zr)
    assert abs(orc.norm(p, psi) - exact) < 1e-12 * exact
    # mass is the plain sum times dV
    assert abs(orc.mass(p, one) - p.cells * p.dx * p.dy * p.dz) < 1e-12 * p.cells * p.dx * p.dy * p.dz


def test_energy_plane_wave(orc):
    for order in (2, 4):
        p = grid_params(order=order)
        amp = 0.8 + 0.1j
        f, th = plane_wave(p, (1, 2, 3), amp)
        lam = sum(sym2(th[a], (p.dx, p.dy, p.dz)[a], order) for a in range(3))
        exact = p.cells * p.dx * p.dy * p.dz * (-lam * abs(amp) ** 2 + 0.5 * abs(amp) ** 4)
        assert abs(orc.energy(p, f) - exact) < 1e-12 * abs(exact)


def test_renormalise(orc):
    p = grid_params(nn=3.0)
    psi = np.full(p.shape, 2.0 + 0j)
    out = orc.renormalise(p, psi, 4.0)
    assert np.allclose(out, math.sqrt(3.0) * 2.0 / 2.0, rtol=1e-15)
