"""The host-array drop-ins of the reference's variables module never touch
the device state, and rkqs follows the host's real_time flag.

* The Fortran variables module (gpe_amd/fortran/variables.f90), in resident
  mode: get_norm, energy, mass, momentum, linelength and renormalise are
  handed another array (the reference's save_energy(t, psi%old) hands the
  stale psi%old, src/gpe.f90:177-178); the state is unchanged bit for bit
  and the values are those of the array handed in
  (tests/fortran/dropin_variables.f90).
* The Python mirror (gpe_amd.variables) and the C ABI's auxiliary field
  (gpe_field_upload, gpe_select_field).
* gpe_set_real_time: a context created for imaginary time and switched to
  real time retries along the real axis, as the reference's rkqs does with
  real_time = .true. (src/solve.f90:195-201): the same accepts, rejects, dt
  and t as the oracle's rkqs."""
import os
import subprocess

import numpy as np
import pytest

from gpe_amd import Context, GpeError, RunParams, StepInfo
from gpe_amd import parameters as P
from gpe_amd import variables as V

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "tests", "fortran", "dropin_variables")


def fields(p):
    k, j, i = np.meshgrid(np.arange(p.nz), np.arange(p.ny), np.arange(p.nx), indexing="ij")
    psi = (1.0 + 0.3 * np.cos(0.7 * i + 0.2 * j)) + 1j * (0.2 * np.sin(0.5 * j - 0.3 * k))
    other = (0.5 * np.sin(0.4 * i) - 0.2) + 1j * (0.8 + 0.1 * np.cos(0.9 * k + 0.1 * i))
    return np.ascontiguousarray(psi), np.ascontiguousarray(other)


def test_fortran_check_program_built():
    assert os.access(EXE, os.X_OK), "built by make -C gpe_amd/fortran (__graft_entry__.build())"


@pytest.mark.gpu
def test_fortran_variables_dropins_leave_the_state(tmp_path):
    r = subprocess.run([EXE], cwd=tmp_path, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    v = {}
    for ln in r.stdout.splitlines():
        key, *vals = ln.split()
        v[key] = np.array([float(x) for x in vals])
    assert v["state_max_change"][0] == 0.0
    assert v["state_energy_after"][0] == v["state_energy_before"][0]
    assert v["state_norm_after"][0] == v["state_norm_before"][0]
    p = RunParams(nx=20, ny=18, nz=16, xr=5.0, yr=4.5, zr=4.0, order=4, bcs=1, eqn_to_solve=1)
    psi, other = fields(p)
    with Context(p) as ctx:
        ctx.upload(other)
        exp = dict(other_energy=ctx.energy(), other_norm=ctx.norm(), other_mass=ctx.mass(),
                   other_momentum=ctx.momentum(), other_linelength=ctx.linelength())
        ctx.upload(psi)
        e_psi = ctx.energy()
    assert abs(v["state_energy_before"][0] - e_psi) <= 1e-12 * abs(e_psi)
    for key, val in exp.items():
        val = np.atleast_1d(val)
        assert np.allclose(v[key], val, rtol=1e-12, atol=1e-12 * np.abs(val).max()), (key, v[key], val)
    # renormalise(var, norm) scales the array handed in: sum |var|^2 nn/norm
    assert np.isclose(v["renormalised_norm_sq"][0], v["other_norm_sq"][0] / v["other_norm"][0], rtol=1e-12)


@pytest.mark.gpu
def test_python_mirror_and_aux_field_leave_the_state():
    p = RunParams(nx=20, ny=18, nz=16, xr=5.0, yr=4.5, zr=4.0, order=4, bcs=1, eqn_to_solve=1)
    psi, other = fields(p)
    ctx = P.setup(p)
    try:
        ctx.upload(psi)
        e0 = ctx.energy()
        halo = np.zeros((p.nz + 4, p.ny + 4, p.nx), complex)
        halo[2:-2, 2:-2] = other
        eo, no, mo = V.energy(halo), V.get_norm(other), V.mass(other)
        ren = other.copy()
        V.renormalise(ren, no)
        assert np.array_equal(ctx.download(), psi) and ctx.energy() == e0
        assert np.allclose(ren, other / np.sqrt(no), rtol=1e-14, atol=1e-15)
        with Context(p) as c2:
            c2.upload(other)
            assert (eo, no, mo) == (c2.energy(), c2.norm(), c2.mass())
        # the auxiliary field is selected explicitly; stepping uses the state
        ctx.upload_aux(other)
        ctx.select(True)
        assert np.array_equal(ctx.download(), other) and ctx.norm() == no
        ctx.step("rk4", 1, StepInfo(0.01 + 0j))
        assert np.array_equal(ctx.download(), other)  # the aux field stays as uploaded
        ctx.select(False)
        assert not np.array_equal(ctx.download(), psi)  # the state moved
    finally:
        P.ctx.close()
        P.ctx = None


@pytest.mark.gpu
def test_select_aux_before_upload_fails():
    p = RunParams(nx=8, ny=8, nz=8, xr=2.0, yr=2.0, zr=2.0)
    with Context(p) as ctx:
        with pytest.raises(GpeError, match="no auxiliary field"):
            ctx.select(True)


@pytest.mark.gpu
@pytest.mark.parametrize("storage", ["c128", "c64", "mixed"])
def test_set_real_time_steers_rkqs_retries(orc, storage):
    """A context created with real_time = 0 (imaginary time) and then set to
    real time: a step too large for tolerance 1e-6 is retried with a real,
    shrunk dt, t advances and im_t stays 0 -- as the oracle's rkqs with
    real_time = 1.  Left at imaginary time the same retry turns dt onto the
    negative imaginary axis (the reference's imaginary-time rule)."""
    p = RunParams(nx=16, ny=16, nz=16, xr=6.0, yr=6.0, zr=6.0, order=4, bcs=1, eqn_to_solve=1, scheme="rk45",
                  real_time=False, storage=storage)
    rng = np.random.default_rng(2)
    psi = np.ascontiguousarray(1.0 + 0.2 * (rng.standard_normal(p.shape) + 1j * rng.standard_normal(p.shape)))
    dt0 = 0.3 + 0j
    with Context(p) as ctx:
        ctx.upload(psi)
        ctx.set_real_time(True)
        info = StepInfo(dt0)
        seq = []
        for _ in range(4):
            r0 = info.n_reject
            ctx.step("rk45", 1, info)
            seq.append((info.n_reject - r0, info.t, info.im_t, info.dt))
        got = ctx.download()
    assert seq[0][0] >= 1 and all(s[2] == 0.0 and s[3].imag == 0.0 and s[3].real > 0 for s in seq)
    assert all(b[1] > a[1] for a, b in zip(seq, seq[1:]))
    if storage != "c128":
        return  # complex64 increments move errmax at 1e-2 relative: the sequence is the c128 one only
    q = p.replace(real_time=True)
    x, dt, t = psi, dt0, (0.0, 0.0)
    for nrej, tt, im_t, dtn in seq:
        x, dt, t, _, orej, _ = orc.rkqs(q, x, dt, t)
        assert orej == nrej and t[1] == 0.0
        assert abs(t[0] - tt) <= 1e-10 * tt and abs(dt - dtn) <= 1e-9 * abs(dtn)
    assert np.linalg.norm(got - x) / np.linalg.norm(x) <= 1e-9
    with Context(p) as ctx:  # left in imaginary time
        ctx.upload(psi)
        info = StepInfo(dt0)
        ctx.step("rk45", 1, info)
        assert info.n_reject >= 1 and info.dt.real == 0.0 and info.dt.imag < 0.0


@pytest.mark.gpu
@pytest.mark.parametrize("order,bcs", [(2, 1), (4, 2)])
def test_python_module_mirror_with_halo_arrays(orc, order, bcs):
    """The Python mirror of the reference's modules, called as the reference
    calls them: derivs.deriv_* and variables.laplacian take the halo array
    var(0:nx1, js-2:je+2, ks-2:ke+2) (src/derivs.f90:20-21,
    src/variables.f90:55-71); solve.solver(var_in, var_out) advances t, im_t
    and dt in `parameters` (src/solve.f90:25-53).  Against the oracle."""
    from gpe_amd import derivs as D
    from gpe_amd import solve as S

    p = RunParams(nx=18, ny=14, nz=12, xr=3.0, yr=2.5, zr=2.0, order=order, bcs=bcs, eqn_to_solve=1,
                  scheme="rk4", tau=0.004)
    psi, _ = fields(p)
    halo = np.zeros((p.nz + 4, p.ny + 4, p.nx), dtype=np.complex128)
    halo[2:-2, 2:-2, :] = psi
    P.setup(p)
    try:
        for fn, op in [(D.deriv_x, 0), (D.deriv_y, 1), (D.deriv_z, 2), (D.deriv_xx, 3), (D.deriv_yy, 4),
                       (D.deriv_zz, 5)]:
            out = np.empty(p.shape, dtype=np.complex128)
            fn(halo, out)
            exp = orc.deriv(p, op, psi)
            assert np.linalg.norm(out - exp) <= 1e-13 * np.linalg.norm(exp), fn.__name__
        lap = V.laplacian(halo)
        exp = orc.deriv(p, 6, psi)
        assert np.linalg.norm(lap - exp) <= 1e-13 * np.linalg.norm(exp)
        x = psi
        for n in range(3):
            var_out = np.empty(p.shape, dtype=np.complex128)
            h = np.zeros_like(halo)
            h[2:-2, 2:-2, :] = x
            S.solver(h, var_out)
            x, (t, im_t) = orc.step(p, "rk4", x, p.dt0(), (t, im_t) if n else (0.0, 0.0))
            assert np.linalg.norm(var_out - x) <= 1e-13 * np.linalg.norm(x)
            assert P.t == t and P.im_t == im_t and P.dt == p.dt0()
            x = var_out
    finally:
        P.ctx.close()
        P.ctx = None
