"""Checkpoint / restart / stop semantics of the Fortran host and the
imaginary-to-real-time switch of its main loop (src/gpe.f90:134-278).

* Stop and restart: removing RUNNING makes the host write its checkpoint
  proc00/end_state.dat and stop (src/gpe.f90:136-159, end_state
  src/io.f90:840-891).  Copied to end_state00.dat (scripts/run.sh -r) and
  read back with restart = .true. (state_restart, src/ic.f90:117-186), the
  run lands bit for bit on the uninterrupted one, for RK4 and for RKF45
  (whose dt is carried by the checkpoint).
* Guards: restart without end_state00.dat, a fresh run with end_state00.dat
  present (src/gpe.f90:116-124, src/ic.f90:96-101), saved_restart's
  filtered state (src/ic.f90:143-159).
* The switch: eqn 4 in imaginary time with renorm converges until the
  relative change of the norm drops below 1e-12; then the host stops
  (stop_imag) or sets real_time, im_t = 0, resets the save counters and
  rotates dt into real time (src/gpe.f90:251-271).  After the switch a
  rejected RKF45 attempt shrinks dt along the real axis (src/solve.f90:
  195-201).  Checked against the same sequence driven through the C ABI and
  against the oracle's rkqs.
The failures that stop before any device call run on the CPU; the rest
need the GPU."""
import os
import shutil
import subprocess

import numpy as np
import pytest

from gpe_amd import Context, RunParams, StepInfo, ics
from gpe_amd.io import read_dens, read_end_state, run_in_text, set_namelist, write_end_state

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "gpe_amd", "fortran", "gpe_hip_main")


def run(cwd, text, env=None):
    (cwd / "run.in").write_text(text)
    e = dict(os.environ)
    e.pop("GPE_REMOVE_RUNNING_AT_P", None)
    e.update(env or {})
    return subprocess.run([EXE], cwd=cwd, capture_output=True, text=True, timeout=300, env=e)


def fresh_dir(tmp_path, name, p, psi):
    d = tmp_path / name
    d.mkdir()
    write_end_state(str(d / "ic_state.dat"), psi, 0, 0.0, 0.0, p.dt0())
    return d


# ---------------------------------------------------------------------------
# CPU: failures before the first device call


def test_restart_without_checkpoint_stops(tmp_path):
    """restart = .true. and no end_state00.dat: ERROR (src/ic.f90:96-101)."""
    p = RunParams(nx=8, ny=8, nz=8, xr=2.0, yr=2.0, zr=2.0)
    r = run(tmp_path, run_in_text(p, 1.0, restart=True))
    assert r.returncode != 0
    msg = (tmp_path / "ERROR").read_text()
    assert "restart=.true. but end_state00.dat does not exist." in msg
    assert not (tmp_path / "RUNNING").exists()


def test_fresh_run_without_initial_state_stops(tmp_path):
    p = RunParams(nx=8, ny=8, nz=8, xr=2.0, yr=2.0, zr=2.0)
    r = run(tmp_path, run_in_text(p, 1.0, restart=False))
    assert r.returncode != 0
    assert "ic_state.dat does not exist" in (tmp_path / "ERROR").read_text()


def test_checkpoint_grid_mismatch_stops(tmp_path):
    p = RunParams(nx=8, ny=8, nz=8, xr=2.0, yr=2.0, zr=2.0)
    write_end_state(str(tmp_path / "end_state00.dat"), np.ones((8, 8, 9), complex), 0, 0.0, 0.0, p.dt0())
    r = run(tmp_path, run_in_text(p, 1.0, restart=True))
    assert r.returncode != 0
    assert "grid does not match" in (tmp_path / "ERROR").read_text()


# ---------------------------------------------------------------------------
# GPU: stop -> checkpoint -> restart


def small_case(scheme):
    p = RunParams(nx=24, ny=20, nz=18, xr=8.0, yr=7.0, zr=6.5, order=4, bcs=1, eqn_to_solve=1, scheme=scheme,
                  tau=0.01 if scheme == "rk4" else 0.02)
    rng = np.random.default_rng(3)
    psi = ics.vortex_ring(p, ics.RingParam("xy", 0.0, 0.0, 0.0, 3.0)) + 0.01
    psi = psi + 1e-3 * (rng.standard_normal(p.shape) + 1j * rng.standard_normal(p.shape))
    return p, np.ascontiguousarray(psi)


@pytest.mark.gpu
@pytest.mark.parametrize("scheme", ["rk4", "rk45"])
@pytest.mark.parametrize("mode", ["resident", "parity"])
def test_stop_checkpoint_restart_is_bitwise(tmp_path, scheme, mode):
    p, psi = small_case(scheme)
    end_time = (12 - 0.5) * p.tau if scheme == "rk4" else 1.5
    k = 5
    text = set_namelist(run_in_text(p, end_time, restart=False), host_mode=mode)
    # A: uninterrupted
    a = fresh_dir(tmp_path, "a", p, psi)
    ra = run(a, text)
    assert ra.returncode == 0, ra.stdout + ra.stderr
    out_a, p_a, t_a, imt_a, dt_a = read_end_state(str(a / "proc00" / "end_state.dat"))
    assert p_a > k + 2
    # B: RUNNING removed at the top of iteration p = k
    b = fresh_dir(tmp_path, "b", p, psi)
    rb = run(b, text, {"GPE_REMOVE_RUNNING_AT_P": str(k)})
    assert rb.returncode == 0, rb.stdout + rb.stderr
    assert "Stop requested." in rb.stdout
    assert not (b / "RUNNING").exists()
    out_b, p_b, t_b, imt_b, dt_b = read_end_state(str(b / "proc00" / "end_state.dat"))
    assert p_b == k and 0.0 < t_b < t_a
    if scheme == "rk45":
        assert dt_b != p.dt0()  # the adapted step travels with the checkpoint
    # C: restart from B's checkpoint (run.sh -r copies proc00/end_state.dat
    # to end_state00.dat)
    c = tmp_path / "c"
    c.mkdir()
    shutil.copy(b / "proc00" / "end_state.dat", c / "end_state00.dat")
    rc = run(c, set_namelist(text, restart=True))
    assert rc.returncode == 0, rc.stdout + rc.stderr
    assert "Getting restart conditions" in rc.stdout
    out_c, p_c, t_c, imt_c, dt_c = read_end_state(str(c / "proc00" / "end_state.dat"))
    assert (p_c, t_c, imt_c, dt_c) == (p_a, t_a, imt_a, dt_a)
    assert np.array_equal(out_c, out_a), float(np.abs(out_c - out_a).max())
    print(f"{scheme} {mode}: stopped at p = {p_b}, t = {t_b:.4f}; restarted run equals the uninterrupted one "
          f"bit for bit at p = {p_c}, t = {t_c:.6f}")


@pytest.mark.gpu
def test_fresh_run_with_checkpoint_present_stops(tmp_path):
    """restart = .false. while end_state00.dat exists: ERROR after the
    initial diagnostics (src/gpe.f90:116-124)."""
    p, psi = small_case("rk4")
    d = fresh_dir(tmp_path, "d", p, psi)
    write_end_state(str(d / "end_state00.dat"), psi, 0, 0.0, 0.0, p.dt0())
    r = run(d, run_in_text(p, 1.0, restart=False))
    assert r.returncode != 0
    assert "restart=.false. but end_state00.dat exists." in (d / "ERROR").read_text()


@pytest.mark.gpu
def test_saved_restart_multiplies_filtered_state(tmp_path):
    """saved_restart: the restarted state is the checkpoint times the field
    read from proc00/end_state_filtered.dat (src/ic.f90:143-159).  The
    reference's reader skips 4 integers and 2 reals, while filtered_surface
    writes nx, ny, nz, p, t, im_t, dt and the field (src/io.f90:608-617), so
    the field it reads starts with dt and drops the last value; the host
    reads it the same way."""
    p, psi = small_case("rk4")
    rng = np.random.default_rng(5)
    filt = np.ascontiguousarray(0.9 + 0.1 * rng.random(p.shape) + 0.05j * rng.random(p.shape))
    fdt = 0.95 + 0.02j
    (tmp_path / "proc00").mkdir()
    write_end_state(str(tmp_path / "end_state00.dat"), psi, 3, 3 * p.tau, 0.0, p.dt0())
    write_end_state(str(tmp_path / "proc00" / "end_state_filtered.dat"), filt, 7, 1.0, 0.0, fdt)
    read_as = np.concatenate([[fdt], filt.ravel()[:-1]]).reshape(p.shape)
    text = set_namelist(run_in_text(p, (3 + 2 - 0.5) * p.tau, restart=True), saved_restart=True)
    r = run(tmp_path, text)
    assert r.returncode == 0, r.stdout + r.stderr
    out, pp, t, _, _ = read_end_state(str(tmp_path / "proc00" / "end_state.dat"))
    assert pp == 5
    with Context(p) as ctx:
        ctx.upload(np.ascontiguousarray(psi * read_as))
        ctx.step("rk4", 2, StepInfo(p.dt0(), 3 * p.tau))
        exp = ctx.download()
    assert np.abs(out - exp).max() <= 1e-14 * np.abs(exp).max()


# ---------------------------------------------------------------------------
# GPU: the imaginary-to-real-time switch


def trap_case(scheme, tau):
    p = RunParams(nx=16, ny=16, nz=16, xr=3.0, yr=3.0, zr=3.0, order=2, bcs=2, eqn_to_solve=4, scheme=scheme,
                  tau=tau, real_time=False, g=1.0, mu=2.5, omx=1.0, omy=1.0, omz=1.0)
    return p, np.ascontiguousarray(ics.fermi(p) + 0.0j)


def ground_state_case(tau=0.01, n=16, xr=3.0, om=1.0, pert=1e-12):
    """The linear trap (g = 0) started from its discrete ground state: the
    product of the 1D ground states of -1/2 D2 + om^2 x^2 / 2 (order 2,
    reflective) is an eigenvector of the 3D operator, and mu = E0 makes it
    stationary, so imaginary time converges at once and RKF45's dt grows to
    where real time has to retry.  A 1e-12 perturbation keeps the errors
    away from zero."""
    p = RunParams(nx=n, ny=n, nz=n, xr=xr, yr=xr, zr=xr, order=2, bcs=2, eqn_to_solve=4, scheme="rk45",
                  tau=tau, real_time=False, g=0.0, mu=0.0, omx=om, omy=om, omz=om)
    dx = 2 * xr / (n - 1)
    x = -xr + dx * np.arange(n)
    d2 = np.zeros((n, n))
    for i in range(n):
        for j, w in ((i - 1, 1.0), (i, -2.0), (i + 1, 1.0)):
            j = -j if j < 0 else (2 * (n - 1) - j if j >= n else j)  # even reflection
            d2[i, j] += w / dx ** 2
    ev, vec = np.linalg.eig(-0.5 * d2 + np.diag(0.5 * om ** 2 * x ** 2))
    k = int(np.argmin(ev.real))
    v = vec[:, k].real / vec[np.argmax(np.abs(vec[:, k])), k].real
    psi = np.einsum("k,j,i->kji", v, v, v).astype(complex)
    rng = np.random.default_rng(1)
    psi = psi + pert * (rng.standard_normal(psi.shape) + 1j * rng.standard_normal(psi.shape))
    return p.replace(mu=3 * float(ev[k].real)), np.ascontiguousarray(psi)


def drive(p, psi, end_time, stop_imag, step, norm, renorm):
    """The eqn-4 part of the reference's main loop (src/gpe.f90:134-278):
    before the loop the norm and renormalisation, then per iteration a step,
    the norm, relnorm, the switch and renorm.  step(psi, dt, t, im_t,
    real_time) -> (psi, dt, t, im_t, n_reject).  Returns the final state,
    counters and the switch iteration."""
    t, im_t, dt, real_time = 0.0, 0.0, p.dt0(), False
    prev = norm(psi)
    psi = renorm(psi, prev)
    pp, p_switch, rej_after = 0, None, 0
    while t + im_t <= end_time:
        psi, dt, t, im_t, nrej = step(psi, dt, t, im_t, real_time)
        if p_switch is not None:
            rej_after += nrej
        nrm = norm(psi)
        rel = abs((nrm - prev) / prev)
        if not real_time:
            if rel < 1e-12:
                if stop_imag:
                    p_switch = pp
                    break
                real_time, im_t = True, 0.0
                dt = complex(abs(dt.imag), abs(dt.real))
                p_switch = pp
            psi = renorm(psi, nrm)
        prev = nrm
        pp += 1
    return psi, pp, t, im_t, dt, p_switch, rej_after


def gpu_drive(p, psi, end_time, stop_imag):
    with Context(p) as ctx:
        ctx.upload(psi)

        def step(_, dt, t, im_t, real_time):
            ctx.set_real_time(real_time)
            info = StepInfo(dt, t, im_t)
            ctx.step(p.scheme, 1, info)
            return None, info.dt, info.t, info.im_t, info.n_reject

        def renorm(_, nrm):
            ctx.renormalise(nrm)

        out = drive(p, None, end_time, stop_imag, step, lambda _: ctx.norm(), renorm)
        return (ctx.download(),) + out[1:]


def oracle_drive(orc, p, psi, end_time, stop_imag):
    def step(psi, dt, t, im_t, real_time):
        q = p.replace(real_time=real_time)
        if p.scheme == "rk45":
            new, dtn, (t, im_t), _, nrej, _ = orc.rkqs(q, psi, dt, (t, im_t))
            return new, dtn, t, im_t, nrej
        new, (t, im_t) = orc.step(q, "rk4", psi, dt, (t, im_t))
        return new, dt, t, im_t, 0

    return drive(p, psi, end_time, stop_imag, step, lambda s: orc.norm(p, s), lambda s, n: orc.renormalise(p, s, n))


def host_run(d, p, psi, end_time, mode, stop_imag=False, **nml):
    write_end_state(str(d / "ic_state.dat"), psi, 0, 0.0, 0.0, p.dt0())
    text = set_namelist(run_in_text(p, end_time, restart=False), renorm=True, stop_imag=stop_imag,
                        host_mode=mode, save_rate=1, **nml)
    r = run(d, text)
    assert r.returncode == 0, r.stdout + r.stderr
    return r


@pytest.mark.gpu
@pytest.mark.parametrize("mode", ["resident", "parity"])
def test_imaginary_to_real_switch_rk45(tmp_path, orc, mode):
    """RKF45 in imaginary time until relnorm < 1e-12, then real time (the
    stationary start of ground_state_case; from a generic state relnorm does
    not fall that far under RKF45, whose changing dt changes the norm's decay
    per step): the
    switch step, the rotated dt, im_t = 0 afterwards, and the retries after
    the switch shrinking dt along the real axis (a retry along the imaginary
    axis would stop t and grow im_t).  Host == C-ABI drive; C-ABI drive ==
    oracle rkqs to rounding."""
    p, psi = ground_state_case()
    end_time = 1.0
    gpu = gpu_drive(p, psi, end_time, False)
    final, pp, t, im_t, dt, p_switch, rej_after = gpu
    assert p_switch is not None and 0 < p_switch < pp - 5, "the run must cross the switch"
    assert rej_after >= 1, "a retry after the switch exercises rkqs in real time"
    assert im_t == 0.0 and dt.imag == 0.0 and dt.real > 0.0 and t > end_time
    # the oracle's rkqs with the reference's real_time flag
    o_final, o_pp, o_t, o_im_t, o_dt, o_switch, o_rej = oracle_drive(orc, p, psi, end_time, False)
    assert (o_pp, o_switch, o_rej) == (pp, p_switch, rej_after)
    # the error estimate is the Cash-Karp error of the 1e-12 perturbation on a
    # stationary state: rounding at 1e-16 moves errmax at ~1e-4 relative and
    # each dt at ~2e-5 (dt ~ errmax^-1/5)
    assert abs(o_t - t) <= 1e-6 * t and o_im_t == 0.0 and abs(o_dt - dt) <= 1e-4 * abs(dt)
    err = float(np.linalg.norm(final - o_final) / np.linalg.norm(o_final))
    assert err <= 1e-9, err
    # the host
    d = tmp_path / "h"
    d.mkdir()
    r = host_run(d, p, psi, end_time, mode)
    assert "Switching to real time" in r.stdout
    out, hp, ht, him_t, hdt = read_end_state(str(d / "proc00" / "end_state.dat"))
    assert (hp, ht, him_t, hdt) == (pp, t, im_t, dt)
    assert np.abs(out - final).max() <= 1e-13 * np.abs(final).max()
    # timestep.dat (t, im_t, |dt|, p): after the switch im_t stays 0 and t grows
    ts = np.array([[float(v) for v in ln.split()] for ln in open(d / "timestep.dat")])
    after = ts[ts[:, 3] > p_switch]
    assert len(after) >= 2 and np.all(after[:, 1] == 0.0) and np.all(np.diff(after[:, 0]) > 0)
    print(f"rk45 {mode}: switch at p = {p_switch}, {rej_after} retries after it, final p = {pp}, "
          f"t = {t:.6f}, dt = {dt.real:.5f}; GPU vs oracle rel L2 {err:.2e}")


@pytest.mark.gpu
@pytest.mark.parametrize("mode", ["resident", "parity"])
def test_imaginary_to_real_switch_rk4_schedule(tmp_path, mode):
    """RK4 across the switch with save_3d: the 3D save at the switch and the
    reset save_rate2 counter (a save at the next iteration) appear in
    p_saved.dat; dt is rotated to (tau, 0) and im_t restarts at 0."""
    p, psi = trap_case("rk4", 0.02)
    end_time = 9.0  # relnorm < 1e-12 near im_t = 8.7 (p = 435); then t runs to 9
    final, pp, t, im_t, dt, p_switch, _ = gpu_drive(p, psi, end_time, False)
    assert p_switch is not None and dt == complex(p.tau, 0.0) and im_t == 0.0
    d = tmp_path / "h"
    d.mkdir()
    host_run(d, p, psi, end_time, mode, save_3d=True)
    out, hp, ht, him_t, hdt = read_end_state(str(d / "proc00" / "end_state.dat"))
    assert (hp, ht, him_t, hdt) == (pp, t, im_t, dt)
    assert np.abs(out - final).max() <= 1e-13 * np.abs(final).max()
    saved = [int(v) for v in open(d / "p_saved.dat").read().split()]
    # before the loop, first iteration (n = 0), at the switch, the iteration
    # after it (n reset to int(t/save_rate2) = 0), the final state
    assert saved == [0, 0, p_switch, p_switch + 1, pp]
    dd = read_dens(str(d / "proc00" / f"dens{p_switch + 1:07d}.dat"))
    assert dd["time"] == 0.0  # t + im_t right after the switch


@pytest.mark.gpu
def test_stop_imag_ends_at_the_switch(tmp_path):
    """stop_imag: the run ends when relnorm < 1e-12, in imaginary time."""
    p, psi = trap_case("rk4", 0.02)
    end_time = 1e3
    final, pp, t, im_t, dt, p_switch, _ = gpu_drive(p, psi, end_time, True)
    assert p_switch == pp and im_t > 0 and dt.real == 0.0
    d = tmp_path / "h"
    d.mkdir()
    r = host_run(d, p, psi, end_time, "resident", stop_imag=True)
    assert "Switching to real time" not in r.stdout
    out, hp, ht, him_t, hdt = read_end_state(str(d / "proc00" / "end_state.dat"))
    assert (hp, ht, him_t, hdt) == (pp, t, im_t, dt)
    assert np.abs(out - final).max() <= 1e-13 * np.abs(final).max()
