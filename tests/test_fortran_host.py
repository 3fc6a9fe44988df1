"""The Fortran host (gpe_amd/fortran): the reference-compatible modules
solve/derivs/variables over the ISO_C_BINDING layer, and a driver with the
reference's main loop.  GPU tests run the driver on the inputs of the
reference-build fixtures and compare its checkpoint proc00/end_state.dat with the reference's
final state; the CPU test checks the error convention (emergency_stop writes
ERROR and removes RUNNING) when no device is available."""
import json
import os
import subprocess

import numpy as np
import pytest

from gpe_amd import Context, RunParams, StepInfo, ics
from gpe_amd.io import read_dens, read_end_state, run_in_text, set_namelist, write_end_state

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "gpe_amd", "fortran", "gpe_hip_main")
GOLD = os.path.join(ROOT, "tests", "golden")


def run_host(tmp_path, p, psi, nsteps, mode="resident"):
    text = run_in_text(p, (nsteps - 0.5) * p.tau).replace("host_mode = 'resident'", f"host_mode = '{mode}'")
    (tmp_path / "run.in").write_text(text)
    write_end_state(str(tmp_path / "end_state00.dat"), psi, 0, 0.0, 0.0, p.dt0())
    r = subprocess.run([EXE], cwd=tmp_path, capture_output=True, text=True, timeout=300)
    return r


def test_host_binary_built():
    assert os.access(EXE, os.X_OK), "build with __graft_entry__.build() (make -C gpe_amd/fortran)"


def test_host_error_convention_without_device(tmp_path):
    """No GPU here: context creation fails and the host stops the way the
    reference does (message, ERROR file, RUNNING removed, non-zero exit)."""
    try:
        import torch  # noqa: F401  (only to see whether a device exists)
        has_gpu = torch.cuda.is_available()
    except Exception:
        has_gpu = False
    if has_gpu:
        pytest.skip("a GPU is present; the failure path is not reachable")
    p = RunParams(nx=16, ny=16, nz=16, xr=4.0, yr=4.0, zr=4.0)
    r = run_host(tmp_path, p, np.ones(p.shape, dtype=np.complex128), 2)
    assert r.returncode != 0
    assert (tmp_path / "ERROR").exists()
    assert "ERROR" in (tmp_path / "ERROR").read_text()
    assert not (tmp_path / "RUNNING").exists()


def load_fixture(name):
    z = np.load(os.path.join(GOLD, f"ref_{name}.npz"), allow_pickle=False)
    p = RunParams(**json.loads(str(z["params"])))
    psi = z["inp"] if "inp" in z.files else ics.config_ring64(seed=0)[1]
    return p, np.ascontiguousarray(psi), z


@pytest.mark.gpu
@pytest.mark.parametrize("name,mode", [("ring64_rk4_200", "resident"), ("eq4_ref_o4_rk4", "parity"),
                                       ("eq1_ref_o2_rk2_u_diss", "parity"), ("eq4_ref_o2_imag_rk4", "resident")])
def test_host_matches_reference_build(tmp_path, name, mode):
    p, psi, z = load_fixture(name)
    nsteps = int(z["p"])
    r = run_host(tmp_path, p, psi, nsteps, mode)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
    out, pp, t, im_t, dt = read_end_state(str(tmp_path / "proc00" / "end_state.dat"))
    assert pp == nsteps
    err = float(np.linalg.norm(out - z["out"]) / np.linalg.norm(z["out"]))
    print(f"Fortran host ({mode}) vs reference build, {name}: rel L2 = {err:.3e}")
    assert err <= 1e-12
    assert abs(t - float(z["t"])) <= 1e-12 and abs(im_t - float(z["im_t"])) <= 1e-12
    assert not (tmp_path / "RUNNING").exists()


@pytest.mark.gpu
@pytest.mark.parametrize("name", ["ring64_rk4_200", "eq4_ref_o4_rk4", "eq1_per_o4_rk45", "eq2_per_o4_euler"])
def test_reference_program_with_gpu_dropin(tmp_path, name):
    """oracle/_ref/<grid>/dropin/ref_driver_gpu is built from the reference's
    own parameters/error/ic/variables modules and its main-loop sequence
    (copy, MPI ghost exchange, solver) with this package's solve and derivs
    modules in place of the reference's: the drop-in inside the reference's
    program.  It must reproduce the reference build's results."""
    p, psi, z = load_fixture(name)
    exe = os.path.join(ROOT, "oracle", "_ref", f"{p.nx}x{p.ny}x{p.nz}", "dropin", "ref_driver_gpu")
    if not os.path.exists(exe):
        pytest.skip("built by oracle/build_ref.sh where the reference sources are present")
    nsteps = int(z["p"])
    if p.scheme == "rk45":
        # the loop ran while t <= end_time: end_time lies in [t_{n-1}, t_n)
        end_time = float(z["trace"][-2, 0]) + 1e-12
    else:
        end_time = (nsteps - 0.5) * p.tau
    (tmp_path / "run.in").write_text(run_in_text(p, end_time))
    write_end_state(str(tmp_path / "in.dat"), psi, 0, 0.0, 0.0, p.dt0())
    r = subprocess.run(["bash", "-c", f"ulimit -s unlimited && exec {exe} in.dat out.dat trace.dat"], cwd=tmp_path,
                       capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    out, pp, t, im_t, dt, extra = read_end_state(str(tmp_path / "out.dat"), 2)
    assert pp == nsteps
    err = float(np.linalg.norm(out - z["out"]) / np.linalg.norm(z["out"]))
    print(f"reference program + GPU drop-in, {name}: rel L2 = {err:.3e}")
    assert err <= 1e-12
    # rk45: each dt comes from an error norm formed by cancellation between
    # the stage increments, so rounding-level differences in f(psi) move dt
    # (and the sum t) at the 1e-11 relative level; fixed steps add exact dts
    t_tol = 1e-10 * abs(float(z["t"])) if p.scheme == "rk45" else 1e-12
    assert abs(t - float(z["t"])) <= t_tol
    assert abs(extra[0] - float(z["norm"])) <= 1e-12 * float(z["norm"])


def fnum(v):
    """A Fortran Ew.d field: with a three-digit exponent the E is dropped
    (0.864549105-315)."""
    try:
        return float(v)
    except ValueError:
        k = max(v.rfind("-"), v.rfind("+"))
        return float(v[:k] + "E" + v[k:])


def read_pdf_file(path):
    head, rows = {}, []
    for line in open(path):
        line = line.strip()  # the headers are right-justified (a10)
        if line.startswith("#"):
            key, *vals = line[1:].split()
            head[key.rstrip(":")] = np.array([fnum(v) for v in vals])
        else:
            rows.append([fnum(v) for v in line.split()])
    return head, np.array(rows)


@pytest.mark.gpu
@pytest.mark.parametrize("name", ["diag_ring64", "diag_per_o4_turb", "diag_ref_o4_trap_line"])
def test_host_diagnostic_files_match_reference(tmp_path, name):
    """The host's diagnostics at the first iteration (save_pdf, save_vcf at
    save_rate3, momentum with the save_rate time series) in the formats of
    save_velocity_pdf, save_vel_corr and save_momentum (src/io.f90:222-311,
    426-445), against the reference's routines on the same state: a run of
    one step writes them for the input state (p = 0).  The files hold 9
    significant digits (e17.9), so values compare to 1e-8 relative; the PDF
    to one count per bin as in the device test."""
    p, psi, z = load_fixture(name)
    nbins = int(z["nbins"])
    text = run_in_text(p, 0.5 * p.tau, nbins=nbins)
    text = set_namelist(text, save_pdf=True, save_vcf=True, save_mom=True, save_rate=1, save_rate3=1.0)
    (tmp_path / "run.in").write_text(text)
    write_end_state(str(tmp_path / "end_state00.dat"), psi, 0, 0.0, 0.0, p.dt0())
    r = subprocess.run([EXE], cwd=tmp_path, capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    mom = np.array([[fnum(v) for v in ln.split()] for ln in open(tmp_path / "momentum.dat")])
    assert mom.shape == (1, 4) and mom[0, 0] == 0.0
    scale = np.abs(z["mom"]).max()
    assert np.all(np.abs(mom[0, 1:] - z["mom"]) <= 1e-8 * scale + 1e-12)
    head, rows = read_pdf_file(tmp_path / "pdf" / "pdf_vel0000000.dat")
    assert rows.shape == (nbins, 9)
    assert np.allclose(head["Mean"], z["vmean"], rtol=1e-8, atol=1e-8 * np.abs(z["vstdev"]).max())
    assert np.allclose(head["StDev"], z["vstdev"], rtol=1e-8)
    assert np.allclose(head["Sum"], 1.0, rtol=1e-8)
    for a in range(3):
        i = np.arange(-nbins // 2 + 1, nbins // 2 + 1)
        assert np.allclose(rows[:, 3 * a], 2.0 * i * z["vmax"][a] / nbins, rtol=1e-8, atol=1e-12)
        dc = np.abs(rows[:, 3 * a + 1] - z["pdf"][a]) * z["counts"][a]
        assert dc.max() <= 1.0 + 1e-3 and dc.sum() <= 2.0 + 1e-3
        sd, mu = z["vstdev"][a], z["vmean"][a]
        gauss = np.exp(-0.5 * (rows[:, 3 * a] - mu) ** 2 / sd ** 2) / (sd * np.sqrt(2 * np.pi))
        assert np.allclose(rows[:, 3 * a + 2], gauss, rtol=1e-6, atol=1e-12)
    vcf = read_rows(tmp_path / "vcf" / "vcf0000000.dat")
    assert np.array_equal(vcf[:, 0], np.arange(p.nx))
    assert np.allclose(vcf[:, 1], z["vcf"], rtol=1e-8, atol=1e-9)


def read_rows(path):
    return np.array([[fnum(v) for v in ln.split()] for ln in open(path) if "#" not in ln])


def schedule(nsteps, tau, save_rate, save_rate2, save_rate3, p_save):
    """The save schedule of the reference's main loop (src/gpe.f90:134-236)
    for a real-time fixed-step run from t = 0: per iteration p the flags
    (time series, save_rate3 diagnostics, save_rate2 3D save, periodic
    end_state), evaluated as the loop does on the accumulated t."""
    t, ps, n, m = 0.0, 0, 0, 0
    out = []
    for it in range(nsteps):
        ev = dict(p=it, t=t, ts=(t % (tau * save_rate)) < tau, diag=t >= save_rate3 * m, dens=t >= save_rate2 * n,
                  end=t >= p_save * ps)
        m += ev["diag"]
        n += ev["dens"]
        ps += ev["end"]
        out.append(ev)
        t += tau
    return out, t


@pytest.mark.gpu
@pytest.mark.parametrize("mode", ["resident", "parity"])
def test_host_save_schedule_and_files(tmp_path, mode):
    """The main loop's outputs follow the reference's schedule: energy.dat
    and psi_time.dat every save_rate steps (modulo(t, |dt| save_rate) < |dt|
    on the accumulated t), condensed_particles (mass.dat, spectra) before the
    loop and with the velocity PDF / VCF at save_rate3 intervals,
    dens files at save_rate2 intervals and at the end, proc00/end_state.dat and
    save.dat at p_save intervals and at the end.  Every value is checked
    against the same run driven step by step through the C ABI."""
    p = RunParams(nx=24, ny=20, nz=18, xr=8.0, yr=7.0, zr=6.5, order=4, bcs=1, eqn_to_solve=1, scheme="rk4", tau=0.01)
    psi = ics.vortex_ring(p, ics.RingParam("xy", 0.0, 0.0, 0.0, 3.0)) + 0.01
    nsteps, nbins = 25, 32
    cfg = dict(save_rate=4, save_rate2=0.07, save_rate3=0.05, p_save=0.12)
    text = run_in_text(p, (nsteps - 0.5) * p.tau, nbins=nbins)
    text = set_namelist(text, save_3d=True, save_pdf=True, save_vcf=True, save_mom=True, save_spectrum=True,
                        save_ll=True, host_mode=mode, **cfg)
    (tmp_path / "run.in").write_text(text)
    write_end_state(str(tmp_path / "end_state00.dat"), psi, 0, 0.0, 0.0, p.dt0())
    r = subprocess.run([EXE], cwd=tmp_path, capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    events, t_end = schedule(nsteps, p.tau, cfg["save_rate"], cfg["save_rate2"], cfg["save_rate3"], cfg["p_save"])
    ts_rows, dens_ps, diag_ps, cond = [], [], [], []
    with Context(p) as ctx:
        ctx.upload(psi)
        cond.append((0, 0.0, ctx.condensed_particles()))  # before the loop (src/gpe.f90:113)
        info = StepInfo(p.dt0())
        for ev in events:
            if ev["ts"]:
                c = ctx.peek(p.nx // 2, p.ny // 2, p.nz // 2)
                ts_rows.append((ev["t"], ctx.energy() / (p.nx * p.ny * p.nz), c, ctx.momentum(), ctx.linelength()))
            if ev["diag"]:
                cond.append((ev["p"], ev["t"], ctx.condensed_particles()))
                diag_ps.append((ev["p"], ctx.velocity_pdf(nbins), ctx.vcf()))
            if ev["dens"]:
                dens_ps.append((ev["p"], ev["t"], ctx.download()))
            ctx.step("rk4", 1, info)
        final = ctx.download()
    # time series
    en = read_rows(tmp_path / "energy.dat")
    pt = read_rows(tmp_path / "psi_time.dat")
    mo = read_rows(tmp_path / "momentum.dat")
    ll = read_rows(tmp_path / "linelength.dat")
    assert len(en) == len(ts_rows) == len(pt) == len(mo) == len(ll) and len(ts_rows) >= 5
    for k, (t, e, c, m, length) in enumerate(ts_rows):
        assert np.isclose(ll[k, 2], length, rtol=1e-8) and length > 0
        assert abs(en[k, 0] - t) <= 1e-8 * max(t, 1e-3) and en[k, 1] == 0.0
        assert np.isclose(en[k, 2], e, rtol=1e-8)
        exp = [c.real, c.imag, abs(c) ** 2, np.arctan2(c.imag + 1e-6, c.real)]
        assert np.allclose(pt[k, 2:], exp, rtol=1e-8, atol=1e-12)
        assert np.allclose(mo[k, 1:], m, rtol=1e-8, atol=1e-8 * np.abs(m).max())
    # save_rate3 diagnostics
    got = sorted(os.listdir(tmp_path / "pdf"))
    assert got == [f"pdf_vel{pp:07d}.dat" for pp, _, _ in diag_ps]
    for pp, st, f in diag_ps:
        head, rows = read_pdf_file(tmp_path / "pdf" / f"pdf_vel{pp:07d}.dat")
        assert np.allclose(head["Mean"], st["vmean"], rtol=1e-8, atol=1e-12)
        assert np.allclose(rows[:, 1], st["pdf"][0], rtol=1e-8, atol=1e-12)
        assert np.allclose(read_rows(tmp_path / "vcf" / f"vcf{pp:07d}.dat")[:, 1], f, rtol=1e-8, atol=1e-9)
    # condensed_particles: mass.dat rows; spectrum, idf and eta_time.dat
    ma = read_rows(tmp_path / "mass.dat")
    et = read_rows(tmp_path / "eta_time.dat")
    assert len(ma) == len(et) == len(cond)
    for k, (pp, t, c) in enumerate(cond):
        assert abs(ma[k, 0] - t) <= 1e-8 * max(t, 1e-3) and ma[k, 1] == 0.0
        assert np.allclose(ma[k, 2:], c["row"], rtol=1e-8, atol=0)
        assert np.allclose(et[k, 1:], c["eta"][:3], rtol=1e-8, atol=1e-300)
        sp = read_rows(tmp_path / "spectrum" / f"spectrum{pp:07d}.dat")
        assert np.array_equal(sp[:, 0], np.arange(1, 8)) and np.array_equal(sp[:, 1], c["nharm"])
        assert np.allclose(sp[:, 2], c["eta"], rtol=1e-8, atol=1e-300)
        idf = read_rows(tmp_path / "idf" / f"idf{pp:07d}.dat")
        assert np.array_equal(idf[:, 0], np.arange(1, p.nx // 2 + 1))
        assert np.allclose(idf[:, 1], c["F"], rtol=1e-8, atol=1e-300)
    # dens files: save_rate2 schedule plus the final state
    dens = sorted(f for f in os.listdir(tmp_path / "proc00") if f.startswith("dens"))
    want = sorted({f"dens{pp:07d}.dat" for pp, _, _ in dens_ps} | {f"dens{nsteps:07d}.dat"})
    assert dens == want
    for pp, t, a in dens_ps:
        d = read_dens(str(tmp_path / "proc00" / f"dens{pp:07d}.dat"))
        assert (d["nx"], d["ny"], d["nz"], d["js"], d["je"], d["ks"], d["ke"]) == (24, 20, 18, 0, 19, 0, 17)
        assert abs(d["time"] - t) <= 1e-15 and np.allclose(d["x"], p.axes()[0], rtol=0, atol=1e-15)
        assert np.abs(d["psi"] - a).max() <= 1e-13
    d = read_dens(str(tmp_path / "proc00" / f"dens{nsteps:07d}.dat"))
    assert np.abs(d["psi"] - final).max() <= 1e-13
    # p_saved.dat: the 3D save before the loop (src/gpe.f90:106-108), the
    # save_rate2 ones (the first at p = 0 again) and the final one
    assert [int(v) for v in open(tmp_path / "p_saved.dat").read().split()] == \
        [0] + [pp for pp, _, _ in dens_ps] + [nsteps]
    out, pp, t, _, _ = read_end_state(str(tmp_path / "proc00" / "end_state.dat"))
    assert pp == nsteps and abs(t - t_end) <= 1e-15 and np.abs(out - final).max() <= 1e-13
    assert "Periodically saved state" in (tmp_path / "save.dat").read_text()
    assert not (tmp_path / "RUNNING").exists()


@pytest.mark.gpu
def test_host_save_sentinel(tmp_path):
    """A SAVE file present during the run makes the host write the 3D state
    after the next step (save_run, src/io.f90:1566-1595), named with the p of
    that iteration, and remove SAVE."""
    p = RunParams(nx=24, ny=20, nz=18, xr=8.0, yr=7.0, zr=6.5, order=2, bcs=2, scheme="rk4", tau=0.01)
    psi = ics.vortex_line(p, ics.LineParam(0.3, 0.2, 0.0, 0.0, 0.0, 0.0, 1.0, "z", False))
    (tmp_path / "run.in").write_text(run_in_text(p, 2.5 * p.tau))
    write_end_state(str(tmp_path / "end_state00.dat"), psi, 0, 0.0, 0.0, p.dt0())
    (tmp_path / "SAVE").write_text("")
    r = subprocess.run([EXE], cwd=tmp_path, capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    assert not (tmp_path / "SAVE").exists()
    assert sorted(os.listdir(tmp_path / "proc00")) == ["dens0000000.dat", "end_state.dat"]
    with Context(p) as ctx:
        ctx.upload(psi)
        ctx.step("rk4", 1, StepInfo(p.dt0()))
        one = ctx.download()
    d = read_dens(str(tmp_path / "proc00" / "dens0000000.dat"))
    assert abs(d["time"] - p.tau) <= 1e-16 and np.abs(d["psi"] - one).max() <= 1e-13


@pytest.mark.gpu
@pytest.mark.parametrize("mode", ["resident", "parity"])
def test_host_imaginary_time_with_imprint(tmp_path, mode):
    """Eqn 4 in imaginary time with renorm and imprint_vl: after each step
    the norm, renormalise and psi <- |psi| vortex_line(vl1) (gpe.f90:241-271,
    imprint_vortex_line src/variables.f90:1148-1162); norm.dat at save_rate.
    Checked against the same sequence through the C ABI."""
    p = RunParams(nx=24, ny=20, nz=18, xr=6.0, yr=5.0, zr=4.5, order=4, bcs=2, eqn_to_solve=4, scheme="rk4",
                  tau=0.005, real_time=False, g=1.0, mu=3.0, omx=0.4, omy=0.4, omz=0.5)
    psi = ics.fermi(p) + 0.05
    vl = ics.LineParam(0.3, -0.2, 0.0, 0.2, 0.1, 4.0, 1.0, "z", False)
    nsteps = 6
    text = run_in_text(p, (nsteps - 0.5) * p.tau)
    text = set_namelist(text, renorm=True, imprint_vl=True, save_rate=2, host_mode=mode)
    text = text.replace("  save_mom = .false.", "  save_mom = .false.\n  vl1 = 0.3, -0.2, 0.0, 0.2, 0.1, 4.0, 1.0, 'z', .false.")
    (tmp_path / "run.in").write_text(text)
    write_end_state(str(tmp_path / "end_state00.dat"), psi, 0, 0.0, 0.0, p.dt0())
    r = subprocess.run([EXE], cwd=tmp_path, capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    norms = []
    with Context(p) as ctx:
        ctx.upload(psi)
        prev = ctx.norm()
        ctx.renormalise(prev)
        norms.append((0.0, prev, 0.0))
        info = StepInfo(p.dt0())
        for it in range(nsteps):
            ctx.step("rk4", 1, info)
            nrm = ctx.norm()
            rel = abs((nrm - prev) / prev)
            if (info.t + info.im_t) % (p.tau * 2) < p.tau:
                norms.append((info.im_t, nrm, rel))
            ctx.renormalise(nrm)
            ctx.imprint_line(vl)
            prev = nrm
        final = ctx.download()
    out, pp, t, im_t, _ = read_end_state(str(tmp_path / "proc00" / "end_state.dat"))
    assert pp == nsteps and abs(im_t - nsteps * p.tau) <= 1e-15
    err = np.abs(out - final).max() / np.abs(final).max()
    print(f"imprint_vl ({mode}): max rel difference {err:.2e}")
    assert err <= 1e-13
    nd = read_rows(tmp_path / "norm.dat")
    assert len(nd) == len(norms)
    for row, (im, nrm, rel) in zip(nd, norms):
        assert np.isclose(row[1], im, rtol=1e-8, atol=1e-12) and np.isclose(row[2], nrm, rtol=1e-8)
        assert np.isclose(row[3], rel, rtol=1e-6, atol=1e-12)
