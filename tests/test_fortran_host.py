"""The Fortran host (gpe_amd/fortran): the reference-compatible modules
solve/derivs/variables over the ISO_C_BINDING layer, and a driver with the
reference's main loop.  GPU tests run the driver on the inputs of the
reference-build fixtures and compare its end_state.dat with the reference's
final state; the CPU test checks the error convention (emergency_stop writes
ERROR and removes RUNNING) when no device is available."""
import json
import os
import subprocess

import numpy as np
import pytest

from gpe_amd import RunParams, ics
from gpe_amd.io import read_end_state, run_in_text, write_end_state

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
    out, pp, t, im_t, dt = read_end_state(str(tmp_path / "end_state.dat"))
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
    """The host's final-state diagnostics (save_pdf, save_vcf, save_mom)
    written in the formats of save_velocity_pdf, save_vel_corr and
    save_momentum (src/io.f90:222-311, 426-445), against the reference's
    routines on the same state.  A run of zero steps keeps the input state.
    The files hold 9 significant digits (e17.9), so values compare to 1e-8
    relative; the PDF to one count per bin as in the device test."""
    p, psi, z = load_fixture(name)
    nbins = int(z["nbins"])
    text = run_in_text(p, -0.5 * p.tau, nbins=nbins)
    text = text.replace("save_pdf = .false.", "save_pdf = .true.").replace("save_vcf = .false.", "save_vcf = .true.")
    text = text.replace("host_mode = 'resident'", "host_mode = 'resident'\n  save_mom = .true.")
    (tmp_path / "run.in").write_text(text)
    write_end_state(str(tmp_path / "end_state00.dat"), psi, 0, 0.0, 0.0, p.dt0())
    r = subprocess.run([EXE], cwd=tmp_path, capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    mom = np.array([[fnum(v) for v in ln.split()] for ln in open(tmp_path / "momentum.dat")])
    assert mom.shape == (1, 4) and mom[0, 0] == 0.0
    scale = np.abs(z["mom"]).max()
    assert np.all(np.abs(mom[0, 1:] - z["mom"]) <= 1e-8 * scale + 1e-12)
    head, rows = read_pdf_file(tmp_path / "pdf_vel0000000.dat")
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
    vcf = np.array([[fnum(v) for v in ln.split()] for ln in open(tmp_path / "vcf0000000.dat") if "#" not in ln])
    assert np.array_equal(vcf[:, 0], np.arange(p.nx))
    assert np.allclose(vcf[:, 1], z["vcf"], rtol=1e-8, atol=1e-9)
