"""Writes tests/golden/ref_*.npz: outputs of the reference Fortran
time-stepping path (compiled unmodified by oracle/build_ref.sh into
oracle/_ref) on seeded inputs.  Each stepping fixture holds the input field,
the run parameters, the reference's final field, t, im_t, dt, p, its norm
and energy (get_norm, energy) and the per-step (t, im_t, dt) trace.  The
ref_diag_* fixtures hold an input field and the reference's momentum,
velocity statistics and PDF (nbins = 64), velocity correlation function and
line length.
The ref_ll_* fixtures hold inputs and the reference's line length for the
branches smooth fields do not reach.  The ref_ic_* fixtures hold the reference's initial conditions (its ics with
the expressions of oracle/ref/ic_cases.txt as ic.in) on a 16 x 12 x 10 grid.

Run where the reference sources and their build (oracle/build_ref.sh) are present:
    python tests/golden/make_ref_fixtures.py
"""
from __future__ import annotations

import dataclasses
import json
import os
import shutil
import subprocess
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from gpe_amd import RunParams, ics  # noqa: E402
from gpe_amd.io import read_end_state, run_in_text, write_end_state  # noqa: E402

REF = os.path.join(ROOT, "oracle", "_ref")
OUT = os.path.dirname(os.path.abspath(__file__))


def smooth_field(p, seed):
    rng = np.random.default_rng(seed)
    x, y, z = p.axes()
    Z, Y, X = np.meshgrid(z, y, x, indexing="ij")
    psi = np.ones(p.shape, dtype=np.complex128)
    for _ in range(4):
        k = 2 * np.pi * rng.integers(-2, 3, size=3) / np.array([p.nx * p.dx, p.ny * p.dy, p.nz * p.dz])
        psi += 0.2 * (rng.standard_normal() + 1j * rng.standard_normal()) * np.exp(1j * (k[0] * X + k[1] * Y + k[2] * Z))
    return np.ascontiguousarray(psi)


def run_reference(name, p, psi, nsteps=None, end_time=None):
    exe = os.path.join(REF, f"{p.nx}x{p.ny}x{p.nz}", "ref_driver")
    if not os.path.exists(exe):
        raise SystemExit(f"{exe} missing: run oracle/build_ref.sh first")
    run = os.path.join(REF, "run_" + name)
    shutil.rmtree(run, ignore_errors=True)
    os.makedirs(run)
    if end_time is None:
        end_time = (nsteps - 0.5) * p.tau  # exactly nsteps steps of the loop t+im_t <= end_time
    with open(os.path.join(run, "run.in"), "w") as f:
        f.write(run_in_text(p, end_time))
    write_end_state(os.path.join(run, "in.dat"), psi, 0, 0.0, 0.0, p.dt0())
    t0 = time.perf_counter()
    # the reference keeps its work arrays on the stack (automatic arrays)
    subprocess.run(["bash", "-c", f"ulimit -s unlimited && exec {exe} in.dat out.dat trace.dat"], cwd=run, check=True,
                   env=dict(os.environ, OMP_NUM_THREADS="1"))
    secs = time.perf_counter() - t0
    out, pp, t, im_t, dt, extra = read_end_state(os.path.join(run, "out.dat"), 2)
    tr = np.fromfile(os.path.join(run, "trace.dat"), dtype="<f8").reshape(-1, 4)
    shutil.rmtree(run)
    print(f"{name}: {pp} steps, t={t:.6f}, im_t={im_t:.6f}, {secs:.1f} s", flush=True)
    return dict(out=out, p=pp, t=t, im_t=im_t, dt=dt, norm=extra[0], energy=extra[1], trace=tr, secs=secs)


def save(name, p, psi, res, store_input=True):
    d = dict(params=json.dumps(dataclasses.asdict(p)), out=res["out"], p=res["p"], t=res["t"], im_t=res["im_t"],
             dt=res["dt"], norm=res["norm"], energy=res["energy"], trace=res["trace"], ref_seconds=res["secs"])
    if store_input:
        d["inp"] = psi
    np.savez_compressed(os.path.join(OUT, f"ref_{name}.npz"), **d)


def small_cases():
    base = dict(nx=24, ny=20, nz=18, xr=8.0, yr=7.0, zr=6.5)
    cases = [
        ("eq1_per_o2_rk4", RunParams(**base, order=2, bcs=1, eqn_to_solve=1, scheme="rk4", tau=0.01), 20, 1),
        ("eq4_ref_o4_rk4", RunParams(**base, order=4, bcs=2, eqn_to_solve=4, scheme="rk4", tau=0.005, g=1.0, mu=2.0,
                                     omx=0.3, omy=0.25, omz=0.2), 20, "trap"),
        ("eq2_per_o4_euler", RunParams(**base, order=4, bcs=1, eqn_to_solve=2, scheme="euler", tau=0.002), 10, 3),
        ("eq1_ref_o2_rk2_u_diss", RunParams(**base, order=2, bcs=2, eqn_to_solve=1, scheme="rk2", tau=0.01, Urhs=0.3,
                                            diss_amp=0.02), 10, 4),
        ("eq4_ref_o2_imag_rk4", RunParams(**base, order=2, bcs=2, eqn_to_solve=4, scheme="rk4", tau=0.005, g=1.0,
                                          mu=2.0, omx=0.3, omy=0.3, omz=0.3, real_time=False), 10, "trap"),
        ("eq3_per_o4_rk4", RunParams(nx=24, ny=20, nz=18, xr=14.0, yr=13.0, zr=15.0, order=4, bcs=1, eqn_to_solve=3,
                                     scheme="rk4", tau=0.01), 10, 6),
    ]
    for name, p, n, seed in cases:
        if seed == "trap":
            psi = np.ascontiguousarray(ics.fermi(p) * ics.vortex_line_z(p) + 0.01)
        else:
            psi = smooth_field(p, seed)
        save(name, p, psi, run_reference(name, p, psi, nsteps=n))
    # rk45 from a too-large first dt (rejections), to t ~ 0.6
    p = RunParams(**base, order=4, bcs=1, eqn_to_solve=1, scheme="rk45", tau=0.4)
    psi = smooth_field(p, 7)
    save("eq1_per_o4_rk45", p, psi, run_reference("eq1_per_o4_rk45", p, psi, end_time=0.6))
    p = RunParams(**base, order=2, bcs=2, eqn_to_solve=4, scheme="rk45", tau=0.05, g=1.0, mu=2.0, omx=0.3, omy=0.25,
                  omz=0.2)
    psi = np.ascontiguousarray(ics.fermi(p) * ics.vortex_line_z(p) + 0.01)
    save("eq4_ref_o2_rk45", p, psi, run_reference("eq4_ref_o2_rk45", p, psi, end_time=0.5))


def run_reference_diag(name, p, psi, nbins):
    """The reference's momentum, velocity PDF and velocity correlation
    function of psi (oracle/ref/ref_diag.f90)."""
    exe = os.path.join(REF, f"{p.nx}x{p.ny}x{p.nz}", "ref_diag")
    if not os.path.exists(exe):
        raise SystemExit(f"{exe} missing: run oracle/build_ref.sh first")
    run = os.path.join(REF, "run_" + name)
    shutil.rmtree(run, ignore_errors=True)
    os.makedirs(run)
    with open(os.path.join(run, "run.in"), "w") as f:
        f.write(run_in_text(p, 0.0, nbins=nbins))
    write_end_state(os.path.join(run, "in.dat"), psi, 0, 0.0, 0.0, p.dt0())
    subprocess.run(["bash", "-c", f"ulimit -s unlimited && exec {exe} in.dat out.dat"], cwd=run, check=True,
                   env=dict(os.environ, OMP_NUM_THREADS="1"))
    raw = open(os.path.join(run, "out.dat"), "rb").read()
    shutil.rmtree(run)
    o = 0
    head = np.frombuffer(raw, "<f8", 12, o); o += 96
    counts = np.frombuffer(raw, "<i4", 3, o); o += 12
    pdf = np.frombuffer(raw, "<f8", 3 * nbins, o).reshape(3, nbins); o += 24 * nbins
    vcf = np.frombuffer(raw, "<f8", p.nx, o); o += 8 * p.nx
    linelength = float(np.frombuffer(raw, "<f8", 1, o)[0]); o += 8
    assert o == len(raw), (o, len(raw))
    print(f"{name}: momentum {head[:3]}, valid {counts}", flush=True)
    return dict(mom=head[0:3].copy(), vmean=head[3:6].copy(), vstdev=head[6:9].copy(), vmax=head[9:12].copy(),
                counts=counts.copy(), pdf=pdf.copy(), vcf=vcf.copy(), linelength=linelength)


def diag_cases():
    base = dict(nx=24, ny=20, nz=18, xr=8.0, yr=7.0, zr=6.5)
    nbins = 64
    cases = []
    p = RunParams(**base, order=2, bcs=1, eqn_to_solve=1)
    cases.append(("diag_per_o2_smooth", p, smooth_field(p, 11)))
    p = RunParams(**base, order=4, bcs=2, eqn_to_solve=4, g=1.0, mu=2.0, omx=0.3, omy=0.25, omz=0.2)
    cases.append(("diag_ref_o4_trap_line", p, np.ascontiguousarray(ics.fermi(p) * ics.vortex_line_z(p) + 0.01)))
    p = RunParams(nx=24, ny=20, nz=18, xr=0.15 * 23, yr=0.15 * 19, zr=0.15 * 17, order=4, bcs=1, eqn_to_solve=1)
    cases.append(("diag_per_o4_turb", p, ics.random_phase(p, seed=3)))
    p, psi = ics.config_ring64(seed=0)
    cases.append(("diag_ring64", p, psi))
    for name, p, psi in cases:
        res = run_reference_diag(name, p, psi, nbins)
        # the 64^3 ring input is regenerated by ics.config_ring64(seed=0)
        extra = {} if name == "diag_ring64" else dict(inp=np.ascontiguousarray(psi))
        np.savez_compressed(os.path.join(OUT, f"ref_{name}.npz"), params=json.dumps(dataclasses.asdict(p)),
                            nbins=nbins, **extra, **res)


# Run parameters of each initial-condition case of oracle/ref/ic_cases.txt:
# grid extents, scal (eqn_params) and the trap of fermi.
IC_GRID = dict(nx=16, ny=12, nz=10)
IC_CASES = {
    "ring_xy": dict(xr=4.0, yr=3.5, zr=3.0, scal=1.0),
    "ring_xz": dict(xr=4.0, yr=3.5, zr=3.0, scal=1.3),
    "ring_yz": dict(xr=4.5, yr=4.0, zr=3.5, scal=1.0),
    "line_x": dict(xr=4.0, yr=3.0, zr=2.5, scal=1.0),
    "line_y_phase": dict(xr=4.0, yr=3.0, zr=2.5, scal=0.8),
    "line_z": dict(xr=4.0, yr=3.0, zr=2.5, scal=1.2),
    "trap_line": dict(xr=6.0, yr=5.0, zr=4.0, scal=1.0, mu=4.0, g=1.5, omx=0.6, omy=0.5, omz=0.7),
    "pade_ring": dict(xr=6.0, yr=5.0, zr=4.5, scal=1.0),
    "pade_pulse": dict(xr=6.0, yr=5.0, zr=4.5, scal=1.0),
    "pair_sphere": dict(xr=15.0, yr=12.0, zr=14.0, scal=1.0),
    "wall_sphere2": dict(xr=6.0, yr=5.0, zr=40.0, scal=1.0),
}


def ic_cases():
    for name, kw in IC_CASES.items():
        kw = dict(kw)
        scal = kw.pop("scal")
        p = RunParams(**IC_GRID, **kw)
        exe = os.path.join(REF, "ic", name, "ref_ic")
        if not os.path.exists(exe):
            raise SystemExit(f"{exe} missing: run oracle/build_ref.sh first")
        run = os.path.join(REF, "run_ic_" + name)
        shutil.rmtree(run, ignore_errors=True)
        os.makedirs(run)
        text = run_in_text(p, 0.0, restart=False).replace("scal = 1.0", f"scal = {scal!r}")
        with open(os.path.join(run, "run.in"), "w") as f:
            f.write(text)
        subprocess.run([exe, "out.dat"], cwd=run, check=True, env=dict(os.environ, OMP_NUM_THREADS="1"))
        out = np.fromfile(os.path.join(run, "out.dat"), dtype="<c16").reshape(p.nz, p.ny, p.nx)
        shutil.rmtree(run)
        print(f"ic {name}: max|psi| {np.abs(out).max():.4f}", flush=True)
        np.savez_compressed(os.path.join(OUT, f"ref_ic_{name}.npz"), params=json.dumps(dataclasses.asdict(p)),
                            scal=scal, out=out)


def ll_cases():
    """Line-length fixtures for the branches smooth fields never reach: white
    noise (cells with a single hit, which reuse the previous cell's second
    hit), a vortex line along a grid edge (the edge branch) and a zero plane
    of Re psi (the face branch).  The grid spacing is 1, so the zeros fall
    exactly on grid points."""
    p = RunParams(nx=24, ny=20, nz=18, xr=11.5, yr=9.5, zr=8.5)
    x, y, z = p.axes()
    Z, Y, X = np.meshgrid(z, y, x, indexing="ij")
    # seed 8: the first cell with a hit has two (the reference's hit slots
    # are undefined before the first such cell)
    rng = np.random.default_rng(8)
    cases = [("noise", rng.standard_normal(p.shape) + 1j * rng.standard_normal(p.shape)),
             ("edge_line", (X - 0.5) + 1j * (Y - 0.5)),
             ("zero_plane", (X - 0.5) + 1j * Y + 0.3j * np.sin(0.7 * Z))]
    for name, psi in cases:
        psi = np.ascontiguousarray(psi, dtype=np.complex128)
        res = run_reference_diag("ll_" + name, p, psi, 64)
        np.savez_compressed(os.path.join(OUT, f"ref_ll_{name}.npz"), params=json.dumps(dataclasses.asdict(p)),
                            inp=psi, linelength=res["linelength"])
        print(f"ll {name}: {res['linelength']!r}", flush=True)


def ring64():
    p, psi = ics.config_ring64(seed=0)
    res = run_reference("ring64_rk4_200", p, psi, nsteps=200)
    save("ring64_rk4_200", p, psi, res, store_input=False)


if __name__ == "__main__":
    which = sys.argv[1:] or ["small", "ring64", "diag", "ic", "ll"]
    if "small" in which:
        small_cases()
    if "ring64" in which:
        ring64()
    if "diag" in which:
        diag_cases()
    if "ic" in which:
        ic_cases()
    if "ll" in which:
        ll_cases()
