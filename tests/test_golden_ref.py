"""Golden vectors from the reference itself: tests/golden/ref_*.npz were
written by the reference Fortran time-stepping path (its derivs, solve,
variables and ic modules compiled unmodified, oracle/build_ref.sh) on seeded
inputs (tests/golden/make_ref_fixtures.py).

CPU tests pin the oracle to them; GPU tests run the HIP library on the same
inputs and compare with the same fixtures (<= 1e-12 relative L2 for
complex128 storage).
"""
import glob
import json
import os

import numpy as np
import pytest

from gpe_amd import RunParams, ics

HERE = os.path.dirname(os.path.abspath(__file__))
FIXTURES = sorted(f for f in glob.glob(os.path.join(HERE, "golden", "ref_*.npz"))
                  if not os.path.basename(f).startswith(("ref_diag_", "ref_ic_", "ref_ll_")))


def load(path):
    z = np.load(path, allow_pickle=False)
    p = RunParams(**json.loads(str(z["params"])))
    if "inp" in z.files:
        psi = z["inp"]
    else:
        _, psi = ics.config_ring64(seed=0)
    return p, psi, z


def rel(a, b):
    return float(np.linalg.norm((a - b).ravel()) / np.linalg.norm(np.ravel(b)))


def oracle_run(orc, p, psi, z):
    """Replays the reference loop with the oracle: fixed steps for
    euler/rk2/rk4, and rkqs steps while t+im_t <= end_time for rk45."""
    nsteps = int(z["p"])
    if p.scheme == "rk45":
        cur, dt, t = psi, p.dt0(), (0.0, 0.0)
        dts = []
        for _ in range(nsteps):
            cur, dt, t, _, _, _ = orc.rkqs(p, cur, dt, t)
            dts.append(dt)
        return cur, t, dt, dts
    if p.scheme == "rk4":
        out, t = orc.rk4_n(p, psi, p.dt0(), nsteps)
        return out, t, p.dt0(), None
    cur, t = psi, (0.0, 0.0)
    for _ in range(nsteps):
        cur, t = orc.step(p, p.scheme, cur, p.dt0(), t)
    return cur, t, p.dt0(), None


def ids(paths):
    return [os.path.basename(x)[4:-4] for x in paths]


@pytest.mark.parametrize("path", FIXTURES, ids=ids(FIXTURES))
def test_oracle_matches_reference_build(orc, path):
    p, psi, z = load(path)
    out, t, dt, dts = oracle_run(orc, p, psi, z)
    err = rel(out, z["out"])
    assert err <= 1e-12, err
    assert abs(t[0] - float(z["t"])) <= 1e-12 and abs(t[1] - float(z["im_t"])) <= 1e-12
    # the embedded error estimate cancels |k|/|err| ~ 1e4, so 1-ulp differences
    # in f(psi) move errmax, and with it dt, at the 1e-12 level
    assert abs(dt - complex(z["dt"])) <= 1e-10 * abs(dt)
    if dts is not None:
        ref_dts = z["trace"][:, 2] + 1j * z["trace"][:, 3]
        assert np.allclose(np.array(dts), ref_dts, rtol=1e-10, atol=0)
    assert abs(orc.norm(p, out) - float(z["norm"])) <= 1e-13 * float(z["norm"])
    e = orc.energy(p, out)
    assert abs(e - float(z["energy"])) <= 1e-12 * abs(float(z["energy"]))


@pytest.mark.gpu
@pytest.mark.parametrize("path", FIXTURES, ids=ids(FIXTURES))
def test_gpu_matches_reference_build(path):
    from gpe_amd import Context, StepInfo

    p, psi, z = load(path)
    nsteps = int(z["p"])
    with Context(p) as ctx:
        ctx.upload(np.ascontiguousarray(psi))
        info = StepInfo(p.dt0())
        ctx.step(p.scheme, nsteps, info)
        got = ctx.download()
        nrm, e = ctx.norm(), ctx.energy()
    err = rel(got, z["out"])
    print(f"{os.path.basename(path)}: GPU vs reference build rel L2 = {err:.3e}")
    assert err <= 1e-12
    # rk45 dts come from a cancellation-limited error norm (see the dt check below)
    t_tol = 1e-10 * max(abs(float(z["t"])), 1.0) if p.scheme == "rk45" else 1e-12
    assert abs(info.t - float(z["t"])) <= t_tol and abs(info.im_t - float(z["im_t"])) <= t_tol
    assert abs(info.dt - complex(z["dt"])) <= 1e-9 * abs(info.dt)
    assert abs(nrm - float(z["norm"])) <= 1e-12 * float(z["norm"])
    assert abs(e - float(z["energy"])) <= 1e-11 * abs(float(z["energy"]))


def bench_field(p):
    """The analytic field of oracle/ref/ref_bench.f90, array (z, y, x)."""
    k, j, i = np.meshgrid(np.arange(p.nz), np.arange(p.ny), np.arange(p.nx), indexing="ij")
    two_pi = 6.283185307179586
    ph = 0.3 * np.sin(two_pi * i / p.nx) + 0.2 * np.cos(two_pi * (2 * j + k) / p.ny)
    return np.ascontiguousarray((1.0 + 0.1 * np.cos(ph)) * np.exp(1j * ph))


def test_reference_decomposition_matches_one_rank_and_oracle(orc, tmp_path):
    """The multi-rank timing build of the reference (its own y-z
    decomposition and ghost exchange over 2 x 2 MPI ranks, oracle/build_ref.sh)
    gives the one-rank build's final norm bit for bit, and the oracle's to
    1e-13, after 5 RK4 steps of a 32^3 periodic 4th-order run."""
    import shutil
    import subprocess

    from gpe_amd.io import run_in_text

    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    exes = {n: os.path.join(root, "oracle", "_ref", f"bench_32_{n}", "ref_bench") for n in ("1x1", "2x2")}
    mpiexec = shutil.which("mpiexec", path="/opt/conda/bin") or shutil.which("mpiexec")
    if not mpiexec or not all(os.access(e, os.X_OK) for e in exes.values()):
        pytest.skip("built by oracle/build_ref.sh where the reference sources and MPICH are present")
    n = 32
    xr = 0.3 * (n - 1) / 2
    p = RunParams(nx=n, ny=n, nz=n, xr=xr, yr=xr, zr=xr, order=4, bcs=1, eqn_to_solve=1, scheme="rk4", tau=0.005)
    (tmp_path / "run.in").write_text(run_in_text(p, 1.0))
    norms = {}
    for name, ranks in (("1x1", 1), ("2x2", 4)):
        r = subprocess.run(["bash", "-c", f"ulimit -s unlimited && exec {mpiexec} -n {ranks} {exes[name]} 5"],
                           cwd=tmp_path, capture_output=True, text=True, timeout=120)
        assert r.returncode == 0, r.stderr[-2000:]
        fields = r.stdout.split()
        assert int(fields[0]) == ranks and int(fields[1]) == n ** 3 and int(fields[2]) == 5
        norms[name] = float(fields[4])
    assert norms["1x1"] == norms["2x2"]
    exp, _ = orc.rk4_n(p, bench_field(p), 0.005 + 0j, 5)
    n_orc = orc.norm(p, exp)
    assert abs(norms["1x1"] - n_orc) <= 1e-13 * n_orc, (norms, n_orc)
