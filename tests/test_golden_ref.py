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
