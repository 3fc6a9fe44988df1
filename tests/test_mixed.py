"""Mixed storage (GPE_MIXED): complex128 state, FP64 stencil and updates,
complex64 RK increments.  Checked against the FP64 oracle.

Tolerances: the increments k_j = dt f(u_j) and the partial Butcher sums are
rounded to FP32, a relative 2^-24 of an O(dt |f|) quantity per stage, so a
few RK steps stay within 5e-8 relative L2 of the oracle (complex64 storage is
at ~1e-6 on the same runs).  The RKF45 error estimate is a small difference of
increments: its scaled max-norm is compared to 1e-2 relative.
"""
import numpy as np
import pytest

from gpe_amd import Context, GpeError, RunParams, StepInfo
from gpe_amd import ics
from gpe_amd.ensemble import Ensemble

from test_gpu_parity import rel, smooth_field

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("order,bcs,eqn", [(2, 1, 1), (4, 2, 4), (4, 1, 2)])
def test_mixed_rk4_matches_oracle(orc, order, bcs, eqn):
    p = RunParams(nx=24, ny=20, nz=18, xr=8.0, yr=7.0, zr=6.0, order=order, bcs=bcs, eqn_to_solve=eqn,
                  g=1.0, mu=1.0, omx=0.1, omy=0.1, omz=0.1, storage="mixed")
    psi = smooth_field(p, seed=2)
    dt = 0.01 + 0j
    exp, t = orc.rk4_n(p, psi, dt, 5)
    with Context(p) as ctx:
        ctx.upload(psi)
        info = StepInfo(dt)
        ctx.step("rk4", 5, info)
        got = ctx.download()
        # state stays complex128: the reductions see the FP64 field
        assert got.dtype == np.complex128
    err = rel(got, exp)
    print(f"mixed rk4 order {order} bcs {bcs} eqn {eqn}: rel L2 {err:.3e}")
    assert err < 5e-8
    assert abs(info.t - t[0]) < 1e-15 and info.n_rhs == 20


def test_mixed_ring64_rk4_far_below_c64(orc):
    """64^3 ring, 20 RK4 steps: mixed storage is orders of magnitude closer to
    the FP64 oracle than complex64 storage."""
    p, psi = ics.config_ring64(seed=1)
    exp, _ = orc.rk4_n(p, psi, 0.01 + 0j, 20)
    errs = {}
    for storage in ("mixed", "c64"):
        with Context(p.replace(storage=storage)) as ctx:
            ctx.upload(psi)
            ctx.step("rk4", 20, StepInfo(0.01 + 0j))
            errs[storage] = rel(ctx.download().astype(np.complex128), exp)
    print(f"ring64 20 steps: mixed {errs['mixed']:.3e}, c64 {errs['c64']:.3e}")
    assert errs["mixed"] < 5e-8
    assert errs["mixed"] < errs["c64"] / 30


@pytest.mark.parametrize("fused", [True, False])
@pytest.mark.parametrize("order,bcs", [(2, 1), (4, 2)])
def test_mixed_rkck_attempt_matches_oracle(orc, order, bcs, fused):
    """Mixed-storage Cash-Karp attempt, fused (three kernels) and as six
    stage kernels, against the oracle; a retry equals a fresh attempt."""
    p = RunParams(nx=24, ny=22, nz=20, xr=8.0, yr=8.0, zr=8.0, order=order, bcs=bcs, eqn_to_solve=1,
                  storage="mixed")
    psi = smooth_field(p, seed=4)
    new_o, err_o = orc.rkck(p, psi, 0.05 + 0j)
    f0 = orc.rhs(p, psi)
    scal = np.abs(psi) + 0.05 * np.abs(f0) + 1e-30
    em_o = np.max(np.abs(err_o) / np.abs(scal)) / 1e-6
    with Context(p) as ctx:
        ctx.set_fused(fused)
        ctx.upload(psi)
        em, a2 = ctx.rk45_attempt(0.05 + 0j, 0.05 + 0j)
        assert abs(em - em_o) <= 1e-2 * em_o
        assert abs(a2 - np.max(np.abs(new_o) ** 2)) <= 1e-8 * a2
        ctx.rk45_accept()
        assert rel(ctx.download(), new_o) < 5e-8
        # a retry after a reject reuses F0 = f(psi0): a second attempt at a
        # smaller dt from a fresh upload and from a retry agree exactly
        ctx.upload(psi)
        ctx.rk45_attempt(0.05 + 0j, 0.05 + 0j)
        em_retry, _ = ctx.rk45_attempt(0.02 + 0j, 0.05 + 0j)
        ctx.rk45_accept()
        got_retry = ctx.download()
        ctx.upload(psi)
        em_fresh, _ = ctx.rk45_attempt(0.02 + 0j, 0.05 + 0j)
        ctx.rk45_accept()
        assert em_retry == em_fresh
        assert np.array_equal(got_retry, ctx.download())


@pytest.mark.parametrize("fused", [True, False])
def test_mixed_rkqs_steps_match_oracle(orc, fused):
    p = RunParams(nx=24, ny=22, nz=20, xr=8.0, yr=8.0, zr=8.0, order=4, bcs=1, eqn_to_solve=1, storage="mixed")
    psi = smooth_field(p, seed=5)
    cur, dt, t = psi, 0.4 + 0j, (0.0, 0.0)
    nrej_o = 0
    for _ in range(4):
        cur, dt, t, _, nrej, _ = orc.rkqs(p, cur, dt, t)
        nrej_o += nrej
    with Context(p) as ctx:
        ctx.set_fused(fused)
        ctx.upload(psi)
        info = StepInfo(0.4 + 0j)
        ctx.step("rk45", 4, info)
        got = ctx.download()
    assert info.n_reject == nrej_o and nrej_o > 0
    assert abs(info.dt - dt) <= 1e-3 * abs(dt)
    assert abs(info.t - t[0]) <= 1e-3 * t[0]
    assert rel(got, cur) < 1e-5


MX_CASES = [((24, 20, 18), 2, 1, 1), ((37, 21, 19), 4, 2, 1), ((70, 33, 12), 4, 1, 2), ((31, 17, 23), 4, 2, 4),
            ((6, 6, 6), 2, 2, 4)]


@pytest.mark.parametrize("shape,order,bcs,eqn", MX_CASES, ids=[f"{s[0]}x{s[1]}x{s[2]}-o{o}-b{b}-e{e}"
                                                                for s, o, b, e in MX_CASES])
def test_mixed_fused_attempt_matches_stage_kernels(shape, order, bcs, eqn):
    """The fused mixed-storage attempt rounds its complex64 increments at the
    same points as the six mixed stage kernels: candidate within 1e-9
    relative L2 (an FP64 contraction difference can move one FP32 rounding
    by an ulp, ~6e-8 of an O(dt f) increment at that cell), the error
    max-norm within 1e-3, for a first attempt and a retry."""
    nx, ny, nz = shape
    kw = dict(g=1.0, mu=3.0, omx=0.2, omy=0.25, omz=0.3) if eqn == 4 else {}
    p = RunParams(nx=nx, ny=ny, nz=nz, xr=0.35 * nx, yr=0.35 * ny, zr=0.35 * nz, order=order, bcs=bcs,
                  eqn_to_solve=eqn, storage="mixed", **kw)
    psi = smooth_field(p, seed=13)
    dt0 = 0.1 + 0j
    res = {}
    for fused in (True, False):
        with Context(p) as ctx:
            ctx.set_fused(fused)
            ctx.upload(psi)
            em1, a21 = ctx.rk45_attempt(dt0, dt0)
            em2, a22 = ctx.rk45_attempt(0.5 * dt0, dt0)
            ctx.rk45_accept()
            res[fused] = (em1, a21, em2, a22, ctx.download())
    f, s = res[True], res[False]
    print(f"mixed fused/stage errmax {f[0]:.6e}/{s[0]:.6e}, {f[2]:.6e}/{s[2]:.6e}; cand rel {rel(f[4], s[4]):.2e}")
    assert rel(f[4], s[4]) < 1e-9
    for i in (0, 2):
        assert abs(f[i] - s[i]) <= 1e-3 * s[i] + 1e-6
    for i in (1, 3):
        assert abs(f[i] - s[i]) <= 1e-9 * s[i]


def test_mixed_euler_rk2_use_the_fp64_path(orc):
    """Euler and RK2 carry no complex64 increments in mixed storage: they are
    the complex128 kernels, at 1e-12."""
    p = RunParams(nx=20, ny=18, nz=16, xr=7.0, yr=7.0, zr=6.0, order=4, bcs=1, eqn_to_solve=1, storage="mixed")
    psi = smooth_field(p, seed=6)
    for scheme in ("euler", "rk2"):
        exp, t = psi, (0.0, 0.0)
        for _ in range(3):
            exp, t = orc.step(p, scheme, exp, 0.01 + 0j, t)
        with Context(p) as ctx:
            ctx.upload(psi)
            ctx.step(scheme, 3, StepInfo(0.01 + 0j))
            assert rel(ctx.download(), exp) < 1e-12


def test_ensemble_rejects_mixed_storage():
    ms = [RunParams(nx=16, ny=16, nz=16, xr=5.0, yr=5.0, zr=5.0, storage="mixed")]
    with pytest.raises(GpeError, match="storage"):
        Ensemble(ms)
