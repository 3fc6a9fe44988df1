"""The benchmark configurations on the HIP path at their own workloads.

* Batched ensemble: 48 members, seeds 0-47, cubic grids drawn from
  {64, 96, 128, 160, 192}^3 (gpe_amd.ensemble.ensemble_config), dx 0.3,
  2nd order, periodic, RK4 dt 0.005: every ensemble mode equals the
  single-context run of each member.
* Two coaxial rings (radii 6 and 10 at z = -10 and +10, moving towards each
  other), 4th order, RKF45 at tolerance 1e-6: rkqs against the oracle's
  rkqs (src/solve.f90:158-244, 248-378) at 64^3, and at the full 384^3 the
  fused Cash-Karp attempt against the six stage kernels.
* Random-phase turbulence with complex64 storage: Cash-Karp attempts
  against the oracle at 40^3 (complex64 rounding bounds the agreement), an
  RKF45 run to a fixed time against the complex128 run that the oracle-driven
  controller pins, and at 512^3 the complex64 attempt (fused and stage
  kernels) against the complex128 one."""
import numpy as np
import pytest

from gpe_amd import Context, StepInfo, ics
from gpe_amd.ensemble import MODES, Ensemble, ensemble_config

pytestmark = pytest.mark.gpu


def rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


@pytest.fixture(scope="module")
def ensemble48():
    members, fields = ensemble_config(48)
    dt = 0.005 + 0j
    exp = []
    for m, f in zip(members, fields):
        with Context(m) as ctx:
            ctx.upload(f)
            ctx.step("rk4", 3, StepInfo(dt))
            exp.append(ctx.download())
    return members, fields, exp


@pytest.mark.parametrize("mode", list(MODES))
def test_ensemble_48_members_every_mode(ensemble48, mode):
    members, fields, exp = ensemble48
    assert len(members) == 48 and {m.nx for m in members} <= {64, 96, 128, 160, 192}
    assert all(m.order == 2 and m.bcs == 1 and abs(m.dx - 0.3) < 1e-15 for m in members)
    with Ensemble(members) as ens:
        for i, f in enumerate(fields):
            ens.upload(i, f)
        ens.step(mode, 0.005 + 0j, 3)
        worst = 0.0
        for i in range(len(members)):
            got = ens.download(i)
            if not np.array_equal(got, exp[i]):
                worst = max(worst, rel(got, exp[i]))
    print(f"48-member ensemble, mode {mode}: worst member rel L2 vs single context {worst:.2e}")
    assert worst <= 1e-14


def test_two_rings_rkqs_matches_oracle(orc):
    """64^3 two-ring IC (dx 0.5), c128 RKF45: 3 rkqs steps from a step too
    large for the tolerance, fused and stage kernels, against the oracle."""
    p, psi = ics.config_two_rings(n=64, dx=0.5)
    assert p.scheme == "rk45" and p.order == 4
    dt0 = 0.3 + 0j
    cur, dt, t, nrej_o = psi, dt0, (0.0, 0.0), 0
    trace_o = []
    for _ in range(3):
        cur, dt, t, em, nrej, _ = orc.rkqs(p, cur, dt, t)
        nrej_o += nrej
        trace_o.append((t[0], dt))
    assert nrej_o >= 1
    for fused in (True, False):
        with Context(p) as ctx:
            ctx.set_fused(fused)
            ctx.upload(psi)
            info = StepInfo(dt0)
            trace = []
            for _ in range(3):
                ctx.step("rk45", 1, info)
                trace.append((info.t, info.dt))
            got = ctx.download()
        assert info.n_reject == nrej_o
        for (tg, dg), (to, do) in zip(trace, trace_o):
            assert abs(tg - to) <= 1e-10 * to and abs(dg - do) <= 1e-9 * abs(do)
        r = rel(got, cur)
        print(f"two rings 64^3 rkqs ({'fused' if fused else 'stage kernels'}) vs oracle: rel L2 {r:.2e}, "
              f"t = {info.t:.6f}, {info.n_reject} rejects")
        assert r <= 1e-12


def test_two_rings_384_fused_attempt_matches_stage_kernels():
    """The full 384^3 two-ring workload: one attempt and a retry at half the
    step, fused against stage kernels (candidate 1e-13, errmax 1e-9)."""
    p, psi = ics.config_two_rings(n=384, dx=0.5)
    res = {}
    for fused in (True, False):
        with Context(p) as ctx:
            ctx.set_fused(fused)
            ctx.upload(psi)
            em1, _ = ctx.rk45_attempt(0.1 + 0j, 0.1 + 0j)
            em2, a2 = ctx.rk45_attempt(0.05 + 0j, 0.1 + 0j)
            ctx.rk45_accept()
            res[fused] = (em1, em2, a2, ctx.download())
    f, s = res[True], res[False]
    r = rel(f[3], s[3])
    print(f"two rings 384^3 fused vs stage: errmax {f[0]:.6e}/{s[0]:.6e}, {f[1]:.6e}/{s[1]:.6e}; rel L2 {r:.2e}")
    assert r <= 1e-13
    for i in (0, 1):
        assert abs(f[i] - s[i]) <= 1e-9 * s[i]
    assert abs(f[2] - s[2]) <= 1e-13 * s[2]


def turb40():
    p, psi = ics.config_turb(n=40, scheme="rk45", tau=0.01)
    return p, psi


@pytest.mark.parametrize("fused", [True, False])
def test_c64_cash_karp_attempt_matches_oracle(orc, fused):
    """complex64 storage (FP32 stencil and increments, FP64 error norm): one
    attempt against the oracle's rkck in complex128.  The candidate carries
    FP32 rounding (~1e-7 relative per stage value); the error estimate is a
    cancellation of the increments (|k| / |err| ~ 1e4-1e5), so errmax agrees
    to about 1e-2 (1.1% measured)."""
    p, psi = turb40()
    dt = 0.05 + 0j
    new_o, err_o = orc.rkck(p, psi, dt)
    scal = np.abs(psi) + abs(dt) * np.abs(orc.rhs(p, psi)) + 1e-30
    em_o = np.max(np.abs(err_o) / scal) / 1e-6
    with Context(p.replace(storage="c64")) as ctx:
        ctx.set_fused(fused)
        ctx.upload(psi)
        em, a2 = ctx.rk45_attempt(dt, dt)
        ctx.rk45_accept()
        got = ctx.download()
    r = rel(got, new_o)
    print(f"c64 attempt ({'fused' if fused else 'stage'}) vs oracle: candidate rel L2 {r:.2e}, "
          f"errmax {em:.6e} vs {em_o:.6e}")
    assert r <= 2e-7
    assert abs(em - em_o) <= 2e-2 * em_o
    assert abs(a2 - np.max(np.abs(new_o) ** 2)) <= 1e-6 * a2


def oracle_integrate(orc, p, psi, T, dt0):
    """The controller of gpe_amd.controller.integrate (rkqs rules, last step
    clamped onto T) with the oracle's rkck attempts and error scale
    (src/solve.f90:158-244)."""
    from gpe_amd.controller import PlainController

    ctrl = PlainController()
    cur, t, dt, steps, rejects = psi, 0.0, complex(dt0, 0.0), 0, 0
    while t < T * (1 - 1e-14):
        if t + dt.real > T:
            dt = complex(T - t, 0.0)
        d0 = dt
        scal = np.abs(cur) + abs(d0) * np.abs(orc.rhs(p, cur)) + 1e-30
        while True:
            new, err = orc.rkck(p, cur, dt)
            em = np.max(np.abs(err) / scal) / p.eps
            if em <= 1.0:
                break
            rejects += 1
            dt = ctrl.shrink(dt, em)
        cur, t, steps = new, t + dt.real, steps + 1
        dt = ctrl.propose(dt, em, None)
    return cur, steps, rejects


def test_c64_rkqs_run_to_fixed_time(orc):
    """RKF45 at tolerance 1e-6 to T = 0.3 (last step clamped).  complex64
    cannot follow the oracle step for step: its error estimate carries
    ~0.005 of FP32 noise in errmax (stage values rounded to 6e-8 of |psi|,
    amplified by the stencil's ~1e2 spectral radius), so each proposed dt
    moves by a few percent where errmax is small.  The chain pins it instead:
    the complex128 run equals the oracle-driven controller to 1e-10, and the
    complex64 run stays within FP32-level distance of it with a similar step
    count."""
    from gpe_amd.controller import PlainController, integrate

    p, psi = turb40()
    T = 0.3
    exp, steps_o, rej_o = oracle_integrate(orc, p.replace(eps=1e-6), psi, T, 0.02)
    out = {}
    for storage in ("c128", "c64"):
        with Context(p.replace(storage=storage, eps=1e-6)) as ctx:
            ctx.upload(psi)
            st = integrate(ctx, T, 0.02, PlainController(), clamp=True)
            out[storage] = (ctx.download().astype(np.complex128), st.n_accept, st.n_reject)
    r128, r64 = rel(out["c128"][0], exp), rel(out["c64"][0], out["c128"][0])
    print(f"RKF45 to T={T}: oracle {steps_o} steps/{rej_o} rejects; c128 {out['c128'][1]}/{out['c128'][2]} "
          f"(rel L2 {r128:.2e}); c64 {out['c64'][1]}/{out['c64'][2]} (rel L2 vs c128 {r64:.2e})")
    assert out["c128"][1:] == (steps_o, rej_o) and r128 <= 1e-10
    assert abs(out["c64"][1] - steps_o) <= max(2, 0.2 * steps_o)
    assert r64 <= 2e-5


def test_c64_full_size_attempt_against_c128():
    """512^3 turbulence: the complex64 attempt (fused and stage kernels)
    against the complex128 fused attempt (oracle-pinned at small sizes)."""
    p, psi = ics.config_turb(n=512, scheme="rk45", tau=0.01)
    res = {}
    for storage, fused in (("c128", True), ("c64", True), ("c64", False)):
        with Context(p.replace(storage=storage)) as ctx:
            ctx.set_fused(fused)
            ctx.upload(psi)
            em, a2 = ctx.rk45_attempt(0.02 + 0j, 0.02 + 0j)
            ctx.rk45_accept()
            res[(storage, fused)] = (em, a2, ctx.download())
    ref = res[("c128", True)]
    for key in (("c64", True), ("c64", False)):
        em, a2, out = res[key]
        r = rel(out, ref[2])
        print(f"512^3 c64 attempt ({'fused' if key[1] else 'stage'}) vs c128: rel L2 {r:.2e}, "
              f"errmax {em:.6e} vs {ref[0]:.6e}")
        assert r <= 5e-7
        assert abs(em - ref[0]) <= 5e-2 * ref[0]
        assert abs(a2 - ref[1]) <= 1e-5 * ref[1]
    assert rel(res[("c64", True)][2], res[("c64", False)][2]) <= 1e-6
