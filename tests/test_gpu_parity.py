"""Parity of the HIP kernels (called through the C ABI of libgpehip.so) with
the CPU oracle on the same seeded inputs.

Tolerances: complex128 storage must agree with the FP64 oracle to 1e-12
relative L2 (single operators to 1e-13); complex64 storage is compared with
the FP64 oracle at the accuracy of FP32 arithmetic (stated per test).
"""
import numpy as np
import pytest

from gpe_amd import Context, GpeError, RunParams, StepInfo
from gpe_amd import ics

pytestmark = pytest.mark.gpu


def rel(a, b):
    return float(np.linalg.norm((a - b).ravel()) / np.linalg.norm(np.ravel(b)))


def rand_field(p, seed=0, amp=1.0):
    rng = np.random.default_rng(seed)
    base = np.exp(1j * rng.uniform(0, 2 * np.pi, p.shape)) * (1 + 0.2 * rng.standard_normal(p.shape))
    return np.ascontiguousarray(amp * base)


def smooth_field(p, seed=0):
    """A smooth non-trivial field: a few random plane waves on a background."""
    rng = np.random.default_rng(seed)
    x, y, z = p.axes()
    Z, Y, X = np.meshgrid(z, y, x, indexing="ij")
    psi = np.ones(p.shape, dtype=np.complex128)
    for _ in range(4):
        k = 2 * np.pi * rng.integers(-2, 3, size=3) / np.array([p.nx * p.dx, p.ny * p.dy, p.nz * p.dz])
        psi += 0.2 * (rng.standard_normal() + 1j * rng.standard_normal()) * np.exp(1j * (k[0] * X + k[1] * Y + k[2] * Z))
    return np.ascontiguousarray(psi)


SHAPES = [(16, 12, 10), (37, 21, 19), (64, 8, 9), (70, 33, 12)]


@pytest.mark.parametrize("shape", SHAPES)
@pytest.mark.parametrize("order", [2, 4])
@pytest.mark.parametrize("bcs", [1, 2])
def test_derivs_match_oracle(orc, shape, order, bcs):
    nx, ny, nz = shape
    p = RunParams(nx=nx, ny=ny, nz=nz, xr=3.1, yr=2.2, zr=1.9, order=order, bcs=bcs)
    f = rand_field(p, seed=nx + ny)
    with Context(p) as ctx:
        for name, op in [("x", 0), ("y", 1), ("z", 2), ("xx", 3), ("yy", 4), ("zz", 5), ("lap", 6)]:
            got = ctx.deriv(name, f)
            exp = orc.deriv(p, op, f)
            assert rel(got, exp) < 1e-13, (name, shape, order, bcs)


@pytest.mark.parametrize("tiles", [(64, 8, 3), (64, 4, 5), (32, 8, 7), (32, 16, 4), (64, 16, 0), (32, 4, 100)])
def test_tile_shapes_agree(orc, tiles):
    p = RunParams(nx=45, ny=29, nz=23, xr=3.0, yr=2.0, zr=2.5, order=4, bcs=1)
    f = rand_field(p, seed=5)
    with Context(p) as ctx:
        ctx.set_tiles(*tiles)
        assert rel(ctx.deriv("lap", f), orc.deriv(p, 6, f)) < 1e-13


@pytest.mark.parametrize("eqn,extra", [
    (1, {}), (1, {"Urhs": 0.3}), (1, {"diss_amp": 0.02}), (1, {"vscale": 1.0, "omx": 0.2, "omy": 0.1, "omz": 0.3}),
    (2, {}), (2, {"Urhs": -0.4}), (3, {}), (4, {"g": 1.0, "mu": 10.0, "omx": 0.05, "omy": 0.07, "omz": 0.05}),
])
@pytest.mark.parametrize("order,bcs", [(2, 1), (4, 2)])
def test_rhs_matches_oracle(orc, eqn, extra, order, bcs):
    p = RunParams(nx=20, ny=18, nz=16, xr=14.0, yr=12.0, zr=13.0, order=order, bcs=bcs, eqn_to_solve=eqn, **extra)
    f = smooth_field(p, seed=eqn)
    with Context(p) as ctx:
        assert rel(ctx.rhs(f), orc.rhs(p, f)) < 1e-13


@pytest.mark.parametrize("scheme", ["euler", "rk2", "rk4"])
@pytest.mark.parametrize("order,bcs,eqn", [(2, 1, 1), (4, 2, 4), (4, 1, 2)])
def test_fixed_steppers_match_oracle(orc, scheme, order, bcs, eqn):
    p = RunParams(nx=24, ny=20, nz=18, xr=8.0, yr=7.0, zr=6.0, order=order, bcs=bcs, eqn_to_solve=eqn,
                  g=1.0, mu=1.0, omx=0.1, omy=0.1, omz=0.1)
    psi = smooth_field(p, seed=2)
    dt = 0.01 + 0j
    exp = psi
    t = (0.0, 0.0)
    for _ in range(5):
        exp, t = orc.step(p, scheme, exp, dt, t)
    with Context(p) as ctx:
        ctx.upload(psi)
        info = StepInfo(dt)
        ctx.step(scheme, 5, info)
        got = ctx.download()
    assert rel(got, exp) < 1e-12
    assert abs(info.t - t[0]) < 1e-15 and info.im_t == t[1]


def test_imaginary_time_rk4(orc):
    p = RunParams(nx=20, ny=20, nz=20, xr=6.0, yr=6.0, zr=6.0, order=4, bcs=2, eqn_to_solve=4, g=1.0, mu=2.0,
                  omx=0.3, omy=0.3, omz=0.3, real_time=False, tau=0.005)
    psi = np.ascontiguousarray(ics.fermi(p) + 0.01)
    exp, t = orc.rk4_n(p, psi, p.dt0(), 10)
    with Context(p) as ctx:
        ctx.upload(psi)
        info = StepInfo(p.dt0())
        ctx.step("rk4", 10, info)
        got = ctx.download()
    assert rel(got, exp) < 1e-12
    assert abs(info.im_t - t[1]) < 1e-15


def test_ring64_rk4_200_steps(orc):
    """The 64^3 plain method: periodic, dx = 0.5, 2nd order, eqn 1, RK4
    dt = 0.01 for 200 steps from the noisy vortex ring; <= 1e-12 relative L2."""
    p, psi = ics.config_ring64(seed=0)
    exp, t = orc.rk4_n(p, psi, 0.01 + 0j, 200)
    with Context(p) as ctx:
        ctx.upload(psi)
        info = StepInfo(0.01 + 0j)
        ctx.step("rk4", 200, info)
        got = ctx.download()
        n_gpu = ctx.norm()
    err = rel(got, exp)
    print(f"ring64 rk4 200 steps: rel L2 = {err:.3e}")
    assert err <= 1e-12
    assert abs(n_gpu - orc.norm(p, exp)) <= 1e-12 * abs(n_gpu)
    assert abs(info.t - t[0]) < 1e-12


@pytest.mark.parametrize("fused", [True, False])
@pytest.mark.parametrize("order,bcs", [(2, 1), (4, 2)])
def test_rkck_attempt_and_rkqs_match_oracle(orc, order, bcs, fused):
    """One Cash-Karp attempt and four rkqs steps from a too-large dt, fused
    (three kernels) and as six stage kernels, against the oracle."""
    p = RunParams(nx=24, ny=22, nz=20, xr=8.0, yr=8.0, zr=8.0, order=order, bcs=bcs, eqn_to_solve=1)
    psi = smooth_field(p, seed=4)
    with Context(p) as ctx:
        ctx.set_fused(fused)
        ctx.upload(psi)
        # single attempt: errmax and the candidate
        new_o, err_o = orc.rkck(p, psi, 0.05 + 0j)
        f0 = orc.rhs(p, psi)
        scal = np.abs(psi) + 0.05 * np.abs(f0) + 1e-30
        em_o = np.max(np.abs(err_o) / np.abs(scal)) / 1e-6
        em, a2 = ctx.rk45_attempt(0.05 + 0j, 0.05 + 0j)
        assert abs(em - em_o) <= 1e-9 * em_o
        assert abs(a2 - np.max(np.abs(new_o) ** 2)) <= 1e-12 * a2
        ctx.rk45_accept()
        assert rel(ctx.download(), new_o) < 1e-13
        # several rkqs steps from a too-large dt (forces rejections)
        ctx.upload(psi)
        info = StepInfo(0.4 + 0j)
        cur, dt, t = psi, 0.4 + 0j, (0.0, 0.0)
        nrej_o = 0
        for _ in range(4):
            cur, dt, t, em, nrej, _ = orc.rkqs(p, cur, dt, t)
            nrej_o += nrej
        ctx.step("rk45", 4, info)
        got = ctx.download()
    assert info.n_reject == nrej_o and nrej_o > 0
    assert abs(info.dt - dt) <= 1e-9 * abs(dt)
    # the accepted dts carry the error norm's cancellation-limited rounding
    assert abs(info.t - t[0]) <= 1e-10 * t[0]
    assert rel(got, cur) < 1e-12


def test_reductions_match_oracle(orc):
    for order, bcs in [(2, 1), (4, 2)]:
        p = RunParams(nx=33, ny=27, nz=19, xr=5.0, yr=4.0, zr=3.0, order=order, bcs=bcs, nn=2.5)
        psi = smooth_field(p, seed=7)
        with Context(p) as ctx:
            ctx.upload(psi)
            assert abs(ctx.norm() - orc.norm(p, psi)) <= 1e-13 * orc.norm(p, psi)
            assert abs(ctx.mass() - orc.mass(p, psi)) <= 1e-13 * orc.mass(p, psi)
            e = orc.energy(p, psi)
            assert abs(ctx.energy() - e) <= 1e-12 * abs(e)
            nrm = ctx.norm()
            ctx.renormalise(nrm)
            assert rel(ctx.download(), orc.renormalise(p, psi, nrm)) < 1e-15


def test_halo_array_upload_and_c64_host(orc):
    p = RunParams(nx=16, ny=12, nz=10, xr=3.0, yr=3.0, zr=3.0, order=4, bcs=1)
    psi = rand_field(p, seed=9)
    halo = np.zeros((p.nz + 4, p.ny + 4, p.nx), dtype=np.complex128)
    halo[2:-2, 2:-2, :] = psi
    with Context(p) as ctx:
        ctx.upload(halo)
        assert np.array_equal(ctx.download(), psi)
        assert rel(ctx.deriv("lap", halo), orc.deriv(p, 6, psi)) < 1e-13
        ctx.upload(psi.astype(np.complex64))
        assert np.array_equal(ctx.download(np.complex64), psi.astype(np.complex64))


def test_c64_storage_drift_against_fp64_oracle(orc):
    """complex64 storage / FP32 stencil against the FP64 oracle: 20 RK4 steps
    of the 64^3 ring agree to 1e-5 relative L2; the norm (FP64 reduction of
    FP32 values) to 1e-6."""
    p, psi = ics.config_ring64(seed=1)
    exp, _ = orc.rk4_n(p, psi, 0.01 + 0j, 20)
    with Context(p.replace(storage="c64")) as ctx:
        ctx.upload(psi)
        ctx.step("rk4", 20, StepInfo(0.01 + 0j))
        got = ctx.download()
        nrm = ctx.norm()
    err = rel(got, exp)
    print(f"c64 drift after 20 steps: {err:.3e}")
    assert err < 1e-5
    assert abs(nrm - orc.norm(p, exp)) < 1e-6 * nrm


def test_deterministic_reruns():
    p, psi = ics.config_ring64(seed=2)
    outs = []
    for _ in range(2):
        with Context(p.replace(scheme="rk45")) as ctx:
            ctx.upload(psi)
            info = StepInfo(0.05 + 0j)
            ctx.step("rk45", 3, info)
            outs.append((ctx.download(), info.dt, info.errmax, ctx.energy(), ctx.norm()))
    assert np.array_equal(outs[0][0], outs[1][0])
    assert outs[0][1:] == outs[1][1:]


def test_error_paths():
    p = RunParams(nx=16, ny=16, nz=16)
    with Context(p) as ctx:
        with pytest.raises(GpeError, match="Unrecognised"):
            ctx._chk(ctx.lib.gpe_step(ctx.h, 9, 1, StepInfo().c), "gpe_step")
        with pytest.raises(GpeError, match="no attempt"):
            ctx.rk45_accept()
        with pytest.raises(GpeError):
            ctx.set_tiles(48, 8, 0)
        with pytest.raises(ValueError):
            ctx.upload(np.zeros((3, 3, 3), dtype=np.complex128))


def test_hamiltonian_closed_form_and_conservation():
    """Conserved energy: plane-wave closed form for eqn 1, and RK4 keeps it to
    O(dt^4) for eqn 4 on a periodic grid (where the discrete Laplacian is
    self-adjoint under plain sums, so the discrete energy is an invariant)."""
    p = RunParams(nx=24, ny=20, nz=18, xr=8.0, yr=7.0, zr=6.5, order=4, bcs=1, eqn_to_solve=1)
    x, y, z = p.axes()
    Z, Y, X = np.meshgrid(z, y, x, indexing="ij")
    kx = 2 * np.pi * 2 / (p.nx * p.dx)
    amp = 0.7 + 0.2j
    f = np.ascontiguousarray(amp * np.exp(1j * kx * X))
    th = kx * p.dx
    lam = (-2 * np.cos(2 * th) + 32 * np.cos(th) - 30) / (12 * p.dx ** 2)
    a2 = abs(amp) ** 2
    exact = p.cells * p.dx * p.dy * p.dz * (-lam * a2 - a2 + 0.5 * a2 * a2)
    with Context(p) as ctx:
        ctx.upload(f)
        assert abs(ctx.hamiltonian() - exact) <= 1e-12 * abs(exact)
    q, psi = ics.config_trap256(n=32)
    q = q.replace(xr=6.0, yr=6.0, zr=6.0, omx=0.3, omy=0.3, omz=0.3, mu=3.0, bcs=1)
    psi = np.ascontiguousarray(ics.fermi(q) * ics.vortex_line_z(q))
    drifts = []
    for dt in (0.004, 0.002):
        with Context(q) as ctx:
            ctx.upload(psi)
            h0 = ctx.hamiltonian()
            ctx.step("rk4", int(round(0.2 / dt)), StepInfo(complex(dt)))
            drifts.append(abs(ctx.hamiltonian() - h0) / abs(h0))
    assert drifts[0] < 1e-6 and drifts[1] < drifts[0] / 8


@pytest.mark.parametrize("shape", [(24, 20, 18), (37, 21, 19), (70, 33, 12), (6, 6, 6)])
@pytest.mark.parametrize("order,bcs,eqn", [(2, 1, 1), (4, 2, 1), (4, 1, 2), (2, 2, 2)])
def test_fused_rk4_matches_oracle_and_stage_kernels(orc, shape, order, bcs, eqn):
    """Fused RK4 (two kernels per step, overlapped tiles, unfolded boundary
    cells) against the FP64 oracle (1e-12) and the four stage kernels (1e-13)."""
    nx, ny, nz = shape
    p = RunParams(nx=nx, ny=ny, nz=nz, xr=0.35 * nx, yr=0.35 * ny, zr=0.35 * nz, order=order, bcs=bcs,
                  eqn_to_solve=eqn)
    psi = smooth_field(p, seed=7)
    dt = 0.01 + 0j
    exp, t = orc.rk4_n(p, psi, dt, 6)
    outs = {}
    for fused in (True, False):
        with Context(p) as ctx:
            ctx.set_fused(fused)
            ctx.upload(psi)
            info = StepInfo(dt)
            ctx.step("rk4", 6, info)
            outs[fused] = ctx.download()
            assert abs(info.t - t[0]) < 1e-15
    assert rel(outs[True], exp) < 1e-12
    assert rel(outs[True], outs[False]) < 1e-13


CK_CASES = [((24, 20, 18), 2, 1, 1), ((37, 21, 19), 4, 2, 1), ((70, 33, 12), 4, 1, 2), ((6, 6, 6), 2, 2, 2),
            ((31, 17, 23), 4, 2, 4), ((40, 35, 9), 2, 1, 4)]


@pytest.mark.parametrize("shape,order,bcs,eqn", CK_CASES, ids=[f"{s[0]}x{s[1]}x{s[2]}-o{o}-b{b}-e{e}"
                                                                for s, o, b, e in CK_CASES])
def test_fused_rkck_attempt_matches_oracle_and_stage_kernels(orc, shape, order, bcs, eqn):
    """Fused Cash-Karp attempt (three kernels of two stages, overlapped
    tiles, unfolded boundary cells; eqn 4 with the trap) against the FP64
    oracle's rkck (candidate 1e-12, error max-norm 1e-9) and the six stage
    kernels (candidate 1e-13, error max-norm 1e-10).  A second attempt at a
    smaller dt (a retry) re-evaluates f(psi0) and must agree too."""
    nx, ny, nz = shape
    kw = dict(g=1.0, mu=3.0, omx=0.2, omy=0.25, omz=0.3) if eqn == 4 else {}
    p = RunParams(nx=nx, ny=ny, nz=nz, xr=0.35 * nx, yr=0.35 * ny, zr=0.35 * nz, order=order, bcs=bcs,
                  eqn_to_solve=eqn, **kw)
    psi = smooth_field(p, seed=11)
    dt0 = 0.1 + 0j
    res = {}
    for fused in (True, False):
        with Context(p) as ctx:
            ctx.set_fused(fused)
            ctx.upload(psi)
            em1, a21 = ctx.rk45_attempt(dt0, dt0)
            em2, a22 = ctx.rk45_attempt(0.5 * dt0, dt0)  # retry at a smaller dt
            ctx.rk45_accept()
            res[fused] = (em1, a21, em2, a22, ctx.download())
    print(f"errmax: first attempt {res[True][0]:.4e}, retry {res[True][2]:.4e}")
    new_o, err_o = orc.rkck(p, psi, 0.5 * dt0)
    f0 = orc.rhs(p, psi)
    scal = np.abs(psi) + dt0 * np.abs(f0) + 1e-30
    em_o = np.max(np.abs(err_o) / np.abs(scal)) / p.eps
    ef, es = res[True], res[False]
    # the error norm is a cancellation of O(dt f) increments: besides the
    # relative bound, allow 1e-9 in errmax units (1e-15 of |psi_scal|)
    assert rel(ef[4], new_o) < 1e-12
    assert abs(ef[2] - em_o) <= 1e-9 * em_o + 1e-9
    assert rel(ef[4], es[4]) < 1e-13
    for i in (0, 2):
        assert abs(ef[i] - es[i]) <= 1e-9 * es[i] + 1e-9, (i, ef[i], es[i])
    for i in (1, 3):
        assert abs(ef[i] - es[i]) <= 1e-13 * es[i]


@pytest.mark.parametrize("scheme,nsteps", [("euler", 1), ("euler", 2), ("euler", 5), ("rk2", 3)])
@pytest.mark.parametrize("order,bcs,eqn,real_time", [(4, 1, 1, True), (2, 2, 4, True), (4, 2, 4, False)])
def test_fused_euler_and_rk2(orc, scheme, nsteps, order, bcs, eqn, real_time):
    """Euler steps fused in pairs and the midpoint RK2 step fused into one
    kernel (the stages-1+2 body of the fused RK4 pair): against the FP64
    oracle (1e-12) and the stage kernels (1e-13), with odd Euler step counts
    (the last step runs alone), the trap, and imaginary time; complex64
    against its stage kernels at FP32 accuracy (1e-6)."""
    kw = dict(g=1.0, mu=2.0, omx=0.3, omy=0.2, omz=0.25) if eqn == 4 else {}
    p = RunParams(nx=30, ny=26, nz=22, xr=0.3 * 30, yr=0.3 * 26, zr=0.3 * 22, order=order, bcs=bcs,
                  eqn_to_solve=eqn, real_time=real_time, tau=0.004, **kw)
    psi = smooth_field(p, seed=17)
    dt = p.dt0()
    exp, t = psi, (0.0, 0.0)
    for _ in range(nsteps):
        exp, t = orc.step(p, scheme, exp, dt, t)
    fused_kind = "rk2_fused" if scheme == "rk2" else "euler2_fused"
    for storage in ("c128", "c64"):
        q = p.replace(storage=storage)
        outs = {}
        for fused in (True, False):
            with Context(q) as ctx:
                ctx.set_fused(fused)
                ctx.upload(psi)
                info = StepInfo(dt)
                ctx.step(scheme, nsteps, info)
                outs[fused] = (ctx.download(), info.t, info.im_t, info.n_accept, info.n_rhs)
                prof = ctx.kernel_profile(lambda: ctx.step(scheme, 2, StepInfo(dt)))
                assert (fused_kind in prof) == fused, (fused, prof.keys())
        f, st = outs[True], outs[False]
        assert f[1:] == st[1:]
        assert abs(f[1] - t[0]) < 1e-15 and abs(f[2] - t[1]) < 1e-15
        if storage == "c128":
            assert rel(f[0], exp) < 1e-12
            assert rel(f[0], st[0]) < 1e-13
        else:
            assert rel(f[0], st[0]) < 1e-6


G_FORMS = [(0.004 + 0j, "real-dt"), (-0.004j, "imaginary-dt"), (0.004 - 0.003j, "complex-dt")]


@pytest.mark.parametrize("eqn", [1, 4])
@pytest.mark.parametrize("dt", [g for g, _ in G_FORMS], ids=[n for _, n in G_FORMS])
def test_fused_step_constant_forms(orc, dt, eqn):
    """The fused kernels fold i*pim*dt into one constant G whose form selects
    the arithmetic of every stage update: imaginary for a real dt, real for an
    imaginary dt, general for a complex dt (the C ABI takes any complex dt).
    Each form, with and without the trap term: complex128 RK4 steps and a
    Cash-Karp attempt against the FP64 oracle (1e-12) and the stage kernels
    (1e-13); complex64 and mixed storage against their own stage kernels at
    the accuracy of their FP32 roundings (1e-6 for complex64 fields; 1e-9
    for mixed storage, whose FP32 increments may round either way when the
    FP64 products they are rounded from differ in the last bit)."""
    kw = dict(g=1.0, mu=3.0, omx=0.2, omy=0.25, omz=0.3) if eqn == 4 else {}
    p = RunParams(nx=30, ny=26, nz=22, xr=0.35 * 30, yr=0.35 * 26, zr=0.35 * 22, order=4, bcs=2 if eqn == 4 else 1,
                  eqn_to_solve=eqn, **kw)
    psi = smooth_field(p, seed=21)
    exp, _ = orc.rk4_n(p, psi, dt, 4)
    new_o, _ = orc.rkck(p, psi, 5 * dt)
    for storage in ("c128", "c64", "mixed"):
        q = p.replace(storage=storage)
        res = {}
        for fused in (True, False):
            with Context(q) as ctx:
                ctx.set_fused(fused)
                if storage != "mixed":
                    ctx.upload(psi)
                    ctx.step("rk4", 4, StepInfo(dt))
                    rk4 = ctx.download()
                    prof = ctx.kernel_profile(lambda: ctx.step("rk4", 1, StepInfo(dt)))
                    assert ("rk4_fused34" in prof) == fused
                else:
                    rk4 = None
                ctx.upload(psi)
                em, a2 = ctx.rk45_attempt(5 * dt, 5 * dt)
                ctx.rk45_accept()
                res[fused] = (rk4, em, ctx.download())
        f, st = res[True], res[False]
        tol = {"c128": 1e-13, "c64": 1e-6, "mixed": 1e-9}[storage]
        print(f"{storage} dt={dt}: RK4 fused/stage "
              f"{rel(f[0], st[0]) if f[0] is not None else 0:.2e}, attempt {rel(f[2], st[2]):.2e}")
        if f[0] is not None:
            assert rel(f[0], st[0]) < tol
        assert rel(f[2], st[2]) < tol
        if storage == "c128":
            assert rel(f[0], exp) < 1e-12
            assert rel(f[2], new_o) < 1e-12
            assert abs(f[1] - st[1]) <= 1e-9 * st[1] + 1e-9


def test_fused_rkck_rkqs_run_and_complex64():
    """rkqs with the fused attempt reproduces the stage kernels' step
    sequence (accepts, rejects, dt) over 12 steps from a too-large dt; with
    complex64 fields the fused and stage attempts agree to FP32 accuracy.
    The tolerance eps = 1e-8 keeps dt (~0.03) inside the Cash-Karp stability
    limit of the grid's highest modes (~0.065 at dx 0.38): at 1e-6 this
    smooth field lets the controller take dt ~0.09, where those modes grow
    ~6x per step from rounding level and any two correct implementations
    part.  The error estimate is ~eps/|dt f| ~1e-7 of the increments it is
    formed from, so its rounding (~1e-9 relative) moves each dt by ~2e-10
    (dt ~ errmax^-0.2): t, dt and the field are compared at 1e-8."""
    p = RunParams(nx=48, ny=40, nz=36, xr=9.0, yr=8.0, zr=7.0, order=4, bcs=1, eqn_to_solve=1, scheme="rk45",
                  eps=1e-8)
    psi = smooth_field(p, seed=12)
    runs = {}
    for fused in (True, False):
        with Context(p) as ctx:
            ctx.set_fused(fused)
            ctx.upload(psi)
            info = StepInfo(0.2 + 0j)
            ctx.step("rk45", 12, info)
            runs[fused] = (info.n_accept, info.n_reject, info.n_rhs, info.t, info.dt, ctx.download())
    f, s = runs[True], runs[False]
    print(f"rkqs: {f[0]} accepted, {f[1]} rejected, t = {f[3]:.12f} / {s[3]:.12f}, field rel {rel(f[5], s[5]):.2e}")
    assert f[:3] == s[:3] and s[1] > 0
    assert abs(f[3] - s[3]) <= 1e-8 * s[3] and abs(f[4] - s[4]) <= 1e-8 * abs(s[4])
    assert rel(f[5], s[5]) < 1e-8
    out = {}
    for fused in (True, False):
        with Context(p.replace(storage="c64")) as ctx:
            ctx.set_fused(fused)
            ctx.upload(psi)
            em, a2 = ctx.rk45_attempt(0.05 + 0j, 0.05 + 0j)
            ctx.rk45_accept()
            out[fused] = (em, a2, ctx.download())
    print("complex64 errmax fused/stage:", out[True][0], out[False][0])
    assert rel(out[True][2], out[False][2]) < 1e-6
    # the error norm is a difference of FP32 increments: its rounding noise is ~1e-3 of it here
    assert abs(out[True][0] - out[False][0]) <= 5e-2 * out[False][0]
    assert abs(out[True][1] - out[False][1]) <= 1e-6 * out[False][1]


def test_fused_rk4_imaginary_time_and_fallback(orc):
    """Complex dt through the fused kernels (eqn 2, imaginary time), and a
    right-hand side with Urhs d/dx, which falls back to the stage kernels."""
    p = RunParams(nx=30, ny=26, nz=22, xr=8.0, yr=7.0, zr=6.0, order=4, bcs=2, eqn_to_solve=2, real_time=False,
                  tau=0.004)
    psi = smooth_field(p, seed=8)
    exp, t = orc.rk4_n(p, psi, p.dt0(), 5)
    with Context(p) as ctx:
        ctx.upload(psi)
        info = StepInfo(p.dt0())
        ctx.step("rk4", 5, info)
        got = ctx.download()
        prof = ctx.kernel_profile(lambda: ctx.step("rk4", 1, StepInfo(p.dt0())))
    assert rel(got, exp) < 1e-12
    assert abs(info.im_t - t[1]) < 1e-15
    assert set(prof) == {"rk4_fused12", "rk4_fused34"}
    pu = p.replace(Urhs=0.3, real_time=True)
    with Context(pu) as ctx:
        ctx.upload(psi)
        profu = ctx.kernel_profile(lambda: ctx.step("rk4", 1, StepInfo(0.01 + 0j)))
    assert "rk4_fused12" not in profu and "rk4_stage3" in profu


@pytest.mark.parametrize("shape", [(30, 26, 22), (37, 21, 19), (6, 6, 6)])
@pytest.mark.parametrize("order,bcs", [(4, 2), (2, 2), (4, 1)])
@pytest.mark.parametrize("real_time", [True, False])
def test_fused_rk4_with_trap(orc, shape, order, bcs, real_time):
    """Eqn 4 with an anisotropic trap through the fused kernels (the trap
    evaluated at the mapped indices of unfolded cells): against the FP64
    oracle (1e-12) and the four stage kernels (1e-13), in real and
    imaginary time."""
    nx, ny, nz = shape
    p = RunParams(nx=nx, ny=ny, nz=nz, xr=0.3 * nx, yr=0.3 * ny, zr=0.3 * nz, order=order, bcs=bcs, eqn_to_solve=4,
                  g=1.0, mu=2.0, omx=0.3, omy=0.2, omz=0.4, real_time=real_time, tau=0.004)
    psi = smooth_field(p, seed=11)
    exp, t = orc.rk4_n(p, psi, p.dt0(), 5)
    outs = {}
    for fused in (True, False):
        with Context(p) as ctx:
            ctx.set_fused(fused)
            ctx.upload(psi)
            ctx.step("rk4", 5, StepInfo(p.dt0()))
            outs[fused] = ctx.download()
            prof = ctx.kernel_profile(lambda: ctx.step("rk4", 1, StepInfo(p.dt0())))
            assert ("rk4_fused12" in prof) == fused
    assert rel(outs[True], exp) < 1e-12
    assert rel(outs[True], outs[False]) < 1e-13


@pytest.mark.gpu
@pytest.mark.parametrize("storage", ["c128", "c64"])
def test_imprint_line_matches_host_generator(storage):
    """gpe_imprint_line: psi <- |psi| vortex_line(vl) against the host
    restatement of vortex_line (itself pinned to the reference build) for
    every direction, a displaced line, a phase-only line and scal != 1.
    1e-13 (c128): device sin/cos/atan2/exp/pow and numpy's differ by a few
    ulp; c64 rounds the stored result (2e-7)."""
    from gpe_amd import ics

    p = RunParams(nx=20, ny=18, nz=16, xr=4.0, yr=3.5, zr=3.0, storage=storage)
    rng = np.random.default_rng(3)
    psi = (1.0 + 0.3 * rng.standard_normal(p.shape) + 0.3j * rng.standard_normal(p.shape)).astype(np.complex128)
    cases = [(ics.LineParam(0.3, -0.5, 0.2, 0.4, 0.3, 6.0, 1.0, "x", False), 1.0),
             (ics.LineParam(-0.2, 0.1, 0.4, 0.5, 0.2, 5.0, -1.0, "y", True), 0.8),
             (ics.LineParam(0.25, -0.35, 0.0, 0.0, 0.0, 0.0, 1.0, "z", False), 1.2)]
    tol = 1e-13 if storage == "c128" else 2e-7
    with Context(p) as ctx:
        for vl, scal in cases:
            ctx.upload(psi)
            ctx.imprint_line(vl, scal)
            got = ctx.download()
            exp = np.abs(psi) * ics.vortex_line(p, vl, scal)
            err = np.abs(got - exp).max() / np.abs(exp).max()
            print(f"{storage} {vl.dir}: {err:.2e}")
            assert err <= tol
        assert ctx.peek(3, 4, 5) == got[5, 4, 3] or storage == "c64"
        assert abs(ctx.peek(3, 4, 5) - got[5, 4, 3]) <= tol * abs(got[5, 4, 3])
        with pytest.raises(GpeError):
            ctx.imprint_line(ics.LineParam(amp1=0.5, ll=0.0))
        with pytest.raises(GpeError):
            ctx.peek(20, 0, 0)


@pytest.mark.gpu
@pytest.mark.parametrize("order,bcs,eqn", [(2, 1, 1), (4, 2, 1), (4, 2, 4), (2, 2, 2)])
def test_fused_rk4_complex64(orc, order, bcs, eqn):
    """Complex64 storage through the fused kernels (FP32 stencil and stages):
    against the complex64 stage kernels and the FP64 oracle at FP32 accuracy
    (6 steps: 2e-6 of the field)."""
    p = RunParams(nx=37, ny=21, nz=19, xr=0.35 * 37, yr=0.35 * 21, zr=0.35 * 19, order=order, bcs=bcs,
                  eqn_to_solve=eqn, storage="c64")
    if eqn == 4:
        p = p.replace(g=1.0, mu=2.0, omx=0.3, omy=0.2, omz=0.4)
    psi = smooth_field(p, seed=12)
    exp, _ = orc.rk4_n(p, psi, 0.01 + 0j, 6)
    outs = {}
    for fused in (True, False):
        with Context(p) as ctx:
            ctx.set_fused(fused)
            ctx.upload(psi.astype(np.complex64))
            ctx.step("rk4", 6, StepInfo(0.01 + 0j))
            outs[fused] = ctx.download()
            prof = ctx.kernel_profile(lambda: ctx.step("rk4", 1, StepInfo(0.01 + 0j)))
            assert ("rk4_fused12" in prof) == fused
    print(f"c64 fused vs oracle {rel(outs[True], exp):.2e}, vs stage kernels {rel(outs[True], outs[False]):.2e}")
    assert rel(outs[True], exp) < 2e-6
    assert rel(outs[True], outs[False]) < 2e-6


def plane_wave_rk4(lam, amp, dt, nsteps):
    """The classic RK4 map of y' = (i/2)(lam + 1 - |y|^2) y: a plane wave of
    discrete Laplacian symbol lam stays a plane wave under eqn 1, with this
    amplitude (every cell has the same |psi|, so the nonlinearity is uniform)."""
    f = lambda y: 0.5j * (lam + 1.0 - abs(y) ** 2) * y  # noqa: E731
    y = complex(amp)
    for _ in range(nsteps):
        k1 = f(y)
        k2 = f(y + 0.5 * dt * k1)
        k3 = f(y + 0.5 * dt * k2)
        k4 = f(y + dt * k3)
        y = y + dt * (k1 / 6 + k2 / 3 + k3 / 3 + k4 / 6)
    return y


@pytest.mark.gpu
@pytest.mark.parametrize("storage,fused,tol", [("c128", True, 1e-13), ("c128", False, 1e-13), ("c64", True, 5e-6),
                                               ("c64", False, 5e-6)])
def test_full_size_plane_wave_closed_form(storage, fused, tol):
    """512^3, 4th order, periodic (the benchmark grid): a plane wave with
    three different wavenumbers evolves exactly as the scalar RK4 map
    predicts in every cell (all tiles, z-chunks and wraps): to 1e-13 with
    complex128 storage, to 5e-6 (FP32 rounding of the stored field and the
    stencil, over 5 steps) with complex64 storage."""
    n, dx = 512, 0.3
    xr = dx * (n - 1) / 2
    p = RunParams(nx=n, ny=n, nz=n, xr=xr, yr=xr, zr=xr, order=4, bcs=1, eqn_to_solve=1, storage=storage)
    m = (3, 5, 7)
    th = [2 * np.pi * mi / n for mi in m]  # k dx
    lam = sum((-2 * np.cos(2 * t) + 32 * np.cos(t) - 30) / (12 * dx * dx) for t in th)
    idx = np.arange(n)
    amp = 0.8 + 0.3j
    ex = np.exp(1j * th[0] * idx)
    ey = np.exp(1j * th[1] * idx)
    ez = np.exp(1j * th[2] * idx)
    psi = amp * (ez[:, None, None] * ey[None, :, None] * ex[None, None, :])
    dt, steps = 0.005, 5
    y = plane_wave_rk4(lam, amp, dt, steps)
    with Context(p) as ctx:
        ctx.set_fused(fused)
        ctx.upload(psi)
        ctx.step("rk4", steps, StepInfo(dt + 0j))
        got = ctx.download()
    exp = psi * (y / amp)
    del psi
    err = float(np.abs(got - exp).max())
    print(f"512^3 plane wave ({storage}, {'fused' if fused else 'stage kernels'}): max |error| = {err:.2e}")
    assert err <= tol


@pytest.mark.gpu
def test_full_size_fused_matches_stage_kernels():
    """512^3 random-phase workload: fused and four-stage RK4 agree to 1e-13
    after 4 steps, and the discrete norm is conserved to the RK4 error."""
    p, psi = ics.config_turb(n=512, scheme="rk4", tau=0.005)
    outs, norms = {}, {}
    for fused in (True, False):
        with Context(p) as ctx:
            ctx.set_fused(fused)
            ctx.upload(psi)
            n0 = ctx.norm()
            ctx.step("rk4", 4, StepInfo(0.005 + 0j))
            outs[fused] = ctx.download()
            norms[fused] = ctx.norm() / n0 - 1.0
    r = rel(outs[True], outs[False])
    print(f"512^3 fused vs stage kernels: rel L2 {r:.2e}; norm drift {norms[True]:.2e}")
    assert r <= 1e-13
    assert abs(norms[True]) < 1e-6 and abs(norms[True] - norms[False]) < 1e-12


@pytest.mark.gpu
@pytest.mark.parametrize("storage", ["c128", "mixed"])
def test_full_size_fused_rkck_matches_stage_kernels(storage):
    """512^3 random-phase workload, one Cash-Karp attempt and a retry at half
    the dt: the fused attempt (three kernels) and the six stage kernels agree
    on the candidate (1e-13 relative L2) and the error max-norm (1e-9)."""
    p, psi = ics.config_turb(n=512, scheme="rk45", tau=0.01)
    p = p.replace(storage=storage, eps=1e-6)
    res = {}
    for fused in (True, False):
        with Context(p) as ctx:
            ctx.set_fused(fused)
            ctx.upload(psi)
            em1, _ = ctx.rk45_attempt(0.02 + 0j, 0.02 + 0j)
            em2, a2 = ctx.rk45_attempt(0.01 + 0j, 0.02 + 0j)
            ctx.rk45_accept()
            res[fused] = (em1, em2, a2, ctx.download())
    f, s = res[True], res[False]
    r = rel(f[3], s[3])
    print(f"512^3 {storage} rkck fused vs stage kernels: errmax {f[0]:.6e}/{s[0]:.6e}, "
          f"{f[1]:.6e}/{s[1]:.6e}; candidate rel L2 {r:.2e}")
    assert r <= 1e-13
    for i in (0, 1):
        assert abs(f[i] - s[i]) <= 1e-9 * s[i] + 1e-9
    assert abs(f[2] - s[2]) <= 1e-13 * s[2]
