"""Batched ensemble (gpe_ens_*): every member must equal the single-context
run of the same member in every mode (one launch per member, batched
persistent launches, hipGraph replay), including graph replays interleaved
with eager work."""
import ctypes as ct

import numpy as np
import pytest

from gpe_amd import Context, RunParams, StepInfo, _lib
from gpe_amd.ensemble import Ensemble, ensemble_config


def members(order, bcs, storage="c128"):
    shapes = [(16, 20, 24), (24, 24, 24), (32, 16, 20), (20, 28, 12), (40, 8, 10)]
    out = []
    for i, (nx, ny, nz) in enumerate(shapes):
        out.append(RunParams(nx=nx, ny=ny, nz=nz, xr=0.3 * (nx - 1) / 2, yr=0.3 * (ny - 1) / 2,
                             zr=0.3 * (nz - 1) / 2, order=order, bcs=bcs, eqn_to_solve=1, storage=storage))
    return out


def fields(ms, seed=0):
    rng = np.random.default_rng(seed)
    return [np.ascontiguousarray(1.0 + 0.2 * (rng.standard_normal(m.shape) + 1j * rng.standard_normal(m.shape)))
            for m in ms]


def single(m, psi, nsteps, dt):
    with Context(m) as ctx:
        ctx.upload(psi)
        ctx.step("rk4", nsteps, StepInfo(dt))
        return ctx.download()


def test_config_sizes_and_cells():
    ms, fs = ensemble_config(n_members=6)
    assert all(m.nx in (64, 96, 128, 160, 192) and m.nx == m.ny == m.nz for m in ms)
    assert all(abs(m.dx - 0.3) < 1e-15 for m in ms)
    assert all(abs(np.mean(np.abs(f) ** 2) - 1.0) < 1e-12 for f in fs)


def test_create_rejects_mixed_members():
    lib = _lib.load()
    a = RunParams(nx=16, ny=16, nz=16, order=2).to_c()
    b = RunParams(nx=16, ny=16, nz=16, order=4).to_c()
    arr = (_lib.gpe_params * 2)(a, b)
    h = ct.c_void_p()
    assert lib.gpe_ens_create(arr, 2, ct.byref(h)) == -1
    assert b"share" in lib.gpe_ens_last_error(None)


@pytest.mark.gpu
@pytest.mark.parametrize("order,bcs", [(2, 1), (4, 2)])
@pytest.mark.parametrize("mode", ["per_member", "batched", "graph", "grouped", "fused", "fused_graph", "per_member_fused"])
def test_modes_match_single_context(order, bcs, mode):
    ms = members(order, bcs)
    fs = fields(ms)
    dt = 0.005 + 0j
    exp = [single(m, f, 7, dt) for m, f in zip(ms, fs)]
    with Ensemble(ms) as ens:
        # small groups (about two members each) visited 3 steps at a time: 7 = 3 + 3 + 1
        ens.set_grouping(2 * 4 * 16 * 24 * 24 * 24, 3)
        assert ens.num_groups() >= 2
        for i, f in enumerate(fs):
            ens.upload(i, f)
        ens.step(mode, dt, 7)
        for i in range(len(ms)):
            got = ens.download(i)
            assert np.array_equal(got, exp[i]) or np.linalg.norm(got - exp[i]) <= 1e-14 * np.linalg.norm(exp[i])


@pytest.mark.gpu
@pytest.mark.parametrize("tiles", [(32, 8), (64, 4), (64, 8), (32, 4)])
def test_graph_replays_with_eager_work_between(tiles):
    """Capture once, replay several times with eager batched steps and host
    syncs in between; every replay must match the eager reference."""
    ms = members(2, 1)
    fs = fields(ms, seed=3)
    dt = 0.004 + 0j
    with Ensemble(ms) as eager, Ensemble(ms) as graphed:
        for e in (eager, graphed):
            e.set_tiles(*tiles, tz=[3, 5, 4, 2, 6])
            for i, f in enumerate(fs):
                e.upload(i, f)
        for _ in range(3):
            eager.step("batched", dt, 6)
            graphed.step("graph", dt, 6)       # one chunk of 6 steps (captured on the first call)
            graphed.sync()
            eager.step("batched", dt, 1)       # eager work between replays
            graphed.step("batched", dt, 1)
            for i in range(len(ms)):
                assert np.array_equal(graphed.download(i), eager.download(i))


@pytest.mark.gpu
def test_c64_ensemble_matches_single_context():
    ms = members(2, 1, storage="c64")
    fs = [f.astype(np.complex64) for f in fields(ms, seed=5)]
    dt = 0.005 + 0j
    with Ensemble(ms) as ens:
        for i, f in enumerate(fs):
            ens.upload(i, f)
        ens.step("graph", dt, 5)
        for i, (m, f) in enumerate(zip(ms, fs)):
            with Context(m) as ctx:
                ctx.upload(f)
                ctx.step("rk4", 5, StepInfo(dt))
                exp = ctx.download(np.complex64)
            got = ens.download(i)
            assert np.linalg.norm(got - exp) <= 1e-6 * np.linalg.norm(exp)


@pytest.mark.gpu
@pytest.mark.parametrize("dt", [-0.005j, 0.004 - 0.003j], ids=["imaginary-dt", "complex-dt"])
@pytest.mark.parametrize("mode", ["fused", "fused_graph", "per_member_fused", "batched"])
def test_modes_match_single_context_other_dt_forms(mode, dt):
    """The fused batched kernels are instantiated per form of the step
    constant i*pim*dt (real, imaginary, complex dt): every form against the
    single-context run of each member (1e-14)."""
    ms = members(4, 2)
    fs = fields(ms, seed=5)
    exp = [single(m, f, 4, dt) for m, f in zip(ms, fs)]
    with Ensemble(ms) as ens:
        for i, f in enumerate(fs):
            ens.upload(i, f)
        ens.step(mode, dt, 4)
        for i in range(len(ms)):
            got = ens.download(i)
            assert np.linalg.norm(got - exp[i]) <= 1e-14 * np.linalg.norm(exp[i]), (mode, dt, i)


@pytest.mark.gpu
def test_fused_modes_alternate_buffers_and_replay():
    """Fused modes keep each member's state in one of two buffers: odd step
    counts, graph replays with eager fused steps between and downloads in
    between must agree bit for bit with eager fused steps alone, and with the
    single-context fused path to 1e-14 (the batched kernels inline the same
    stage code into another kernel, which may contract differently)."""
    ms = members(2, 1)
    fs = fields(ms, seed=9)
    dt = 0.004 + 0j
    with Ensemble(ms) as ens, Ensemble(ms) as eager:
        for e in (ens, eager):
            for i, f in enumerate(fs):
                e.upload(i, f)
        total = 0
        for mode, n in (("fused", 3), ("fused_graph", 6), ("fused_graph", 7), ("fused", 1), ("fused_graph", 4)):
            ens.step(mode, dt, n)
            eager.step("fused", dt, n)
            total += n
            for i, (m, f) in enumerate(zip(ms, fs)):
                got = ens.download(i)
                assert np.array_equal(got, eager.download(i)), (mode, n, i)
                exp = single(m, f, total, dt)
                assert np.linalg.norm(got - exp) <= 1e-14 * np.linalg.norm(exp), (mode, n, i)


@pytest.mark.gpu
def test_fused_modes_reject_optional_terms():
    ms = [m.replace(eqn_to_solve=4, g=1.0, mu=2.0, omx=0.1, omy=0.1, omz=0.1) for m in members(2, 2)]
    with Ensemble(ms) as ens:
        with pytest.raises(RuntimeError, match="fused modes"):
            ens.step("fused", 0.004 + 0j, 1)
