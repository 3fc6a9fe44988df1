"""State diagnostics: momentum, velocity statistics and PDF, velocity
correlation function.  The CPU oracle (oracle/diag_oracle.py) is pinned
against the reference's own routines (momentum, get_pdf_velocity, get_pdf,
get_vcf compiled unmodified, tests/golden/ref_diag_*.npz); the GPU (C ABI)
is compared with both.

Tolerances: momentum to 1e-12 of the quadrature of the integrand's absolute
value (a component can cancel to ~0); mean and stdev to 2 N u (N values
summed in a different order, u the unit roundoff) of the stdev scale; vmax
to 1e-12 relative; counts exactly; the PDF to within one count per bin and
two counts in all (a value within an ulp of a bin edge may fall on either
side); the VCF to 1e-10.
"""
import glob
import json
import os

import numpy as np
import pytest

from gpe_amd import Context, RunParams, ics

HERE = os.path.dirname(os.path.abspath(__file__))
FIXTURES = sorted(glob.glob(os.path.join(HERE, "golden", "ref_diag_*.npz")))


def load(path):
    z = np.load(path)
    p = RunParams(**json.loads(str(z["params"])))
    if "inp" in z.files:
        psi = np.ascontiguousarray(z["inp"])
    else:
        p, psi = ics.config_ring64(seed=0)
    return p, psi, z


def ids(paths):
    return [os.path.basename(x)[9:-4] for x in paths]


def momentum_scale(p, psi):
    from oracle import diag_oracle as dg

    out = []
    for d in dg.grad(p, psi):
        f = np.abs((psi.real - 1.0) * np.conj(d).imag) + np.abs(psi.imag * d.real)
        out.append(dg.integrate(p, f))
    return np.array(out)


def check_stats(got, ref, nbins):
    assert np.array_equal(np.asarray(got["counts"]), np.asarray(ref["counts"]))
    sd, mean = np.asarray(ref["vstdev"]), np.asarray(ref["vmean"])
    # sums of N values in different orders differ by up to ~N u (sum of |v|)
    nu = 2.0 * np.asarray(ref["counts"], dtype=np.float64) * np.finfo(np.float64).eps
    print(f"  vmean diff {np.asarray(got['vmean']) - mean!r}, vstdev rel {np.asarray(got['vstdev']) / sd - 1!r}")
    assert np.all(np.abs(got["vmean"] - mean) <= nu * (np.abs(mean) + sd))
    assert np.all(np.abs(np.asarray(got["vstdev"]) / sd - 1) <= nu * (1 + (mean / sd) ** 2))
    assert np.allclose(got["vmax"], ref["vmax"], rtol=1e-12, atol=0)
    total = np.asarray(ref["counts"], dtype=np.float64)[:, None]
    dc = np.abs(np.asarray(got["pdf"]) - np.asarray(ref["pdf"])) * total
    assert dc.max() <= 1.0 + 1e-6 and dc.sum(axis=1).max() <= 2.0 + 1e-6


def test_quadrature_weights_integrate_cubics_exactly():
    from oracle import diag_oracle as dg

    p = RunParams(nx=11, ny=9, nz=8, xr=2.0, yr=1.5, zr=1.0)
    x, y, z = p.axes()
    Z, Y, X = np.meshgrid(z, y, x, indexing="ij")
    f = (1 + X + X ** 2 + X ** 3) * (2 - Y ** 3) * (1 + Z ** 2)
    exact = (2 * p.xr + 2 * p.xr ** 3 / 3) * (2 * 2 * p.yr) * (2 * p.zr + 2 * p.zr ** 3 / 3)
    assert abs(dg.integrate(p, f) - exact) < 1e-12 * abs(exact)


def test_momentum_of_a_plane_wave_closed_form(orc):
    """psi = 1 + a exp(i k x) on a periodic grid: the integrand's oscillating
    parts sum to zero over whole periods, leaving -a^2 k_h per unit volume,
    k_h = sin(k dx)/dx (2nd order) the discrete derivative symbol."""
    from oracle import diag_oracle as dg

    p = RunParams(nx=24, ny=8, nz=8, xr=6.0, yr=2.0, zr=2.0, order=2, bcs=1)
    # periodic grid: period n dx; three whole waves along x
    k = 2 * np.pi * 3 / (p.nx * p.dx)
    x, y, z = p.axes()
    a = 0.3
    psi = np.ascontiguousarray(np.broadcast_to(1 + a * np.exp(1j * k * x)[None, None, :], p.shape)).astype(np.complex128)
    kh = np.sin(k * p.dx) / p.dx
    m = dg.momentum(p, psi)
    # the end-corrected weights sum to (n-1) per axis, so the volume is (n-1)^3 dx dy dz ...
    w = dg.qweights(p.nx)
    e = np.exp(1j * k * x)
    # ... but along x the oscillation meets the non-uniform end weights: integrate exactly
    integrand_x = (np.real(1 + a * e) - 1) * np.imag(np.conj(1j * kh * a * e)) - np.imag(a * e) * np.real(1j * kh * a * e)
    expect_x = (w * integrand_x).sum() * p.dx * (p.ny - 1) * p.dy * (p.nz - 1) * p.dz
    assert abs(m[0] - expect_x) < 1e-12 * max(1.0, abs(expect_x))
    assert abs(m[1]) < 1e-13 and abs(m[2]) < 1e-13


@pytest.mark.parametrize("path", FIXTURES, ids=ids(FIXTURES))
def test_oracle_matches_reference_diagnostics(orc, path):
    from oracle import diag_oracle as dg

    p, psi, z = load(path)
    nbins = int(z["nbins"])
    m = dg.momentum(p, psi)
    assert np.all(np.abs(m - z["mom"]) <= 1e-12 * momentum_scale(p, psi))
    check_stats(dg.velocity_pdf(p, psi, nbins), z, nbins)
    assert np.allclose(dg.vcf(p, psi), z["vcf"], rtol=1e-10, atol=1e-12)


@pytest.mark.gpu
@pytest.mark.parametrize("path", FIXTURES, ids=ids(FIXTURES))
@pytest.mark.parametrize("storage", ["c128", "mixed"])
def test_gpu_diagnostics_match_reference(orc, path, storage):
    p, psi, z = load(path)
    nbins = int(z["nbins"])
    with Context(p.replace(storage=storage)) as ctx:
        ctx.upload(psi)
        m = ctx.momentum()
        st = ctx.velocity_pdf(nbins)
        f = ctx.vcf()
    print(f"{os.path.basename(path)}: momentum {m!r}, ref {z['mom']!r}, diff {m - z['mom']!r}, "
          f"tol {1e-12 * momentum_scale(p, psi)!r}; vmean {st['vmean']!r} ref {z['vmean']!r}")
    assert np.all(np.abs(m - z["mom"]) <= 1e-12 * momentum_scale(p, psi))
    check_stats(st, z, nbins)
    assert np.allclose(f, z["vcf"], rtol=1e-10, atol=1e-12)


@pytest.mark.gpu
def test_gpu_diagnostics_c64_and_larger_grid(orc):
    """512^2 x 64 random-phase field: GPU against the oracle (c128), and the
    complex64 state at FP32 accuracy."""
    from oracle import diag_oracle as dg

    p = RunParams(nx=96, ny=80, nz=48, xr=0.3 * 95 / 2, yr=0.3 * 79 / 2, zr=0.3 * 47 / 2, order=4, bcs=1)
    psi = ics.random_phase(p, seed=5)
    nbins = 128
    exp_m, exp_st, exp_f = dg.momentum(p, psi), dg.velocity_pdf(p, psi, nbins), dg.vcf(p, psi)
    with Context(p) as ctx:
        ctx.upload(psi)
        m, st, f = ctx.momentum(), ctx.velocity_pdf(nbins), ctx.vcf()
    assert np.all(np.abs(m - exp_m) <= 1e-12 * momentum_scale(p, psi))
    check_stats(st, exp_st, nbins)
    assert np.allclose(f, exp_f, rtol=1e-10, atol=1e-12)
    with Context(p.replace(storage="c64")) as ctx:
        ctx.upload(psi)
        m64, st64 = ctx.momentum(), ctx.velocity_pdf(nbins)
    assert np.all(np.abs(m64 - exp_m) <= 1e-5 * momentum_scale(p, psi))
    assert np.allclose(st64["vstdev"], exp_st["vstdev"], rtol=1e-4)


@pytest.mark.gpu
@pytest.mark.parametrize("nbins", [2, 1364, 1366, 4096])
def test_gpu_velocity_pdf_bin_counts(nbins):
    """The histogram either side of nbins 1364, the largest count whose bin
    edges the kernel tabulates in LDS (above it each comparison forms its
    edge), up to the largest count accepted, and nbins 2 (one bin each side
    of zero): against the oracle's get_pdf."""
    from oracle import diag_oracle as dg

    p = RunParams(nx=40, ny=36, nz=32, xr=0.3 * 39 / 2, yr=0.3 * 35 / 2, zr=0.3 * 31 / 2, order=4, bcs=1)
    psi = ics.random_phase(p, seed=7)
    exp = dg.velocity_pdf(p, psi, nbins)
    with Context(p) as ctx:
        ctx.upload(psi)
        st = ctx.velocity_pdf(nbins)
    check_stats(st, exp, nbins)


@pytest.mark.gpu
@pytest.mark.parametrize("nx", [40, 100, 256, 300, 512, 1000, 1500])
def test_gpu_vcf_line_lengths(nx):
    """The velocity correlation over every lag-block width of the kernel
    (K = 1..16 lags per lane, nx up to 1024) and the direct-sum kernel beyond
    (nx 1500), against the oracle's get_vcf to 1e-10."""
    from oracle import diag_oracle as dg

    ny, nz = 7, 6
    p = RunParams(nx=nx, ny=ny, nz=nz, xr=0.3 * (nx - 1) / 2, yr=0.3 * (ny - 1) / 2, zr=0.3 * (nz - 1) / 2, order=4,
                  bcs=1)
    rng = np.random.default_rng(nx)
    x = np.arange(nx)
    psi = np.ascontiguousarray(np.exp(1j * (2 * np.pi * 3 * x / nx)[None, None, :]) *
                               (1.0 + 0.2 * rng.standard_normal((nz, ny, nx))))
    exp_f = dg.vcf(p, psi)
    with Context(p) as ctx:
        ctx.upload(psi)
        f = ctx.vcf()
    assert np.allclose(f, exp_f, rtol=1e-10, atol=1e-12)


LL_FIXTURES = sorted(glob.glob(os.path.join(HERE, "golden", "ref_ll_*.npz")))


def ll_cases():
    """(name, params, psi, reference line length) of every fixture that holds one."""
    out = []
    for path in FIXTURES + LL_FIXTURES:
        p, psi, z = load(path)
        out.append((os.path.basename(path)[4:-4], p, psi, float(z["linelength"])))
    return out


def test_oracle_linelength_matches_reference_bitwise(orc):
    """The C restatement visits cells in the reference's order with the same
    operation order, so it reproduces the reference's sum exactly: smooth
    fields (vortex ring, trapped line, turbulence), white noise (682 cells
    with one hit that reuse the previous cell's second hit), a line along a
    grid edge (the 1/4 rule) and a zero plane (the 1/2 rule)."""
    from oracle import oracle as orc_

    seen = set()
    for name, p, psi, ref in ll_cases():
        v, counts = orc_.linelength(p, psi)
        print(f"{name}: {v!r} vs {ref!r}, counts {counts}")
        assert v == ref
        seen.update(k for k in range(4) if counts[k])
    assert seen == {0, 1, 2, 3}  # every branch is exercised


@pytest.mark.gpu
@pytest.mark.parametrize("storage", ["c128", "mixed"])
def test_gpu_linelength_matches_reference(storage):
    """Per cell the device reproduces the reference's arithmetic (contraction
    off); only the order of the final sum differs: 1e-13 relative."""
    for name, p, psi, ref in ll_cases():
        with Context(p.replace(storage=storage)) as ctx:
            ctx.upload(psi)
            got = ctx.linelength()
        print(f"{name} {storage}: {got!r} vs {ref!r}")
        assert abs(got - ref) <= 1e-13 * max(ref, 1.0)


@pytest.mark.gpu
def test_gpu_linelength_of_a_ring_and_long_lines(orc):
    """A vortex ring of radius 20 on a 128^3 grid (length close to 2 pi 20),
    a white-noise field on a ragged grid with lines longer than one chunk of
    64 cells, both against the restatement."""
    from oracle import oracle as orc_

    p = RunParams(nx=128, ny=128, nz=128, xr=31.75, yr=31.75, zr=31.75)
    psi = ics.vortex_ring(p, ics.RingParam("xz", 0.3, 0.2, -0.1, 20.0))
    ref, _ = orc_.linelength(p, psi)
    assert abs(ref - 2 * np.pi * 20) < 0.5
    rng = np.random.default_rng(4)
    q = RunParams(nx=150, ny=37, nz=21, xr=7.45, yr=1.8, zr=1.0)
    noise = rng.standard_normal(q.shape) + 1j * rng.standard_normal(q.shape)
    ref2, counts = orc_.linelength(q, noise)
    assert counts[1] > 0
    for pp, f, r in ((p, psi, ref), (q, noise, ref2)):
        with Context(pp) as ctx:
            ctx.upload(f)
            got = ctx.linelength()
        assert abs(got - r) <= 1e-13 * r
