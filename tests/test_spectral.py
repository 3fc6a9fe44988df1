"""Spectral diagnostics: condensed_particles (src/io.f90:315-404) and
spectrum (src/io.f90:632-725).  The reference's `fft` needs FFTW2, which the
image lacks, so the CPU restatement (oracle/spectral_oracle.py) is pinned by
closed forms here, and the device (hipFFT + gfx950 kernels, through the C
ABI) is compared with it.

Tolerances (device vs oracle): shell counts exact; eta, F, the k-grid sum
and the mass.dat row 1e-11 relative for complex128 (sums of up to N terms in
a different order, FFTs of different algorithms), 1e-4 for complex64."""
import numpy as np
import pytest

from gpe_amd import Context, RunParams, ics


def grid(n=(12, 10, 8), **kw):
    return RunParams(nx=n[0], ny=n[1], nz=n[2], xr=3.0, yr=2.5, zr=2.0, **kw)


def plane_wave(p, m, amp=1.0):
    """amp exp(-2 pi i (mx i/nx + my j/ny + mz k/nz)): the backward
    transform puts all of it at mode index m."""
    i, j, k = np.arange(p.nx), np.arange(p.ny), np.arange(p.nz)
    ph = m[0] * i[None, None, :] / p.nx + m[1] * j[None, :, None] / p.ny + m[2] * k[:, None, None] / p.nz
    return amp * np.exp(-2j * np.pi * ph)


def test_fft_normalisation_and_parseval():
    from oracle import spectral_oracle as so

    p = grid()
    rng = np.random.default_rng(0)
    psi = rng.standard_normal(p.shape) + 1j * rng.standard_normal(p.shape)
    a = so.fft_backward(psi)
    assert np.isclose((np.abs(a) ** 2).sum(), (np.abs(psi) ** 2).sum(), rtol=1e-13)
    # uniform state: everything in the zero mode, |a0|^2 = N
    a = so.fft_backward(np.ones(p.shape, dtype=complex))
    assert np.isclose(abs(a[0, 0, 0]) ** 2, p.nx * p.ny * p.nz, rtol=1e-14)


@pytest.mark.parametrize("m", [(1, 0, 0), (0, 3, 1), (2, 2, 2), (5, 4, 3), (7, 1, 0), (9, 0, 4)])
def test_spectrum_of_a_plane_wave(m):
    """One mode holds |amp|^2 N: it lands in the shell of its k2 (the
    reference's index map) and in F from kk = ceil(sqrt(k2)) on."""
    from oracle import spectral_oracle as so

    p = grid((16, 12, 10))
    N = p.nx * p.ny * p.nz
    eta, nharm, F = so.spectrum(p, plane_wave(p, m, 0.5))
    k2 = int(so.mode_sq(p.nx)[m[0]] + so.mode_sq(p.ny)[m[1]] + so.mode_sq(p.nz)[m[2]])
    shell = int(np.floor(abs(np.log2(np.sqrt(k2) / (2 * np.pi)))))
    want = np.zeros(7)
    want[shell] = 0.25 * N / nharm[shell]
    assert np.allclose(eta, want, rtol=1e-12, atol=1e-12)
    kk = np.arange(1, p.nx // 2 + 1)
    assert np.allclose(F, np.where(kk * kk >= k2, 0.25 * N, 0.0), rtol=1e-12, atol=1e-12)


def test_mode_map_and_count():
    """Even n: |frequency|; odd n: index (n+1)/2 maps to ((n+1)/2)^2 (the
    reference's i <= nx1/2 + 1 test).  Every non-zero mode of a small grid
    falls in one of the 7 shells (sqrt(k2)/(2 pi) within 2^-7 .. 2^7)."""
    from oracle import spectral_oracle as so

    assert list(so.mode_sq(6)) == [0, 1, 4, 9, 4, 1]
    assert list(so.mode_sq(7)) == [0, 1, 4, 9, 16, 4, 1]
    p = grid((16, 12, 9))
    _, nharm, _ = so.spectrum(p, np.ones(p.shape, dtype=complex))
    assert nharm.sum() == p.nx * p.ny * p.nz - 1


def test_tot_sum_closed_forms():
    """(k2 + r)/(k2^2 + 2 r k2) = (1/k2 + 1/(k2 + 2r))/2, and the 4th-order
    symbol is 4 s^2/d^2 + (4/3) s^4/d^2 with s = sin(k d/2)."""
    from oracle import spectral_oracle as so

    p = grid((9, 8, 7))
    rho = 0.37
    dk = 2.0 * np.pi / (p.nx - 1)

    def sym(n, d):
        s = np.sin((-np.pi + np.arange(n) * dk) * d / 2)
        return 4 * s ** 2 / d ** 2 + (4.0 / 3.0) * s ** 4 / d ** 2

    k2 = sym(p.nx, p.dx)[None, None, :] + sym(p.ny, p.dy)[None, :, None] + sym(p.nz, p.dz)[:, None, None]
    k2 = k2[np.abs(k2) > 1e-12]
    want = 0.5 * (1.0 / k2 + 1.0 / (k2 + 2 * rho)).sum()
    assert np.isclose(so.tot_sum(p, rho), want, rtol=1e-12)


def test_condensed_row_of_a_uniform_state():
    """psi = 1: n0 = N, M = N dx dy dz, n0/V = 1, temp = (M - N)/tot."""
    from oracle import spectral_oracle as so

    p = grid()
    N = p.nx * p.ny * p.nz
    row = so.condensed(p, np.ones(p.shape, dtype=complex))
    M = N * p.dx * p.dy * p.dz
    assert np.isclose(row[1], 1.0, rtol=1e-14) and np.isclose(row[3], N, rtol=1e-14)
    assert np.isclose(row[2], M, rtol=1e-14)
    assert np.isclose(row[4], (M - N) / so.tot_sum(p, 1.0), rtol=1e-13)


def random_field(p, seed):
    rng = np.random.default_rng(seed)
    return (1.0 + 0.4 * (rng.standard_normal(p.shape) + 1j * rng.standard_normal(p.shape))).astype(np.complex128)


@pytest.mark.gpu
@pytest.mark.parametrize("shape", [(16, 12, 10), (33, 20, 17), (64, 64, 64)])
@pytest.mark.parametrize("storage", ["c128", "c64"])
def test_gpu_condensed_and_spectrum_match_oracle(shape, storage):
    from oracle import spectral_oracle as so

    p = RunParams(nx=shape[0], ny=shape[1], nz=shape[2], xr=6.0, yr=5.0, zr=4.0, storage=storage)
    psi = random_field(p, 2) + 0.3 * ics.vortex_ring(p, ics.RingParam("xy", r0=2.0))
    row, (eta, nharm, F) = so.condensed(p, psi), so.spectrum(p, psi)
    with Context(p) as ctx:
        ctx.upload(psi)
        got = ctx.condensed_particles()
        row_only = ctx.condensed_particles(spectrum=False)["row"]
    tol = 1e-11 if storage == "c128" else 1e-4
    print(f"{shape} {storage}: row rel {np.abs(got['row'] / row - 1).max():.2e}, "
          f"eta rel {np.abs(got['eta'] / np.where(eta == 0, 1, eta) - np.where(eta == 0, 0, 1)).max():.2e}")
    assert np.array_equal(got["nharm"], nharm)
    assert np.allclose(got["row"], row, rtol=tol, atol=0)
    assert np.array_equal(row_only, got["row"])
    assert np.allclose(got["eta"], eta, rtol=tol, atol=0)
    assert np.allclose(got["F"], F, rtol=tol, atol=tol * F.max())


@pytest.mark.gpu
def test_gpu_spectrum_of_plane_waves():
    from oracle import spectral_oracle as so

    p = RunParams(nx=32, ny=24, nz=20, xr=6.0, yr=5.0, zr=4.0)
    with Context(p) as ctx:
        for m in [(1, 0, 0), (3, 5, 2), (16, 12, 10), (0, 0, 0)]:
            psi = plane_wave(p, m, 0.7)
            ctx.upload(psi)
            got = ctx.condensed_particles()
            eta, nharm, F = so.spectrum(p, psi)
            assert np.array_equal(got["nharm"], nharm)
            assert np.allclose(got["eta"], eta, rtol=1e-11, atol=1e-9)
            assert np.allclose(got["F"], F, rtol=1e-11, atol=1e-9)
            assert np.isclose(got["row"][3], 0.49 * p.nx * p.ny * p.nz if m == (0, 0, 0) else 0.0, rtol=1e-12,
                              atol=1e-9)
