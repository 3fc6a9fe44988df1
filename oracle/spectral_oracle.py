"""CPU restatement of the reference's spectral diagnostics:
condensed_particles (src/io.f90:315-404) and spectrum (src/io.f90:632-725)
over the normalised backward transform of `fft` (src/ic.f90:905-1001),
with numpy's FFT in FP64.

The reference's `fft` needs FFTW2, which is not in the image, so these
functions are pinned by closed forms (tests/test_spectral.py): plane waves,
Parseval, the mode count, the partial-fraction form of the k-grid sum and the
closed form of the 4th-order symbol.

TEST INFRASTRUCTURE ONLY: imported by tests/, never by the product package.
"""
from __future__ import annotations

import numpy as np

PI = 3.1415926535897932384626433832795
NSHELLS = 7


def fft_backward(psi: np.ndarray) -> np.ndarray:
    """a = sum_x psi(x) exp(+2 pi i k.x / n) / sqrt(N) (FFTW_BACKWARD, then
    the division by sqrt(nx ny nz) of `fft`), array (z, y, x)."""
    n = psi.size
    return np.fft.ifftn(psi) * (n / np.sqrt(float(n)))


def mode_sq(n: int) -> np.ndarray:
    """ii2 of spectrum (src/io.f90:651-668): i^2 for i <= (n-1)/2 + 1, else
    (n - i)^2."""
    i = np.arange(n)
    m = np.where(i <= (n - 1) // 2 + 1, i, n - 1 - i + 1)
    return (m * m).astype(np.int64)


def k2_grid(p) -> np.ndarray:
    return mode_sq(p.nz)[:, None, None] + mode_sq(p.ny)[None, :, None] + mode_sq(p.nx)[None, None, :]


def spectrum(p, psi: np.ndarray):
    """Returns (eta, nharm, F): per shell m = 1..7 (|log2(sqrt(k2)/(2 pi))|
    in [m-1, m)) the mean |a|^2 (0 for an empty sum) and the mode count, and
    F(kk) = sum of |a|^2 over modes with sqrt(k2) <= kk, kk = 1..nx/2."""
    a2 = np.abs(fft_backward(psi)) ** 2
    k2 = k2_grid(p)
    nz = k2 != 0
    a2, k2 = a2[nz], k2[nz]
    rk = np.sqrt(k2.astype(np.float64))
    lk = np.abs(np.log(0.5 * rk / PI) / np.log(2.0))
    eta = np.zeros(NSHELLS)
    nharm = np.zeros(NSHELLS, dtype=np.int64)
    for m in range(1, NSHELLS + 1):
        sel = (lk < m) & (lk >= m - 1)
        eta[m - 1] = a2[sel].sum()
        nharm[m - 1] = int(sel.sum())
    eta = np.where(eta == 0.0, 0.0, eta / np.maximum(nharm, 1))
    F = np.array([a2[rk <= kk].sum() for kk in range(1, p.nx // 2 + 1)])
    return eta, nharm, F


def tot_sum(p, rho0: float) -> float:
    """The k-grid sum of condensed_particles (src/io.f90:352-369):
    kx = -pi + i dk with dk = 2 pi/(nx - 1) on every axis, k2 the 4th-order
    symbol -(1/12) sum_a (-2 cos(2 k_a d_a) + 32 cos(k_a d_a) - 30)/d_a^2,
    terms (k2 + rho0)/(k2^2 + 2 rho0 k2) for k2 != 0."""
    dk = 2.0 * PI / (p.nx - 1)

    def axis(n, d):
        kk = -PI + np.arange(n) * dk
        return (-2.0 * np.cos(2.0 * kk * d) + 32.0 * np.cos(kk * d) - 30.0) / (d * d)

    sx, sy, sz = axis(p.nx, p.dx), axis(p.ny, p.dy), axis(p.nz, p.dz)
    k2 = -(1.0 / 12.0) * (sx[None, None, :] + sy[None, :, None] + sz[:, None, None])
    k2 = k2[k2 != 0.0]
    return float(((k2 + rho0) / (k2 * k2 + 2.0 * rho0 * k2)).sum())


def condensed(p, psi: np.ndarray) -> np.ndarray:
    """The mass.dat values after (time, im_t) (src/io.f90:375-386):
    M/(8 xr yr zr), n0/V, M, n0, temp, temp2, H/V."""
    V = float(p.nx * p.ny * p.nz)
    a = fft_backward(psi)
    n0 = float(abs(a[0, 0, 0]) ** 2)
    M = float((np.abs(psi) ** 2).sum() * p.dx * p.dy * p.dz)
    tot = tot_sum(p, n0 / V)
    temp = (M - n0) / tot
    E0 = (1.0 / (2.0 * V)) * (M ** 2 + (M - n0) ** 2)
    H = E0 + temp * (V - 1.0)
    vol = 8.0 * p.xr * p.yr * p.zr
    temp2 = ((M / vol) - (n0 / V)) / tot
    return np.array([M / vol, n0 / V, M, n0, temp, temp2, H / V])
