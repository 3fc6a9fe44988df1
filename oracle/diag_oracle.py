"""CPU restatement of the reference's state diagnostics: momentum, velocity
statistics and PDF, velocity correlation function.  Derivatives come from the
C oracle (orc_deriv, pinned by its own closed-form tests); the rest is numpy
in FP64.

TEST INFRASTRUCTURE ONLY: imported by tests/, never by the product package.
"""
from __future__ import annotations

import numpy as np

from . import oracle as orc


def qweights(n: int) -> np.ndarray:
    """End-corrected quadrature weights of integrate_x/y/z
    (src/variables.f90:49-51, 713-823): 3/8, 7/6, 23/24 at both ends."""
    w = np.ones(n)
    for e, c in ((0, 3.0 / 8.0), (1, 7.0 / 6.0), (2, 23.0 / 24.0)):
        w[e] = c
        w[n - 1 - e] = c
    return w


def integrate(p, f: np.ndarray) -> float:
    """integrate_x, then _y, then _z of a real field (z, y, x) with the
    weights and the dx, dy, dz factors (src/variables.f90:713-823)."""
    dx, dy, dz = p.dx, p.dy, p.dz
    fx = (f * qweights(p.nx)[None, None, :]).sum(axis=2) * dx
    fy = (fx * qweights(p.ny)[None, :]).sum(axis=1) * dy
    return float((fy * qweights(p.nz)).sum() * dz)


def grad(p, psi: np.ndarray):
    """deriv_x, deriv_y, deriv_z (src/derivs.f90:25-142)."""
    return [orc.deriv(p, a, psi) for a in range(3)]


def momentum(p, psi: np.ndarray) -> np.ndarray:
    """momentum (src/variables.f90:615-694):
    (Re psi - 1) Im(d psi*) - Im psi Re(d psi), integrated per axis."""
    out = []
    for d in grad(p, psi):
        dstar = np.conj(d)  # deriv of conjg(psi): the derivative weights are real
        f = (psi.real - 1.0) * dstar.imag - psi.imag * d.real
        out.append(integrate(p, f))
    return np.array(out)


def velocity(p, psi: np.ndarray, masked: bool = True) -> np.ndarray:
    """The velocity of get_pdf_velocity / get_vcf (src/variables.f90:393-400,
    478-483): -i/2 (psi* g - psi g*) / |psi|^2 per component, where the
    reference's gradient arrays g, g* are declared real (lines 387, 474), so
    they hold Re(d psi) = Re(d psi*); the expression is then
    -Im(psi) Re(d psi) / |psi|^2.  masked: 1e3 where |psi|^2 <= 0.01."""
    a2 = np.abs(psi) ** 2
    vel = []
    for d in grad(p, psi):
        g = d.real
        v = (-0.5j * (np.conj(psi) * g - psi * g)).real / a2
        if masked:
            v = np.where(a2 > 0.01, v, 1e3)
        vel.append(v)
    return np.array(vel)


def velocity_pdf(p, psi: np.ndarray, nbins: int) -> dict:
    """get_pdf_velocity + get_pdf (src/variables.f90:377-463): per component
    the valid count, mean, stdev = sqrt(<v^2> - mean^2), vmax = max(|max|,
    |min|) and the fractions in (2(i-1) vmax/nbins, 2 i vmax/nbins]."""
    vel = velocity(p, psi, masked=True)
    counts, mean, sd, vmax, pdf = [], [], [], [], []
    for v in vel:
        good = v[v < 0.5e3]
        m = good.sum() / good.size
        counts.append(good.size)
        mean.append(m)
        sd.append(np.sqrt((good ** 2).sum() / good.size - m * m))
        vm = max(abs(good.max()), abs(good.min()))
        vmax.append(vm)
        i = np.arange(-nbins // 2 + 1, nbins // 2 + 1)
        lo = 2.0 * (i - 1).astype(np.float64) * vm / nbins
        hi = 2.0 * i.astype(np.float64) * vm / nbins
        hist = np.array([np.count_nonzero((good > a) & (good <= b)) for a, b in zip(lo, hi)], dtype=np.float64)
        pdf.append(hist / hist.sum())
    return dict(counts=np.array(counts), vmean=np.array(mean), vstdev=np.array(sd), vmax=np.array(vmax),
                pdf=np.array(pdf))


def vcf(p, psi: np.ndarray) -> np.ndarray:
    """get_vcf (src/variables.f90:467-507): Q(r) = sum over i + r < nx and
    (y, z) of v_x(i) v_x(i+r) / ((nx - r) ny nz), f = Q / Q(0); unmasked v_x."""
    vx = velocity(p, psi, masked=False)[0]  # (z, y, x)
    nx = p.nx
    q = np.array([(vx[:, :, : nx - r] * vx[:, :, r:]).sum() for r in range(nx)])
    q = q / ((nx - np.arange(nx)) * p.ny * p.nz)
    return q / q[0]
