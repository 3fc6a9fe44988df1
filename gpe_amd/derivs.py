"""Mirror of the reference's module `derivs` (src/derivs.f90:20-21):
deriv_x, deriv_y, deriv_z, deriv_xx, deriv_yy, deriv_zz with the
reference's (f, fx) argument convention, computed by libgpehip.so."""
from __future__ import annotations

import numpy as np

from . import parameters as P


def _d(op: str, f: np.ndarray, out: np.ndarray) -> None:
    out[...] = P.context().deriv(op, np.ascontiguousarray(f))


def deriv_x(f, fx):
    _d("x", f, fx)


def deriv_y(f, fy):
    _d("y", f, fy)


def deriv_z(f, fz):
    _d("z", f, fz)


def deriv_xx(f, fxx):
    _d("xx", f, fxx)


def deriv_yy(f, fyy):
    _d("yy", f, fyy)


def deriv_zz(f, fzz):
    _d("zz", f, fzz)
