"""Mirror of the hot-path routines of the reference's module `variables`:
laplacian (src/variables.f90:55-71), get_norm (537-559), energy (563-587),
mass (591-611), renormalise (698-709).  Computed by libgpehip.so."""
from __future__ import annotations

import numpy as np

from . import parameters as P


def laplacian(in_var: np.ndarray) -> np.ndarray:
    return P.context().deriv("lap", np.ascontiguousarray(in_var))


def _on_state(in_var: np.ndarray):
    ctx = P.context()
    ctx.upload(np.ascontiguousarray(in_var))
    return ctx


def get_norm(in_var: np.ndarray) -> float:
    return _on_state(in_var).norm()


def energy(in_var: np.ndarray) -> float:
    return _on_state(in_var).energy()


def mass(in_var: np.ndarray) -> float:
    return _on_state(in_var).mass()


def renormalise(var: np.ndarray, norm: float) -> None:
    ctx = _on_state(var)
    ctx.renormalise(norm)
    var[...] = ctx.download(var.dtype)
