"""Mirror of the hot-path routines of the reference's module `variables`:
laplacian (src/variables.f90:55-71), get_norm (537-559), energy (563-587),
mass (591-611), renormalise (698-709).  Computed by libgpehip.so on the
context's auxiliary field: the array handed in is never copied into the
device state, so a stale or auxiliary array (the reference's
save_energy(t, psi%old)) leaves a device-resident run untouched."""
from __future__ import annotations

from contextlib import contextmanager

import numpy as np

from . import parameters as P


def laplacian(in_var: np.ndarray) -> np.ndarray:
    return P.context().deriv("lap", np.ascontiguousarray(in_var))


@contextmanager
def _on_aux(in_var: np.ndarray):
    ctx = P.context()
    ctx.upload_aux(np.ascontiguousarray(in_var))
    ctx.select(True)
    try:
        yield ctx
    finally:
        ctx.select(False)


def get_norm(in_var: np.ndarray) -> float:
    with _on_aux(in_var) as ctx:
        return ctx.norm()


def energy(in_var: np.ndarray) -> float:
    with _on_aux(in_var) as ctx:
        return ctx.energy()


def mass(in_var: np.ndarray) -> float:
    with _on_aux(in_var) as ctx:
        return ctx.mass()


def renormalise(var: np.ndarray, norm: float) -> None:
    with _on_aux(var) as ctx:
        ctx.renormalise(norm)
        var[...] = ctx.download(var.dtype)
