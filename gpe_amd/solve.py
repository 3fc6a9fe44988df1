"""Mirror of the reference's module `solve` (src/solve.f90:15-21): the one
public routine `solver(var_in, var_out)` advances the field by one step of
the configured scheme and updates `parameters.t`, `im_t` and `dt` exactly as
the reference does.  The work runs in libgpehip.so."""
from __future__ import annotations

import numpy as np

from . import _lib
from . import parameters as P
from .context import StepInfo


def solver(var_in: np.ndarray, var_out: np.ndarray) -> None:
    """solver(var_in, var_out) (src/solve.f90:25-53).  var_in is the
    interior (nz, ny, nx) field or the reference's halo array
    (nz+4, ny+4, nx); var_out the interior field (written)."""
    ctx = P.context()
    if P.run.scheme not in _lib.SCHEMES:
        raise _lib.GpeError("ERROR: Unrecognised time stepping scheme.")
    info = StepInfo(P.dt, P.t, P.im_t)
    ctx.set_real_time(P.real_time)  # rkqs reads it on a retry (src/solve.f90:195-201)
    ctx.upload(np.ascontiguousarray(var_in))
    ctx.step(P.run.scheme, 1, info)
    var_out[...] = ctx.download(var_out.dtype)
    P.t, P.im_t, P.dt = info.t, info.im_t, info.dt
