"""Module state of the time-stepping path, mirroring the reference's module
`parameters` (src/parameters.f90:15-131): the run parameters and the
globals `dt` (complex), `t`, `im_t` and `p` that `solve.solver` advances,
and `real_time`, which a driver sets when imaginary-time relaxation
converges (src/gpe.f90:251-271).

`setup(params)` plays the role of the reference's start-up (namelist read,
get_grid, dt from tau: src/gpe.f90:76-80) and creates the device context.
"""
from __future__ import annotations

from .context import Context
from .params import RunParams

run: RunParams | None = None
dt: complex = 0j
t: float = 0.0
im_t: float = 0.0
p: int = 0
real_time: bool = True
errcon: float = 0.0
ctx: Context | None = None


def setup(params: RunParams) -> Context:
    global run, dt, t, im_t, p, ctx, errcon, real_time
    if ctx is not None:
        ctx.close()
    run = params
    dt = params.dt0()
    t = 0.0
    im_t = 0.0
    p = 0
    real_time = bool(params.real_time)
    errcon = (5.0 / params.safety) ** (1.0 / params.dt_increase)
    ctx = Context(params)
    return ctx


def context() -> Context:
    if ctx is None:
        raise RuntimeError("gpe_amd.parameters.setup(params) has not been called")
    return ctx
