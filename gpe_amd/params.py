"""Run-time parameters of the time-stepping path.

Mirrors the reference's module `parameters` and its namelists run_params /
eqn_params (src/parameters.f90:25-55) plus the compile-time grid size of
src/parameters.in:22-29, which here is a run-time value.  `read_namelist`
reads the reference's `run.in` format so one input file drives both codes.
"""
from __future__ import annotations

import dataclasses
import re
from dataclasses import dataclass, field

from . import _lib

STORAGE = {"c128": _lib.GPE_C128, "c64": _lib.GPE_C64, "mixed": _lib.GPE_MIXED}


@dataclass
class RunParams:
    nx: int = 64
    ny: int = 64
    nz: int = 64
    xr: float = 15.75
    yr: float = 15.75
    zr: float = 15.75
    tau: float = 0.01
    end_time: float = 1.0
    scheme: str = "rk4"
    eqn_to_solve: int = 1
    bcs: int = 1
    order: int = 2
    real_time: bool = True
    Urhs: float = 0.0
    diss_amp: float = 0.0
    g: float = 1.0
    mu: float = 1.0
    nn: float = 1.0
    omx: float = 0.0
    omy: float = 0.0
    omz: float = 0.0
    vscale: float = 0.0     # coefficient of V in eqn 1 (0 = the reference's eqn 1)
    storage: str = "c128"   # "c128", "c64" or "mixed" (c128 state, c64 RK increments)
    # adaptive time stepping (src/parameters.f90:89-92)
    eps: float = 1e-6
    safety: float = 0.9
    dt_decrease: float = -0.25
    dt_increase: float = -0.20

    # ---- grid (src/ic.f90:32-69)
    @property
    def dx(self) -> float:
        return (self.xr + self.xr) / (self.nx - 1)

    @property
    def dy(self) -> float:
        return (self.yr + self.yr) / (self.ny - 1)

    @property
    def dz(self) -> float:
        return (self.zr + self.zr) / (self.nz - 1)

    def axes(self):
        import numpy as np

        x = -self.xr + np.arange(self.nx, dtype=np.float64) * self.dx
        y = -self.yr + np.arange(self.ny, dtype=np.float64) * self.dy
        z = -self.zr + np.arange(self.nz, dtype=np.float64) * self.dz
        return x, y, z

    @property
    def shape(self):
        """numpy shape of a field: (nz, ny, nx), C order = x fastest."""
        return (self.nz, self.ny, self.nx)

    @property
    def cells(self) -> int:
        return self.nx * self.ny * self.nz

    def dt0(self) -> complex:
        """Initial dt: (tau, 0) in real time, (0, -tau) in imaginary time (src/gpe.f90:76-80)."""
        return complex(self.tau, 0.0) if self.real_time else complex(0.0, -self.tau)

    def to_c(self) -> _lib.gpe_params:
        if self.storage not in STORAGE:
            raise ValueError("storage must be 'c128', 'c64' or 'mixed'")
        return _lib.gpe_params(
            self.nx, self.ny, self.nz, self.order, self.bcs, self.eqn_to_solve, int(bool(self.real_time)),
            STORAGE[self.storage],
            self.xr, self.yr, self.zr, self.Urhs, self.diss_amp, self.g, self.mu, self.nn,
            self.omx, self.omy, self.omz, self.vscale, self.eps, self.safety, self.dt_decrease, self.dt_increase)

    def replace(self, **kw) -> "RunParams":
        return dataclasses.replace(self, **kw)


def xr_for_spacing(n: int, dx: float) -> float:
    """Half width giving spacing dx on n points: dx = 2 xr/(n-1) (src/ic.f90:43-46).
    The periodic period is then n*dx."""
    return dx * (n - 1) / 2.0


_FIELDS = {f.name for f in dataclasses.fields(RunParams)}


def read_namelist(text: str, base: RunParams | None = None) -> RunParams:
    """Parse the reference's run.in namelists (&run_params, &eqn_params, ...)
    into RunParams.  Unknown keys (I/O and misc parameters) are ignored."""
    p = dataclasses.replace(base) if base else RunParams()
    body = re.sub(r"!.*", "", text)
    for grp in re.findall(r"&\w+(.*?)/", body, flags=re.S):
        for key, val in re.findall(r"(\w+)\s*=\s*('[^']*'|\"[^\"]*\"|[^,\n]+)", grp):
            k = key.strip()
            v = val.strip().rstrip(",").strip()
            if k not in _FIELDS:
                continue
            cur = getattr(p, k)
            if isinstance(cur, bool):
                setattr(p, k, v.lower().strip(".") in ("true", "t"))
            elif isinstance(cur, int):
                setattr(p, k, int(v))
            elif isinstance(cur, float):
                setattr(p, k, float(v.lower().replace("_pr", "").replace("d", "e")))
            else:
                setattr(p, k, v.strip("'\"").strip())
    return p
