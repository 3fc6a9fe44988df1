"""Step-size control for the adaptive Cash-Karp path, on top of the C ABI's
gpe_rk45_attempt / gpe_rk45_accept.

`PlainController` is the reference's controller (rkqs, src/solve.f90:173-229,
Numerical Recipes' elementary controller): on reject
dt <- sign(max(|0.9 dt errmax^-0.25|, 0.1 |dt|)); on accept
dt_next = 0.9 dt errmax^-0.2 if errmax > errcon else 5 dt.  The library's
gpe_step(GPE_RK45) runs the same rules in C++; this module adds what a
time-to-T study needs: a clamp of the last step onto T, per-step traces and
a pluggable proposer (the learned controller in `gpe_amd.learned`).
"""
from __future__ import annotations

import math
import time
from dataclasses import dataclass, field

import numpy as np


class PlainController:
    name = "plain"

    def __init__(self, safety=0.9, dt_decrease=-0.25, dt_increase=-0.2):
        self.safety = safety
        self.dt_decrease = dt_decrease
        self.dt_increase = dt_increase
        self.errcon = (5.0 / safety) ** (1.0 / dt_increase)

    def reset(self):
        pass

    def shrink(self, dt: complex, errmax: float) -> complex:
        f = errmax ** self.dt_decrease
        mag = max(abs(complex(self.safety * dt.real * f, self.safety * dt.imag * f)), 0.1 * abs(dt))
        return complex(mag, 0.0) if dt.imag == 0.0 else complex(0.0, -mag)

    def propose(self, dt: complex, errmax: float, state: "StepState") -> complex:
        if errmax > self.errcon:
            f = errmax ** self.dt_increase
            return complex(self.safety * dt.real * f, self.safety * dt.imag * f)
        return 5.0 * dt


@dataclass
class StepState:
    """History the controllers may use (most recent last)."""
    errmax: list = field(default_factory=list)
    dt: list = field(default_factory=list)
    # the first attempt of each accepted step (dt tried, its errmax)
    dt_first: list = field(default_factory=list)
    errmax_first: list = field(default_factory=list)
    psi2max: float = 0.0
    energy_drift: float = 0.0


@dataclass
class RunStats:
    method: str
    t: float = 0.0
    n_accept: int = 0
    n_reject: int = 0
    n_rhs: int = 0
    wall_s: float = 0.0
    # rows (t, dt, errmax, n_rej_this_step, psi2max, energy_drift, dt_first, errmax_first):
    # the accepted step and the first attempt of that step
    trace: list = field(default_factory=list)


def integrate(ctx, T: float, dt0: float, controller, clamp: bool = True, energy: bool = False,
              max_steps: int = 10 ** 7, energy_every: int = 32) -> RunStats:
    """Advance the device state of `ctx` from t = 0 to T with Cash-Karp
    attempts and `controller`.  With clamp the last step lands on T exactly;
    without it the loop is the reference's `do while (t <= T)` (overshoot).
    energy = True evaluates the energy every `energy_every` accepted steps
    (the drift input of the learned controller; it changes slowly, and each
    evaluation is a stencil pass over the field).  Wall time covers attempts, controller and
    syncs, not IC or I/O."""
    st = RunStats(method=getattr(controller, "name", "controller"))
    hist = StepState()
    controller.reset()
    E0 = ctx.energy() if energy else 0.0
    dt = complex(dt0, 0.0)
    t = 0.0
    ctx.sync()
    t_start = time.perf_counter()
    while (t < T * (1 - 1e-14)) if clamp else (t <= T):
        if st.n_accept >= max_steps:
            raise RuntimeError("integrate: step limit reached")
        if clamp and t + dt.real > T:
            dt = complex(T - t, 0.0)
        dt_entry = dt
        rej = 0
        first = None
        while True:
            errmax, a2 = ctx.rk45_attempt(dt, dt_entry)
            if first is None:
                first = (dt.real, errmax)
            st.n_rhs += 5 if rej else 6  # f(psi) is reused on retries
            if errmax <= 1.0:
                break
            rej += 1
            if not math.isfinite(errmax):
                raise RuntimeError("integrate: non-finite error estimate")
            dt = controller.shrink(dt, errmax)
        ctx.rk45_accept()
        t += dt.real
        st.n_accept += 1
        st.n_reject += rej
        hist.errmax.append(errmax)
        hist.dt.append(dt.real)
        hist.dt_first.append(first[0])
        hist.errmax_first.append(first[1])
        hist.psi2max = a2
        if energy and st.n_accept % energy_every == 0:
            E = ctx.energy()
            hist.energy_drift = (E - E0) / abs(E0) if E0 else 0.0
        st.trace.append((t, dt.real, errmax, rej, a2, hist.energy_drift, first[0], first[1]))
        dt = controller.propose(dt, errmax, hist)
    ctx.sync()
    st.wall_s = time.perf_counter() - t_start
    st.t = t
    return st


def trace_array(stats: RunStats) -> np.ndarray:
    return np.array(stats.trace, dtype=np.float64).reshape(-1, 8)
