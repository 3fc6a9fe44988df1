"""Batched ensemble of independent members (libgpehip.so gpe_ens_*), the
48-member workload and a tile autotuner.

Workload: 48 independent random-phase runs (IC as the 512^3 turbulence
workload: Gaussian shell k0 = 1, sigma = 0.2, mean |psi|^2 = 1), seeds 0-47,
cubic grids drawn uniformly from {64, 96, 128, 160, 192}^3 with
numpy.random.default_rng(2024), periodic, dx = 0.3, 2nd-order FD, eqn 1,
RK4 with dt = 0.005 for 1000 steps.

    python -m gpe_amd.ensemble [--steps 1000] [--members 48] [--no-tune]

prints one JSON line per mode (one launch per member, batched persistent
launches, hipGraph replay, cache-resident groups, and for complex128 the
fused RK4 once per member and of all members in two persistent launches per
step, eager and in a hipGraph) with total cell-steps/s, the batching and
fusion gains apart, and the autotuner log.
"""
from __future__ import annotations

import argparse
import ctypes as ct
import itertools
import json
import sys
import time

import numpy as np

from . import _lib, ics
from .params import RunParams, xr_for_spacing

MODES = {"per_member": 0, "batched": 1, "graph": 2, "grouped": 3, "fused": 4, "fused_graph": 5,
         "per_member_fused": 6}
SIZES = (64, 96, 128, 160, 192)


class Ensemble:
    def __init__(self, members: list[RunParams]):
        self.lib = _lib.load()
        self.members = members
        arr = (_lib.gpe_params * len(members))(*[m.to_c() for m in members])
        h = ct.c_void_p()
        rc = self.lib.gpe_ens_create(arr, len(members), ct.byref(h))
        if rc != 0:
            raise _lib.GpeError(f"gpe_ens_create failed ({rc}): {self.lib.gpe_ens_last_error(None).decode()}")
        self.h = h
        self.dtype = np.complex128 if members[0].storage == "c128" else np.complex64
        self.hdt = _lib.GPE_HOST_C128 if members[0].storage == "c128" else _lib.GPE_HOST_C64

    def _chk(self, rc, what):
        if rc != 0:
            raise _lib.GpeError(f"{what} failed ({rc}): {self.lib.gpe_ens_last_error(self.h).decode()}")

    def close(self):
        if getattr(self, "h", None) is not None and self.h.value:
            self.lib.gpe_ens_destroy(self.h)
            self.h = ct.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def upload(self, m: int, psi: np.ndarray):
        a = np.ascontiguousarray(psi, dtype=self.dtype)
        if a.shape != self.members[m].shape:
            raise ValueError("member field shape mismatch")
        self._chk(self.lib.gpe_ens_upload(self.h, m, a.ctypes.data, self.hdt), "gpe_ens_upload")

    def download(self, m: int) -> np.ndarray:
        out = np.empty(self.members[m].shape, dtype=self.dtype)
        self._chk(self.lib.gpe_ens_download(self.h, m, out.ctypes.data, self.hdt), "gpe_ens_download")
        return out

    def set_tiles(self, tx: int, ty: int, tz=None):
        arr = None if tz is None else (ct.c_int * len(self.members))(*[int(v) for v in tz])
        self._chk(self.lib.gpe_ens_set_tiles(self.h, tx, ty, arr), "gpe_ens_set_tiles")

    def set_grouping(self, budget_bytes: int, steps_per_visit: int):
        self._chk(self.lib.gpe_ens_set_grouping(self.h, budget_bytes, steps_per_visit), "gpe_ens_set_grouping")

    def num_groups(self) -> int:
        n = self.lib.gpe_ens_num_groups(self.h)
        self._chk(0 if n >= 0 else n, "gpe_ens_num_groups")
        return n

    def step(self, mode: str, dt: complex, nsteps: int):
        self._chk(self.lib.gpe_ens_step(self.h, MODES[mode], dt.real, dt.imag, nsteps), "gpe_ens_step")

    def sync(self):
        self._chk(self.lib.gpe_ens_sync(self.h), "gpe_ens_sync")

    @property
    def stream(self) -> int:
        return self.lib.gpe_ens_stream(self.h)

    @property
    def cells(self) -> int:
        return sum(m.cells for m in self.members)


def ensemble_config(n_members: int = 48, draw_seed: int = 2024, storage: str = "c128", dx: float = 0.3):
    """Members (RunParams) and their ICs."""
    rng = np.random.default_rng(draw_seed)
    sizes = rng.choice(SIZES, size=n_members)
    members, fields = [], []
    for seed, n in enumerate(sizes):
        n = int(n)
        xr = xr_for_spacing(n, dx)
        p = RunParams(nx=n, ny=n, nz=n, xr=xr, yr=xr, zr=xr, tau=0.005, scheme="rk4", eqn_to_solve=1, bcs=1,
                      order=2, storage=storage)
        members.append(p)
        fields.append(ics.random_phase(p, seed=seed))
    return members, fields


class Timer:
    def __init__(self):
        import os
        self.h = ct.CDLL(os.path.join("/opt/rocm", "lib", "libamdhip64.so"))

    def time(self, ens: Ensemble, fn) -> float:
        e0, e1 = ct.c_void_p(), ct.c_void_p()
        self.h.hipEventCreate(ct.byref(e0))
        self.h.hipEventCreate(ct.byref(e1))
        self.h.hipDeviceSynchronize()
        self.h.hipEventRecord(e0, ct.c_void_p(ens.stream))
        fn()
        self.h.hipEventRecord(e1, ct.c_void_p(ens.stream))
        self.h.hipEventSynchronize(e1)
        ms = ct.c_float()
        self.h.hipEventElapsedTime(ct.byref(ms), e0, e1)
        self.h.hipEventDestroy(e0)
        self.h.hipEventDestroy(e1)
        return ms.value


def autotune(ens: Ensemble, dt: complex, steps: int = 10, shapes=((32, 8), (64, 4), (64, 8), (32, 4)),
             tzs=(4, 8, 16, 32), size_tzs=(4, 8, 16, 32, 64, 96, 192), log=None):
    """Tile search for the batched launch: global (tx, ty, tz), then tz per
    grid size.  The state is advanced while tuning, so callers re-upload.
    Returns (tx, ty, tz_per_member, log)."""
    timer = Timer()
    log = [] if log is None else log

    def measure(tx, ty, tz):
        ens.set_tiles(tx, ty, tz)
        ens.step("graph", dt, steps)  # capture + warm
        ms = timer.time(ens, lambda: ens.step("graph", dt, steps))
        return ms / steps

    best = None
    for (tx, ty), tz in itertools.product(shapes, tzs):
        tzm = [min(tz, m.nz) for m in ens.members]
        ms = measure(tx, ty, tzm)
        log.append({"tx": tx, "ty": ty, "tz": tz, "ms_per_step": ms})
        if best is None or ms < best[0]:
            best = (ms, tx, ty, tzm)
    ms0, tx, ty, tzm = best
    # per-size refinement of tz
    for n in sorted({m.nz for m in ens.members}):
        for tz in [t for t in size_tzs if t <= n]:
            trial = [min(tz, m.nz) if m.nz == n else t for m, t in zip(ens.members, tzm)]
            ms = measure(tx, ty, trial)
            log.append({"tx": tx, "ty": ty, "size": n, "tz": tz, "ms_per_step": ms})
            if ms < ms0:
                ms0, tzm = ms, trial
    ens.set_tiles(tx, ty, tzm)
    return tx, ty, tzm, log


def time_modes(ens: Ensemble, fields, modes, steps: int, dt: complex = 0.005 + 0j, warm: int = 2):
    """Per mode: ms per step over `steps` steps (HIP events on the ensemble's
    stream) after `warm` untimed ones, from the same initial fields."""
    timer = Timer()
    cells = ens.cells
    esz = 16 if ens.dtype == np.complex128 else 8
    rows = []
    for mode in modes:
        for m, f in enumerate(fields):
            ens.upload(m, f)
        ens.step(mode, dt, warm)  # warm-up (and graph capture)
        for m, f in enumerate(fields):
            ens.upload(m, f)
        t0 = time.perf_counter()
        ms = timer.time(ens, lambda: ens.step(mode, dt, steps))
        wall = time.perf_counter() - t0
        # field accesses per cell-step: 13 for the stage kernels, 5 fused
        acc = 5 if "fused" in mode else 13
        rows.append({"ensemble": mode, "steps": steps, "ms_per_step": ms / steps,
                     "gcell_steps_s": cells * steps / (ms * 1e-3) / 1e9, "wall_s": wall, "accesses": acc,
                     "hbm_frac": cells * acc * esz * steps / (ms * 1e-3) / 8e12,
                     "equivalent_13_access_frac": cells * 13 * esz * steps / (ms * 1e-3) / 8e12})
    return rows


def gains(rows):
    """Batching gain (one persistent launch per stage or pair for all members
    against one launch per member, same kernels) apart from the fusion gain
    (fused against stage kernels, one launch per member each)."""
    r = {x["ensemble"]: x["gcell_steps_s"] for x in rows}
    out = {}
    if "per_member" in r:
        out.update({f"{k}_vs_per_member": r[k] / r["per_member"] for k in ("batched", "graph", "grouped") if k in r})
    if "per_member_fused" in r:
        out.update({f"{k}_vs_per_member_fused": r[k] / r["per_member_fused"]
                    for k in ("fused", "fused_graph") if k in r})
        if "per_member" in r:
            out["fusion_per_member_fused_vs_per_member"] = r["per_member_fused"] / r["per_member"]
    return out


def run_study(n_members: int, steps: int, tune: bool, storage: str = "c128"):
    members, fields = ensemble_config(n_members, storage=storage)
    dt = 0.005 + 0j
    with Ensemble(members) as ens:
        cells = ens.cells
        print(json.dumps({"ensemble": "config", "members": n_members, "cells": cells,
                          "sizes": [m.nx for m in members]}), flush=True)
        if tune:
            tx, ty, tzm, log = autotune(ens, dt)
            print(json.dumps({"ensemble": "autotune", "tx": tx, "ty": ty,
                              "tz_by_size": {str(m.nx): t for m, t in zip(members, tzm)}, "log": log}), flush=True)
        print(json.dumps({"ensemble": "groups", "count": ens.num_groups()}), flush=True)
        modes = ["per_member", "batched", "graph", "grouped"]
        if storage == "c128":
            modes += ["per_member_fused", "fused", "fused_graph"]
        rows = time_modes(ens, fields, modes, steps, dt)
        for row in rows:
            print(json.dumps(row), flush=True)
        base = rows[0]["gcell_steps_s"]
        print(json.dumps({"ensemble": "speedup_vs_per_member",
                          **{r["ensemble"]: r["gcell_steps_s"] / base for r in rows}}), flush=True)
        print(json.dumps({"ensemble": "gains", **gains(rows)}), flush=True)
    return rows


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--members", type=int, default=48)
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--no-tune", action="store_true")
    ap.add_argument("--storage", default="c128")
    a = ap.parse_args(argv)
    run_study(a.members, a.steps, not a.no_tune, a.storage)
    return 0


if __name__ == "__main__":
    sys.exit(main())
