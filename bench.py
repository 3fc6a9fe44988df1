#!/usr/bin/env python3
"""Benchmark of the GPE time-stepping hot path on one MI355X.

Workload (the headline metric's configuration): 512^3 periodic grid,
dx = 0.3, 4th-order centred FD, eqn 1 (i/2(lap + (1-|psi|^2)psi)), classic
RK4 with dt = 0.005, random-phase turbulence IC (Gaussian shell spectrum,
k0 = 1, sigma = 0.2, seed 0, mean |psi|^2 = 1), complex128 fields.  A step
is one RK4 step of the whole grid (two fused kernels of two stages each;
--unfused: the four stage kernels); inputs are resident in HBM when the
timed region starts.

Prints one JSON line: value = Gcell-steps/s, plus the roofline of the
dominant kernel (the one with the largest share of the step's time), timed
live with HIP events around each launch on the context's stream, the
four-stage-kernel run (13 field accesses per cell, the bandwidth-bound
formulation), the complex64- and mixed-storage runs, the time per Euler and
RK2 step of the fused and stage kernels (euler_rk2), the time per RKF45
(Cash-Karp) attempt of the stage and fused kernels in complex128 and mixed
storage on the same grid (rkf45_attempt), the
time to T = 2 of plain complex128, mixed-storage, learned and complex64
RKF45 with each tolerance tuned to the same 1e-6 error (time_to_t), the
48-member ensemble in every mode (ensemble), and a CPU
baseline (the FP64 oracle, OpenMP, on a bounded sample of the same
workload: --cpu-steps RK4 steps of the same grid and IC), and the
reference's own Fortran build timed on the host cores (cpu_reference:
--ref-steps RK4 steps of the same grid over 16 MPI ranks with the
reference's y-z decomposition, where oracle/_ref holds that build).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E, published
ROOT = os.path.dirname(os.path.abspath(__file__))

from gpe_amd.hiprt import Hip, copy_probe  # noqa: E402


def timed_steps(ctx, hip, scheme, steps, info):
    """K steps between two events on the context's stream, bracketed by device syncs."""
    hip.device_sync()
    e0, e1 = hip.event(), hip.event()
    hip.record(e0, ctx.stream)
    ctx.step(scheme, steps, info)
    hip.record(e1, ctx.stream)
    ms = hip.elapsed_ms(e0, e1)
    hip.device_sync()
    return ms


def kernel_profile(ctx, scheme, steps, info):
    """Per-kind launch count, average duration (ms) and algorithmic bytes per launch."""
    return ctx.kernel_profile(lambda: ctx.step(scheme, steps, info))


def load_traffic(name):
    """HBM bytes per launch measured with rocprofv3 --pmc (FETCH_SIZE, WRITE_SIZE),
    committed under profiles/ as JSON; None if not measured."""
    path = os.path.join(ROOT, "profiles", name)
    if not os.path.exists(path):
        return None
    with open(path) as f:
        return json.load(f)["traffic_bytes_per_launch"]


def run_storage(p, psi, args, hip, fused=True):
    from gpe_amd import Context, StepInfo

    with Context(p) as ctx:
        ctx.set_fused(fused)
        if args.tiles:
            ctx.set_tiles(*[int(v) for v in args.tiles.split(",")])
        ctx.upload(psi)
        info = StepInfo(p.dt0())
        ctx.step("rk4", args.warmup, info)
        ms = timed_steps(ctx, hip, "rk4", args.steps, info)
        prof = kernel_profile(ctx, "rk4", args.steps, info)
        tiles = ctx.tiles()
    return ms, prof, tiles


def rkf45_attempts(p, psi, count):
    """RKF45 (rkqs, tolerance 1e-6, dt0 = 0.01) on the same 512^3 workload for
    the Cash-Karp variants: ms per attempt from the per-launch event timing of
    `count` accepted steps after two untimed ones, and per-kernel figures."""
    from gpe_amd import Context, StepInfo
    from gpe_amd.studies import variant

    out = {}
    for spec in ("c128/stages", "c128", "mixed/stages", "mixed"):
        storage, fused = variant(spec)
        with Context(p.replace(storage=storage, scheme="rk45", eps=1e-6)) as ctx:
            ctx.set_fused(fused)
            ctx.upload(psi)
            info = StepInfo(0.01 + 0j)
            ctx.step("rk45", 2, info)
            a0 = info.n_accept + info.n_reject
            prof = ctx.kernel_profile(lambda: ctx.step("rk45", count, info))
            na = info.n_accept + info.n_reject - a0
        ms = sum(v["avg_ms"] * v["launches"] for v in prof.values()) / max(na, 1)
        attempt_bytes = sum(v["bytes"] * v["launches"] for v in prof.values()) / max(na, 1)
        out[spec] = {"ms_per_attempt": round(ms, 4), "attempts": na,
                     "hbm_frac": round(attempt_bytes / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
                     "kernels": kernel_table(prof)}
        print(f"[bench] rkf45 {spec}: {ms:.3f} ms/attempt", file=sys.stderr, flush=True)
    return out


def euler_rk2_steps(p, psi, steps):
    """Euler and midpoint RK2 (src/solve.f90:57-105) on the same 512^3
    workload: ms per step of the fused kernels (RK2: one kernel per step;
    Euler: one kernel per two steps) and of the stage kernels, from the
    per-launch event timing of `steps` steps after two untimed ones."""
    from gpe_amd import Context, StepInfo

    out = {}
    for scheme in ("euler", "rk2"):
        out[scheme] = {}
        for fused in (True, False):
            with Context(p) as ctx:
                ctx.set_fused(fused)
                ctx.upload(psi)
                info = StepInfo(p.dt0())
                ctx.step(scheme, 2, info)
                prof = ctx.kernel_profile(lambda: ctx.step(scheme, steps, info))
            ms = sum(v["avg_ms"] * v["launches"] for v in prof.values()) / steps
            step_bytes = sum(v["bytes"] * v["launches"] for v in prof.values()) / steps
            out[scheme]["fused" if fused else "stages"] = {
                "ms_per_step": round(ms, 4), "gcell_steps_s": round(p.cells / (ms * 1e-3) / 1e9, 3),
                "hbm_frac": round(step_bytes / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4), "kernels": kernel_table(prof)}
            print(f"[bench] {scheme} {'fused' if fused else 'stage kernels'}: {ms:.3f} ms/step", file=sys.stderr,
                  flush=True)
    return out


def ensemble_leg(steps):
    """The 48-member ensemble workload (members from {64..192}^3, seeds
    0-47, 2nd order, RK4 dt 0.005): ms per step and total cell-steps/s of
    every mode over `steps` steps, and the batching gain (one persistent
    launch for all members against one launch per member, same kernels)
    apart from the fusion gain."""
    from gpe_amd.ensemble import MODES, Ensemble, ensemble_config, gains, time_modes

    members, fields = ensemble_config(48)
    with Ensemble(members) as ens:
        rows = time_modes(ens, fields, list(MODES), steps)
        cells = ens.cells
    out = {"workload": "48 random-phase members, seeds 0-47, n from {64,96,128,160,192}^3, dx 0.3, 2nd-order FD, "
                       "RK4 dt 0.005, complex128", "cells": cells, "steps": steps,
           "modes": {r["ensemble"]: {"ms_per_step": round(r["ms_per_step"], 4),
                                     "gcell_steps_s": round(r["gcell_steps_s"], 3),
                                     "field_accesses": r["accesses"], "hbm_frac": round(r["hbm_frac"], 4)}
                     for r in rows},
           "gains": {k: round(v, 4) for k, v in gains(rows).items()}}
    print(f"[bench] ensemble: " + ", ".join(f"{k} {v['gcell_steps_s']:.1f}" for k, v in out["modes"].items()),
          file=sys.stderr, flush=True)
    return out


def time_to_t_leg(p, psi, T):
    """RKF45 from the same 512^3 IC to T (last step clamped onto T): each
    method's tolerance tuned so its relative L2 error against an FP64 run at
    tolerance 1e-10 lands at or just under 1e-6; wall time (attempts,
    controller, syncs) and the speed-up over the fastest plain complex128
    RKF45 (the reference's controller, fused Cash-Karp attempt).  Also
    classic RK4 (fused, complex128) at a fixed step T/N tuned the same way."""
    from gpe_amd.learned import DEFAULT_MODEL, LearnedController, load_model
    from gpe_amd.studies import time_to_t_tuned

    pr = p.replace(scheme="rk45", tau=0.01)
    model = load_model(DEFAULT_MODEL)
    methods = [("plain-c128", "c128", None), ("plain-mixed", "mixed", None),
               ("learned-mixed", "mixed", lambda eps: LearnedController(model, pr.replace(eps=eps))),
               ("plain-c64", "c64", None), ("rk4-c128", "rk4:c128", None)]
    t0 = time.perf_counter()
    runs = time_to_t_tuned(pr, psi, T, 0.01, methods)
    out = {"workload": f"random-phase turbulence {p.nx}^3, dx 0.3, 4th-order FD, eqn 1, RKF45 to T={T:g}",
           "error_target": 1e-6, "baseline": "plain-c128", "runs": runs,
           "leg_wall_s": round(time.perf_counter() - t0, 2)}

    def knob(v):
        return f"dt {v['dt']:.3g}" if "dt" in v else f"eps {v['eps']:.2g}"

    print("[bench] time to T: " + ", ".join(f"{k} {v['wall_s']:.2f} s ({knob(v)}, err "
                                           f"{v.get('rel_l2_vs_oracle', 0):.2e})" for k, v in runs.items()),
          file=sys.stderr, flush=True)
    return out


def dominant(prof):
    """The kernel kind with the largest share of the step's time."""
    return max(prof, key=lambda k: prof[k]["avg_ms"] * prof[k]["launches"])


def roofline(prof, kind, traffic=None):
    k = prof[kind]
    achieved = k["bytes"] / (k["avg_ms"] * 1e-3) / 1e9
    return {"bound": "hbm", "kernel": kind, "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic, "avg_launch_ms": round(k["avg_ms"], 5),
            "algorithmic_bytes_per_launch": k["bytes"]}


def kernel_table(prof):
    return {k: {"avg_ms": round(v["avg_ms"], 5), "launches": v["launches"],
                "GB_s": round(v["bytes"] / (v["avg_ms"] * 1e-3) / 1e9, 1)} for k, v in prof.items()}


def cpu_baseline(p, psi, steps):
    """The FP64 oracle (plain C restatement of the reference, OpenMP) on
    `steps` RK4 steps of the same workload."""
    from oracle import oracle as orc

    orc.build()
    t0 = time.perf_counter()
    orc.rk4_n(p, psi, p.dt0(), steps)
    s = time.perf_counter() - t0
    return {"value": round(p.cells * steps / s / 1e9, 6), "unit": "Gcell-steps/s", "cores": orc.threads(),
            "kind": "port", "sample": f"{steps} RK4 step(s) of the same {p.nx}^3 workload ({s:.1f} s)"}


def cpu_reference(steps, n=512, ranks=(4, 4)):
    """The reference's own Fortran time-stepping path (oracle/_ref, compiled
    from its unmodified sources), timed on the host cores:
      * where oracle/_ref/bench_<n>_<py>x<pz>/ref_bench is built, `steps` RK4
        steps of the n^3 workload (periodic, dx 0.3, 4th order, eqn 1,
        dt 0.005) over py * pz MPI ranks with the reference's own y-z
        decomposition and ghost exchange, from an analytic initial field
        (oracle/ref/ref_bench.f90; the time per step does not depend on it);
      * else 40 RK4 steps of the 64^3 vortex-ring workload on one rank
        (oracle/_ref/64x64x64/ref_driver).
    None where neither build is present."""
    import shutil
    import subprocess
    import tempfile

    from gpe_amd import RunParams, ics
    from gpe_amd.io import run_in_text, write_end_state

    here = os.path.dirname(os.path.abspath(__file__))
    np_ = ranks[0] * ranks[1]
    big = os.path.join(here, "oracle", "_ref", f"bench_{n}_{ranks[0]}x{ranks[1]}", "ref_bench")
    mpiexec = shutil.which("mpiexec", path="/opt/conda/bin") or shutil.which("mpiexec")
    run = tempfile.mkdtemp(prefix="gpe_ref_")
    try:
        if os.access(big, os.X_OK) and mpiexec:
            xr = 0.3 * (n - 1) / 2
            p = RunParams(nx=n, ny=n, nz=n, xr=xr, yr=xr, zr=xr, order=4, bcs=1, eqn_to_solve=1, scheme="rk4",
                          tau=0.005)
            with open(os.path.join(run, "run.in"), "w") as f:
                f.write(run_in_text(p, 1.0))
            r = subprocess.run(["bash", "-c", f"ulimit -s unlimited && exec {mpiexec} -n {np_} {big} {steps}"],
                               cwd=run, capture_output=True, text=True, env=dict(os.environ, OMP_NUM_THREADS="1"))
            if r.returncode == 0:
                f = r.stdout.split()
                cells, nst, s = int(f[1]), int(f[2]), float(f[3])
                return {"value": round(cells * nst / s / 1e9, 6), "unit": "Gcell-steps/s", "cores": np_,
                        "kind": "reference",
                        "sample": f"{nst} RK4 steps of the {n}^3 workload (analytic initial field), reference "
                                  f"build, {np_} MPI ranks ({ranks[0]}x{ranks[1]} y-z decomposition), {s:.1f} s"}
        exe = os.path.join(here, "oracle", "_ref", "64x64x64", "ref_driver")
        if not os.access(exe, os.X_OK):
            return None
        steps = 40
        p, psi = ics.config_ring64(seed=0)
        with open(os.path.join(run, "run.in"), "w") as f:
            f.write(run_in_text(p, (steps - 0.5) * p.tau))
        write_end_state(os.path.join(run, "in.dat"), psi, 0, 0.0, 0.0, p.dt0())
        t0 = time.perf_counter()
        r = subprocess.run(["bash", "-c", f"ulimit -s unlimited && exec {exe} in.dat out.dat trace.dat"], cwd=run,
                           capture_output=True, text=True, env=dict(os.environ, OMP_NUM_THREADS="1"))
        s = time.perf_counter() - t0
    finally:
        shutil.rmtree(run, ignore_errors=True)
    if r.returncode != 0:
        return None
    return {"value": round(p.cells * steps / s / 1e9, 6), "unit": "Gcell-steps/s", "cores": 1, "kind": "reference",
            "sample": f"{steps} RK4 steps of the 64^3 vortex-ring workload, reference build, 1 MPI rank ({s:.1f} s)"}


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--n", type=int, default=512, help="grid points per axis (512 = the benchmark workload)")
    ap.add_argument("--tiles", default="", help="tx,ty,tz override of the stencil tiling")
    ap.add_argument("--cpu-steps", type=int, default=4, help="RK4 steps of the CPU baseline sample (0: skip)")
    ap.add_argument("--ref-steps", type=int, default=2,
                    help="RK4 steps of the reference build's 512^3 multi-rank sample (0: skip; the 64^3 "
                         "one-rank fallback runs 40)")
    ap.add_argument("--no-c64", action="store_true", help="skip the complex64-storage run")
    ap.add_argument("--no-mixed", action="store_true", help="skip the mixed-storage run")
    ap.add_argument("--euler-rk2-steps", type=int, default=10,
                    help="timed Euler and RK2 steps, fused and stage kernels (0: skip)")
    ap.add_argument("--rk45-steps", type=int, default=3,
                    help="timed RKF45 steps per Cash-Karp variant (c128 and mixed, stage and fused; 0: skip)")
    ap.add_argument("--ttt-T", type=float, default=2.0,
                    help="final time of the matched-error time-to-T leg (0: skip)")
    ap.add_argument("--ens-steps", type=int, default=5, help="timed steps per mode of the 48-member ensemble leg "
                                                            "(0: skip)")
    ap.add_argument("--unfused", action="store_true", help="four stage kernels per RK4 step instead of two fused")
    ap.add_argument("--no-unfused", action="store_true", help="skip the four-stage-kernel comparison run")
    ap.add_argument("--copy-probe-mib", type=int, default=0,
                    help="launch one plain copy kernel of this size first (counter calibration)")
    args = ap.parse_args(argv)

    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {args.gpus} is not supported "
              "(the path would shard as z-slabs with 2-plane halos, not built)", file=sys.stderr)
        return 2

    from gpe_amd import ics

    t0 = time.perf_counter()
    p, psi = ics.config_turb(n=args.n, storage="c128", seed=0, scheme="rk4", tau=0.005)
    print(f"[bench] IC {args.n}^3 built in {time.perf_counter() - t0:.1f} s", file=sys.stderr, flush=True)
    hip = Hip()
    if args.copy_probe_mib:
        copy_probe(hip, args.copy_probe_mib << 20)

    ms, prof, tiles = run_storage(p, psi, args, hip, fused=not args.unfused)
    print(f"[bench] c128: {ms / args.steps:.3f} ms/step", file=sys.stderr, flush=True)
    value = p.cells * args.steps / (ms * 1e-3) / 1e9
    step_bytes = sum(v["bytes"] for k, v in prof.items() if k.startswith("rk4"))
    dom = dominant(prof)
    rf = roofline(prof, dom, load_traffic(f"traffic_{dom}_c128_{args.n}.json"))
    result = {
        "metric": "RK4 Gcell-steps/s + %HBM peak at 512³; time-to-T, learned vs plain RKF45",
        "value": round(value, 4),
        "unit": "Gcell-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms / args.steps, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"random-phase turbulence {args.n}^3 periodic, dx=0.3, 4th-order FD, eqn 1, "
                               "RK4 dt=0.005, complex128 fields",
                   "grid": [args.n] * 3, "order": 4, "bcs": "periodic", "scheme": "rk4", "dt": 0.005,
                   "storage": "c128", "kernels_per_step": "4 stage kernels" if args.unfused else "2 fused (stages 1+2, 3+4)",
                   "tiles": list(tiles)},
        "roofline": rf,
        "step_hbm_frac": round(step_bytes * args.steps / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
        # the bytes a stage-by-stage RK4 must stream (13 field accesses per
        # cell) over this step time, as a fraction of peak: above 1 means no
        # stage-by-stage implementation could match the step at any bandwidth
        "step_equivalent_hbm_frac": round(13 * 16 * p.cells * args.steps / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
        "kernels": kernel_table(prof),
    }
    if not args.unfused and not args.no_unfused:
        msu, profu, _ = run_storage(p, psi, args, hip, fused=False)
        print(f"[bench] c128 four stage kernels: {msu / args.steps:.3f} ms/step", file=sys.stderr, flush=True)
        bytes_u = sum(v["bytes"] for k, v in profu.items() if k.startswith("rk4"))
        domu = dominant(profu)
        result["unfused"] = {"value": round(p.cells * args.steps / (msu * 1e-3) / 1e9, 4),
                             "ms_per_step": round(msu / args.steps, 4),
                             "kernels_per_step": "4 stage kernels (13 field accesses per cell)",
                             "roofline": roofline(profu, domu, load_traffic(f"traffic_{domu}_c128_{args.n}.json")),
                             "step_hbm_frac": round(bytes_u * args.steps / (msu * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
                             "kernels": kernel_table(profu)}
    if not args.no_c64:
        p64 = p.replace(storage="c64")
        ms64, prof64, _ = run_storage(p64, psi, args, hip)
        print(f"[bench] c64: {ms64 / args.steps:.3f} ms/step", file=sys.stderr, flush=True)
        result["c64"] = {"value": round(p.cells * args.steps / (ms64 * 1e-3) / 1e9, 4),
                         "ms_per_step": round(ms64 / args.steps, 4), "dtype": "f32 stencil, f64 reductions",
                         "roofline": roofline(prof64, dominant(prof64),
                                              load_traffic(f"traffic_{dominant(prof64)}_c64_{args.n}.json")),
                         "kernels": kernel_table(prof64)}
    if not args.no_mixed:
        pm = p.replace(storage="mixed")
        msm, profm, _ = run_storage(pm, psi, args, hip)
        print(f"[bench] mixed: {msm / args.steps:.3f} ms/step", file=sys.stderr, flush=True)
        result["mixed"] = {"value": round(p.cells * args.steps / (msm * 1e-3) / 1e9, 4),
                           "ms_per_step": round(msm / args.steps, 4),
                           "dtype": "f64 state and stencil, complex64 RK increments",
                           "roofline": roofline(profm, dominant(profm),
                                                load_traffic(f"traffic_{dominant(profm)}_mixed_{args.n}.json")),
                           "kernels": kernel_table(profm)}
    if args.euler_rk2_steps > 0:
        result["euler_rk2"] = euler_rk2_steps(p, psi, args.euler_rk2_steps)
    if args.rk45_steps > 0:
        result["rkf45_attempt"] = rkf45_attempts(p, psi, args.rk45_steps)
    if args.ttt_T > 0:
        result["time_to_t"] = time_to_t_leg(p, psi, args.ttt_T)
    if args.ens_steps > 0:
        result["ensemble"] = ensemble_leg(args.ens_steps)
    if args.cpu_steps > 0:
        result["cpu_baseline"] = cpu_baseline(p, psi, args.cpu_steps)
    else:
        result["cpu_baseline"] = None
    result["cpu_reference"] = cpu_reference(args.ref_steps, n=args.n) if args.ref_steps > 0 else None
    print(json.dumps(result), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
