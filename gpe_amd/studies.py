"""Accuracy and speed studies of the storage and step-control variants, on one GPU:

  drift      complex64 storage vs complex128 on the trapped-vortex workload
             (256^3 reflective, dx = 0.2, 4th order, eqn 4, RK4 dt = 0.002,
             2000 steps): drift of the norm and of the conserved energy
             (gpe_hamiltonian) for each, and the c64 field's distance from the
             FP64 (c128) run.
  time_to_t_tuned  RKF45 to a fixed final time with each method's
             tolerance tuned so its error lands at 1e-6 (the matched-error
             comparison), against plain complex128 with the fused attempt;
             also classic RK4 at a fixed step tuned the same way.
  time_to_t  RKF45 to a fixed final time on the random-phase workload:
             an FP64 reference at a tight tolerance, then plain c128 (six
             stage kernels per attempt, and fused: three kernels), c64 and
             mixed storage (c128 state, c64 increments) at 1e-6, and the
             learned controller on mixed storage when a model is given;
             wall time, steps, rejections, relative L2 error at T and the
             speed-up over plain c128.

  attempts   RKF45 attempts per storage with per-kernel times (the RKCK
             profiling target).
  diag       wall time per call of the state diagnostics (momentum, velocity
             PDF, VCF, condensed particles + spectra, line length) on the
             random-phase workload, with the algorithmic bytes each moves
             (the diagnostics profiling target).

Each prints one JSON object per line.  Usage:
    python -m gpe_amd.studies drift [--n 256] [--steps 2000]
    python -m gpe_amd.studies time_to_t [--n 512] [--T 20]
    python -m gpe_amd.studies time_to_t_tuned [--n 512] [--T 20]
    python -m gpe_amd.studies attempts [--n 512] [--attempts 6] [--storages c128/stages,c128,mixed]
    python -m gpe_amd.studies diag [--n 512] [--reps 5]
"""
from __future__ import annotations

import argparse
import json
import sys
import time

import numpy as np

from . import ics
from .context import Context, StepInfo
from .controller import PlainController, RunStats, integrate


def rel_l2(a, b) -> float:
    return float(np.linalg.norm((a - b).ravel()) / np.linalg.norm(b.ravel()))


def emit(d):
    print(json.dumps(d), flush=True)


def drift(n: int, steps: int, every: int):
    out = {}
    finals = {}
    for storage in ("c128", "c64"):
        p, psi = ics.config_trap256(storage=storage, n=n)
        with Context(p) as ctx:
            ctx.upload(psi)
            n0, e0 = ctx.norm(), ctx.hamiltonian()
            info = StepInfo(p.dt0())
            series = [(0, n0, e0)]
            t0 = time.perf_counter()
            done = 0
            while done < steps:
                k = min(every, steps - done)
                ctx.step("rk4", k, info)
                done += k
                series.append((done, ctx.norm(), ctx.hamiltonian()))
            ctx.sync()
            wall = time.perf_counter() - t0
            finals[storage] = ctx.download()
        nf, ef = series[-1][1], series[-1][2]
        out[storage] = dict(norm0=n0, energy0=e0, norm_drift=(nf - n0) / n0, energy_drift=(ef - e0) / abs(e0),
                            max_norm_drift=max(abs(s[1] - n0) / n0 for s in series),
                            max_energy_drift=max(abs(s[2] - e0) / abs(e0) for s in series), wall_s=wall,
                            gcell_steps_s=p.cells * steps / wall / 1e9)
        emit(dict(study="drift", storage=storage, n=n, steps=steps, **out[storage]))
    s64, s128 = out["c64"], out["c128"]
    emit(dict(study="drift_c64_vs_fp64", n=n, steps=steps,
              field_rel_l2=rel_l2(finals["c64"].astype(np.complex128), finals["c128"]),
              norm_rel_diff=abs(s64["norm0"] * (1 + s64["norm_drift"]) - s128["norm0"] * (1 + s128["norm_drift"]))
              / (s128["norm0"] * (1 + s128["norm_drift"])),
              energy_rel_diff=abs(s64["energy0"] * (1 + s64["energy_drift"]) - s128["energy0"] * (1 + s128["energy_drift"]))
              / abs(s128["energy0"] * (1 + s128["energy_drift"]))))


def variant(spec: str):
    """'c128', 'c64', 'mixed': that storage with the fused kernels where they
    apply (the library default); a '/stages' suffix forces the stage kernels."""
    storage, _, how = spec.partition("/")
    if how not in ("", "stages"):
        raise ValueError(f"unknown variant {spec!r}")
    return storage, how != "stages"


def run_rkf45(p, psi, T, dt0, eps, storage, controller=None, energy=False, name=None, fused=True):
    pp = p.replace(storage=storage, eps=eps)
    ctrl = controller or PlainController()
    with Context(pp) as ctx:
        ctx.set_fused(fused)
        ctx.upload(psi)
        st = integrate(ctx, T, dt0, ctrl, clamp=True, energy=energy)
        final = ctx.download()
    st.method = name or f"{ctrl.name}-{storage}{'' if fused else '/stages'}-eps{eps:g}"
    return st, final


def run_rk4_fixed(p, psi, T, dt, storage, fused=True, name=None):
    """Classic RK4 at a fixed step onto T: N = ceil(T/dt) steps of T/N, issued
    in one gpe_step call (no host round trip per step).  Wall time covers the
    steps and the final sync, not IC or I/O."""
    import math

    n = max(1, math.ceil(T / dt * (1 - 1e-12)))
    h = T / n
    pp = p.replace(storage=storage, scheme="rk4", tau=h)
    with Context(pp) as ctx:
        ctx.set_fused(fused)
        ctx.upload(psi)
        info = StepInfo(dt=complex(h, 0.0))
        ctx.sync()
        t0 = time.perf_counter()
        ctx.step("rk4", n, info)
        ctx.sync()
        wall = time.perf_counter() - t0
        final = ctx.download()
    st = RunStats(method=name or f"rk4-{storage}{'' if fused else '/stages'}-dt{h:.4g}", t=info.t,
                  n_accept=n, n_rhs=4 * n, wall_s=wall)
    return st, final, h


def time_to_t(n: int, T: float, dt0: float, oracle_eps: float, eps_list, learned_model=None,
              storages=("c128/stages", "c128", "c64", "mixed/stages", "mixed")):
    """The baseline of the speed-ups is plain c128 with the six stage kernels
    per attempt ('c128/stages') when it is among the variants."""
    p, psi = ics.config_turb(n=n, storage="c128", seed=0, scheme="rk45", tau=dt0)
    ref, ref_final = run_rkf45(p, psi, T, dt0, oracle_eps, "c128", name=f"oracle-c128-eps{oracle_eps:g}")
    emit(dict(study="time_to_t", n=n, T=T, method=ref.method, wall_s=ref.wall_s, steps=ref.n_accept,
              rejects=ref.n_reject, rhs=ref.n_rhs))
    rows = []
    runs = []
    for eps in eps_list:
        for storage in storages:
            runs.append((storage, eps, None, False))
        if learned_model is not None:
            from .learned import LearnedController
            runs.append(("mixed", eps, LearnedController(learned_model, p.replace(eps=eps)), True))
    for spec, eps, ctrl, energy in runs:
        storage, fused = variant(spec)
        st, fin = run_rkf45(p, psi, T, dt0, eps, storage, ctrl, energy, fused=fused)
        err = rel_l2(fin.astype(np.complex128), ref_final)
        row = dict(study="time_to_t", n=n, T=T, method=st.method, storage=spec, eps=eps, wall_s=st.wall_s,
                   steps=st.n_accept, rejects=st.n_reject, rhs=st.n_rhs, rel_l2_vs_oracle=err)
        rows.append(row)
        emit(row)
    for eps in eps_list:
        plain = [r for r in rows if r["eps"] == eps and r["method"].startswith("plain")]
        base = next((r for r in plain if r["storage"] == "c128/stages"), None) or \
            next((r for r in plain if r["storage"] == "c128"), None)
        if base is None:
            continue
        emit(dict(study="time_to_t_speedup", n=n, T=T, eps=eps, baseline=base["method"],
                  **{r["method"]: dict(speedup=base["wall_s"] / r["wall_s"], rel_l2_vs_oracle=r["rel_l2_vs_oracle"],
                                       within_1e6=r["rel_l2_vs_oracle"] <= 1e-6)
                     for r in rows if r["eps"] == eps}))
    return rows


def tune_to_error(run, target: float = 1e-6, eps0: float = 1e-6, slope: float = 0.8, max_runs: int = 5,
                  slope_max: float = 2.0, max_factor: float = 10.0):
    """Tolerance of a method tuned so that its error at T lands at or just
    under `target`: a secant search in log(eps) on log(error), starting from
    error ~ eps^slope.  run(eps) -> (stats, error).  Returns (eps, stats,
    error, tries) of the largest tolerance tried whose error is within the
    target (tries: every (eps, error, wall_s) evaluated).  The knob may be
    any quantity the error grows with (a fixed step: slope ~4 for RK4); a
    non-finite error (an unstable step) counts as an error of 1e3."""
    import math

    tries = []
    e_prev = None
    eps = eps0
    best = None
    for k in range(max_runs):
        st, err = run(eps)
        if not math.isfinite(err):
            err = 1e3
        tries.append((eps, err, st.wall_s))
        if err <= target and (best is None or eps > best[0]):
            best = (eps, st, err)
        if best is not None and target / 1.3 <= err <= target:
            break
        le, lr = math.log(eps), math.log(err)
        if e_prev is not None and abs(le - e_prev[0]) > 1e-9:
            s = (lr - e_prev[1]) / (le - e_prev[0])
            if best is None and s < 0.15 and le < e_prev[0]:
                break  # the error no longer follows the tolerance: a rounding floor above the target
            slope = min(max(s, 0.3), slope_max)
        e_prev = (le, lr)
        # aim a little under the target so the landing run stays within it,
        # moving eps by at most max_factor per run
        step = (math.log(0.85 * target) - lr) / slope
        lim = math.log(max_factor)
        eps = math.exp(le + min(max(step, -lim), lim))
    if best is None:
        return None, None, None, tries
    return best[0], best[1], best[2], tries


def time_to_t_tuned(p, psi, T: float, dt0: float, methods, target: float = 1e-6, oracle_eps: float = 1e-10,
                    baseline: str = "plain-c128"):
    """Wall time to T at a matched error: the FP64 reference at oracle_eps,
    then each method (name, variant spec, controller factory or None) with
    its tolerance tuned to land at `target` relative L2 against it.  A spec
    'rk4:<variant>' is classic RK4 at a fixed step instead of RKF45; its
    step (T/N, starting from dt0) is tuned the same way.  The speed-up is
    against `baseline` (the fastest plain complex128 RKF45: the reference's
    controller with the fused attempt)."""
    ref, ref_final = run_rkf45(p, psi, T, dt0, oracle_eps, "c128", name=f"oracle-c128-eps{oracle_eps:g}")
    print(f"[time_to_t] reference eps {oracle_eps:g}: {ref.n_accept} steps, {ref.wall_s:.2f} s", file=sys.stderr,
          flush=True)
    out = {"oracle": dict(method=ref.method, eps=oracle_eps, wall_s=ref.wall_s, steps=ref.n_accept,
                          rejects=ref.n_reject, rhs=ref.n_rhs)}
    for name, spec, make_ctrl in methods:
        if spec.startswith("rk4:"):
            storage, fused = variant(spec[4:])

            def run_dt(dt):
                st, fin, h = run_rk4_fixed(p, psi, T, dt, storage, fused=fused)
                err = rel_l2(fin.astype(np.complex128), ref_final)
                print(f"[time_to_t] {name} dt {h:.4g}: {st.n_accept} steps, {st.wall_s:.2f} s, rel L2 {err:.3g}",
                      file=sys.stderr, flush=True)
                st.dt = h
                return st, err

            # steps of T/N: error ~ dt^4; an unstable step blows up, so move by at most 2x per run
            dt, st, err, tries = tune_to_error(run_dt, target, eps0=dt0, slope=4.0, slope_max=6.0, max_factor=2.0)
            tr = [dict(dt=e, rel_l2=r, wall_s=w) for e, r, w in tries]
            if st is None:
                e, r, w = min(tries, key=lambda x: x[1])
                out[name] = dict(storage=spec, scheme="rk4", dt=e, wall_s=w, rel_l2_vs_oracle=r, within_target=False,
                                 tuning=tr)
            else:
                out[name] = dict(storage=spec, scheme="rk4", dt=st.dt, wall_s=st.wall_s, steps=st.n_accept,
                                 rejects=0, rhs=st.n_rhs, rel_l2_vs_oracle=err, within_target=True, tuning=tr)
            continue
        storage, fused = variant(spec)

        def run(eps):
            ctrl = make_ctrl(eps) if make_ctrl else None
            st, fin = run_rkf45(p, psi, T, dt0, eps, storage, ctrl, energy=ctrl is not None, fused=fused)
            err = rel_l2(fin.astype(np.complex128), ref_final)
            print(f"[time_to_t] {name} eps {eps:.3g}: {st.n_accept} steps, {st.n_reject} rejects, "
                  f"{st.wall_s:.2f} s, rel L2 {err:.3g}", file=sys.stderr, flush=True)
            return st, err

        eps, st, err, tries = tune_to_error(run, target)
        tr = [dict(eps=e, rel_l2=r, wall_s=w) for e, r, w in tries]
        if st is None:  # the error target is out of the method's reach (complex64 rounding)
            e, r, w = min(tries, key=lambda x: x[1])
            out[name] = dict(storage=spec, eps=e, wall_s=w, rel_l2_vs_oracle=r, within_target=False, tuning=tr)
        else:
            out[name] = dict(storage=spec, eps=eps, wall_s=st.wall_s, steps=st.n_accept, rejects=st.n_reject,
                             rhs=st.n_rhs, rel_l2_vs_oracle=err, within_target=True, tuning=tr)
    base = out.get(baseline)
    for name, r in out.items():
        if name != "oracle" and base:
            r["speedup"] = base["wall_s"] / r["wall_s"] if r["within_target"] else None
    return out


def attempts(n: int, count: int, storages, copy_mib: int = 0):
    """`count` RKF45 steps (rkqs: attempts until one is accepted) of the
    random-phase workload at tolerance 1e-6 for each storage, with the
    library's per-kernel timing: time per attempt and per RKCK stage kernel.
    Also the target of the RKCK rocprofv3 profiles (copy_mib launches the
    counter-calibration copy first)."""
    from .hiprt import Hip, copy_probe

    if copy_mib:
        copy_probe(Hip(), copy_mib << 20)
    p, psi = ics.config_turb(n=n, storage="c128", seed=0, scheme="rk45", tau=0.01)
    for spec in storages:
        storage, fused = variant(spec)
        with Context(p.replace(storage=storage, eps=1e-6)) as ctx:
            ctx.set_fused(fused)
            ctx.upload(psi)
            info = StepInfo(p.dt0())
            ctx.step("rk45", 2, info)
            a0 = info.n_accept + info.n_reject
            t0 = time.perf_counter()
            prof = ctx.kernel_profile(lambda: ctx.step("rk45", count, info))
            wall = time.perf_counter() - t0
            na = info.n_accept + info.n_reject - a0
            ker = {k: dict(avg_ms=v["avg_ms"], launches=v["launches"],
                           GB_s=v["bytes"] / (v["avg_ms"] * 1e-3) / 1e9) for k, v in prof.items()}
            emit(dict(study="attempts", n=n, storage=spec, steps=count, attempts=na,
                      ms_per_attempt=sum(v["avg_ms"] * v["launches"] for v in prof.values()) / max(na, 1),
                      wall_ms_per_attempt=1e3 * wall / max(na, 1), kernels=ker))


# Algorithmic HBM bytes per cell of each diagnostic (complex128 state):
#   momentum   one stencil pass over psi                                    16
#   vel_pdf    velocity pass with its statistics (psi 16 + v 24), histogram (24) 64
#   vcf        velocity pass (16 + 24), one read of v_x (8)                  48
#   condensed  copy for the FFT (16 + 16), spectrum read (16); the FFT's
#              own traffic (library) is not counted                         48
#   linelength one pass over psi (the corners of a cell are shared; the
#              leading-cell pass re-reads only the starts of some lines)    16
DIAG_BYTES = dict(momentum=16, vel_pdf=64, vcf=48, condensed=48, linelength=16)


def diag(n: int, reps: int):
    """Wall time per call of each state diagnostic on the random-phase
    workload (the state has vortex lines everywhere)."""
    p, psi = ics.config_turb(n=n, storage="c128", seed=0, scheme="rk4", tau=0.005)
    calls = dict(momentum=lambda c: c.momentum(), vel_pdf=lambda c: c.velocity_pdf(256), vcf=lambda c: c.vcf(),
                 condensed=lambda c: c.condensed_particles(), linelength=lambda c: c.linelength())
    cells = p.nx * p.ny * p.nz
    with Context(p) as ctx:
        ctx.upload(psi)
        for name, fn in calls.items():
            fn(ctx)
            t0 = time.perf_counter()
            for _ in range(reps):
                out = fn(ctx)
            ms = 1e3 * (time.perf_counter() - t0) / reps
            gbs = DIAG_BYTES[name] * cells / (ms * 1e-3) / 1e9
            extra = dict(value=out) if name == "linelength" else {}
            emit(dict(study="diag", n=n, diagnostic=name, ms=round(ms, 4), bytes_per_cell=DIAG_BYTES[name],
                      GB_s=round(gbs, 1), hbm_frac=round(gbs / 8000.0, 4), **extra))


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("study", choices=["drift", "time_to_t", "time_to_t_tuned", "attempts", "diag"])
    ap.add_argument("--reps", type=int, default=5)
    ap.add_argument("--n", type=int, default=None)
    ap.add_argument("--steps", type=int, default=2000)
    ap.add_argument("--every", type=int, default=100)
    ap.add_argument("--T", type=float, default=20.0)
    ap.add_argument("--dt0", type=float, default=0.01)
    ap.add_argument("--oracle-eps", type=float, default=1e-10)
    ap.add_argument("--eps", default="1e-6")
    ap.add_argument("--model", default=None, help="learned controller weights (.npz)")
    ap.add_argument("--storages", default=None,
                    help="comma-separated variants: c128, c64, mixed, each optionally /stages "
                         "(default: time_to_t c128/stages,c128,c64,mixed/stages,mixed; "
                         "attempts c128/stages,c128,mixed/stages,mixed,c64)")
    ap.add_argument("--methods", default=None,
                    help="time_to_t_tuned: comma-separated subset of plain-c128, plain-c128/stages, plain-mixed, "
                         "learned-mixed, plain-c64, rk4-c128 (default: all)")
    ap.add_argument("--attempts", type=int, default=6)
    ap.add_argument("--copy-probe-mib", type=int, default=0)
    a = ap.parse_args(argv)
    if a.study == "drift":
        drift(a.n or 256, a.steps, a.every)
    elif a.study == "diag":
        diag(a.n or 512, a.reps)
    elif a.study == "time_to_t_tuned":
        from .learned import DEFAULT_MODEL, LearnedController, load_model

        n = a.n or 512
        p, psi = ics.config_turb(n=n, storage="c128", seed=0, scheme="rk45", tau=a.dt0)
        model = load_model(a.model or DEFAULT_MODEL)
        methods = [("plain-c128", "c128", None), ("plain-c128/stages", "c128/stages", None),
                   ("plain-mixed", "mixed", None),
                   ("learned-mixed", "mixed", lambda eps: LearnedController(model, p.replace(eps=eps))),
                   ("plain-c64", "c64", None), ("rk4-c128", "rk4:c128", None)]
        if a.methods:
            keep = a.methods.split(",")
            methods = [m for m in methods if m[0] in keep]
        out = time_to_t_tuned(p, psi, a.T, a.dt0, methods, oracle_eps=a.oracle_eps)
        emit(dict(study="time_to_t_tuned", n=n, T=a.T, error_target=1e-6, baseline="plain-c128", runs=out))
    elif a.study == "attempts":
        attempts(a.n or 512, a.attempts, tuple((a.storages or "c128/stages,c128,mixed/stages,mixed,c64").split(",")),
                 a.copy_probe_mib)
    else:
        model = None
        if a.model:
            from .learned import load_model
            model = load_model(a.model)
        time_to_t(a.n or 512, a.T, a.dt0, a.oracle_eps, [float(v) for v in a.eps.split(",")], model,
                  tuple((a.storages or "c128/stages,c128,c64,mixed/stages,mixed").split(",")))
    return 0


if __name__ == "__main__":
    sys.exit(main())
