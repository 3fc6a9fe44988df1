"""Learned step-size proposer for the adaptive Cash-Karp path.

An FP32 MLP (10 -> 64 -> 64 -> 1, tanh; HIP kernels gpe_mlp_*) proposes
log(dt_next / dt) from the recent history of the solve:

    4 error ratios (errmax of the last 4 accepted steps, log),
    4 step sizes (last 4 accepted dt, log of dt / dt_ref with dt_ref the
      inverse spectral radius of the linear RHS operator of the grid),
    max |psi|^2 of the current state (log),
    the relative drift of the reference energy since t = 0 (signed log).

Acceptance is still decided by the embedded Cash-Karp estimate
(errmax <= 1); after a rejection the reference's shrink rule is used, and a
non-finite proposal falls back to the reference's growth rule (the plain
controller is the fallback and the arbiter).

Training data are RKF45 traces the solver generates itself on random-phase
runs (the 512^3 turbulence workload's IC on smaller grids, 2nd and 4th
order, several seeds), with an exploring controller that perturbs the plain
proposal so the data cover steps larger and smaller than the plain rule's.
The target for the step after step n is the step size that would have given
an error ratio rho on its first attempt, dt* = dt_try (rho / errmax_try)^(1/5).

    python -m gpe_amd.learned all [--model gpe_amd/models/proposer.npz]
        generate traces, train on the GPU, measure on held-out traces how much
        of the plain rule's step-size error the network removes, evaluate on
        the two-ring workload
"""
from __future__ import annotations

import argparse
import ctypes as ct
import json
import math
import os
import sys
import time

import numpy as np

from . import _lib, ics
from .context import Context
from .controller import PlainController, integrate, trace_array
from .params import RunParams

N_IN = 10
# error ratio the proposals aim at on the first attempt: the plain rule's own
# level (safety^5 = 0.9^5, dt_next = 0.9 dt errmax^-1/5 lands there when the
# error constant does not change), so the two controllers trade the same
# error per step and differ only in how well they anticipate its change
RHO = 0.9 ** 5
CLIP = (math.log(0.1), math.log(5.0))
PROPOSE_CLIP = (math.log(0.5), math.log(2.0))
HERE = os.path.dirname(os.path.abspath(__file__))
DEFAULT_MODEL = os.path.join(HERE, "models", "proposer.npz")


# ---------------------------------------------------------------------------
# MLP on the GPU

class MLP:
    def __init__(self, n_in: int = N_IN):
        self.lib = _lib.load()
        h = ct.c_void_p()
        rc = self.lib.gpe_mlp_create(n_in, ct.byref(h))
        if rc != 0:
            raise _lib.GpeError(f"gpe_mlp_create failed: {self.lib.gpe_mlp_last_error(None).decode()}")
        self.h = h
        self.n_in = n_in
        self.np = self.lib.gpe_mlp_num_params(h)

    def _chk(self, rc, what):
        if rc != 0:
            raise _lib.GpeError(f"{what} failed ({rc}): {self.lib.gpe_mlp_last_error(self.h).decode()}")

    def close(self):
        if getattr(self, "h", None) is not None and self.h.value:
            self.lib.gpe_mlp_destroy(self.h)
            self.h = ct.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_params(self, p: np.ndarray):
        p = np.ascontiguousarray(p, dtype=np.float32)
        self._chk(self.lib.gpe_mlp_set_params(self.h, p.ctypes.data, p.size), "gpe_mlp_set_params")

    def get_params(self) -> np.ndarray:
        p = np.empty(self.np, dtype=np.float32)
        self._chk(self.lib.gpe_mlp_get_params(self.h, p.ctypes.data, p.size), "gpe_mlp_get_params")
        return p

    def forward(self, X: np.ndarray) -> np.ndarray:
        X = np.ascontiguousarray(X, dtype=np.float32).reshape(-1, self.n_in)
        y = np.empty(X.shape[0], dtype=np.float32)
        self._chk(self.lib.gpe_mlp_forward(self.h, X.ctypes.data, X.shape[0], y.ctypes.data), "gpe_mlp_forward")
        return y

    def grad(self, X, Y):
        X = np.ascontiguousarray(X, dtype=np.float32)
        Y = np.ascontiguousarray(Y, dtype=np.float32)
        g = np.empty(self.np, dtype=np.float32)
        loss = np.empty(1, dtype=np.float32)
        self._chk(self.lib.gpe_mlp_grad(self.h, X.ctypes.data, Y.ctypes.data, X.shape[0], g.ctypes.data,
                                        loss.ctypes.data), "gpe_mlp_grad")
        return g, float(loss[0])

    def train(self, X, Y, steps: int, batch: int, lr: float) -> np.ndarray:
        X = np.ascontiguousarray(X, dtype=np.float32)
        Y = np.ascontiguousarray(Y, dtype=np.float32)
        losses = np.empty(steps, dtype=np.float32)
        self._chk(self.lib.gpe_mlp_train(self.h, X.ctypes.data, Y.ctypes.data, X.shape[0], steps, batch, lr,
                                         losses.ctypes.data), "gpe_mlp_train")
        return losses


def init_params(n_in: int = N_IN, seed: int = 0) -> np.ndarray:
    """Glorot-uniform weights, zero biases, small output layer."""
    rng = np.random.default_rng(seed)

    def glorot(fo, fi):
        a = math.sqrt(6.0 / (fi + fo))
        return rng.uniform(-a, a, size=(fo, fi))

    H = 64
    parts = [glorot(H, n_in).ravel(), np.zeros(H), glorot(H, H).ravel(), np.zeros(H), 0.1 * glorot(1, H).ravel(),
             np.zeros(1)]
    return np.concatenate(parts).astype(np.float32)


# ---------------------------------------------------------------------------
# features and data

def dt_ref(p: RunParams) -> float:
    """Inverse spectral radius of the linear part of the RHS on this grid."""
    c = 4.0 if p.order == 2 else 64.0 / 12.0
    s = c * (1 / p.dx ** 2 + 1 / p.dy ** 2 + 1 / p.dz ** 2)
    scale = 1.0 if p.eqn_to_solve == 2 else 0.5
    return 1.0 / (scale * s)


def raw_features(errs, dts, psi2max: float, drift: float, dtr: float) -> np.ndarray:
    e = list(errs[-4:])
    d = list(dts[-4:])
    e = [e[0]] * (4 - len(e)) + e
    d = [d[0]] * (4 - len(d)) + d
    f = [math.log(max(v, 1e-8)) for v in e] + [math.log(v / dtr) for v in d]
    f += [math.log(max(psi2max, 1e-30)), math.copysign(math.log1p(abs(drift) * 1e6), drift)]
    return np.array(f, dtype=np.float64)


def error_exponent(traces) -> float:
    """Effective order p of errmax ~ dt^p near the operating point, fitted by
    least squares over consecutive first attempts of the traces (steps whose
    dt changed by more than 5%)."""
    num = den = 0.0
    for tr in traces:
        dl = np.diff(np.log(tr[:-1, 6]))
        el = np.diff(np.log(np.maximum(tr[:-1, 7], 1e-12)))
        m = np.abs(dl) > 0.05
        num += float(np.sum(dl[m] * el[m]))
        den += float(np.sum(dl[m] * dl[m]))
    p = num / den if den > 0 else 5.0
    return min(max(p, 5.0), 20.0)


def dataset_from_trace(tr: np.ndarray, dtr: float, rho: float = RHO, p: float = 5.0):
    """(X, Y) pairs of one trace: features after accepted step n, target from
    the first attempt of step n+1 (the last, clamped step is left out):
    the step that would have given errmax = rho, with errmax ~ dt^p."""
    X, Y = [], []
    for n in range(len(tr) - 2):
        x = raw_features(tr[:n + 1, 2], tr[:n + 1, 1], tr[n, 4], tr[n, 5], dtr)
        dt_try, e_try = tr[n + 1, 6], tr[n + 1, 7]
        target = dt_try * (rho / max(e_try, 1e-8)) ** (1.0 / p)
        y = min(max(math.log(target / tr[n, 1]), CLIP[0]), CLIP[1])
        X.append(x)
        Y.append(y)
    return np.array(X), np.array(Y)


class ExploringController(PlainController):
    """The plain controller with its proposals scaled by exp(u), u uniform in
    [-spread, spread] (seeded): trace generation only."""
    name = "exploring"

    def __init__(self, seed: int, spread: float = 0.35):
        super().__init__()
        self.rng = np.random.default_rng(seed)
        self.spread = spread

    def propose(self, dt, errmax, state):
        base = super().propose(dt, errmax, state)
        return base * math.exp(self.rng.uniform(-self.spread, self.spread))


class LearnedController(PlainController):
    """Proposes dt_next = dt exp(mlp(features)), clipped to [dt/2, 2 dt] and
    without growth right after a rejected attempt; the reference's rules
    shrink after a rejection and replace a non-finite proposal.

    The plain controller is the judge.  After every accepted step the first
    attempt's errmax at the tried dt gives the step that would have hit
    errmax = rho (the error constant errmax / dt^5 is observed whichever
    proposal was tried), so both the network's and the plain rule's
    proposals for that step are scored in log dt.  The network proposes only
    while its mean squared error over the last `window` steps is below the
    plain rule's; otherwise, and for the first `window` steps, the plain rule
    proposes.  With explore > 0 the proposal is perturbed (on-policy trace
    generation)."""
    name = "learned"

    def __init__(self, model: dict, params: RunParams, mlp: MLP | None = None, explore: float = 0.0, seed: int = 0,
                 window: int = 16):
        super().__init__(params.safety, params.dt_decrease, params.dt_increase)
        self.model = model
        self.dtr = dt_ref(params)
        self.mlp = mlp or MLP(N_IN)
        self.mlp.set_params(model["params"])
        self.mu = model["mu"]
        self.sd = model["sd"]
        self.rho = float(model.get("rho", RHO))
        self.explore = explore
        self.rng = np.random.default_rng(seed)
        self.window = window
        self.reset()

    def reset(self):
        self.n_fallback = 0        # non-finite proposals
        self.n_plain_steps = 0     # steps proposed by the plain rule (the judge preferred it)
        self.rejected = False
        self.pending = None        # (dt, y_net, y_plain) of the last proposal
        self.score = []            # (squared error network, squared error plain) per judged step

    def shrink(self, dt, errmax):
        self.rejected = True
        return super().shrink(dt, errmax)

    def _judge(self, state):
        """Score the previous proposals against the step just taken."""
        if self.pending is None or not state.dt_first:
            return
        dt_prev, y_net, y_plain = self.pending
        target = math.log(state.dt_first[-1] * (self.rho / max(state.errmax_first[-1], 1e-12)) ** 0.2 / dt_prev)
        self.score = (self.score + [((target - y_net) ** 2, (target - y_plain) ** 2)])[-self.window:]

    def propose(self, dt, errmax, state):
        self._judge(state)
        plain = super().propose(dt, errmax, state)
        y_plain = math.log(abs(plain) / abs(dt))
        x = (raw_features(state.errmax, state.dt, state.psi2max, state.energy_drift, self.dtr) - self.mu) / self.sd
        y = float(self.mlp.forward(x[None, :])[0])
        rejected, self.rejected = self.rejected, False
        if not math.isfinite(y):
            self.n_fallback += 1
            self.pending = None
            return plain
        hi = 0.0 if rejected else PROPOSE_CLIP[1]
        y = min(max(y, PROPOSE_CLIP[0]), hi)
        self.pending = (abs(dt), y, y_plain)
        trusted = len(self.score) == self.window and \
            sum(a for a, _ in self.score) < sum(b for _, b in self.score)
        if not trusted:
            self.n_plain_steps += 1
            return plain
        if self.explore > 0:
            y += self.rng.uniform(-self.explore, self.explore)
        return dt * math.exp(y)


def predictability(data, model, rho: float = RHO) -> dict:
    """How much of the next step's error the network anticipates, on traces
    (RunParams, trace) the model did not see.  The target is the step that
    would have hit errmax = rho on the next first attempt; the plain rule
    (src/solve.f90:214-220) proposes log(0.9 errmax^-1/5) (capped at 5), which
    is that target exactly when the error constant errmax / dt^5 does not
    change from one step to the next.  Returns the mean squared errors of both
    proposals in log dt and gain = 1 - mse_network / mse_plain, the share of
    the plain rule's error that the network removes (0: it anticipates
    nothing beyond the last step's error; 1: it predicts every change)."""
    errcon = (5.0 / 0.9) ** (1.0 / -0.2)
    mlp = MLP(N_IN)
    mlp.set_params(model["params"])
    ys, yp, yn = [], [], []
    for p, tr in data:
        X, Y = dataset_from_trace(tr, dt_ref(p), rho, 5.0)
        e = tr[:len(Y), 2]
        plain = np.where(e > errcon, math.log(0.9) - 0.2 * np.log(np.maximum(e, 1e-30)), math.log(5.0))
        net = mlp.forward(((X - model["mu"]) / model["sd"]).astype(np.float32)).astype(np.float64)
        ys.append(Y)
        yp.append(np.minimum(plain, CLIP[1]))
        yn.append(np.clip(net, PROPOSE_CLIP[0], PROPOSE_CLIP[1]))
    Y, P, N = np.concatenate(ys), np.concatenate(yp), np.concatenate(yn)
    mse_p, mse_n = float(np.mean((Y - P) ** 2)), float(np.mean((Y - N) ** 2))
    return {"samples": int(len(Y)), "mse_plain": mse_p, "mse_network": mse_n, "gain": 1.0 - mse_n / mse_p}


def save_model(path: str, model: dict) -> None:
    os.makedirs(os.path.dirname(path), exist_ok=True)
    np.savez(path, params=model["params"], mu=model["mu"], sd=model["sd"], rho=model["rho"],
             meta=json.dumps(model.get("meta", {})))


def load_model(path: str) -> dict:
    z = np.load(path, allow_pickle=False)
    return {"params": z["params"], "mu": z["mu"], "sd": z["sd"], "rho": float(z["rho"]),
            "meta": json.loads(str(z["meta"]))}


# ---------------------------------------------------------------------------
# pipeline

TRACE_RUNS = [  # (n, order, seed, eps, T, exploring spread)
    (128, 4, 0, 1e-6, 20.0, 0.0), (128, 4, 1, 1e-6, 20.0, 0.35), (128, 4, 2, 3e-6, 20.0, 0.35),
    (96, 4, 3, 1e-6, 20.0, 0.5), (128, 2, 4, 1e-6, 20.0, 0.35), (96, 2, 5, 3e-7, 20.0, 0.35),
    (64, 4, 6, 1e-6, 20.0, 0.5), (128, 4, 7, 1e-7, 10.0, 0.35),
]


def generate_traces(runs=TRACE_RUNS, log=print, model=None):
    """RKF45 traces of the random-phase workload; with a model, the learned
    controller drives the runs (exploring by `spread`), otherwise the plain
    controller does (exploring by `spread`)."""
    data = []
    for n, order, seed, eps, T, spread in runs:
        p, psi = ics.config_turb(n=n, seed=seed, scheme="rk45", tau=0.01)
        p = p.replace(order=order, eps=eps)
        if model is not None:
            ctrl = LearnedController(model, p, explore=spread, seed=seed)
        else:
            ctrl = ExploringController(seed, spread) if spread > 0 else PlainController()
        with Context(p) as ctx:
            ctx.upload(psi)
            st = integrate(ctx, T, 0.01, ctrl, clamp=True, energy=True)
        tr = trace_array(st)
        data.append((p, tr))
        log(json.dumps({"learned": "trace", "policy": "learned" if model is not None else "plain", "n": n,
                        "order": order, "seed": seed, "eps": eps, "T": T, "spread": spread, "steps": st.n_accept,
                        "rejects": st.n_reject, "wall_s": st.wall_s}))
    return data


def train_pipeline(runs=TRACE_RUNS, rounds: int = 2, steps: int = 4000, log=print):
    """Pass 1: traces of the plain/exploring controller, train.  Later
    passes: traces of the learned controller itself (small exploration),
    added to the data set, retrain (on-policy aggregation)."""
    data = generate_traces(runs, log=log)
    model = None
    for r in range(rounds):
        if r > 0:
            data += generate_traces([(n, o, s + 100 * r, e, T, 0.1) for n, o, s, e, T, _ in runs], log=log,
                                    model=model)
        X, Y = build_dataset(data)
        model = train_model(X, Y, steps=steps, log=log)
        model["meta"]["round"] = r
        model["meta"]["error_exponent"] = build_dataset.last_exponent
        log(json.dumps({"learned": "round", "round": r, "error_exponent": build_dataset.last_exponent}))
    return model


def build_dataset(data, rho=RHO, seed=0):
    pexp = error_exponent([tr for _, tr in data])
    Xs, Ys = [], []
    for p, tr in data:
        X, Y = dataset_from_trace(tr, dt_ref(p), rho, pexp)
        Xs.append(X)
        Ys.append(Y)
    X = np.concatenate(Xs)
    Y = np.concatenate(Ys)
    perm = np.random.default_rng(seed).permutation(len(X))
    build_dataset.last_exponent = pexp
    return X[perm], Y[perm]


def train_model(X, Y, steps=4000, batch=256, lr=2e-3, seed=0, log=print):
    mu = X.mean(0)
    sd = X.std(0) + 1e-6
    Xn = ((X - mu) / sd).astype(np.float32)
    mlp = MLP(N_IN)
    mlp.set_params(init_params(N_IN, seed))
    t0 = time.perf_counter()
    losses = mlp.train(Xn, Y.astype(np.float32), steps, batch, lr)
    wall = time.perf_counter() - t0
    _, full = mlp.grad(Xn, Y.astype(np.float32))
    base = float(np.mean((Y - Y.mean()) ** 2))
    log(json.dumps({"learned": "train", "samples": len(X), "steps": steps, "batch": batch, "lr": lr,
                    "loss_first": float(losses[0]), "loss_last": float(losses[-1]), "loss_full": full,
                    "variance_of_target": base, "wall_s": wall}))
    return {"params": mlp.get_params(), "mu": mu, "sd": sd, "rho": RHO,
            "meta": {"samples": int(len(X)), "loss": full, "target_variance": base}}


def held_out_predictability(model, log=print):
    """predictability() on plain-controller traces the model did not see: the
    turbulence workload with a new seed (128^3, T = 20) and the two-ring
    workload (128^3, dx 0.5, T = 40), whose error constant trends smoothly."""
    out = {}
    for name, (p, psi, T) in {"turbulence_seed9": (*ics.config_turb(n=128, seed=9, scheme="rk45", tau=0.01), 20.0),
                              "two_rings": (*ics.config_two_rings(n=128, dx=0.5), 40.0)}.items():
        p = p.replace(eps=1e-6)
        with Context(p) as ctx:
            ctx.upload(psi)
            st = integrate(ctx, T, 0.01, PlainController(), clamp=True, energy=True)
        tr = trace_array(st)
        lc = np.log(np.maximum(tr[:, 2], 1e-30)) - 5.0 * np.log(tr[:, 1])
        d = np.diff(lc[20:])
        r = predictability([(p, tr)], model)
        r["increment_autocorrelation"] = float(np.corrcoef(d[:-1], d[1:])[0, 1])
        r["plain_steps"], r["plain_rejects"] = st.n_accept, st.n_reject
        out[name] = r
        log(json.dumps({"learned": "predictability", "workload": name, **r}))
    return out


def evaluate(model, n=384, dx=0.5, T=40.0, eps=1e-6, oracle_eps=1e-10, log=print):
    """Two coaxial rings (radius 6 and 10, z = -10 and +10, moving towards
    each other), 4th order, RKF45 to T: plain vs learned, both against an
    FP64 reference at a tight tolerance."""
    from .studies import rel_l2

    p, psi = ics.config_two_rings(n=n, dx=dx)
    rows = []
    finals = {}
    for name, ctrl_fn, e in [("oracle", lambda: PlainController(), oracle_eps),
                             ("plain", lambda: PlainController(), eps),
                             ("learned", lambda: LearnedController(model, p), eps)]:
        pp = p.replace(eps=e)
        ctrl = ctrl_fn()
        with Context(pp) as ctx:
            ctx.upload(psi)
            st = integrate(ctx, T, 0.01, ctrl, clamp=True, energy=(name == "learned"))
            finals[name] = ctx.download()
        row = {"learned": "evaluate", "method": name, "n": n, "T": T, "eps": e, "steps": st.n_accept,
               "rejects": st.n_reject, "rhs": st.n_rhs, "wall_s": st.wall_s}
        if name == "learned":
            row["plain_rule_steps"] = ctrl.n_plain_steps
        rows.append(row)
    for r in rows:
        r["rel_l2_vs_oracle"] = rel_l2(finals[r["method"]], finals["oracle"]) if r["method"] != "oracle" else 0.0
        log(json.dumps(r))
    pl = next(r for r in rows if r["method"] == "plain")
    le = next(r for r in rows if r["method"] == "learned")
    log(json.dumps({"learned": "gain", "wall_ratio_plain_over_learned": pl["wall_s"] / le["wall_s"],
                    "rhs_ratio": pl["rhs"] / le["rhs"], "steps_ratio": pl["steps"] / le["steps"],
                    "error_ratio_learned_over_plain": le["rel_l2_vs_oracle"] / max(pl["rel_l2_vs_oracle"], 1e-300)}))
    return rows


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("what", choices=["all", "train", "evaluate"])
    ap.add_argument("--model", default=DEFAULT_MODEL)
    ap.add_argument("--n", type=int, default=384)
    ap.add_argument("--T", type=float, default=40.0)
    ap.add_argument("--steps", type=int, default=4000)
    a = ap.parse_args(argv)
    if a.what in ("all", "train"):
        model = train_pipeline(steps=a.steps)
        save_model(a.model, model)
    if a.what in ("all", "evaluate"):
        model = load_model(a.model)
        held_out_predictability(model)
        evaluate(model, n=a.n, T=a.T)
    return 0


if __name__ == "__main__":
    sys.exit(main())
