"""Learned Newton warm start: a per-vertex MLP that predicts each step's
position increment from local state, trained on trajectories of the plain
solver.

The reference starts every Newton solve from x = x_n
(projects/FastIPC/NEWTON_IPC.py:958-960).  Here the network's prediction is
tried first: it is applied like a Newton direction, i.e. clipped by the CCD
and inversion step bound (:456-468), halved while an edge crosses a triangle
(:983-985) and kept only if it lowers the incremental potential (:986-992);
otherwise the step starts from x_n as in the reference.  The Newton
convergence test is unchanged, so the warm start changes how many iterations
a step takes, not what a converged step satisfies.

Network: 16 features -> 4 x 256 SiLU -> 3, BF16 matrix-core products with
FP32 accumulation (``fipc_mlp_*``), Adam.  Features and the prediction are
scaled by s = 0.1 dt bbox-diagonal.  The correction c = x_{n+1} - xTilde to
the inertial position xTilde = x_n + dt v + dt^2 g changes slowly from step
to step, so the start is xTilde + c_prev + s net(f) and the target is the
change (c - c_prev) / s: an untrained (zero) network repeats the previous
correction, which is second-order extrapolation of the trajectory.

Command line:

  python -m fastipc_amd.warmstart train --family slab --out W.npz
  python -m fastipc_amd.warmstart eval  --family slab --weights W.npz
"""
from __future__ import annotations

import argparse
import ctypes as C
import json
import math
import os
import sys
import time

import numpy as np

from . import _lib

N_IN, WIDTH, DEPTH, N_OUT = 16, 256, 4, 3
HERE = os.path.dirname(os.path.abspath(__file__))
WEIGHTS_DIR = os.path.join(HERE, "weights")


def scale_for(scene) -> float:
    """Length scale of features and prediction: 0.1 dt bbox-diagonal."""
    return 0.1 * scene.dt * scene.bbox_diag


def bf16_round(w: np.ndarray) -> np.ndarray:
    """float32 -> nearest BF16 (ties to even), returned as float32."""
    u = np.ascontiguousarray(w, np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    return (u.astype(np.uint32) << 16).view(np.float32)


class WarmStartMLP:
    """The network and its device-resident training set."""

    def __init__(self, seed: int = 0, params: np.ndarray | None = None):
        _lib.load()
        h = C.c_void_p()
        _lib.call("fipc_mlp_create", C.byref(h), N_IN, WIDTH, DEPTH, N_OUT, seed)
        self.h = h
        n = C.c_long()
        _lib.call("fipc_mlp_nparams", h, C.byref(n))
        self.nparams = n.value
        if params is not None:
            self.set_params(params)

    def close(self):
        if self.h:
            _lib.call("fipc_mlp_destroy", self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ---- parameters
    def params(self) -> np.ndarray:
        w = np.zeros(self.nparams, np.float32)
        _lib.call("fipc_mlp_get_params", self.h, _lib.fptr(w))
        return w

    def set_params(self, w):
        w = np.ascontiguousarray(w, np.float32)
        assert w.size == self.nparams
        _lib.call("fipc_mlp_set_params", self.h, _lib.fptr(w))

    def save(self, path: str, meta: dict | None = None):
        """Parameters rounded to BF16 (the precision the products use), stored
        as the upper 16 bits of float32, with JSON metadata."""
        w = bf16_round(self.params())
        np.savez(path, bf16=(w.view(np.uint32) >> 16).astype(np.uint16),
                 meta=np.array(json.dumps(meta or {})))

    @classmethod
    def load(cls, path: str) -> "WarmStartMLP":
        d = np.load(path, allow_pickle=False)
        w = (d["bf16"].astype(np.uint32) << 16).view(np.float32)
        m = cls(params=w)
        m.meta = json.loads(str(d["meta"]))
        return m

    # ---- evaluation
    def forward(self, X: np.ndarray) -> np.ndarray:
        X = np.ascontiguousarray(X, np.float32)
        Y = np.zeros((len(X), N_OUT), np.float32)
        _lib.call("fipc_mlp_forward", self.h, len(X), _lib.fptr(X), _lib.fptr(Y))
        return Y

    def grad(self, X, T):
        X = np.ascontiguousarray(X, np.float32)
        T = np.ascontiguousarray(T, np.float32)
        g = np.zeros(self.nparams, np.float32)
        loss = C.c_float()
        _lib.call("fipc_mlp_grad", self.h, len(X), _lib.fptr(X), _lib.fptr(T), C.byref(loss), _lib.fptr(g))
        return loss.value, g

    # ---- training set and training
    def dataset_size(self) -> int:
        n = C.c_long()
        _lib.call("fipc_mlp_dataset_size", self.h, C.byref(n))
        return n.value

    def dataset_add(self, X, T):
        X = np.ascontiguousarray(X, np.float32)
        T = np.ascontiguousarray(T, np.float32)
        _lib.call("fipc_mlp_dataset_add", self.h, len(X), _lib.fptr(X), _lib.fptr(T))

    def dataset(self):
        n = self.dataset_size()
        X = np.zeros((n, N_IN), np.float32)
        T = np.zeros((n, N_OUT), np.float32)
        if n:
            _lib.call("fipc_mlp_dataset_get", self.h, 0, n, _lib.fptr(X), _lib.fptr(T))
        return X, T

    def train(self, iters: int, batch: int = 8192, lr: float = 1e-3, seed: int = 0) -> np.ndarray:
        losses = np.zeros(iters, np.float32)
        _lib.call("fipc_mlp_train", self.h, iters, batch, lr, seed, _lib.fptr(losses))
        return losses


class LearnedWarmStart:
    """Hook for NewtonIPC.step(warm_start=...): writes the predicted increment
    into the solver's direction vector on the device."""

    def __init__(self, mlp: WarmStartMLP, scale: float):
        self.mlp = mlp
        self.scale = scale

    def __call__(self, sim):
        _lib.call("fipc_ws_predict", sim.ctx, self.mlp.h, self.scale)
        return True


class Capture:
    """Hook for NewtonIPC.step(capture=...): records (features, target) rows of
    nsample random free vertices per step into the network's training set, a
    share ``near_frac`` of them drawn among the vertices within d-hat of a
    contact primitive."""

    def __init__(self, mlp: WarmStartMLP, scale: float, nsample: int = 4096, seed: int = 0, near_frac: float = 0.0):
        self.mlp, self.scale, self.nsample, self.seed = mlp, scale, nsample, seed
        self.near_frac = near_frac
        self.k = 0

    def begin(self, sim):
        _lib.call("fipc_ws_capture_begin", sim.ctx, self.mlp.h, self.scale, self.nsample,
                  self.seed * 1000003 + self.k, float(self.near_frac))
        self.k += 1

    def end(self, sim):
        _lib.call("fipc_ws_capture_end", sim.ctx, self.mlp.h, self.scale)


# ----------------------------------------------------------------------------
# scene families: the training variations and the held-out evaluation scenes
# ----------------------------------------------------------------------------
def family_scene(family: str, seed: int, size: str = "full"):
    """Training and evaluation scenes.  slab: the compressed slab with seeded
    closing rate (0.1-0.3 heights/s), twist amplitude (10-45 deg) and frequency
    (0.5-1.5 Hz) at 20x32x20 ("small"), 30x48x30 ("eval") or 50x80x50
    ("full") vertices; "bench" is the 50x80x50 benchmark scene itself (0.2,
    30 deg, 1 Hz), held out of training.  cubes: the 3x3x3 grid of cubes with
    seeded rotations and jitter, 8^3-vertex cubes or 16^3 ("bench")."""
    from . import scene as S
    rng = np.random.default_rng(seed)
    if family == "slab":
        if size == "bench":
            return S.compressed_slab(50, 80, 50)
        close = rng.uniform(0.1, 0.3)
        twist = rng.uniform(10.0, 45.0)
        omega = 2 * math.pi * rng.uniform(0.5, 1.5)
        dims = {"small": (20, 32, 20), "eval": (30, 48, 30), "full": (50, 80, 50)}[size]
        return S.compressed_slab(*dims, close_rate=close, twist_deg=twist, omega=omega)
    if family == "cubes":
        n = 16 if size in ("bench", "full") else 8
        return S.cube_grid(n, 3, seed=seed)
    raise ValueError(family)


def capture_trajectories(mlp, family, seeds, steps, nsample=4096, size="full", log=print, warm=None, near_frac=0.0):
    """Training rows from trajectories of the plain solver, or, with ``warm``
    (a WarmStartMLP), of the solver warm-started by that model (on-policy
    rows: the states the warm start itself produces)."""
    from .newton import NewtonIPC
    stats = []
    for sd in seeds:
        sc = family_scene(family, sd, size)
        sim = NewtonIPC(sc)
        cap = Capture(mlp, scale_for(sc), nsample, seed=sd, near_frac=near_frac)
        hook = LearnedWarmStart(warm, scale_for(sc)) if warm is not None else None
        t0 = time.perf_counter()
        its = []
        for _ in range(steps):
            st = sim.step(warm_start=hook, capture=cap)
            its.append(st.newton_iters)
        sim.close()
        stats.append(float(np.mean(its)))
        log(f"captured seed {sd}: {steps} steps, {np.mean(its):.2f} Newton iterations/step, "
            f"{time.perf_counter() - t0:.1f} s, training set {mlp.dataset_size()} rows")
    return stats


def error_stats(P, T, tol=0.1):
    """Prediction errors in units of the scale: RMS, and quantiles of the
    per-vertex max-norm error, which the Newton test (max |p| < tol, tol = 0.1
    in these units) sees."""
    e = np.abs(P - T).max(axis=1)
    return {"rms": float(np.sqrt(np.mean((P - T) ** 2))), "p99": float(np.quantile(e, 0.99)),
            "p999": float(np.quantile(e, 0.999)), "max": float(e.max()), "frac_above_tol": float(np.mean(e > tol))}


def validation(mlp, family, seed, steps, size="bench", nsample=16384, skip=10, log=print):
    """Held-out trajectory of the plain solver: errors of the zero network
    (second-order extrapolation) and of the trained network."""
    from .newton import NewtonIPC
    holder = WarmStartMLP(seed=1)
    sc = family_scene(family, seed, size)
    sim = NewtonIPC(sc)
    cap = Capture(holder, scale_for(sc), nsample, seed=seed)
    for k in range(steps):
        sim.step(capture=cap if k >= skip else None)
    sim.close()
    X, T = holder.dataset()
    holder.close()
    out = {"rows": int(len(X)), "zero": error_stats(np.zeros_like(T), T), "trained": error_stats(mlp.forward(X), T)}
    log("validation " + json.dumps(out))
    return out


def run_eval(mlp, family, seeds, steps, warmup, size, log=print, precision=0,
             modes=("plain", "extrapolation", "learned")):
    """Newton iterations per step and steps/s on the same scenes for three
    starts: x_n (plain, the reference), the zero network (second-order
    extrapolation) and the trained network; final positions of the warm-started
    runs are compared with the plain run."""
    from .newton import NewtonIPC
    zero = WarmStartMLP(params=np.zeros(mlp.nparams, np.float32))
    out = []
    for sd in seeds:
        row = {"seed": sd}
        xs = {}
        for mode in modes:
            sc = family_scene(family, sd, size)
            sim = NewtonIPC(sc, precision=precision)
            net = {"plain": None, "extrapolation": zero, "learned": mlp}[mode]
            hook = LearnedWarmStart(net, scale_for(sc)) if net is not None else None
            for _ in range(warmup):
                sim.step(warm_start=hook)
            _lib.call("fipc_sync", sim.ctx)
            t0 = time.perf_counter()
            its, acc, stalled = [], 0, []
            for _ in range(steps):
                st = sim.step(warm_start=hook)
                its.append(st.newton_iters)
                stalled.append(bool(st.stalled))
                acc += int(st.warm_accepted)
                if sim.check_collision():
                    raise RuntimeError("intersection after a step")
            _lib.call("fipc_sync", sim.ctx)
            wall = time.perf_counter() - t0
            its_a, st_a = np.array(its), np.array(stalled)
            # steps that converged (a stalled step ends when the line search
            # can no longer lower the energy, however good the start was)
            conv = its_a[~st_a]
            row[mode] = {"newton_per_step": float(np.mean(its)), "steps_per_s": steps / wall,
                         "warm_accepted": acc, "stalled_steps": int(st_a.sum()),
                         "newton_per_converged_step": float(conv.mean()) if len(conv) else None,
                         "newton_per_step_by_quarter": [round(float(q.mean()), 2)
                                                        for q in np.array_split(its_a, 4) if len(q)]}
            xs[mode] = sim.x
            sim.close()
        for mode in ("extrapolation", "learned"):
            if mode in xs and "plain" in xs:
                row[mode]["max_position_difference_over_bbox"] = float(np.abs(xs[mode] - xs["plain"]).max() / sc.bbox_diag)
        log(json.dumps(row))
        out.append(row)
    zero.close()
    return out


def main(argv=None):
    ap = argparse.ArgumentParser(prog="python -m fastipc_amd.warmstart")
    sub = ap.add_subparsers(dest="cmd", required=True)
    tr = sub.add_parser("train")
    tr.add_argument("--family", choices=("slab", "cubes"), default="slab")
    tr.add_argument("--seeds", default="0-15")
    tr.add_argument("--steps", type=int, default=40)
    tr.add_argument("--nsample", type=int, default=4096)
    tr.add_argument("--iters", type=int, default=3000)
    tr.add_argument("--batch", type=int, default=8192)
    tr.add_argument("--lr", type=float, default=1e-3)
    tr.add_argument("--size", default="full")
    tr.add_argument("--val-size", default="full")
    tr.add_argument("--val-seed", type=int, default=1000)
    tr.add_argument("--out", required=True)
    tr.add_argument("--init", default=None, help="continue from these weights")
    tr.add_argument("--onpolicy", action="store_true",
                    help="also capture trajectories warm-started by the initial model")
    tr.add_argument("--near-frac", type=float, default=0.0,
                    help="share of each step's samples drawn near contacts")
    ev = sub.add_parser("eval")
    ev.add_argument("--family", choices=("slab", "cubes"), default="slab")
    ev.add_argument("--weights", required=True)
    ev.add_argument("--seeds", default="1000-1001")
    ev.add_argument("--size", default="bench")
    ev.add_argument("--steps", type=int, default=20)
    ev.add_argument("--warmup", type=int, default=10)
    ev.add_argument("--precision", type=int, default=0)
    ev.add_argument("--modes", default="plain,extrapolation,learned",
                    help="comma-separated starts to run: plain, extrapolation, learned")
    a = ap.parse_args(argv)

    def seeds(spec):
        lo, _, hi = spec.partition("-")
        return list(range(int(lo), int(hi or lo) + 1))

    def log(m):
        print(m, flush=True)

    if a.cmd == "train":
        mlp = WarmStartMLP.load(a.init) if a.init else WarmStartMLP(seed=12345)
        t0 = time.perf_counter()
        capture_trajectories(mlp, a.family, seeds(a.seeds), a.steps, a.nsample, size=a.size, log=log,
                             near_frac=a.near_frac)
        if a.onpolicy:
            # the rows of the first pass stay in the dataset; the model that
            # steps is a copy of the initial weights
            behaviour = WarmStartMLP(params=mlp.params())
            capture_trajectories(mlp, a.family, seeds(a.seeds), a.steps, a.nsample, size=a.size, log=log,
                                 warm=behaviour, near_frac=a.near_frac)
            behaviour.close()
        t1 = time.perf_counter()
        losses = []
        chunk = 250
        for k in range(0, a.iters, chunk):
            l = mlp.train(min(chunk, a.iters - k), a.batch, a.lr, seed=k + 1)
            losses.extend(l.tolist())
            log(f"train iters {k + len(l)}: batch loss {float(np.mean(l[-50:])):.5f}")
        t2 = time.perf_counter()
        X, T = mlp.dataset()
        base = float(np.mean(T ** 2))
        val = validation(mlp, a.family, a.val_seed, a.steps, size=a.val_size, log=log)
        meta = {"family": a.family, "seeds": a.seeds, "steps": a.steps, "nsample": a.nsample, "iters": a.iters,
                "batch": a.batch, "lr": a.lr, "rows": int(len(X)), "final_loss": float(np.mean(losses[-100:])),
                "zero_prediction_loss": base, "capture_s": t1 - t0, "train_s": t2 - t1,
                "size": a.size, "validation_seed": a.val_seed, "validation_size": a.val_size, "validation": val,
                "init": a.init, "onpolicy": bool(a.onpolicy), "near_frac": a.near_frac}
        mlp.save(a.out, meta)
        log(json.dumps(meta))
    else:
        mlp = WarmStartMLP.load(a.weights)
        run_eval(mlp, a.family, seeds(a.seeds), a.steps, a.warmup, a.size, log=log, precision=a.precision,
                 modes=tuple(m for m in a.modes.split(",") if m))


if __name__ == "__main__":
    sys.exit(main())
