"""Train the learned warm start on traces the solver generates, and evaluate it.

    python -m jitterphysics2_amd.lws_train [--samples 2000000] [--out jitterphysics2_amd/models/lws_heap.npz]

Traces: a heap of dynamic spheres/boxes/capsules in a static container (the 131,072-body heap
generator, seed 0x4A2D0004, a different scene from the benchmark's 0x4A2D0002) is stepped with
the previous-frame warm start from the impacts (step 240) to the settled heap (step 900); every
20 steps a probe records, for every contact point, the 24 features at the warm start of the first
substep and the accumulated impulse after 64 plain PGS iterations (the target), and an equal
seeded share of each probe's points is kept.  Samples are shuffled with the seed, split 90 / 10, the
normalisation is fitted on the training part, and the network is trained on the GPU by the
library's BF16 MFMA forward/backward kernels with Adam.  Finally the correction is shrunk by the
factor (1, 1/2, ..., 1/16, or 0) whose probe residuals on the training heap are lowest relative to
the previous-frame warm start (see tune_shrinkage).

Evaluation (`evaluate`): on a heap state, one probe step runs the first substep from identical
states with the previous-frame warm start, the learned one and a cold (zero) start, and reports
R_k = ||e_k|| / ||e_0(cold)|| and the first k with R_k <= 1e-3.
"""
from __future__ import annotations

import argparse
import os
import time

import numpy as np

from . import scenes
from .gpu import GpuWorld
from .lws import (PROBE_COLD, PROBE_LEARNED, PROBE_PREVIOUS, LearnedWarmStart, collect_traces, impulse_scale,
                  iterations_to, relative_residuals)

DEFAULT_MODEL = os.path.join(os.path.dirname(os.path.abspath(__file__)), "models", "lws_heap.npz")


def heap_world(n_bodies: int, seed: int, warmup: int, warm_start: int = 0):
    bodies, shapes, verts, cfg, dt = scenes.heap(n_bodies, seed=seed)
    cfg["warm_start"] = warm_start
    gw = GpuWorld(bodies, shapes, verts, cfg)
    for _ in range(warmup):
        gw.step(dt)
    return gw, dt, int(np.sum(bodies["motion_type"] == 0))


def gather(samples: int, n_bodies: int, seed: int, first: int, last: int, every: int, iterations: int, log=print):
    """Probe every `every` steps from step `first` to `last` of a heap (impacts, then settling) and
    keep an equal seeded share of each probe's contact points, `samples` in total."""
    gw, dt, _ = heap_world(n_bodies, seed, first - 1)
    probes = max(1, (last - first) // every + 1)
    share = samples // probes + 1
    rng = np.random.default_rng(seed)
    fs, ts = [], []
    for p in range(probes):
        f, t = collect_traces(gw, dt, steps=1 if p == 0 else every, every=1 if p == 0 else every,
                              iterations=iterations)
        keep = rng.permutation(len(f))[:share]
        fs.append(f[keep])
        ts.append(t[keep])
        log(f"# traces: step {gw.stats()['steps']}: {len(f)} contact points, kept {len(keep)}")
    f, t = np.concatenate(fs), np.concatenate(ts)
    return f[:samples], t[:samples], gw, dt


def tune_shrinkage(model: LearnedWarmStart, gw: GpuWorld, dt: float, alphas=(1.0, 0.5, 0.25, 0.125, 0.0625),
                   iterations: int = 16, log=print) -> tuple[float, list]:
    """Shrink the learned correction by alpha (target_scale * alpha), choosing the alpha whose probe
    residuals R_4, R_8, R_16 relative to the previous-frame warm start (same probe, identical states)
    have the smallest geometric mean; alpha = 0 (the previous-frame warm start itself) when none is
    below 1.  Each probe advances `gw` by one step."""
    base = model.target_scale.copy()
    table, best = [], (1.0, 0.0)
    for a in alphas:
        model.set_normalisation(model.mean, model.inv_std, base * a)
        gw.set_warm_start_model(model)
        r = gw.step_probe(dt, iterations, PROBE_PREVIOUS | PROBE_LEARNED | PROBE_COLD)
        R = relative_residuals(r["residual"])
        ks = [k for k in (4, 8, 16) if k <= iterations]
        ratio = float(np.exp(np.mean([np.log(R[1, k] / R[0, k]) for k in ks])))
        table.append({"alpha": a, "ratio": ratio})
        log(f"# shrinkage {a}: learned / previous residual {ratio:.4f}")
        if ratio < best[0]:
            best = (ratio, a)
    model.set_normalisation(model.mean, model.inv_std, base * best[1])
    return best[1], table


def _normalised_targets(model: LearnedWarmStart, f: np.ndarray, t: np.ndarray) -> np.ndarray:
    sc = impulse_scale(f)
    d = (t - f[:, 16:19]).astype(np.float32)
    with np.errstate(divide="ignore", invalid="ignore"):
        y = d / (model.target_scale[None, :] * sc[:, None])
    return np.where(sc[:, None] > 0, y, 0.0)


def train(features: np.ndarray, targets: np.ndarray, seed: int = 0x4A2D0004, epochs: int = 24, batch: int = 8192,
          lr: float = 2e-3, log=print) -> tuple[LearnedWarmStart, dict]:
    rng = np.random.default_rng(seed)
    perm = rng.permutation(len(features))
    f, t = features[perm], targets[perm]
    n_train = int(0.9 * len(f))
    model = LearnedWarmStart(seed=seed)
    model.fit_normalisation(f[:n_train], t[:n_train])
    model.set_data(f, t)
    val0 = model.evaluate(n_train, len(f) - n_train)
    t0 = time.time()
    losses = []
    # three phases with decreasing step size
    for frac, scale in ((0.5, 1.0), (0.3, 0.3), (0.2, 0.1)):
        e = max(1, int(round(epochs * frac)))
        losses += model.train(0, n_train, epochs=e, batch=batch, lr=lr * scale).tolist()
        log(f"# train: epoch {len(losses)} loss {losses[-1]:.5f} val {model.evaluate(n_train, len(f) - n_train):.5f}")
    train_s = time.time() - t0
    val = model.evaluate(n_train, len(f) - n_train)
    # the previous-frame warm start predicts y = 0: its validation error is the mean squared target
    info = {"samples": len(f), "train": n_train, "val": len(f) - n_train, "val_mse_untrained": val0,
            "val_mse": val, "train_loss": losses, "train_seconds": train_s,
            "val_mse_previous": float(np.mean(_normalised_targets(model, f[n_train:], t[n_train:]) ** 2))}
    return model, info


def evaluate(gw: GpuWorld, dt: float, iterations: int = 128, tol: float = 1e-3) -> dict:
    """Probe step: R_k for previous / learned / cold warm starts on identical states."""
    r = gw.step_probe(dt, iterations, PROBE_PREVIOUS | PROBE_LEARNED | PROBE_COLD)
    R = relative_residuals(r["residual"])
    out = {"points": r["points"], "iterations": iterations}
    for m, name in ((0, "previous"), (1, "learned"), (2, "cold")):
        out[name] = {"iters_to_tol": iterations_to(R[m], tol),
                     "R": {str(k): float(R[m, k]) for k in (0, 4, 8, 16, 32, 64, iterations) if k <= iterations}}
    return out


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--samples", type=int, default=2_000_000)
    ap.add_argument("--bodies", type=int, default=131072)
    ap.add_argument("--first", type=int, default=240)
    ap.add_argument("--last", type=int, default=900)
    ap.add_argument("--every", type=int, default=20)
    ap.add_argument("--eval-warmup", type=int, default=700)
    ap.add_argument("--epochs", type=int, default=24)
    ap.add_argument("--out", default=DEFAULT_MODEL)
    args = ap.parse_args()
    t0 = time.time()
    f, t, gtrain, dt = gather(args.samples, args.bodies, 0x4A2D0004, args.first, args.last, args.every, 64)
    print(f"# gathered {len(f)} samples in {time.time() - t0:.1f}s", flush=True)
    model, info = train(f, t, epochs=args.epochs)
    print(f"# trained: val mse {info['val_mse']:.5f} (previous-frame warm start {info['val_mse_previous']:.5f}, "
          f"untrained {info['val_mse_untrained']:.5f}) in {info['train_seconds']:.1f}s", flush=True)
    alpha, _ = tune_shrinkage(model, gtrain, dt)
    gtrain.close()
    print(f"# shrinkage alpha = {alpha}", flush=True)
    os.makedirs(os.path.dirname(os.path.abspath(args.out)), exist_ok=True)
    model.save(args.out)
    gw, dt, _ = heap_world(args.bodies, 0x4A2D0002, args.eval_warmup)
    gw.set_warm_start_model(model)
    ev = evaluate(gw, dt, iterations=256)
    for name in ("previous", "learned", "cold"):
        print(f"# {name}: iters to 1e-3 {ev[name]['iters_to_tol']}  R {ev[name]['R']}", flush=True)
    return 0


if __name__ == "__main__":
    raise SystemExit(main())
