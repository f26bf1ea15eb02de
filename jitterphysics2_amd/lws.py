"""Learned warm start: host side of the per-contact MLP (jh_lws_* in include/jitterhip.h).

The network, its training (forward, backward, Adam) and its use inside the step run in the HIP
library on BF16 MFMA; this module moves arrays across the C-ABI, fits the feature normalisation,
collects training traces from probe steps and turns probe residuals into iteration counts.
The reference has no counterpart: it warm-starts every contact from the previous frame's
impulse (src/Jitter2/Dynamics/Contact.cs:802-817), which stays the fallback here.
"""
from __future__ import annotations

import ctypes

import numpy as np

from ._abi import LWS_FEATURES, LWS_PARAMS, PROBE_COLD, PROBE_LEARNED, PROBE_PREVIOUS, check, load_library, ptr

FEATURE_NAMES = (
    "normal.x", "normal.y", "normal.z", "tangent1.x", "tangent1.y", "tangent1.z",
    "tangent2.x", "tangent2.y", "tangent2.z", "v.normal", "v.tangent1", "v.tangent2",
    "penetration", "mass.normal", "mass.tangent1", "mass.tangent2",
    "previous.normal", "previous.tangent1", "previous.tangent2",
    "substep", "friction", "bias", "penalty_bias", "new_contact",
)
assert len(FEATURE_NAMES) == LWS_FEATURES


def impulse_scale(features: np.ndarray) -> np.ndarray:
    """Per-contact impulse scale of the network output (float32, as the device computes it):
    |MassNormal| * (|v.n| + |v.t1| + |v.t2| + |PenaltyBias| + 0.05); predicted impulse =
    previous + y * target_scale * impulse_scale."""
    f = np.asarray(features, np.float32)
    a = np.abs
    return a(f[:, 13]) * ((((a(f[:, 9]) + a(f[:, 10])) + a(f[:, 11])) + a(f[:, 22])) + np.float32(0.05))


class LearnedWarmStart:
    """One jh_lws model: parameters, normalisation and (optionally) resident training data."""

    def __init__(self, seed: int = 0x4A2D0004):
        self.lib = load_library()
        self.handle = ctypes.c_void_p()
        check(self.lib, self.lib.jh_lws_create(ctypes.c_uint64(seed), ctypes.byref(self.handle)), "jh_lws_create")
        self.mean = np.zeros(LWS_FEATURES, np.float32)
        self.inv_std = np.ones(LWS_FEATURES, np.float32)
        self.target_scale = np.ones(3, np.float32)

    def close(self) -> None:
        if getattr(self, "handle", None) and self.handle.value:
            self.lib.jh_lws_destroy(self.handle)
            self.handle = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # parameters and normalisation
    def params(self) -> np.ndarray:
        p = np.zeros(LWS_PARAMS, np.float32)
        check(self.lib, self.lib.jh_lws_get_params(self.handle, ptr(p), LWS_PARAMS), "jh_lws_get_params")
        return p

    def set_params(self, p: np.ndarray) -> None:
        p = np.ascontiguousarray(p, np.float32)
        check(self.lib, self.lib.jh_lws_set_params(self.handle, ptr(p), len(p)), "jh_lws_set_params")

    def set_normalisation(self, mean, inv_std, target_scale) -> None:
        self.mean = np.ascontiguousarray(mean, np.float32)
        self.inv_std = np.ascontiguousarray(inv_std, np.float32)
        self.target_scale = np.ascontiguousarray(target_scale, np.float32)
        check(self.lib, self.lib.jh_lws_set_normalisation(self.handle, ptr(self.mean), ptr(self.inv_std),
                                                          ptr(self.target_scale)), "jh_lws_set_normalisation")

    def fit_normalisation(self, features: np.ndarray, targets: np.ndarray) -> None:
        """Feature mean / inverse standard deviation (constant features get scale 1) and the output
        scale = standard deviation of (target - previous impulse) / impulse_scale over the contacts
        with a non-zero impulse scale, floored at 1e-6."""
        f = np.asarray(features, np.float64)
        sc = impulse_scale(features).astype(np.float64)
        ok = sc > 0
        d = (np.asarray(targets, np.float64) - f[:, 16:19])[ok] / sc[ok, None]
        sd = f.std(axis=0)
        inv = np.where(sd > 1e-12, 1.0 / np.maximum(sd, 1e-12), 1.0)
        ts = np.maximum(d.std(axis=0), 1e-6) if len(d) else np.ones(3)
        self.set_normalisation(f.mean(axis=0), inv, ts)

    # data and training
    def set_data(self, features: np.ndarray, targets: np.ndarray) -> None:
        f = np.ascontiguousarray(features, np.float32)
        t = np.ascontiguousarray(targets, np.float32)
        assert f.shape[1] == LWS_FEATURES and t.shape == (len(f), 3)
        check(self.lib, self.lib.jh_lws_set_data(self.handle, ptr(f), ptr(t), len(f)), "jh_lws_set_data")
        self.n_data = len(f)

    def train(self, begin: int, count: int, epochs: int, batch: int = 8192, lr: float = 1e-3) -> np.ndarray:
        loss = np.zeros(epochs, np.float64)
        check(self.lib, self.lib.jh_lws_train(self.handle, begin, count, epochs, batch, ctypes.c_float(lr), ptr(loss)),
              "jh_lws_train")
        return loss

    def evaluate(self, begin: int, count: int) -> float:
        m = ctypes.c_double(0)
        check(self.lib, self.lib.jh_lws_evaluate(self.handle, begin, count, ctypes.byref(m)), "jh_lws_evaluate")
        return m.value

    def gradient(self, begin: int, count: int) -> tuple[np.ndarray, float]:
        g = np.zeros(LWS_PARAMS, np.float32)
        loss = ctypes.c_double(0)
        check(self.lib, self.lib.jh_lws_gradient(self.handle, begin, count, ptr(g), ctypes.byref(loss)),
              "jh_lws_gradient")
        return g, loss.value

    def predict(self, features: np.ndarray) -> np.ndarray:
        """Network outputs y (n, 3); predicted impulse = features[:, 16:19] + y * target_scale *
        impulse_scale(features) (before the trust-region limit the world applies)."""
        f = np.ascontiguousarray(features, np.float32)
        out = np.zeros((len(f), 3), np.float32)
        check(self.lib, self.lib.jh_lws_predict(self.handle, ptr(f), len(f), ptr(out)), "jh_lws_predict")
        return out

    def save(self, path: str) -> None:
        np.savez(path, params=self.params(), mean=self.mean, inv_std=self.inv_std, target_scale=self.target_scale)

    @classmethod
    def load(cls, path: str) -> "LearnedWarmStart":
        d = np.load(path)  # plain arrays only (no pickle)
        m = cls()
        m.set_params(d["params"])
        m.set_normalisation(d["mean"], d["inv_std"], d["target_scale"])
        return m


def relative_residuals(residual: np.ndarray) -> np.ndarray:
    """R_k = ||e_k|| / ||e_0 (cold)|| per probe mode, from jh_probe residual sums of squares."""
    r = np.asarray(residual, np.float64)
    cold0 = r[2, 0]
    return np.sqrt(r / cold0) if cold0 > 0 else np.zeros_like(r)


def iterations_to(rel: np.ndarray, tol: float = 1e-3):
    """First k with R_k <= tol (None if the probe never got there)."""
    k = np.nonzero(np.asarray(rel) <= tol)[0]
    return int(k[0]) if len(k) else None


def collect_traces(gw, dt: float, steps: int, every: int, iterations: int = 64, capacity: int = 1 << 21):
    """Step `gw` for `steps` steps, probing every `every`-th step: features at the warm start of
    the first substep and the accumulated impulse after `iterations` previous-frame-warm-started
    iterations (the training targets).  Returns (features, targets)."""
    fs, ts = [], []
    for i in range(steps):
        if (i + 1) % every == 0:
            r = gw.step_probe(dt, iterations, PROBE_PREVIOUS, capacity)
            fs.append(r["features"])
            ts.append(r["targets"])
        else:
            gw.step(dt)
    if not fs:
        return np.zeros((0, LWS_FEATURES), np.float32), np.zeros((0, 3), np.float32)
    return np.concatenate(fs), np.concatenate(ts)


__all__ = ["FEATURE_NAMES", "LearnedWarmStart", "impulse_scale", "relative_residuals", "iterations_to", "collect_traces",
           "PROBE_PREVIOUS", "PROBE_LEARNED", "PROBE_COLD"]
