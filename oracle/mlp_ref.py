"""CPU oracle of the learned-warm-start network (test infrastructure only).

The reference has no network; this restates the device MLP of jitterphysics2_amd/csrc/jh_lws.h
in float64 numpy: 24 -> 128 -> 128 -> 3 with exact GELU (x * Phi(x)), mean squared error over the
three outputs.  With ``quantise=True`` the device's BF16 rounding points are reproduced (round to
nearest even from float32): inputs, weights, hidden activations H1 and H2, and the backward
operands dY, dZ2, dZ1 that feed the BF16 matrix products; bias gradients are summed from the
unrounded dY, dZ2, dZ1 as on the device.  With ``quantise=False`` it is the exact float64 network,
pinned against central finite differences in tests/test_lws_oracle.py.  infer_warm_start restates
what the in-step inference (lws::k_infer) does with the network output: trust clamp, impulse scale,
normal impulse >= 0, non-finite fallback, SkipWarmStart rows untouched.
"""
from __future__ import annotations

import numpy as np
from scipy.special import erf

IN, HID, OUT = 24, 128, 3
SHAPES = (("W1", (HID, IN)), ("b1", (HID,)), ("W2", (HID, HID)), ("b2", (HID,)), ("W3", (OUT, HID)), ("b3", (OUT,)))
N_PARAMS = sum(int(np.prod(s)) for _, s in SHAPES)


def bf16(x) -> np.ndarray:
    """float32 -> bfloat16 -> float64, round to nearest even (NaN -> quiet NaN)."""
    u = np.asarray(x, np.float32).view(np.uint32).astype(np.uint64)
    nan = (u & 0x7FFFFFFF) > 0x7F800000
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    r = np.where(nan, np.uint64(0x7FC00000), r).astype(np.uint32)
    return r.view(np.float32).astype(np.float64)


def unpack(p: np.ndarray) -> dict:
    out, o = {}, 0
    for name, shape in SHAPES:
        n = int(np.prod(shape))
        out[name] = np.asarray(p[o:o + n], np.float64).reshape(shape)
        o += n
    return out


def pack(d: dict) -> np.ndarray:
    return np.concatenate([np.asarray(d[name], np.float64).reshape(-1) for name, _ in SHAPES])


def gelu(x):
    return 0.5 * x * (1.0 + erf(x / np.sqrt(2.0)))


def gelu_grad(x):
    return 0.5 * (1.0 + erf(x / np.sqrt(2.0))) + x * np.exp(-0.5 * x * x) / np.sqrt(2.0 * np.pi)


def normalise(features, mean, inv_std) -> np.ndarray:
    """Device normalisation (f - mean) * inv_std in float32."""
    f = np.asarray(features, np.float32)
    return (f - np.asarray(mean, np.float32)) * np.asarray(inv_std, np.float32)


def forward(p: np.ndarray, x: np.ndarray, quantise: bool = True) -> dict:
    q = bf16 if quantise else (lambda a: np.asarray(a, np.float64))
    P = unpack(p)
    X = q(x)
    W1, W2, W3 = q(P["W1"]), q(P["W2"]), q(P["W3"])
    Z1 = X @ W1.T + P["b1"]
    H1 = q(gelu(Z1))
    Z2 = H1 @ W2.T + P["b2"]
    H2 = q(gelu(Z2))
    Y = H2 @ W3.T + P["b3"]
    return {"X": X, "W1": W1, "W2": W2, "W3": W3, "Z1": Z1, "H1": H1, "Z2": Z2, "H2": H2, "Y": Y}


def loss_and_grad(p: np.ndarray, x: np.ndarray, t: np.ndarray, quantise: bool = True):
    """Mean squared error over n x 3 outputs and its gradient (flat, parameter order of SHAPES)."""
    q = bf16 if quantise else (lambda a: np.asarray(a, np.float64))
    F = forward(p, x, quantise)
    n = len(x)
    E = F["Y"] - np.asarray(t, np.float64)
    loss = float(np.mean(E * E))
    dY = 2.0 * E / (3.0 * n)
    dYq = q(dY)
    dZ2 = (dYq @ F["W3"]) * gelu_grad(F["Z2"])
    dZ2q = q(dZ2)
    dZ1 = (dZ2q @ F["W2"]) * gelu_grad(F["Z1"])
    dZ1q = q(dZ1)
    g = {"W3": dYq.T @ F["H2"], "b3": dY.sum(0), "W2": dZ2q.T @ F["H1"], "b2": dZ2.sum(0),
         "W1": dZ1q.T @ F["X"], "b1": dZ1.sum(0)}
    return loss, pack(g)


def init_params(seed: int) -> np.ndarray:
    """Same seeded initialisation as jh_lws_create: SplitMix64 stream, u = (z >> 11) * 2^-53,
    value = (2u - 1) / sqrt(fan_in), layer by layer (weights then bias)."""
    state = [seed & 0xFFFFFFFFFFFFFFFF]
    M = 0xFFFFFFFFFFFFFFFF

    def nxt():
        state[0] = (state[0] + 0x9E3779B97F4A7C15) & M
        z = state[0]
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & M
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & M
        return z ^ (z >> 31)

    out = []
    for count, fan in ((HID * IN + HID, IN), (HID * HID + HID, HID), (OUT * HID + OUT, HID)):
        b = np.float32(1.0) / np.sqrt(np.float32(fan))
        for _ in range(count):
            u = (nxt() >> 11) * (1.0 / 9007199254740992.0)
            out.append(np.float32(np.float64(2.0 * u - 1.0) * np.float64(b)))
    return np.array(out, np.float32)


TRUST = 3.0  # lws::kTrust


def impulse_scale(features) -> np.ndarray:
    """lws::impulse_scale (jh_lws.h) in float32, left to right:
    |MassNormal| * (((|v.n| + |v.t1|) + |v.t2|) + |PenaltyBias|) + 0.05)."""
    f = np.asarray(features, np.float32)
    a = np.abs
    return a(f[:, 13]) * ((((a(f[:, 9]) + a(f[:, 10])) + a(f[:, 11])) + a(f[:, 22])) + np.float32(0.05))


def infer_warm_start(p, mean, inv_std, tscale, features, flags, quantise=True) -> np.ndarray:
    """The warm start lws::k_infer (jh_lws.h) writes for contact rows with raw features [n][24] and
    contact flags [n] (bit 2 SkipWarmStart): rows that skip the warm start keep the previous impulse
    (features 16..18); the others take previous + clamp(y, -3, 3) * (tscale * impulse_scale) in float32,
    with the clamp as fminf(fmaxf(y, -3), 3) (a NaN output clamps to -3), the normal impulse clamped
    at 0, and the previous impulse kept when any component is non-finite."""
    f = np.asarray(features, np.float32)
    prev = f[:, 16:19].copy()
    use = (np.asarray(flags, np.uint32) & 4) == 0
    y = forward(p, normalise(f, mean, inv_std), quantise)["Y"].astype(np.float32)
    y = np.fmin(np.fmax(y, np.float32(-TRUST)), np.float32(TRUST))
    with np.errstate(over="ignore", invalid="ignore"):
        sc = impulse_scale(f)
        step = y * (np.asarray(tscale, np.float32)[None, :] * sc[:, None])
        pred = prev + step
    ok = use & np.all(np.isfinite(pred), axis=1)
    out = prev.copy()
    out[ok, 0] = np.maximum(pred[ok, 0], np.float32(0))
    out[ok, 1:] = pred[ok, 1:]
    return out
