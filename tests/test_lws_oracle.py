"""CPU pins of the learned-warm-start oracle: the float64 MLP against finite differences, BF16
rounding known answers, and the probe (features, residual, trace targets) on physical closed forms."""
import numpy as np
import pytest

from jitterphysics2_amd import scenes
from oracle import mlp_ref
from tests.orc import OracleWorld


def test_bf16_round_to_nearest_even():
    x = np.array([1.0, 1.0 + 2.0 ** -8, 1.0 + 3 * 2.0 ** -9, 1.0 + 2.0 ** -7 + 2.0 ** -8, -0.0, 65504.0, 1e-40],
                 np.float32)
    got = mlp_ref.bf16(x)
    # halfway cases go to the even mantissa; 1 + 2^-7 + 2^-8 is halfway between 1 + 2^-7 (odd) and 1 + 2^-6
    want = np.array([1.0, 1.0, 1.0 + 2.0 ** -7, 1.0 + 2.0 ** -6, -0.0, 65536.0, np.float32(1e-40)], np.float64)
    assert got[:6].tolist() == want[:6].tolist()
    assert np.signbit(got[4])
    assert got[6] == 2.0 ** -133  # subnormal spacing of bfloat16 is 2^-133; 1e-40 is nearest to one step
    assert np.isnan(mlp_ref.bf16(np.float32([np.nan])))[0]


def test_mlp_gradient_finite_differences():
    # exact float64 network: analytic gradient == central differences
    rng = np.random.default_rng(1)
    p = mlp_ref.init_params(7).astype(np.float64)
    x = rng.normal(size=(6, 24))
    t = rng.normal(size=(6, 3))
    loss, g = mlp_ref.loss_and_grad(p, x, t, quantise=False)
    idx = np.concatenate([rng.choice(range(a, b), 8, replace=False) for a, b in
                          ((0, 3072), (3072, 3200), (3200, 19584), (19584, 19712), (19712, 20096), (20096, 20099))
                          if b - a >= 8] + [np.arange(20096, 20099)])
    h = 1e-6
    for i in idx:
        pp, pm = p.copy(), p.copy()
        pp[i] += h
        pm[i] -= h
        fd = (mlp_ref.loss_and_grad(pp, x, t, False)[0] - mlp_ref.loss_and_grad(pm, x, t, False)[0]) / (2 * h)
        assert abs(fd - g[i]) <= 1e-6 * max(1.0, abs(fd)), (i, fd, g[i])


def test_mlp_quantised_close_to_exact():
    rng = np.random.default_rng(2)
    p = mlp_ref.init_params(3)
    x = rng.normal(size=(256, 24))
    ye = mlp_ref.forward(p, x, quantise=False)["Y"]
    yq = mlp_ref.forward(p, x, quantise=True)["Y"]
    assert np.abs(ye - yq).max() < 0.05 * np.abs(ye).max()


def test_init_params_match_layout():
    p = mlp_ref.init_params(0x4A2D0004)
    assert p.dtype == np.float32 and len(p) == 20099
    P = mlp_ref.unpack(p)
    assert np.abs(P["W1"]).max() <= 1 / np.sqrt(24) and np.abs(P["W2"]).max() <= 1 / np.sqrt(128)


def _resting_box(steps=150):
    sb = scenes.SceneBuilder()
    g = sb.body((0.0, -0.5, 0.0), motion=scenes.MOTION_STATIC, friction=0.6)
    sb.box(g, (20.0, 1.0, 20.0))
    sb.bodies[g]["inverse_mass"] = 0.0
    sb.bodies[g]["inverse_inertia"] = (0, 0, 0, 0, 0, 0)
    b = sb.body((0.0, 0.5, 0.0), friction=0.6)
    sb.box(b, (1.0, 1.0, 1.0))
    bodies, shapes, verts = sb.arrays()
    cfg = scenes.default_config()
    cfg["substep_count"] = 4
    cfg["solver_iterations"] = 8
    cfg["allow_deactivation"] = 0  # the box keeps being solved
    ow = OracleWorld(bodies, shapes, verts, cfg)
    for _ in range(steps):
        ow.step(0.01)
    return ow


def test_probe_resting_box_targets_and_residual():
    ow = _resting_box()
    keys = np.array([[0, 1]], np.uint32)
    r = ow.step_probe(0.01, keys, 64, 1 | 4, capacity=16)
    assert r["points"] == 4 and len(r["features"]) == 4
    f, t = r["features"], r["targets"]
    # the ground normal, friction of the pair, no penetration beyond the slop
    assert np.allclose(np.abs(f[:, 1]), 1.0, atol=1e-5) and np.allclose(f[:, 20], 0.6)
    # converged normal impulses carry the weight over one substep: sum = m g dt / substeps
    assert abs(t[:, 0].sum() - 9.81 * 0.01 / 4) < 0.01 * 9.81 * 0.01 / 4
    res = r["residual"]
    R = np.sqrt(res / res[2, 0])
    assert R[2, 0] == 1.0
    assert R[0, 64] < 1e-3 and R[2, 64] < 1e-3  # both warm starts converge on a single resting box
    assert R[0, 0] < R[2, 0]                     # the previous-frame impulse is a better start than zero


def test_probe_does_not_change_the_step():
    bodies, shapes, verts, cfg, dt = scenes.pyramid(rows=4)
    a, b = OracleWorld(bodies, shapes, verts, cfg), OracleWorld(bodies, shapes, verts, cfg)
    for _ in range(20):
        a.step(dt)
        b.step(dt)
    keys = np.array(sorted(map(tuple, a.pairs())), np.uint32)
    a.step(dt, keys)
    b.step_probe(dt, keys, 8, 5, capacity=100)
    for f in ("position", "orientation", "velocity", "angular_velocity"):
        assert np.array_equal(a.bodies()[f], b.bodies()[f])


# ---- infer_warm_start: the in-step inference's post-processing (lws::k_infer) ------------------
def _rows(n=200, seed=3):
    rng = np.random.default_rng(seed)
    f = rng.normal(size=(n, 24)).astype(np.float32)
    f[:, 13] = np.abs(f[:, 13]) + 0.1  # effective normal masses
    f[:, 16] = np.abs(f[:, 16])        # previous normal impulses
    return f, np.zeros(n, np.uint32)


def _const_net(b3):
    P = mlp_ref.unpack(mlp_ref.init_params(5))
    P["W3"][:] = 0.0
    P["b3"][:] = b3
    return mlp_ref.pack(P).astype(np.float32)


def test_infer_zero_output_keeps_previous_impulse():
    f, flags = _rows()
    out = mlp_ref.infer_warm_start(_const_net([0, 0, 0]), np.zeros(24), np.ones(24), np.ones(3), f, flags)
    assert np.array_equal(out, f[:, 16:19])


def test_infer_clamps_to_trust_region_and_normal_impulse():
    # y = (10, -10, 0.5) clamps to (3, -3, 0.5); the normal impulse is not allowed below zero
    f, flags = _rows()
    ts = np.array([0.2, 0.1, 0.3], np.float32)
    out = mlp_ref.infer_warm_start(_const_net([10, -10, 0.5]), np.zeros(24), np.ones(24), ts, f, flags)
    a = np.abs(f.astype(np.float64))
    sc = a[:, 13] * (a[:, 9] + a[:, 10] + a[:, 11] + a[:, 22] + 0.05)
    np.testing.assert_allclose(out[:, 0], np.maximum(f[:, 16] + 3 * 0.2 * sc, 0), rtol=1e-5)
    np.testing.assert_allclose(out[:, 1], f[:, 17] - 3 * 0.1 * sc, rtol=1e-5, atol=1e-6)
    np.testing.assert_allclose(out[:, 2], f[:, 18] + 0.5 * 0.3 * sc, rtol=1e-5, atol=1e-6)
    out = mlp_ref.infer_warm_start(_const_net([-10, 0, 0]), np.zeros(24), np.ones(24), ts, f, flags)
    assert np.all(out[:, 0] >= 0) and np.any(out[:, 0] == 0)


def test_infer_skip_rows_and_overflow_keep_previous():
    f, flags = _rows()
    flags[::3] = 4  # SkipWarmStart
    out = mlp_ref.infer_warm_start(_const_net([1, 1, 1]), np.zeros(24), np.ones(24), np.ones(3), f, flags)
    assert np.array_equal(out[::3], f[::3, 16:19]) and not np.array_equal(out[1::3], f[1::3, 16:19])
    # target scale 3e38: the step overflows to inf wherever the impulse scale exceeds ~1.13
    big = np.full(3, 3e38, np.float32)
    out = mlp_ref.infer_warm_start(_const_net([1, 1, 1]), np.zeros(24), np.ones(24), big, f, np.zeros_like(flags))
    sc = mlp_ref.impulse_scale(f)
    over = sc * np.float32(3e38) > np.finfo(np.float32).max
    assert over.any() and (~over).any()
    assert np.array_equal(out[over], f[over, 16:19]) and np.all(np.isfinite(out))
