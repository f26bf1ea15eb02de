"""Headroom of the learned warm start, on the CPU oracle: warm-starting a settled heap's first substep
with the network's training target itself (the impulses after 64 previous-frame-warm-started
iterations — what a perfect network would predict) must reach a relative residual of 1e-2 in far
fewer iterations than the previous-frame warm start: the ceiling the trained network is measured
against (CPU only)."""
import numpy as np

from jitterphysics2_amd import scenes
from jitterphysics2_amd.lws import PROBE_COLD, PROBE_LEARNED, PROBE_PREVIOUS, iterations_to, relative_residuals
from tests import orc
from tests.orc import OracleWorld


def test_perfect_prediction_of_the_target_needs_far_fewer_iterations():
    orc.set_threads(8)
    try:
        b, s, v, cfg, dt = scenes.heap(2048)
        w1, w2 = OracleWorld(b, s, v, cfg), OracleWorld(b, s, v, cfg)
        for _ in range(240):
            w1.step(dt)
            w2.step(dt)
        a = w1.arbiters()
        pairs = w1.pairs().astype(np.uint64)
        keys = np.concatenate([pairs, np.stack([a["shape1"], a["shape2"]], 1).astype(np.uint64)])
        kk = np.unique(keys[:, 0] << np.uint64(32) | keys[:, 1])
        order = np.stack([(kk >> np.uint64(32)).astype(np.uint32), (kk & np.uint64(0xffffffff)).astype(np.uint32)], 1)
        cap = 4 * len(order) + 64
        r1 = w1.step_probe(dt, order, 64, PROBE_PREVIOUS | PROBE_COLD, capacity=cap)
        r2 = w2.step_probe(dt, order, 64, PROBE_PREVIOUS | PROBE_LEARNED | PROBE_COLD, capacity=cap,
                           learned=r1["targets"])
    finally:
        orc.set_threads(1)
    R1, R2 = relative_residuals(r1["residual"]), relative_residuals(r2["residual"])
    assert r1["points"] > 2000
    # both probes start from the same state: their previous-frame rows agree
    assert np.allclose(R1[0], R2[0], rtol=1e-12, atol=0)
    k_prev, k_ideal = iterations_to(R1[0], 1e-2), iterations_to(R2[1], 1e-2)
    d = np.abs(r1["targets"] - r1["features"][:, 16:19]).mean(0)
    lam = np.abs(r1["features"][:, 16]).mean()
    print(f"settled 2,048-body heap, {r1['points']} contact points: iterations to 1e-2 previous {k_prev}, "
          f"perfect prediction {k_ideal}; R_4 {R1[0, 4]:.4f} / {R2[1, 4]:.4f}; mean correction {d} against "
          f"mean normal impulse {lam:.4g}")
    assert k_prev is not None and k_ideal is not None
    assert 2 * k_ideal <= k_prev
    assert R2[1, 4] < 0.5 * R1[0, 4]
