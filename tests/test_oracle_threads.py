"""The oracle's parallel parts (narrowphase per pair, shape boxes, per-body sweeps, contact refresh)
write only their own records and the contact cache is filled in pair order afterwards, so the oracle's
results are bit-identical for any thread count (CPU)."""
import numpy as np
import pytest

from jitterphysics2_amd import scenes
from tests import orc
from tests.orc import OracleWorld


def _run(scene, steps, threads):
    orc.set_threads(threads)
    try:
        bodies, shapes, verts, cfg, dt = scene
        ow = OracleWorld(bodies, shapes, verts, cfg)
        for _ in range(steps):
            ow.step(dt)
        return ow.bodies(), ow.arbiters(), ow.manifolds(), ow.epa_calls()
    finally:
        orc.set_threads(1)


@pytest.mark.parametrize("scene,steps", [
    (lambda: scenes.heap(5000, seed=4), 40),                        # > 4096 bodies: parallel sweeps too
    (lambda: scenes.full_form(n_bodies=240, cells=32, worlds=4, feature=8.0, pitch=2.4), 40),  # edge filter
    (lambda: scenes.bullets(300, arena=30.0), 30),                  # speculative sweeps
])
def test_thread_count_does_not_change_results(scene, steps):
    sc = scene()
    b1, a1, m1, e1 = _run(sc, steps, 1)
    b4, a4, m4, e4 = _run(sc, steps, 4)
    for f in ("position", "orientation", "velocity", "angular_velocity"):
        assert np.array_equal(b1[f], b4[f]), f
    assert np.array_equal(a1, a4) and np.array_equal(m1, m4) and e1 == e4
    assert len(a1) > 0
