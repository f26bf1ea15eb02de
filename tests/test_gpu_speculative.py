"""GPU parity of speculative contacts (swept broadphase boxes, the sweep queue, k_sweep, one-point
contacts with SolveMode.Angular removed) against the oracle, bit for bit in FP32, replaying the GPU's
solve order in the oracle as the other step tests do."""
import numpy as np
import pytest

from jitterphysics2_amd import scenes
from jitterphysics2_amd.gpu import GpuWorld
from tests.orc import OracleWorld
from tests.test_gpu_step import assert_bodies_equal, run_replay
from tests.test_speculative_oracle import slab_scene

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("solver", ["default", "flow"])
def test_bullets_bitexact(monkeypatch, solver):
    # projectiles at 15-60 m/s against slabs 0.2 thick, every body with speculative contacts on
    if solver != "default":
        monkeypatch.setenv("JH_SOLVER", solver)
    gw, ow = run_replay(scenes.bullets(400, seed=21), steps=60, check_every=6)
    assert gw.stats()["arbiters"] > 0


def test_bullets_mixed_flags_and_mesh_filter_bitexact():
    # half of the bodies speculative; triangle-mesh floor with the TriangleEdgeCollisionFilter
    run_replay(scenes.bullets(300, seed=22, speculative=0.5, mesh=True), steps=45, check_every=5)


def test_speculative_manifold_flags_on_gpu():
    bodies, shapes, verts, cfg, dt = slab_scene(True)
    gw = GpuWorld(bodies, shapes, verts, cfg)
    ow = OracleWorld(bodies, shapes, verts, cfg)
    gw.step(dt)
    ow.step(dt)
    m = gw.manifolds()
    assert len(m) == 1 and m["flags"][0] == 2 and m["count"][0] == 1
    np.testing.assert_array_equal(m["point_a"], ow.manifolds()["point_a"])
    np.testing.assert_array_equal(m["penetration"], ow.manifolds()["penetration"])
    assert_bodies_equal(gw.bodies(), ow.bodies(), "slab step")


def test_speculative_tunnelling_prevented_on_gpu():
    for spec in (False, True):
        bodies, shapes, verts, cfg, dt = slab_scene(spec, speed=90.0)
        gw = GpuWorld(bodies, shapes, verts, cfg)
        ys = []
        for _ in range(6):
            gw.step(dt)
            ys.append(float(gw.bodies()["position"][1, 1]))
        assert (min(ys) > 0.1) if spec else (min(ys) < -0.5), (spec, ys)
