"""jh_world_stabilize (World.Stabilize, World.Step.cs:318-392) on the GPU: bit-exact against the oracle
replaying the GPU's colour order on a settled heap and on a jointed scene (velocities, contact and
constraint impulses; positions and orientations untouched), the reference's own outcome tests
(Api/WorldTests.cs: a constraint error is solved without moving the bodies; MotorTimeStepTests.cs:
the world's substep count is used) and its argument checks."""
import math

import numpy as np
import pytest

from jitterphysics2_amd import constraints as C
from jitterphysics2_amd import scenes
from jitterphysics2_amd._abi import default_config
from jitterphysics2_amd.gpu import GpuWorld
from tests.orc import OracleWorld
from tests.test_gpu_step import assert_arbiters_equal, assert_bodies_equal

pytestmark = pytest.mark.gpu


def with_state(bodies, state):
    out = bodies.copy()
    for f in ("position", "orientation", "velocity", "angular_velocity", "delta_velocity", "delta_angular_velocity"):
        out[f] = state[f]
    return out


@pytest.mark.parametrize("iterations,relaxations", [(8, 2), (1, 0)])
def test_stabilize_heap_bitexact(iterations, relaxations):
    bodies, shapes, verts, cfg, dt = scenes.heap(900, seed=5)
    gw = GpuWorld(bodies, shapes, verts, cfg)
    for _ in range(60):
        gw.step(dt)
    b0, a0 = gw.bodies(), gw.arbiters()
    assert len(a0) > 100
    gw.stabilize(dt, iterations, relaxations)
    keys, _ = gw.solve_order()
    ow = OracleWorld(with_state(bodies, b0), shapes, verts, cfg)
    ow.set_arbiters(a0)
    assert ow.stabilize(dt, iterations, relaxations, keys) == 0
    b1 = gw.bodies()
    assert_bodies_equal(b1, ow.bodies(), "after stabilize")
    assert_arbiters_equal(gw.arbiters(), ow.arbiters(), "after stabilize")
    assert np.array_equal(b1["position"], b0["position"]) and np.array_equal(b1["orientation"], b0["orientation"])
    assert not np.array_equal(b1["velocity"], b0["velocity"])
    # the world steps on from the stabilized state exactly as the oracle does
    gw.step(dt)
    ow.step(dt, gw.solve_order()[0])
    assert_bodies_equal(gw.bodies(), ow.bodies(), "step after stabilize")


def test_stabilize_joints_bitexact():
    bodies, shapes, verts, cfg, dt, cons = scenes.joints(24, seed=41)
    gw = GpuWorld(bodies, shapes, verts, cfg)
    gw.set_constraints(cons)
    for _ in range(30):
        gw.step(dt)
    b0, a0, c0 = gw.bodies(), gw.arbiters(), gw.constraints()
    gw.stabilize(dt, 6, 2)
    ow = OracleWorld(with_state(bodies, b0), shapes, verts, cfg)
    ow.set_arbiters(a0)
    ow.set_constraints(c0)
    ow.set_constraint_order(gw.constraint_order()[0])
    assert ow.stabilize(dt, 6, 2, gw.solve_order()[0]) == 0
    assert_bodies_equal(gw.bodies(), ow.bodies(), "joints after stabilize")
    assert_arbiters_equal(gw.arbiters(), ow.arbiters(), "joints after stabilize")
    assert np.array_equal(gw.constraints()["state"], ow.constraints()["state"])
    assert np.array_equal(gw.bodies()["position"], b0["position"])


def test_stabilize_constraint_error_without_moving():  # WorldTests.cs:258-294
    sb = scenes.SceneBuilder()
    sb.sphere(sb.body((0.0, 0.0, 0.0)), 1.0)
    sb.sphere(sb.body((0.0, 0.0, 0.0)), 1.0)
    bodies, shapes, verts = sb.arrays()
    cfg = default_config()
    cfg["gravity"] = (0.0, 0.0, 0.0)
    socket = C.ball_socket(bodies, 0, 1, (0.0, 0.0, 0.0))
    bodies["position"][1] = (1.0, 0.0, 0.0)
    gw = GpuWorld(bodies, shapes, verts, cfg)
    gw.set_constraints(C.as_array([socket]))
    gw.stabilize(1 / 60, 4, 2)
    b = gw.bodies()
    assert np.array_equal(b["position"], bodies["position"])
    assert np.sum(b["velocity"].astype(np.float64) ** 2) > 0.0
    assert np.sum(gw.constraints()["state"][0][0:3].astype(np.float64) ** 2) > 0.0


def test_stabilize_uses_the_world_substep_count():  # MotorTimeStepTests.cs:51-65
    sb = scenes.SceneBuilder()
    nb = sb.body((0.0, 0.0, 0.0), motion=scenes.MOTION_STATIC)
    sb.bodies[nb]["inverse_mass"] = 0.0
    sb.bodies[nb]["inverse_inertia"] = (0, 0, 0, 0, 0, 0)
    body = sb.body((0.0, 0.0, 0.0))
    sb.sphere(body, 0.5)
    bodies, shapes, verts = sb.arrays()
    bodies["inverse_mass"][body] = 1.0
    bodies["inverse_inertia"][body] = (1, 0, 0, 1, 0, 1)
    cfg = default_config()
    cfg["gravity"] = (0.0, 0.0, 0.0)
    cfg["substep_count"] = 4
    gw = GpuWorld(bodies, shapes, verts, cfg)
    gw.set_constraints(C.as_array([C.angular_motor(bodies, nb, body, (1.0, 0.0, 0.0), velocity=1000.0,
                                                   max_torque=2.0)]))
    gw.stabilize(0.1, 2)
    b = gw.bodies()[body]
    assert abs(b["angular_velocity"][0] - 0.2) < 1e-6
    assert np.array_equal(b["orientation"], [0, 0, 0, 1])


def test_stabilize_argument_checks():  # WorldTests.cs:184-232
    bodies, shapes, verts, cfg, dt = scenes.pyramid(rows=3)
    gw = GpuWorld(bodies, shapes, verts, cfg)
    st = lambda d, i, r: int(gw.lib.jh_world_stabilize(gw.handle, d, i, r))
    assert st(-1 / 60, 1, 0) == -1
    assert st(0.0, 1, 0) == 0
    for bad in (math.nan, math.inf, -math.inf):
        assert st(bad, 1, 0) == -2
    assert st(0.0, 0, 0) == -1 and st(0.0, 1, -1) == -1
    assert st(1 / 60, 0, 0) == -1 and st(1 / 60, 1, -1) == -1
    assert int(gw.lib.jh_world_stabilize(None, 1 / 60, 1, 0)) == -6
    b0 = gw.bodies()
    gw.stabilize(1 / 60, 2, 0)  # before any step: no contacts held, nothing moves
    assert np.array_equal(gw.bodies()["velocity"], b0["velocity"])
