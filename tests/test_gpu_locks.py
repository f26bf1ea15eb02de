"""Motion locks (RigidBody.AllowedMotion) on the GPU, bit-exact against the oracle replaying the GPU's
solve order: a stack of boxes moving in the XY plane only (PlaneXY), spheres that may not rotate,
capsules that may only move vertically and rotate freely, falling onto each other and a ground box,
under every solver kernel; a jointed chain with locked links (constraints with a directional inverse
mass and a restricted inverse inertia); a lock change through jh_world_update_bodies (cleared contact
cache and constraint warm start); RawData's lock bits."""
import math

import numpy as np
import pytest

from jitterphysics2_amd import constraints as C
from jitterphysics2_amd import scenes
from jitterphysics2_amd._abi import (LOCK_ANGULAR, LOCK_ANGULAR_Z, LOCK_LINEAR_X, LOCK_LINEAR_Y, LOCK_LINEAR_Z,
                                     MOTION_STATIC, default_config)
from jitterphysics2_amd.gpu import GpuWorld
from tests.orc import OracleWorld
from tests.test_gpu_step import assert_arbiters_equal, assert_bodies_equal

pytestmark = pytest.mark.gpu

PLANE_XY_LOCKS = 63 ^ (LOCK_LINEAR_X | LOCK_LINEAR_Y | LOCK_ANGULAR_Z)


def locked_scene(seed=11):
    rng = np.random.default_rng(seed)
    sb = scenes.SceneBuilder()
    g = sb.body((0.0, -0.5, 0.0), motion=MOTION_STATIC, friction=0.6)
    sb.box(g, (40.0, 1.0, 40.0))
    for k in range(12):
        a = float(rng.uniform(-0.3, 0.3))
        b = sb.body((float(rng.uniform(-3, 3)), 0.6 + 1.1 * k, 0.0), orientation=(0.0, 0.0, math.sin(a), math.cos(a)),
                    friction=0.6)
        sb.box(b, tuple(float(x) for x in rng.uniform(0.6, 1.4, 3)))
        sb.bodies[b]["locked_axes"] = PLANE_XY_LOCKS
    for k in range(6):
        b = sb.body((float(rng.uniform(-4, 4)), 2.0 + 1.5 * k, float(rng.uniform(-1, 1))),
                    velocity=(float(rng.uniform(-1, 1)), 0.0, float(rng.uniform(-1, 1))), friction=0.4)
        sb.sphere(b, float(rng.uniform(0.3, 0.6)))
        sb.bodies[b]["locked_axes"] = LOCK_ANGULAR
    for k in range(4):
        b = sb.body((float(rng.uniform(-2, 2)), 14.0 + 1.2 * k, float(rng.uniform(-0.5, 0.5))),
                    angular_velocity=tuple(float(x) for x in rng.uniform(-2, 2, 3)))
        sb.capsule(b, 0.3, 1.0)
        sb.bodies[b]["locked_axes"] = LOCK_LINEAR_X | LOCK_LINEAR_Z
    bodies, shapes, verts = sb.arrays()
    cfg = default_config()
    cfg["allow_deactivation"] = 0
    cfg["substep_count"] = 4
    cfg["solver_iterations"] = 8
    cfg["relaxation_iterations"] = 2
    return bodies, shapes, verts, cfg, 0.01


def lockstep(gw, ow, dt, steps, what, cons=False):
    for s in range(steps):
        gw.step(dt)
        ow.step(dt, gw.solve_order()[0])
        assert_bodies_equal(gw.bodies(), ow.bodies(), f"{what}, step {s}")
        assert_arbiters_equal(gw.arbiters(), ow.arbiters(), f"{what}, step {s}")
        if cons:
            assert np.array_equal(gw.constraints()["state"], ow.constraints()["state"]), f"{what}, step {s}"


@pytest.mark.parametrize("solver", ["world", "flow", "barrier", "lanes", "tiles"])
def test_locked_bodies_bitexact(monkeypatch, solver):
    monkeypatch.setenv("JH_SOLVER", solver)
    bodies, shapes, verts, cfg, dt = locked_scene()
    gw = GpuWorld(bodies, shapes, verts, cfg)
    ow = OracleWorld(bodies, shapes, verts, cfg)
    lockstep(gw, ow, dt, 150, solver)
    st = gw.bodies()
    planar = bodies["locked_axes"] == PLANE_XY_LOCKS
    assert np.array_equal(st["position"][planar, 2], bodies["position"][planar, 2])  # z never moves
    q = st["orientation"][planar]
    assert not q[:, 0].any() and not q[:, 1].any()  # rotation about z only
    no_spin = bodies["locked_axes"] == LOCK_ANGULAR
    assert np.array_equal(st["orientation"][no_spin], bodies["orientation"][no_spin])
    assert len(gw.arbiters()) > 10


def test_locked_chain_and_lock_change_bitexact():
    sb = scenes.SceneBuilder()
    anchor = sb.body((0.0, 8.0, 0.0), motion=MOTION_STATIC)
    links = [anchor]
    for k in range(5):
        b = sb.body((0.6 * (k + 1), 8.0, 0.0))
        sb.sphere(b, 0.25)
        links.append(b)
    g = sb.body((0.0, -0.5, 0.0), motion=MOTION_STATIC)
    sb.box(g, (20.0, 1.0, 20.0))
    bodies, shapes, verts = sb.arrays()
    bodies["locked_axes"][links[2]] = LOCK_LINEAR_Z | LOCK_ANGULAR
    bodies["locked_axes"][links[4]] = LOCK_LINEAR_X
    cons = []
    for a, b in zip(links[:-1], links[1:]):
        mid = 0.5 * (bodies["position"][a].astype(np.float32) + bodies["position"][b].astype(np.float32))
        cons.append(C.ball_socket(bodies, a, b, tuple(mid)))
    cons.append(C.hinge_angle(bodies, links[1], links[2], (0.0, 0.0, 1.0)))
    cfg = default_config()
    cfg["allow_deactivation"] = 0
    cfg["substep_count"] = 4
    gw = GpuWorld(bodies, shapes, verts, cfg)
    ow = OracleWorld(bodies, shapes, verts, cfg)
    arr = C.as_array(cons)
    gw.set_constraints(arr)
    ow.set_constraints(arr)
    ow.set_constraint_order(gw.constraint_order()[0])
    lockstep(gw, ow, 0.01, 60, "locked chain", cons=True)
    # unlock one link and lock another: contact cache and constraint warm start of both are reset
    rec = bodies[[links[2], links[3]]].copy()
    st = gw.bodies()
    for f in ("position", "orientation", "velocity", "angular_velocity", "delta_velocity", "delta_angular_velocity"):
        rec[f] = st[f][[links[2], links[3]]]
    rec["locked_axes"] = (0, LOCK_LINEAR_Y)
    gw.update_bodies([links[2], links[3]], rec)
    ow.update_bodies([links[2], links[3]], rec)
    assert_bodies_equal(gw.bodies(), ow.bodies(), "after the lock change")
    assert np.array_equal(gw.constraints()["state"], ow.constraints()["state"])
    lockstep(gw, ow, 0.01, 60, "after the lock change", cons=True)


def test_rigidbodydata_lock_bits():
    bodies, shapes, verts, cfg, dt = locked_scene()
    gw = GpuWorld(bodies, shapes, verts, cfg)
    gw.step(dt)
    rec = gw.rigidbodydata(128)
    flags = np.ascontiguousarray(rec[:, 120:124]).view(np.int32)[:, 0]
    assert np.array_equal((flags >> 4) & 63, bodies["locked_axes"])
    gw.set_rigidbodydata(rec)  # the same locks: accepted
    i = int(np.nonzero(bodies["locked_axes"] == LOCK_ANGULAR)[0][0])
    r = rec.copy()
    r[i, 120:124] = np.array([flags[i] ^ (LOCK_LINEAR_Y << 4)], np.int32).view(np.uint8)
    assert int(gw.lib.jh_world_set_rigidbodydata(gw.handle, r.ctypes.data, len(r), 128)) == -5
