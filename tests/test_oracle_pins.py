"""Independent pins of oracle functions that the reference's golden data does not reach: the
point-cloud and triangle bounding boxes, RigidBody.Update (forces, torques, damping, gravity and
their one-step lag) and the contact cache's slot selection.  Each expectation comes from geometry or
kinematics computed here in float64 (not from the oracle's formula retyped), or from a reference
test mirrored case for case.  CPU only."""
import math

import numpy as np
import pytest

from jitterphysics2_amd import scenes
from jitterphysics2_amd._abi import BODY_DTYPE, SHAPE_DTYPE
from tests.orc import OracleWorld, lib
from tests.test_oracle import aabb, build_manifold, contact_op, mpr, quat_rotate, shape, two_body_state

rng = np.random.default_rng(20261204)


def random_quat():
    q = rng.normal(size=4)
    return q / np.linalg.norm(q)


def rot_matrix(q):
    return np.array([quat_rotate(q, e) for e in np.eye(3)]).T


# ---- PointCloudShape.CalculateBoundingBox (PointCloudShape.cs:140-154) -------------------------
def cloud_shape(points, shift):
    """Point cloud with a shift: GetCenter = mean (irrelevant here), cached box = support box along
    +-x/y/z including the shift (CalcInitBox, PointCloudShape.cs:158-182)."""
    lo, hi = points.min(0) + shift, points.max(0) + shift
    s = np.zeros(1, SHAPE_DTYPE)
    s["type"] = 3
    s["params"] = np.concatenate([points.mean(0) + shift, shift, lo, hi]).astype(np.float32)
    s["vertex_count"] = len(points)
    return s


def test_point_cloud_cached_box_is_the_axis_support_box():
    # CalcInitBox takes the six axis supports; pins the box the scene builder stores
    from tests.test_oracle import support
    pts = (rng.normal(size=(32, 3)) * [0.6, 0.3, 0.9]).astype(np.float32)
    shift = np.array([0.2, -0.1, 0.05], np.float32)
    s = cloud_shape(pts, shift)
    p = s["params"][0]
    for k in range(3):
        e = np.zeros(3, np.float32)
        e[k] = 1
        assert support(s, e, pts)[k] == p[9 + k]
        assert support(s, -e, pts)[k] == p[6 + k]


def test_point_cloud_aabb_is_the_rotated_cached_box():
    for _ in range(40):
        pts = (rng.normal(size=(32, 3)) * rng.uniform(0.2, 1.0, size=3)).astype(np.float32)
        shift = rng.uniform(-0.3, 0.3, size=3).astype(np.float32)
        s = cloud_shape(pts, shift)
        q, p = random_quat(), rng.normal(size=3) * 4
        mn, mx = aabb(s, q, p, )
        lo, hi = s["params"][0][6:9].astype(np.float64), s["params"][0][9:12].astype(np.float64)
        corners = np.array([[(hi if c else lo)[k] for k, c in enumerate(bits)] for bits in np.ndindex(2, 2, 2)])
        world = np.array([quat_rotate(q, c) for c in corners]) + p
        # the tightest box of the rotated cached box (its 8 corners) ...
        np.testing.assert_allclose(mx, world.max(0), atol=3e-5)
        np.testing.assert_allclose(mn, world.min(0), atol=3e-5)
        # ... which contains every rotated, shifted vertex
        v = np.array([quat_rotate(q, x) for x in pts.astype(np.float64) + shift]) + p
        assert np.all(v <= mx + 3e-5) and np.all(v >= mn - 3e-5)


def aabb_v(s, q, p, verts):
    from jitterphysics2_amd._abi import ptr
    out = np.zeros(6, np.float32)
    lib().orc_aabb(0, ptr(s), ptr(np.ascontiguousarray(verts, np.float32).ravel()),
                   ptr(np.asarray(q, np.float32)), ptr(np.asarray(p, np.float32)), ptr(out))
    return out[:3], out[3:]


def test_triangle_aabb_is_the_rotated_vertex_box_with_margin():
    # TriangleShape.CalculateBoundingBox (TriangleShape.cs:111-134): rotated vertices, -/+ 0.01
    s = np.zeros(1, SHAPE_DTYPE)
    s["type"] = 4
    s["vertex_count"] = 3
    for _ in range(40):
        tri = rng.normal(size=(3, 3)).astype(np.float32) * 3
        q, p = random_quat(), rng.normal(size=3) * 10
        mn, mx = aabb_v(s, q, p, tri)
        w = np.array([quat_rotate(q, v) for v in tri.astype(np.float64)]) + p
        np.testing.assert_allclose(mn, w.min(0) - 0.01, atol=3e-5)
        np.testing.assert_allclose(mx, w.max(0) + 0.01, atol=3e-5)
    # a degenerate (flat, axis-aligned) triangle still has a box of thickness 0.02
    tri = np.array([[0, 0, 0], [1, 0, 0], [0, 0, 1]], np.float32)
    mn, mx = aabb_v(s, (0, 0, 0, 1), (0, 0, 0), tri)
    assert abs((mx[1] - mn[1]) - 0.02) < 1e-7


# ---- RigidBody.Update (RigidBody.cs:716-764) against kinematics --------------------------------
def free_body(inv_mass=0.5, inv_inertia=2.0, gravity=True, lin_damp=1.0, ang_damp=1.0):
    """A sphere far from anything else (no pairs): only integration and RigidBody.Update act.  The
    ABI carries float32 inputs, so the cases use values exact in binary (dt = 1/64, g = -10) or read
    the float32 value back (damping multipliers)."""
    sb = scenes.SceneBuilder()
    b = sb.body((0.0, 50.0, 0.0))
    sb.sphere(b, 0.5)
    bodies, shapes, verts = sb.arrays()
    bodies["inverse_mass"] = inv_mass
    bodies["inverse_inertia"] = (inv_inertia, 0, 0, inv_inertia, 0, inv_inertia)
    bodies["affected_by_gravity"] = 1 if gravity else 0
    bodies["linear_damping_multiplier"] = lin_damp
    bodies["angular_damping_multiplier"] = ang_damp
    cfg = scenes.default_config()
    cfg["substep_count"] = 4
    cfg["gravity"] = (0.0, -10.0, 0.0)
    return bodies, shapes, verts, cfg


def test_free_fall_is_semi_implicit_euler():
    # with DeltaVelocity = g h from the previous step, NS substeps of v += g h, p += v h give
    # p = p0 + v0 t + g t^2 / 2 + g t h / 2 and v = v0 + g t  (t = N dt, h = dt / S)
    bodies, shapes, verts, cfg = free_body()
    dt, S, N = 1 / 64, 4, 90
    h = dt / S
    g = np.array([0, -10.0, 0])
    v0 = np.array([1.0, 2.0, -0.5])
    bodies["velocity"] = v0
    bodies["delta_velocity"] = g * h
    ow = OracleWorld(bodies, shapes, verts, cfg, fp64=True)
    for _ in range(N):
        ow.step(dt)
    st = ow.state64()[0]
    t = N * dt
    np.testing.assert_allclose(st[7:10], v0 + g * t, rtol=1e-9, atol=1e-9)
    np.testing.assert_allclose(st[0:3], np.array([0, 50, 0]) + v0 * t + g * t * t / 2 + g * t * h / 2,
                               rtol=1e-9, atol=1e-8)


def test_damping_is_geometric_per_step():
    # with no forces the velocity is multiplied once per step (after the substeps): v_N = v0 d^N and
    # the position advances by dt * v0 * (1 - d^N) / (1 - d)
    bodies, shapes, verts, cfg = free_body(gravity=False, lin_damp=0.9, ang_damp=0.8)
    d, da = float(np.float32(0.9)), float(np.float32(0.8))
    dt, N = 1 / 64, 25
    v0, w0 = np.array([3.0, -1.0, 2.0]), np.array([0.0, float(np.float32(0.4)), 0.0])
    bodies["velocity"], bodies["angular_velocity"] = v0, w0
    ow = OracleWorld(bodies, shapes, verts, cfg, fp64=True)
    for _ in range(N):
        ow.step(dt)
    st = ow.state64()[0]
    np.testing.assert_allclose(st[7:10], v0 * d ** N, rtol=1e-12)
    np.testing.assert_allclose(st[10:13], w0 * da ** N, rtol=1e-12)
    np.testing.assert_allclose(st[0:3], np.array([0, 50, 0]) + dt * v0 * (1 - d ** N) / (1 - d), rtol=1e-10)


def test_constant_spin_rotates_by_angle_times_time():
    # a constant angular velocity about a fixed axis composes to the exact rotation (cos, sin w-hat)
    bodies, shapes, verts, cfg = free_body(gravity=False)
    w = np.array([0.3, -1.2, 0.8], np.float32).astype(np.float64)
    bodies["angular_velocity"] = w
    ow = OracleWorld(bodies, shapes, verts, cfg, fp64=True)
    dt, N = 1 / 64, 120
    for _ in range(N):
        ow.step(dt)
    q = ow.state64()[0][3:7]
    th = np.linalg.norm(w) * N * dt
    exact = np.concatenate([math.sin(th / 2) * w / np.linalg.norm(w), [math.cos(th / 2)]])
    # StableMath's polynomials (StableMath.cs:69-106) in double, composed over 480 substeps
    np.testing.assert_allclose(q * np.sign(q[3]), exact * np.sign(exact[3]), atol=1e-12)


def test_force_and_torque_act_one_step_later_and_reset():
    # ForceImpulseTests.cs:112-140: a force given before step N becomes DeltaVelocity at the end of
    # step N (and is then cleared), so the velocity changes in step N + 1, by F m^-1 dt in total
    # (S substeps of F m^-1 h); a torque likewise through the world inverse inertia (a sphere: 2 I)
    bodies, shapes, verts, cfg = free_body(gravity=False, inv_mass=0.5, inv_inertia=2.0)
    ow = OracleWorld(bodies, shapes, verts, cfg, fp64=True)
    dt = 1 / 64
    ow.step(dt)  # warmup: no forces yet
    b = bodies.copy()  # the state read back, with the body's other properties (mass, damping ...)
    for f in ("position", "orientation", "velocity", "angular_velocity", "delta_velocity", "delta_angular_velocity"):
        b[f] = ow.bodies()[f]
    b["force"] = (100.0, 0.0, 0.0)
    b["force"] += (20.0, -10.0, 0.0)  # AddForce accumulates
    b["torque"] = (0.0, 0.0, 3.0)
    ow.set_bodies(b)
    ow.step(dt)
    after1 = ow.bodies()
    assert np.array_equal(after1["force"][0], [0, 0, 0]) and np.array_equal(after1["torque"][0], [0, 0, 0])
    assert np.array_equal(after1["velocity"][0], [0, 0, 0])  # not yet applied
    ow.step(dt)
    st = ow.state64()[0]
    np.testing.assert_allclose(st[7:10], np.array([120.0, -10.0, 0.0]) * 0.5 * dt, rtol=1e-12)
    np.testing.assert_allclose(st[10:13], np.array([0.0, 0.0, 3.0]) * 2.0 * dt, rtol=1e-12)
    ow.step(dt)  # the force was applied once only
    np.testing.assert_allclose(ow.state64()[0][7:10], np.array([120.0, -10.0, 0.0]) * 0.5 * dt, rtol=1e-12)


def test_static_and_unaffected_bodies_ignore_gravity():
    bodies, shapes, verts, cfg = free_body(gravity=False)
    ow = OracleWorld(bodies, shapes, verts, cfg, fp64=True)
    for _ in range(10):
        ow.step(1 / 64)
    np.testing.assert_allclose(ow.state64()[0][0:3], [0, 50, 0], atol=0)


# ---- Regression/PersistentContactManifoldSelectionTests.cs ------------------------------------
def rotated_face_scenario():
    """CreateRotatedFaceContactScenario (:5-27): two BoxShape(2, 2, 2) (half extents 1), B at
    (0, 1.9, 0) turned 45 degrees about y; NarrowPhase.Collision, then BuildManifold."""
    a, b = shape("box", 1, 1, 1), shape("box", 1, 1, 1)
    qa = (0, 0, 0, 1)
    qb = (0, math.sin(math.pi / 8), 0, math.cos(math.pi / 8))
    pa, pb = (0, 0, 0), (0, 1.9, 0)
    r = mpr(a, b, qa, pa, qb, pb)
    assert r["hit"]
    n, mA, mB = build_manifold(a, b, qa, pa, qb, pb, r["pa"], r["pb"], r["normal"])
    return n, mA, mB, r["normal"]


def register(mA, mB, normal):
    b, a = two_body_state(static_first=False)
    for i in range(len(mA)):
        contact_op(b, a, 0, p1=mA[i], p2=mB[i], n=normal)
    used = [(int(a["usage_mask"][0]) >> s) & 1 for s in range(4)]
    # body 1 sits at the origin: world contact position = RelativePosition1
    return [a["contacts"][0, s]["relative_position1"].astype(np.float64) for s in range(4) if used[s]]


def assert_same_point_set(expected, actual, eps_sq):
    assert len(actual) == len(expected)
    matched = [False] * len(actual)
    for e in expected:
        m = next((i for i, x in enumerate(actual) if not matched[i] and np.sum((x - e) ** 2) <= eps_sq), -1)
        assert m >= 0, e
        matched[m] = True


def test_same_four_contacts_whatever_the_manifold_order():  # :109-136
    n, mA, mB, normal = rotated_face_scenario()
    assert n == 4
    first = register(mA, mB, normal)
    rot = [(i + 2) % n for i in range(n)]  # RotateManifold(manifold, 2)
    second = register(mA[rot], mB[rot], normal)
    assert len(first) == 4 and len(second) == 4
    assert_same_point_set(first, second, 1e-8)


def test_already_reduced_manifold_is_kept():  # :138-161
    n, mA, mB, normal = rotated_face_scenario()
    assert n == 4
    assert_same_point_set([p.astype(np.float64) for p in mA], register(mA, mB, normal), 1e-8)


def test_full_cache_reuses_the_matching_contact():  # :163-196
    normal = np.array([0, 1, 0], np.float32)
    pts = np.array([[-1, 0, -1], [1, 0, -1], [1, 0, 1], [-1, 0, 1]], np.float32)
    b, a = two_body_state(static_first=False)
    for p in pts:
        contact_op(b, a, 0, p1=p, p2=p - 0.1 * normal, n=normal)
    original = [a["contacts"][0, s]["relative_position1"].astype(np.float64) for s in range(4)]
    contact_op(b, a, 0, p1=pts[1], p2=pts[1] - 0.1 * normal, n=normal)
    assert int(a["usage_mask"][0]) & 15 == 15
    updated = [a["contacts"][0, s]["relative_position1"].astype(np.float64) for s in range(4)]
    assert_same_point_set(original, updated, 1e-10)
