"""Pins the oracle (CPU restatement) against the reference's own known-answer tests and against what
the mathematics fixes: closed forms, invariants and brute force.  CPU only.

Each test cites the reference test or formula it follows.  A plausible restatement mistake (a dropped
term, a wrong sign, a swapped operand) makes at least one of these fail.
"""
import ctypes
import math
from fractions import Fraction

import numpy as np
import pytest

from jitterphysics2_amd import scenes
from jitterphysics2_amd._abi import ARBITER_DTYPE, BODY_DTYPE, SHAPE_DTYPE, ptr
from tests.orc import OracleWorld, lib

f32 = np.float32


def shape(kind, *params, verts_count=0, voff=0):
    s = np.zeros(1, SHAPE_DTYPE)
    s["type"] = {"box": 0, "sphere": 1, "capsule": 2, "cloud": 3}[kind]
    p = np.zeros(12, np.float32)
    p[:len(params)] = params
    s["params"] = p
    s["vertex_offset"] = voff
    s["vertex_count"] = verts_count
    return s


def support(s, d, verts=None, fp64=0):
    out = np.zeros(3, np.float32)
    v = None if verts is None else np.ascontiguousarray(verts, np.float32).ravel()
    lib().orc_support(fp64, ptr(s), ptr(v), ptr(np.asarray(d, np.float32)), ptr(out))
    return out


def mpr(sa, sb, qa, pa, qb, pb, verts=None, fp64=0, threshold=0.02):
    out = np.zeros(36, np.float32)
    v = None if verts is None else np.ascontiguousarray(verts, np.float32).ravel()
    arr = lambda x: np.ascontiguousarray(x, np.float32)
    hit = lib().orc_mpr_manifold(fp64, ptr(sa), ptr(sb), ptr(v), ptr(arr(qa)), ptr(arr(pa)), ptr(arr(qb)), ptr(arr(pb)),
                                 ctypes.c_float(threshold), ptr(out))
    n = int(out[11])
    return {"hit": bool(hit), "pa": out[0:3], "pb": out[3:6], "normal": out[6:9], "depth": float(out[9]),
            "epa": bool(out[10]), "count": n, "mA": out[12:12 + 3 * n].reshape(-1, 3),
            "mB": out[24:24 + 3 * n].reshape(-1, 3)}


def quat_rotate(q, v):
    x, y, z, w = (float(c) for c in q)
    qv = np.array([x, y, z])
    t = 2.0 * np.cross(qv, v)
    return np.asarray(v, float) + w * t + np.cross(qv, t)


# ---- support maps: Api/SupportMapTests.cs ---------------------------------------------------
def test_support_sphere_along_normalized_direction():  # SupportMapTests.cs:6-14
    assert np.array_equal(support(shape("sphere", 2.0), (3, 0, 0)), [2, 0, 0])


def test_support_box_corner_by_sign():  # SupportMapTests.cs:16-24
    assert np.array_equal(support(shape("box", 1, 2, 3), (-4, 5, -6)), [-1, 2, -3])
    # SignBit maps -0 to -1 (MathHelper.cs:28-32)
    assert np.array_equal(support(shape("box", 1, 2, 3), (-0.0, 0.0, 1.0)), [-1, 2, 3])


def test_support_capsule_endpoint_plus_offset():  # SupportMapTests.cs:26-34
    assert np.array_equal(support(shape("capsule", 0.5, 2.0), (0, 3, 0)), [0, 2.5, 0])
    # Math.Sign(0) = 0: no segment offset in the equatorial plane (CapsuleShape.cs:86-88)
    assert np.array_equal(support(shape("capsule", 0.5, 2.0), (1, 0, 0)), [0.5, 0, 0])


def test_vertex_support_last_maximum_wins():  # SupportMapTests.cs:78-93
    verts = np.array([[1, 0, 0], [1, 1, 0], [2, 0, 0], [2, 1, 0]], np.float32)
    s = shape("cloud", 0, 0, 0, 0, 0, 0, verts_count=4)
    assert np.array_equal(support(s, (1, 0, 0), verts), [2, 1, 0])


# ---- AABBs: closed forms of each CalculateBoundingBox ---------------------------------------
def aabb(s, q, p):
    out = np.zeros(6, np.float32)
    lib().orc_aabb(0, ptr(s), None, ptr(np.asarray(q, np.float32)), ptr(np.asarray(p, np.float32)), ptr(out))
    return out[:3], out[3:]


def test_aabb_box_rotated_matches_corner_extent():
    rng = np.random.default_rng(1)
    h = np.array([0.3, 0.7, 1.1], np.float32)
    for _ in range(50):
        q = rng.normal(size=4)
        q /= np.linalg.norm(q)
        p = rng.normal(size=3) * 5
        mn, mx = aabb(shape("box", *h), q, p)
        corners = np.array([quat_rotate(q, h * np.array(s)) for s in np.ndindex(2, 2, 2) for s in [np.array(s) * 2 - 1]])
        np.testing.assert_allclose(mx, p + corners.max(0), atol=2e-5)
        np.testing.assert_allclose(mn, p + corners.min(0), atol=2e-5)


def test_aabb_capsule_and_sphere():
    q = np.array([0, 0, math.sin(0.25), math.cos(0.25)])
    mn, mx = aabb(shape("capsule", 0.3, 0.8), q, (1, 2, 3))
    axis = quat_rotate(q, [0, 1, 0]) * 0.8
    np.testing.assert_allclose(mx, np.array([1, 2, 3]) + 0.3 + np.abs(axis), atol=1e-6)
    mn, mx = aabb(shape("sphere", 0.5), (0, 0, 0, 1), (1, 2, 3))
    assert np.array_equal(mn, [0.5, 1.5, 2.5]) and np.array_equal(mx, [1.5, 2.5, 3.5])


# ---- MPR / EPA: Behavior/ConvexPolytopeTests.cs ---------------------------------------------
def sat_min_depth(ha, hb, q, p):
    """MinimumBoxDepth (ConvexPolytopeTests.cs:178-196) in float64: 15 SAT axes of two boxes."""
    axes_a = [np.eye(3)[i] for i in range(3)]
    axes_b = [quat_rotate(q, a) for a in axes_a]
    best = np.inf

    def sup_box(h, d):
        return np.sign(np.where(d == 0, 1, d)) * h

    def depth(axis):
        if np.dot(axis, axis) < 1e-12:
            return np.inf
        axis = axis / np.linalg.norm(axis)
        da = np.dot(sup_box(ha, axis), axis)
        loc = quat_rotate(np.array([-q[0], -q[1], -q[2], q[3]]), -axis)
        sb = quat_rotate(q, sup_box(hb, loc)) + p
        v_pos = da - np.dot(sb, axis)
        da_n = np.dot(sup_box(ha, -axis), -axis)
        loc2 = quat_rotate(np.array([-q[0], -q[1], -q[2], q[3]]), axis)
        sb2 = quat_rotate(q, sup_box(hb, loc2)) + p
        v_neg = da_n - np.dot(sb2, -axis)
        return min(v_pos, v_neg)

    for a in axes_a + axes_b:
        best = min(best, depth(a))
    for a in axes_a:
        for b in axes_b:
            best = min(best, depth(np.cross(a, b)))
    return best


@pytest.mark.parametrize("ha,hb,q,p", [
    # MprEpa_CollinearSupportRegressionMatchesBoxSat (ConvexPolytopeTests.cs:199-211), FP32 literals
    ((0.3, 1.2, 0.7), (0.3, 1.2, 0.7), (0, 0.99999815, 0, -0.0019342888), (0, 0.43866748, 0)),
    # MprEpa_ObliqueBoxRegressionMatchesBoxSat (:213-225)
    ((2, 0.25, 1), (2, 0.25, 1), (-0.28864738, -0.1558499, 0.15853547, -0.931268), (-0.7871254, 0.20686495, -0.014221248)),
    # Collision_DuplicateSupportRegressionMatchesBoxSat inputs (:164-176)
    ((2, 0.25, 1), (0.5, 0.5, 0.5), (-0.34783155, 0.69148487, -0.5953642, -0.21541446), (0.53918016, 0.12698002, -1.1122025)),
])
@pytest.mark.parametrize("fp64", [0, 1])
def test_mpr_epa_box_regressions_match_sat(ha, hb, q, p, fp64):
    r = mpr(shape("box", *ha), shape("box", *hb), (0, 0, 0, 1), (0, 0, 0), q, p, fp64=fp64)
    assert r["hit"]
    expected = sat_min_depth(np.array(ha), np.array(hb), np.array(q, float), np.array(p, float))
    assert abs(r["depth"] - expected) <= 1e-4
    # depth equals the Minkowski support along the returned normal (AssertMprDepth, :227-234)
    n = r["normal"].astype(float)
    sa = support(shape("box", *ha), n)
    loc = quat_rotate(np.array([-q[0], -q[1], -q[2], q[3]]), -n)
    sb = quat_rotate(q, support(shape("box", *hb), loc)) + np.array(p)
    assert abs(np.dot(sa - sb, n) - r["depth"]) <= 1e-4


@pytest.mark.parametrize("kind,h", [("cube", (0.5, 0.5, 0.5)), ("flat box", (1.5, 0.1, 0.7))])
@pytest.mark.parametrize("threshold", [0.0, 0.02])
def test_mpr_epa_floor_contacts_exact(kind, h, threshold):  # MprEpa_FloorContactsRemainExact (:236-264)
    floor = shape("box", 20, 0.5, 20)
    s = shape("box", *h)
    wrong = 0
    for depth in (0.001, 0.005, 0.01, 0.025, 0.05, 0.1):
        for yaw in range(0, 90, 9):
            a = yaw * math.pi / 180
            q = (0, math.sin(a / 2), 0, math.cos(a / 2))
            for x in (-2, 0, 2):
                for z in (-2, 0, 2):
                    pos = (x * 0.37, h[1] - depth, z * 0.53)
                    r = mpr(s, floor, q, pos, (0, 0, 0, 1), (0, -0.5, 0), threshold=threshold)
                    if r["hit"] and abs(r["depth"] - depth) <= 1e-4 and abs(r["normal"][1]) >= 0.9999:
                        continue
                    wrong += 1
    assert wrong == 0


def test_mpr_sphere_sphere_closed_form():
    for d in (0.1, 0.55, 0.9):
        r = mpr(shape("sphere", 0.5), shape("sphere", 0.5), (0, 0, 0, 1), (0, 0, 0), (0, 0, 0, 1), (d, 0, 0))
        assert r["hit"]
        assert abs(r["depth"] - (1.0 - d)) < 1e-5
        # the normal points from B's deepest point (d - 0.5) to A's (0.5) (NarrowPhase.cs:891-898): +x,
        # the direction a positive impulse pushes body 2 (Contact.cs:663-667)
        np.testing.assert_allclose(r["normal"], [1, 0, 0], atol=1e-6)
        np.testing.assert_allclose(r["pa"], [0.5, 0, 0], atol=1e-6)
        np.testing.assert_allclose(r["pb"], [d - 0.5, 0, 0], atol=1e-6)
    assert not mpr(shape("sphere", 0.5), shape("sphere", 0.5), (0, 0, 0, 1), (0, 0, 0), (0, 0, 0, 1), (1.01, 0, 0))["hit"]


def test_mpr_depth_always_matches_its_normal():  # MprEpa_DepthAlwaysMatchesItsNormal (:266-318)
    shapes = [shape("box", 0.5, 0.5, 0.5), shape("box", 2, 0.25, 1), shape("capsule", 0.3, 0.5), shape("sphere", 0.5)]
    rng = np.random.default_rng(11)
    overlapping = wrong = 0
    for i in range(1500):
        a = shapes[rng.integers(len(shapes))]
        b = shapes[rng.integers(len(shapes))]
        axis = rng.normal(size=3)
        axis /= np.linalg.norm(axis)
        ang = rng.uniform(0, 6)
        q = (*(axis * math.sin(ang / 2)), math.cos(ang / 2))
        direction = rng.normal(size=3)
        direction /= np.linalg.norm(direction)
        pos = direction * rng.uniform(0.4, 1.8)
        r = mpr(a, b, (0, 0, 0, 1), (0, 0, 0), q, pos)
        if not r["hit"]:
            continue
        overlapping += 1
        n = r["normal"].astype(float)
        sa = support(a, n)
        loc = quat_rotate(np.array([-q[0], -q[1], -q[2], q[3]]), -n)
        sb = quat_rotate(q, support(b, loc)) + pos
        if abs(np.dot(sa - sb, n) - r["depth"]) > 1e-3:
            wrong += 1
    assert overlapping > 400
    assert wrong == 0


# ---- manifolds: Regression/CollisionManifoldTests.cs ----------------------------------------
def test_equal_face_boxes_four_corner_contacts():  # EqualFaceBoxes_ProduceFourCornerContacts (:167-190)
    r = mpr(shape("box", 1, 1, 1), shape("box", 1, 1, 1), (0, 0, 0, 1), (0, 0, 0), (0, 0, 0, 1), (0, 1.9, 0))
    assert r["hit"] and r["count"] == 4
    for a, b in zip(r["mA"], r["mB"]):
        assert abs(a[1] - 1.0) < 1e-4 and abs(b[1] - 0.9) < 1e-4
        assert abs(abs(a[0]) - 1.0) < 1e-4 and abs(abs(a[2]) - 1.0) < 1e-4
        assert abs(np.dot(a - b, r["normal"]) - r["depth"]) < 1e-4
    assert len({tuple(np.round(a, 3)) for a in r["mA"]}) == 4


def test_rotated_face_boxes_include_intersection_vertices():  # RotatedFaceBoxes_IncludeIntersectionVertices (:221-250)
    a = math.pi / 4
    r = mpr(shape("box", 1, 1, 1), shape("box", 1, 1, 1), (0, 0, 0, 1), (0, 0, 0), (0, math.sin(a / 2), 0, math.cos(a / 2)),
            (0, 1.9, 0))
    assert r["hit"] and r["count"] == 4
    for pa in r["mA"]:
        assert abs(pa[1] - 1.0) < 2e-4
        assert max(abs(pa[0]), abs(pa[2])) <= 1.0 + 2e-4


def build_manifold(sa, sb, qa, pa, qb, pb, pA, pB, n, verts=None):
    out = np.zeros(25, np.float32)
    v = None if verts is None else np.ascontiguousarray(verts, np.float32).ravel()
    f = lambda x: np.ascontiguousarray(x, np.float32)
    lib().orc_build_manifold(0, ptr(sa), ptr(sb), ptr(v), ptr(f(qa)), ptr(f(pa)), ptr(f(qb)), ptr(f(pb)), ptr(f(pA)),
                             ptr(f(pB)), ptr(f(n)), ptr(out))
    c = int(out[0])
    return c, out[1:1 + 3 * c].reshape(-1, 3), out[13:13 + 3 * c].reshape(-1, 3)


@pytest.mark.parametrize("radius", [0.5, 1.0])
@pytest.mark.parametrize("sphere_first", [False, True])
def test_sphere_against_floor_keeps_original_contact(radius, sphere_first):  # :105-127
    floor, sph = shape("box", 100, 100, 100), shape("sphere", radius)
    fpos, spos = (0, -100, 0), (3, radius - 0.005, -2)
    fpt, spt = (3, 0, -2), (3, -0.005, -2)
    if sphere_first:
        c, mA, mB = build_manifold(sph, floor, (0, 0, 0, 1), spos, (0, 0, 0, 1), fpos, spt, fpt, (0, -1, 0))
        ea, eb = spt, fpt
    else:
        c, mA, mB = build_manifold(floor, sph, (0, 0, 0, 1), fpos, (0, 0, 0, 1), spos, fpt, spt, (0, 1, 0))
        ea, eb = fpt, spt
    assert c == 1
    assert np.abs(mA[0] - ea).max() < 1e-5 and np.abs(mB[0] - eb).max() < 1e-5


@pytest.mark.parametrize("prism_first", [False, True])
def test_triangular_face_against_floor(prism_first):  # TriangularFaceAgainstFloor (:192-219)
    floor = shape("box", 2, 1, 2)
    verts = np.array([[-1, -1, -1], [1, -1, -1], [0, -1, 1], [-1, 1, -1], [1, 1, -1], [0, 1, 1]], np.float32)
    prism = shape("cloud", 0, 0, 0, 0, 0, 0, verts_count=6)
    fpos, ppos = (0, -1, 0), (0, 0.995, 0)
    fpt, ppt = (0, 0, 0), (0, -0.005, 0)
    if prism_first:
        c, mA, mB = build_manifold(prism, floor, (0, 0, 0, 1), ppos, (0, 0, 0, 1), fpos, ppt, fpt, (0, -1, 0), verts)
        floor_pts, a3, b3 = mB, ppt, fpt
    else:
        c, mA, mB = build_manifold(floor, prism, (0, 0, 0, 1), fpos, (0, 0, 0, 1), ppos, fpt, ppt, (0, 1, 0), verts)
        floor_pts, a3, b3 = mA, fpt, ppt
    assert c == 4
    expected = [(-1, 0, -1), (1, 0, -1), (0, 0, 1), fpt]
    used = set()
    for e in expected:
        d = np.sum((floor_pts - np.array(e)) ** 2, axis=1)
        j = int(np.argmin(d))
        assert d[j] <= 1e-10 and j not in used
        used.add(j)
    assert np.abs(mA[3] - a3).max() < 1e-5 and np.abs(mB[3] - b3).max() < 1e-5


# ---- contact solver: Regression/ContactSubstepTests.cs, Behavior/ContactLifecycleTests.cs ----
def contact_op(bodies, arb, op, dt=0.01, substeps=1, p1=None, p2=None, n=None, fp64=0):
    f = lambda x: None if x is None else np.ascontiguousarray(x, np.float32)
    r = lib().orc_contact_op(fp64, ptr(bodies), ptr(arb), op, ctypes.c_float(dt), substeps, ptr(f(p1)), ptr(f(p2)),
                             ptr(f(n)))
    assert r == 0


def two_body_state(static_first=True):
    b = np.zeros(2, BODY_DTYPE)
    b["orientation"] = (0, 0, 0, 1)
    b["linear_damping_multiplier"] = 1
    b["angular_damping_multiplier"] = 1
    b[0]["motion_type"] = 2 if static_first else 0
    b[1]["motion_type"] = 0
    for i in range(2):
        b[i]["inverse_mass"] = 1.0
        b[i]["inverse_inertia"] = (1, 0, 0, 1, 0, 1)
    a = np.zeros(1, ARBITER_DTYPE)
    a["body1"], a["body2"] = 0, 1
    a["speculative_relaxation_factor"] = 0.9
    return b, a


@pytest.mark.parametrize("fp64", [0, 1])
def test_rotation_within_step_penalty_bias(fp64):  # RotationWithinStep_... (ContactSubstepTests.cs:11-47)
    b, a = two_body_state()
    mass, inertia = scenes.sphere_mass(0.5)
    inv = scenes._sym_inverse(*inertia)
    b[1]["inverse_mass"] = 1 / mass
    b[1]["inverse_inertia"] = inv
    b[1]["position"] = (0, 0.45, 0)
    b[1]["velocity"] = (0, -1, 0)
    contact_op(b, a, 0, p1=(0, 0, 0), p2=(0, -0.05, 0), n=(0, 1, 0), fp64=fp64)
    assert a["usage_mask"][0] == 1
    b[1]["orientation"] = (0, 0, math.sin(0.15), math.cos(0.15))
    dt, sub = 0.01, 3
    contact_op(b, a, 1, dt=dt, substeps=sub, fp64=fp64)
    pen = 0.5 * math.cos(0.3) - 0.45
    expected = 0.2 * (1 / dt) * (pen - 0.01)
    c0 = a["contacts"][0, 0]
    assert abs(c0["penalty_bias"] - expected) < 1e-5
    np.testing.assert_allclose(c0["relative_position2"], [0, -0.5, 0], atol=1e-5)
    contact_op(b, a, 2, fp64=fp64)
    assert abs(b[1]["velocity"][1] - expected) < 1e-5
    assert np.linalg.norm(b[1]["angular_velocity"]) < 1e-5


def qmul(a, b):
    ax, ay, az, aw = a
    bx, by, bz, bw = b
    return (aw * bx + bw * ax + ay * bz - az * by, aw * by + bw * ay + az * bx - ax * bz,
            aw * bz + bw * az + ax * by - ay * bx, aw * bw - (ax * bx + ay * by + az * bz))


def rot_matrix(q):
    return np.array([quat_rotate(q, e) for e in np.eye(3)]).T


@pytest.mark.parametrize("apply_bias", [True, False])
def test_contact_conserves_momentum(apply_bias):  # ContactIteration_ConservesLinearAndAngularMomentum (ContactLifecycleTests.cs:15-70)
    b, a = two_body_state(static_first=False)
    rx = lambda t: (math.sin(t / 2), 0, 0, math.cos(t / 2))
    ry = lambda t: (0, math.sin(t / 2), 0, math.cos(t / 2))
    rz = lambda t: (0, 0, math.sin(t / 2), math.cos(t / 2))
    q1, q2 = qmul(rx(0.4), rz(0.7)), qmul(ry(0.6), rx(0.3))
    masses, inertias = (2.0, 3.0), ((3, 5, 7), (4, 6, 9))
    for i, (pos, q, v, w) in enumerate((((-0.6, 0.3, 0.4), q1, (2, 3, -1), (0.2, -0.4, 0.1)),
                                        ((0.8, -0.2, -0.5), q2, (-1, -2, 0.5), (-0.3, 0.1, 0.5)))):
        b[i]["position"], b[i]["orientation"], b[i]["velocity"], b[i]["angular_velocity"] = pos, q, v, w
        b[i]["inverse_mass"] = 1 / masses[i]
        I = inertias[i]
        b[i]["inverse_inertia"] = (1 / I[0], 0, 0, 1 / I[1], 0, 1 / I[2])
    a["friction"] = 0.75
    n = np.array([0.3, 1, -0.4])
    n /= np.linalg.norm(n)
    point = (0.1, 0.4, 0.5)
    contact_op(b, a, 0, p1=point, p2=point, n=n)
    contact_op(b, a, 1, dt=0.01)
    c = a["contacts"][0, 0]
    c["penalty_bias"] = 1.0
    a["contacts"][0, 0] = c
    world_I = [rot_matrix(b[i]["orientation"]) @ np.diag(inertias[i]) @ rot_matrix(b[i]["orientation"]).T
               for i in range(2)]

    def momentum(bb):
        lin = sum(masses[i] * bb[i]["velocity"].astype(float) for i in range(2))
        ang = sum(np.cross(bb[i]["position"], masses[i] * bb[i]["velocity"].astype(float)) +
                  world_I[i] @ bb[i]["angular_velocity"] for i in range(2))
        return lin, ang

    l0, g0 = momentum(b)
    w0 = b[0]["angular_velocity"].copy()
    for _ in range(8):
        contact_op(b, a, 2 if apply_bias else 3)
        l1, g1 = momentum(b)
        assert np.abs(l1 - l0).max() < 10 * 1e-5
        assert np.abs(g1 - g0).max() < 10 * 1e-5
    acc = a["contacts"][0, 0]["accumulated"]
    assert acc[0] > 1e-5 and abs(acc[1]) + abs(acc[2]) > 1e-5
    assert np.abs(b[0]["angular_velocity"] - w0).max() > 1e-5


def test_friction_clamped_by_pre_update_normal_impulse():  # Contact.cs:1207-1216
    b, a = two_body_state()
    b[1]["position"] = (0, 0.5, 0)
    b[1]["velocity"] = (3.0, -0.5, 0)
    a["friction"] = 0.6
    contact_op(b, a, 0, p1=(0, 0, 0), p2=(0, 0, 0), n=(0, 1, 0))
    c = a["contacts"][0, 0]
    c["accumulated"] = (0.2, 0, 0)
    a["contacts"][0, 0] = c
    contact_op(b, a, 1)
    acc_before = a["contacts"][0, 0]["accumulated"].copy()
    contact_op(b, a, 2)
    acc = a["contacts"][0, 0]["accumulated"]
    assert abs(acc[1]) <= 0.6 * acc_before[0] + 1e-7 and abs(acc[2]) <= 0.6 * acc_before[0] + 1e-7
    assert abs(abs(acc[1]) - 0.6 * acc_before[0]) < 1e-6  # sliding: at the cone


def test_add_contact_slot_persistence():  # Contact.cs:329-416 (PersistentContactManifoldSelectionTests semantics)
    b, a = two_body_state()
    b[1]["position"] = (0, 0.5, 0)
    pts = [(-0.5, 0, -0.5), (0.5, 0, -0.5), (0.5, 0, 0.5), (-0.5, 0, 0.5)]
    for p in pts:
        contact_op(b, a, 0, p1=p, p2=p, n=(0, 1, 0))
    assert a["usage_mask"][0] == 15
    # re-adding a cached point within BreakThreshold keeps the slot set and refreshes that slot only
    before = a["contacts"][0]["relative_position1"].copy()
    contact_op(b, a, 0, p1=(0.5, 0.001, 0.5), p2=(0.5, 0, 0.5), n=(0, 1, 0))
    assert a["usage_mask"][0] == 15
    changed = np.any(a["contacts"][0]["relative_position1"] != before, axis=1)
    assert changed.tolist() == [False, False, True, False]
    # with all four slots used, a new point replaces exactly one slot (SortCachedPoints, Contact.cs:529-553)
    rel = a["contacts"][0]["relative_position1"].copy()
    contact_op(b, a, 0, p1=(0.0, 0, 0.0), p2=(0, 0, 0), n=(0, 1, 0))
    assert a["usage_mask"][0] == 15
    area_after = a["contacts"][0]["relative_position1"]
    assert np.sum(np.any(area_after != rel, axis=1)) == 1


def test_update_position_breaks_separated_contacts():  # Contact.cs:762-797
    b, a = two_body_state()
    b[1]["position"] = (0, 0.5, 0)
    contact_op(b, a, 0, p1=(0, 0, 0), p2=(0, 0, 0), n=(0, 1, 0))
    contact_op(b, a, 4)
    assert a["usage_mask"][0] & 15 == 1
    b[1]["position"] = (0, 0.5 + 0.0021, 0)  # normal separation > BreakThreshold / 10
    contact_op(b, a, 4)
    assert a["usage_mask"][0] & 15 == 0
    b2, a2 = two_body_state()
    b2[1]["position"] = (0, 0.5, 0)
    contact_op(b2, a2, 0, p1=(0, 0, 0), p2=(0, 0, 0), n=(0, 1, 0))
    b2[1]["position"] = (0.0201, 0.5, 0)  # tangential drift > BreakThreshold
    contact_op(b2, a2, 4)
    assert a2["usage_mask"][0] & 15 == 0


# ---- trigonometry and integration: LinearMath/StableMath.cs, MathHelper.cs:47-78 -----------
def test_stable_sincos_bounds():
    L = lib()
    s, c = ctypes.c_float(), ctypes.c_float()
    worst = 0.0
    for a in np.linspace(-20, 20, 4001, dtype=np.float32):
        L.orc_sincos_f(float(a), ctypes.byref(s), ctypes.byref(c))
        worst = max(worst, abs(s.value - math.sin(float(a))), abs(c.value - math.cos(float(a))))
    assert worst < 5e-6


def test_stable_atan2_bounds():
    L = lib()
    rng = np.random.default_rng(5)
    for y, x in rng.normal(size=(2000, 2)):
        assert abs(L.orc_atan2_d(float(y), float(x)) - math.atan2(y, x)) < 5e-9  # 17th-order series on |x| <= tan(pi/8)
        assert abs(L.orc_atan2_f(float(y), float(x)) - math.atan2(y, x)) < 1e-6
    assert L.orc_atan2_f(0.0, 0.0) == 0.0 and abs(L.orc_atan2_f(1.0, 0.0) - math.pi / 2) < 1e-7


def test_rotation_quaternion_matches_exact():
    L = lib()
    q = np.zeros(4, np.float32)
    for w in ([0.0, 0.0, 1e-5], [3.0, -1.0, 2.0], [50.0, 0.0, 0.0], [0.01, 0.02, -0.03]):
        w = np.array(w, np.float32)
        dt = 1.0 / 240
        L.orc_rotation_quaternion_f(ptr(w), ctypes.c_float(dt), ptr(q))
        ang = np.linalg.norm(w.astype(float))
        axis = w / ang if ang > 0 else np.array([1, 0, 0])
        exact = np.array([*(axis * math.sin(ang * dt / 2)), math.cos(ang * dt / 2)])
        np.testing.assert_allclose(q, exact, atol=2e-7)


# ---- exact predicates: Behavior/ExactPredicateTests.cs + DyadicReference.cs ------------------
def exact_sign(a, b, c, p):
    F = [[Fraction(float(v)) for v in x] for x in (a, b, c, p)]
    a, b, c, p = F
    u = [a[i] - b[i] for i in range(3)]
    v = [a[i] - c[i] for i in range(3)]
    d = [p[i] - a[i] for i in range(3)]
    cr = [u[1] * v[2] - u[2] * v[1], u[2] * v[0] - u[0] * v[2], u[0] * v[1] - u[1] * v[0]]
    det = cr[0] * d[0] + cr[1] * d[1] + cr[2] * d[2]
    return (det > 0) - (det < 0)


def test_exact_side_matches_rational_arithmetic():
    L = lib()
    rng = np.random.default_rng(8)
    for i in range(600):
        a, b, c = (rng.normal(size=3).astype(np.float32) for _ in range(3))
        if i % 3 == 0:  # nearly coplanar point: the filtered predicate cannot decide
            t, s = rng.uniform(size=2)
            p = (a + t * (b - a) + s * (c - a)).astype(np.float32)
        elif i % 3 == 1:  # exactly coplanar: a vertex
            p = b.copy()
        else:
            p = rng.normal(size=3).astype(np.float32)
        got = L.orc_exact_side(ptr(a), ptr(b), ptr(c), ptr(p))
        assert got == exact_sign(a, b, c, p)


# ---- pair set and whole-step behaviour ----------------------------------------------------
def test_sweep_and_prune_equals_brute_force():
    b, s, v, cfg, dt = scenes.heap(800, seed=4)
    w = OracleWorld(b, s, v, cfg)
    for _ in range(40):
        w.step(dt)
    assert np.array_equal(w.pairs(brute=True), w.pairs())
    assert len(w.pairs()) > 100
    # worlds sharing coordinates: the sweep runs world by world, brute force filters by world id
    b, s, v, cfg, dt = scenes.batched_worlds(n_worlds=12, min_bodies=16, max_bodies=120, seed=2)
    w = OracleWorld(b, s, v, cfg)
    for _ in range(10):
        w.step(dt)
    assert np.array_equal(w.pairs(brute=True), w.pairs()) and len(w.pairs()) > 100


def test_pyramid_stays_standing():  # Behavior/StackingTests.cs:36-50 (moves < 1 m)
    b, s, v, cfg, dt = scenes.pyramid(rows=10)
    w = OracleWorld(b, s, v, cfg)
    top0 = b["position"][-1].copy()
    for _ in range(300):
        w.step(dt)
    assert np.linalg.norm(w.bodies()["position"][-1] - top0) < 1.0


def test_resting_box_normal_impulse_carries_weight():
    """A unit box at rest on the ground: per substep the normal impulses sum to m*g*h (closed form)."""
    bodies, shapes, verts = scenes.two_bodies(("box", (20, 1, 20)), ("box", (1, 1, 1)), (0, -0.5, 0), (0, 0.495, 0),
                                              static_a=True)
    cfg = scenes.default_config()
    cfg["substep_count"] = 2
    w = OracleWorld(bodies, shapes, verts, cfg)
    dt = 0.01
    for _ in range(200):
        w.step(dt)
    arbs = w.arbiters()
    assert len(arbs) == 1 and arbs[0]["usage_mask"] & 15 == 15
    total = arbs[0]["contacts"]["accumulated"][:, 0].sum()
    assert abs(total - 9.81 * dt / 2) < 0.05 * 9.81 * dt / 2
    assert abs(w.bodies()["velocity"][1]).max() < 1e-3


def test_fp64_and_fp32_oracles_agree_short_horizon():
    b, s, v, cfg, dt = scenes.pyramid(rows=5)
    w32 = OracleWorld(b, s, v, cfg)
    w64 = OracleWorld(b, s, v, cfg, fp64=True)
    for _ in range(50):
        w32.step(dt)
        w64.step(dt)
    d = np.abs(w32.state64()[:, :3] - w64.state64()[:, :3]).max()
    assert d < 1e-4
