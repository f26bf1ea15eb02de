"""Constraint records (jh_constraint) built the way the reference's Initialize methods set them up.

Each builder takes the body array (positions, orientations at creation time) and returns one record
whose params / state hold the Data fields of the reference's constraint after Initialize with its
default softness and bias (Constraint.cs:150-164).  FP32 arithmetic in the reference's order:
JVector.ConjugatedTransform (JVector.cs:348-361), JQuaternion.Multiply (JQuaternion.cs:353-368).
"""
from __future__ import annotations

import math

import numpy as np

from ._abi import (CONSTRAINT_BALL_SOCKET, CONSTRAINT_DISTANCE_LIMIT, CONSTRAINT_DTYPE, CONSTRAINT_FIXED_ANGLE,
                   CONSTRAINT_HINGE_ANGLE)

f32 = np.float32

# Constraint.cs:150-164
DEFAULT_ANGULAR_SOFTNESS = f32(0.001)
DEFAULT_ANGULAR_BIAS = f32(0.2)
DEFAULT_ANGULAR_LIMIT_SOFTNESS = f32(0.001)
DEFAULT_ANGULAR_LIMIT_BIAS = f32(0.1)
DEFAULT_LINEAR_SOFTNESS = f32(0.00001)
DEFAULT_LINEAR_BIAS = f32(0.2)


def _rot_inv(v, q):
    """JVector.ConjugatedTransform(v, q) in FP32."""
    x, y, z = (f32(c) for c in v)
    qx, qy, qz, qw = (f32(c) for c in q)
    two = f32(2.0)
    n0 = two * (qz * y - qy * z)
    n1 = two * (qx * z - qz * x)
    n2 = two * (qy * x - qx * y)
    n00 = qz * n1 - qy * n2
    n11 = qx * n2 - qz * n0
    n22 = qy * n0 - qx * n1
    return np.array([x + qw * n0 + n00, y + qw * n1 + n11, z + qw * n2 + n22], np.float32)


def _qmul(a, b):
    """JQuaternion.Multiply in FP32, (x, y, z, w)."""
    i1, j1, k1, r1 = (f32(c) for c in a)
    i2, j2, k2, r2 = (f32(c) for c in b)
    w = r1 * r2 - (i1 * i2 + j1 * j2 + k1 * k2)
    x = r1 * i2 + r2 * i1 + j1 * k2 - k1 * j2
    y = r1 * j2 + r2 * j1 + k1 * i2 - i1 * k2
    z = r1 * k2 + r2 * k1 + i1 * j2 - j1 * i2
    return np.array([x, y, z, w], np.float32)


def _conj(q):
    return np.array([-f32(q[0]), -f32(q[1]), -f32(q[2]), f32(q[3])], np.float32)


def _record(kind, b1, b2):
    c = np.zeros(1, CONSTRAINT_DTYPE)[0]
    c["type"] = kind
    c["body1"] = b1
    c["body2"] = b2
    c["enabled"] = 1
    return c


def _local_anchor(bodies, b, anchor):
    d = np.asarray(anchor, np.float32) - bodies["position"][b].astype(np.float32)
    return _rot_inv(d, bodies["orientation"][b])


def ball_socket(bodies, b1, b2, anchor):
    """BallSocket.Initialize(anchor) (BallSocket.cs:53-71)."""
    c = _record(CONSTRAINT_BALL_SOCKET, b1, b2)
    c["params"][0:3] = _local_anchor(bodies, b1, anchor)
    c["params"][3:6] = _local_anchor(bodies, b2, anchor)
    c["params"][6] = DEFAULT_LINEAR_BIAS
    c["params"][7] = DEFAULT_LINEAR_SOFTNESS
    return c


def hinge_angle(bodies, b1, b2, axis, limit=(-math.pi, math.pi)):
    """HingeAngle.Initialize(axis, limit) (HingeAngle.cs:100-126); limit (from, to) in radians, the
    default the full range (AngularLimit.Full)."""
    c = _record(CONSTRAINT_HINGE_ANGLE, b1, b2)
    lo, hi = f32(limit[0]), f32(limit[1])
    p = c["params"]
    p[0] = lo / f32(2.0)
    p[1] = hi / f32(2.0)
    p[2] = f32(0.25) * lo + f32(0.25) * hi
    p[3] = DEFAULT_ANGULAR_BIAS
    p[4] = DEFAULT_ANGULAR_LIMIT_BIAS
    p[5] = DEFAULT_ANGULAR_LIMIT_SOFTNESS
    p[6] = DEFAULT_ANGULAR_SOFTNESS
    a = np.asarray(axis, np.float32)
    a = a * (f32(1.0) / np.sqrt((a * a).sum(dtype=np.float32)))
    p[7:10] = _rot_inv(a, bodies["orientation"][b2])
    p[10:14] = _qmul(_conj(bodies["orientation"][b2]), bodies["orientation"][b1])
    return c


def fixed_angle(bodies, b1, b2):
    """FixedAngle.Initialize() (FixedAngle.cs:60-75)."""
    c = _record(CONSTRAINT_FIXED_ANGLE, b1, b2)
    c["params"][0] = DEFAULT_ANGULAR_BIAS
    c["params"][1] = DEFAULT_ANGULAR_SOFTNESS
    c["params"][2:6] = _qmul(_conj(bodies["orientation"][b2]), bodies["orientation"][b1])
    return c


def distance_limit(bodies, b1, b2, anchor1, anchor2, limit=(1e-6, -1e-6)):
    """DistanceLimit.Initialize(anchor1, anchor2, limit) (DistanceLimit.cs:102-128); the default limit
    LinearLimit.Fixed = (1e-6, -1e-6) (Limit.cs:91-92) keeps the initial distance."""
    c = _record(CONSTRAINT_DISTANCE_LIMIT, b1, b2)
    p = c["params"]
    p[0:3] = _local_anchor(bodies, b1, anchor1)
    p[3:6] = _local_anchor(bodies, b2, anchor2)
    p[6] = DEFAULT_LINEAR_BIAS
    p[7] = DEFAULT_LINEAR_SOFTNESS
    d = np.asarray(anchor2, np.float32) - np.asarray(anchor1, np.float32)
    p[8] = np.sqrt((d * d).sum(dtype=np.float32))
    p[9], p[10] = f32(limit[0]), f32(limit[1])
    return c


def hinge_joint(bodies, b1, b2, anchor, axis, limit=(-math.pi, math.pi)):
    """HingeJoint (Dynamics/Joints/HingeJoint.cs:58-63): HingeAngle about the axis, then BallSocket at
    the anchor (the order the joint creates them, which is the solve order)."""
    return [hinge_angle(bodies, b1, b2, axis, limit), ball_socket(bodies, b1, b2, anchor)]


def weld_joint(bodies, b1, b2, anchor):
    """WeldJoint (Dynamics/Joints/WeldJoint.cs:38-43): FixedAngle, then BallSocket at the anchor."""
    return [fixed_angle(bodies, b1, b2), ball_socket(bodies, b1, b2, anchor)]


def as_array(records) -> np.ndarray:
    return np.array(records, dtype=CONSTRAINT_DTYPE) if len(records) else np.zeros(0, CONSTRAINT_DTYPE)
