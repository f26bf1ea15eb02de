"""Constraint records (jh_constraint) built the way the reference's Initialize methods set them up.

Each builder takes the body array (positions, orientations at creation time) and returns one record
whose params / state hold the Data fields of the reference's constraint after Initialize with its
default softness and bias (Constraint.cs:150-164).  FP32 arithmetic in the reference's order:
JVector.ConjugatedTransform (JVector.cs:348-361), JQuaternion.Multiply (JQuaternion.cs:353-368).
"""
from __future__ import annotations

import math

import numpy as np

from ._abi import (CONSTRAINT_ANGULAR_MOTOR, CONSTRAINT_ANGULAR_SERVO, CONSTRAINT_BALL_SOCKET, CONSTRAINT_CONE_LIMIT,
                   CONSTRAINT_DISTANCE_LIMIT, CONSTRAINT_DTYPE, CONSTRAINT_ELLIPTICAL_CONE_LIMIT,
                   CONSTRAINT_FIXED_ANGLE, CONSTRAINT_HINGE_ANGLE,
                   CONSTRAINT_LINEAR_AXIS_SERVO, CONSTRAINT_LINEAR_MOTOR, CONSTRAINT_LINEAR_SERVO,
                   CONSTRAINT_POINT_ON_LINE, CONSTRAINT_POINT_ON_PLANE, CONSTRAINT_TWIST_ANGLE, CONSTRAINT_TWIST_SERVO)

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


def _rot(v, q):
    """JVector.Transform(v, q) in FP32."""
    x, y, z = (f32(c) for c in v)
    qx, qy, qz, qw = (f32(c) for c in q)
    two = f32(2.0)
    n0 = two * (qy * z - qz * y)
    n1 = two * (qz * x - qx * z)
    n2 = two * (qx * y - qy * x)
    n00 = qy * n2 - qz * n1
    n11 = qz * n0 - qx * n2
    n22 = qx * n1 - qy * n0
    return np.array([x + qw * n0 + n00, y + qw * n1 + n11, z + qw * n2 + n22], np.float32)


def _normalize(v):
    a = np.asarray(v, np.float32)
    return a * (f32(1.0) / np.sqrt((a * a).sum(dtype=np.float32)))


def _create_orthonormal(v):
    """MathHelper.CreateOrthonormal (MathHelper.cs:204-255) in FP32."""
    v = np.asarray(v, np.float32)
    ax, ay, az = np.abs(v)
    if ax <= ay and ax <= az:
        y, z = v[2], -v[1]
        inv = f32(1.0) / np.sqrt(y * y + z * z)
        return np.array([0.0, y * inv, z * inv], np.float32)
    if ay <= az:
        x, z = -v[2], v[0]
        inv = f32(1.0) / np.sqrt(x * x + z * z)
        return np.array([x * inv, 0.0, z * inv], np.float32)
    x, y = v[1], -v[0]
    inv = f32(1.0) / np.sqrt(x * x + y * y)
    return np.array([x * inv, y * inv, 0.0], np.float32)


def _from_to(a, b):
    """JQuaternion.CreateFromToRotation (JQuaternion.cs:84-101) in FP32, (x, y, z, w)."""
    a = np.asarray(a, np.float32)
    b = np.asarray(b, np.float32)
    d = a[0] * b[0] + a[1] * b[1] + a[2] * b[2]
    if d < -f32(1.0) + f32(1e-6):
        ax = _create_orthonormal(a)
        return np.array([ax[0], ax[1], ax[2], 0.0], np.float32)
    s = np.sqrt((f32(1.0) + d) * f32(2.0))
    inv = f32(1.0) / s
    c = np.array([a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2], a[0] * b[1] - a[1] * b[0]], np.float32)
    return np.array([c[0] * inv, c[1] * inv, c[2] * inv, s * f32(0.5)], np.float32)


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


ANGULAR_FIXED = (1e-6, -1e-6)  # AngularLimit.Fixed (Limit.cs:38-39)


def twist_angle(bodies, b1, b2, axis1, axis2, limit=ANGULAR_FIXED):
    """TwistAngle.Initialize(axis1, axis2, limit) (TwistAngle.cs:76-115)."""
    c = _record(CONSTRAINT_TWIST_ANGLE, b1, b2)
    p = c["params"]
    q1, q2 = bodies["orientation"][b1], bodies["orientation"][b2]
    u1 = _rot_inv(_normalize(axis1), q1)
    B = _rot_inv(_normalize(axis2), q2)
    q_rel = _qmul(_conj(q2), q1)
    u1_in_b2 = _rot(u1, q_rel)
    p[0:3] = B
    p[3:7] = _qmul(_from_to(u1_in_b2, B), q_rel)
    lo, hi = f32(limit[0]), f32(limit[1])
    p[7], p[8] = lo, hi
    p[9] = f32(0.5) * lo + f32(0.5) * hi
    p[10] = DEFAULT_ANGULAR_BIAS
    p[11] = DEFAULT_ANGULAR_SOFTNESS
    return c


def cone_limit(bodies, b1, b2, axis1, axis2=None, limit=(0.0, math.pi / 4)):
    """ConeLimit.Initialize(axisBody1, axisBody2, limit) (ConeLimit.cs:82-107); with one axis both bodies
    use it (rest angle 0)."""
    c = _record(CONSTRAINT_CONE_LIMIT, b1, b2)
    p = c["params"]
    p[0:3] = _rot_inv(_normalize(axis1), bodies["orientation"][b1])
    p[3:6] = _rot_inv(_normalize(axis1 if axis2 is None else axis2), bodies["orientation"][b2])
    p[6] = DEFAULT_ANGULAR_BIAS
    p[7] = DEFAULT_ANGULAR_SOFTNESS
    p[8], p[9] = f32(limit[0]), f32(limit[1])
    return c


def elliptical_cone_limit(bodies, b1, b2, axis1, axis2, reference, limit_reference, limit_perpendicular):
    """EllipticalConeLimit.Initialize(axisBody1, axisBody2, referenceDirectionBody1, limitAlongReference,
    limitAlongPerpendicular) (EllipticalConeLimit.cs:92-126): body 2's axis stays inside the elliptical
    cone about body 1's axis whose semiaxes along the reference direction and perpendicular to it are the
    limits' maxima (and, if both minima are positive, outside the inner ellipse of the minima).  Limits
    are (min, max) radians with 0 <= min <= max < pi, max > 0, minima both zero or both positive."""
    lims = [(f32(limit_reference[0]), f32(limit_reference[1])),
            (f32(limit_perpendicular[0]), f32(limit_perpendicular[1]))]
    for mn, mx in lims:
        if not (math.isfinite(mn) and math.isfinite(mx) and mn >= 0 and mx > 0 and mx < f32(math.pi) and mn <= mx):
            raise ValueError("elliptical cone limits must satisfy 0 <= min <= max < pi and max > 0")
    if (lims[0][0] == 0) != (lims[1][0] == 0):
        raise ValueError("the inner semiaxes must either both be zero or both be positive")
    a1 = _normalize(axis1)
    a2 = _normalize(axis2)
    r = _normalize(reference)
    d = f32(a1[0] * r[0] + a1[1] * r[1] + a1[2] * r[2])
    r = np.array([r[0] - a1[0] * d, r[1] - a1[1] * d, r[2] - a1[2] * d], np.float32)
    if float(r[0] * r[0] + r[1] * r[1] + r[2] * r[2]) < 1e-12:
        raise ValueError("the reference direction must not be parallel to the cone axis")
    r = _normalize(r)
    c = _record(CONSTRAINT_ELLIPTICAL_CONE_LIMIT, b1, b2)
    p = c["params"]
    p[0:3] = _rot_inv(a1, bodies["orientation"][b1])
    p[3:6] = _rot_inv(a2, bodies["orientation"][b2])
    p[6:9] = _rot_inv(r, bodies["orientation"][b1])
    p[9], p[10] = lims[0]
    p[11], p[12] = lims[1]
    p[13] = DEFAULT_ANGULAR_SOFTNESS
    p[14] = DEFAULT_ANGULAR_BIAS
    return c


def _motor(kind, bodies, b1, b2, axis1, axis2, velocity, max_force):
    c = _record(kind, b1, b2)
    p = c["params"]
    p[0:3] = _rot_inv(_normalize(axis1), bodies["orientation"][b1])
    p[3:6] = _rot_inv(_normalize(axis2), bodies["orientation"][b2])
    p[6], p[7] = f32(velocity), f32(max_force)
    return c


def angular_motor(bodies, b1, b2, axis1, axis2=None, velocity=0.0, max_torque=0.0):
    """AngularMotor.Initialize(axis1, axis2) (AngularMotor.cs:67-83) with TargetVelocity and MaximumTorque;
    HingeJoint passes the hinge axis for both (HingeJoint.cs:68-69)."""
    return _motor(CONSTRAINT_ANGULAR_MOTOR, bodies, b1, b2, axis1, axis1 if axis2 is None else axis2, velocity,
                  max_torque)


def linear_motor(bodies, b1, b2, axis1, axis2=None, velocity=0.0, max_force=0.0):
    """LinearMotor.Initialize(axis1, axis2) (LinearMotor.cs:97-113) with TargetVelocity and MaximumForce."""
    return _motor(CONSTRAINT_LINEAR_MOTOR, bodies, b1, b2, axis1, axis1 if axis2 is None else axis2, velocity,
                  max_force)


LINEAR_FULL = (-math.inf, math.inf)  # LinearLimit.Full (Limit.cs:85-86)
DEFAULT_LINEAR_LIMIT_SOFTNESS = f32(0.0001)
DEFAULT_LINEAR_LIMIT_BIAS = f32(0.2)


def point_on_line(bodies, b1, b2, axis, anchor1, anchor2, limit=(1e-6, -1e-6)):
    """PointOnLine.Initialize(axis, anchor1, anchor2, limit) (PointOnLine.cs:92-117): anchor 2 stays on the
    line through anchor 1 along the axis (body-1 frame), its travel along the line inside the limit (the
    three-argument overload uses LinearLimit.Fixed, PointOnLine.cs:69-72)."""
    c = _record(CONSTRAINT_POINT_ON_LINE, b1, b2)
    p = c["params"]
    p[0:3] = _rot_inv(_normalize(axis), bodies["orientation"][b1])
    p[3:6] = _local_anchor(bodies, b1, anchor1)
    p[6:9] = _local_anchor(bodies, b2, anchor2)
    p[9] = DEFAULT_LINEAR_BIAS
    p[10] = DEFAULT_LINEAR_LIMIT_BIAS
    p[11] = DEFAULT_LINEAR_SOFTNESS
    p[12] = DEFAULT_LINEAR_LIMIT_SOFTNESS
    p[13], p[14] = f32(limit[0]), f32(limit[1])
    return c


def point_on_plane(bodies, b1, b2, axis, anchor1, anchor2, limit=(1e-6, -1e-6)):
    """PointOnPlane.Initialize(axis, anchor1, anchor2, limit) (PointOnPlane.cs:89-112): the distance of
    anchor 2 from the plane through anchor 1 with normal axis (body-1 frame) inside the limit."""
    c = _record(CONSTRAINT_POINT_ON_PLANE, b1, b2)
    p = c["params"]
    p[0:3] = _rot_inv(_normalize(axis), bodies["orientation"][b1])
    p[3:6] = _local_anchor(bodies, b1, anchor1)
    p[6:9] = _local_anchor(bodies, b2, anchor2)
    p[9] = DEFAULT_LINEAR_BIAS
    p[10] = DEFAULT_LINEAR_SOFTNESS
    p[11], p[12] = f32(limit[0]), f32(limit[1])
    return c


def prismatic_joint(bodies, b1, b2, center, axis, limit=LINEAR_FULL, pinned=True, motor=None, servo=None):
    """PrismaticJoint (Dynamics/Joints/PrismaticJoint.cs:95-140): PointOnLine through the centre, then
    FixedAngle (pinned) or HingeAngle about the axis (full range), then (motor = (velocity, max force))
    a LinearMotor along the axis or (servo = (target position, max force), JointDrive.Servo) a
    LinearAxisServo along it."""
    out = [point_on_line(bodies, b1, b2, axis, center, center, limit)]
    out.append(fixed_angle(bodies, b1, b2) if pinned else hinge_angle(bodies, b1, b2, axis))
    if motor is not None:
        out.append(linear_motor(bodies, b1, b2, axis, axis, motor[0], motor[1]))
    elif servo is not None:
        out.append(linear_axis_servo(bodies, b1, b2, axis, center, center, servo[0], servo[1]))
    return out


def universal_joint(bodies, b1, b2, center, axis1, axis2, motor=None, servo=None):
    """UniversalJoint (Dynamics/Joints/UniversalJoint.cs:30-72): TwistAngle between the two rotation axes
    (fixed), BallSocket at the centre, then (motor = (velocity, max torque)) an AngularMotor or
    (servo = max torque, JointDrive.Servo) an AngularServo holding the initial relative orientation."""
    out = [twist_angle(bodies, b1, b2, axis1, axis2), ball_socket(bodies, b1, b2, center)]
    if motor is not None:
        out.append(angular_motor(bodies, b1, b2, axis1, axis2, motor[0], motor[1]))
    elif servo is not None:
        out.append(angular_servo(bodies, b1, b2, max_torque=servo))
    return out


def ragdoll_joint(bodies, b1, b2, anchor, axis, cone=math.pi / 4, twist=(-math.pi / 6, math.pi / 6)):
    """A ragdoll shoulder: BallSocket at the anchor, ConeLimit of the bone axis, TwistAngle about it."""
    return [ball_socket(bodies, b1, b2, anchor), cone_limit(bodies, b1, b2, axis, limit=(0.0, cone)),
            twist_angle(bodies, b1, b2, axis, axis, twist)]


def hinge_joint(bodies, b1, b2, anchor, axis, limit=(-math.pi, math.pi), motor=None, servo=None):
    """HingeJoint (Dynamics/Joints/HingeJoint.cs:58-77): HingeAngle about the axis, then BallSocket at
    the anchor, then (motor = (velocity, max torque)) an AngularMotor about the axis or (servo = (target
    angle, max torque), JointDrive.Servo) a TwistServo about it — the order the joint creates them,
    which is the solve order."""
    out = [hinge_angle(bodies, b1, b2, axis, limit), ball_socket(bodies, b1, b2, anchor)]
    if motor is not None:
        out.append(angular_motor(bodies, b1, b2, axis, axis, motor[0], motor[1]))
    elif servo is not None:
        out.append(twist_servo(bodies, b1, b2, axis, target=servo[0], max_torque=servo[1]))
    return out


def angular_servo(bodies, b1, b2, target=None, max_torque=0.0):
    """AngularServo.Initialize() / Initialize(target) (AngularServo.cs:53-73): drives body 2's orientation
    relative to body 1 toward target (x y z w; default the current relative orientation) with at most
    max_torque (+inf: unlimited)."""
    c = _record(CONSTRAINT_ANGULAR_SERVO, b1, b2)
    p = c["params"]
    if target is None:
        target = _qmul(_conj(bodies["orientation"][b1]), bodies["orientation"][b2])
    p[0:4] = np.asarray(target, np.float32)
    p[4] = f32(max_torque)
    p[5] = DEFAULT_ANGULAR_SOFTNESS
    p[6] = DEFAULT_ANGULAR_BIAS
    return c


def linear_servo(bodies, b1, b2, anchor1, anchor2=None, target=(0.0, 0.0, 0.0), max_force=0.0):
    """LinearServo.Initialize(anchor1, anchor2) (LinearServo.cs:55-71) with TargetPosition (body-1 frame
    offset from anchor 1 to anchor 2) and MaximumForce."""
    c = _record(CONSTRAINT_LINEAR_SERVO, b1, b2)
    p = c["params"]
    p[0:3] = _local_anchor(bodies, b1, anchor1)
    p[3:6] = _local_anchor(bodies, b2, anchor1 if anchor2 is None else anchor2)
    p[6:9] = np.asarray(target, np.float32)
    p[9] = f32(max_force)
    p[10] = DEFAULT_LINEAR_SOFTNESS
    p[11] = DEFAULT_LINEAR_BIAS
    return c


def linear_axis_servo(bodies, b1, b2, axis, anchor1, anchor2, target=0.0, max_force=0.0):
    """LinearAxisServo.Initialize(axis, anchor1, anchor2, targetPosition) (LinearAxisServo.cs:61-78): drives
    the signed separation of anchor 2 from anchor 1 along axis (body-1 frame) toward target."""
    c = _record(CONSTRAINT_LINEAR_AXIS_SERVO, b1, b2)
    p = c["params"]
    p[0:3] = _rot_inv(_normalize(axis), bodies["orientation"][b1])
    p[3:6] = _local_anchor(bodies, b1, anchor1)
    p[6:9] = _local_anchor(bodies, b2, anchor2)
    p[9] = f32(target)
    p[10] = f32(max_force)
    p[11] = DEFAULT_LINEAR_SOFTNESS
    p[12] = DEFAULT_LINEAR_BIAS
    return c


def twist_servo(bodies, b1, b2, axis1, axis2=None, target=0.0, max_torque=0.0):
    """TwistServo.Initialize(axis1, axis2, targetAngle) (TwistServo.cs:69-88): drives the twist of body 2
    about axis2 relative to the initial pose toward target (radians in [-pi, pi])."""
    if not -math.pi <= target <= math.pi:
        raise ValueError("twist servo target must be in [-pi, pi]")
    c = _record(CONSTRAINT_TWIST_SERVO, b1, b2)
    p = c["params"]
    q1, q2 = bodies["orientation"][b1], bodies["orientation"][b2]
    local2 = _rot_inv(_normalize(axis1 if axis2 is None else axis2), q2)
    first_in_b2 = _rot_inv(_normalize(axis1), q2)
    p[0:3] = local2
    p[3:7] = _qmul(_from_to(first_in_b2, local2), _qmul(_conj(q2), q1))
    p[7] = f32(target)
    p[8] = f32(max_torque)
    p[9] = DEFAULT_ANGULAR_SOFTNESS
    p[10] = DEFAULT_ANGULAR_BIAS
    return c


def weld_joint(bodies, b1, b2, anchor):
    """WeldJoint (Dynamics/Joints/WeldJoint.cs:38-43): FixedAngle, then BallSocket at the anchor."""
    return [fixed_angle(bodies, b1, b2), ball_socket(bodies, b1, b2, anchor)]


def as_array(records) -> np.ndarray:
    return np.array(records, dtype=CONSTRAINT_DTYPE) if len(records) else np.zeros(0, CONSTRAINT_DTYPE)
