"""numpy layouts of the C-ABI structs declared in include/jitterhip.h, and the library loader.

The dtypes mirror the C structs byte for byte (checked against ``jh_struct_size`` exported by
the library).  Field meanings follow the reference types they restate:
``jh_body`` ↔ RigidBodyData + RigidBody fields (src/Jitter2/Dynamics/RigidBody.cs:107-250, 398-409),
``jh_contact``/``jh_arbiter`` ↔ Contact/ContactData (src/Jitter2/Dynamics/Contact.cs:30-118, 617-661).
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

MOTION_DYNAMIC, MOTION_KINEMATIC, MOTION_STATIC = 0, 1, 2
LOCK_LINEAR_X, LOCK_LINEAR_Y, LOCK_LINEAR_Z, LOCK_ANGULAR_X, LOCK_ANGULAR_Y, LOCK_ANGULAR_Z = 1, 2, 4, 8, 16, 32
LOCK_LINEAR, LOCK_ANGULAR = 7, 56
SHAPE_BOX, SHAPE_SPHERE, SHAPE_CAPSULE, SHAPE_POINTCLOUD, SHAPE_TRIANGLE, SHAPE_CONVEXHULL = 0, 1, 2, 3, 4, 5

JH_OK = 0
ERRORS = {
    -1: "invalid argument",
    -2: "non-finite value",
    -3: "capacity exceeded",
    -4: "HIP runtime error",
    -5: "unsupported feature",
    -6: "invalid state",
    -7: "no HIP device",
    -8: "device wait timed out or step cancelled",
}
JH_ERR_TIMEOUT = -8

CONFIG_DTYPE = np.dtype([
    ("gravity", "<f4", 3),
    ("substep_count", "<i4"),
    ("solver_iterations", "<i4"),
    ("relaxation_iterations", "<i4"),
    ("speculative_relaxation_factor", "<f4"),
    ("enable_auxiliary_contact_points", "<i4"),
    ("persistent_contact_manifold", "<i4"),
    ("pair_capacity", "<i4"),
    ("warm_start", "<i4"),
    ("narrowphase_filter", "<i4"),
    ("speculative_velocity_threshold", "<f4"),
    ("allow_deactivation", "<i4"),
    ("reserved", "<i4", 2),
])

FILTER_NONE, FILTER_TRIANGLE_EDGE = 0, 1

WARM_START_PREVIOUS, WARM_START_LEARNED = 0, 1
LWS_FEATURES, LWS_PARAMS, PROBE_MAX_ITERATIONS = 24, 20099, 256
LWS_PREDICT_MAX_ROWS = 65536  # JH_LWS_PREDICT_MAX_ROWS
PROBE_PREVIOUS, PROBE_LEARNED, PROBE_COLD = 1, 2, 4  # jh_probe.modes bits

PROBE_DTYPE = np.dtype([
    ("iterations", "<i4"),
    ("modes", "<i4"),
    ("capacity", "<i4"),
    ("reserved", "<i4"),
    ("features", "<u8"),
    ("targets", "<u8"),
    ("learned", "<u8"),
    ("points", "<i8"),
    ("rows", "<i8"),
    ("residual", "<f8", (3, PROBE_MAX_ITERATIONS + 1)),
])

BODY_DTYPE = np.dtype([
    ("position", "<f4", 3),
    ("orientation", "<f4", 4),
    ("velocity", "<f4", 3),
    ("angular_velocity", "<f4", 3),
    ("delta_velocity", "<f4", 3),
    ("delta_angular_velocity", "<f4", 3),
    ("inverse_inertia", "<f4", 6),
    ("inverse_mass", "<f4"),
    ("friction", "<f4"),
    ("restitution", "<f4"),
    ("linear_damping_multiplier", "<f4"),
    ("angular_damping_multiplier", "<f4"),
    ("force", "<f4", 3),
    ("torque", "<f4", 3),
    ("motion_type", "<i4"),
    ("affected_by_gravity", "<i4"),
    ("world_id", "<i4"),
    ("enable_speculative_contacts", "<i4"),
    ("deactivation_threshold_linear_sq", "<f4"),
    ("deactivation_threshold_angular_sq", "<f4"),
    ("deactivation_time", "<f4"),
    ("locked_axes", "<i4"),
])

SHAPE_DTYPE = np.dtype([
    ("shape_id", "<u8"),
    ("body", "<i4"),
    ("type", "<i4"),
    ("params", "<f4", 12),
    ("vertex_offset", "<i4"),
    ("vertex_count", "<i4"),
])

CONTACT_DTYPE = np.dtype([
    ("flag", "<u4"),
    ("bias", "<f4"),
    ("penalty_bias", "<f4"),
    ("normal", "<f4", 3),
    ("tangent1", "<f4", 3),
    ("tangent2", "<f4", 3),
    ("mass_normal_tangent", "<f4", 3),
    ("accumulated", "<f4", 3),
    ("position1", "<f4", 3),
    ("position2", "<f4", 3),
    ("relative_position1", "<f4", 3),
    ("relative_position2", "<f4", 3),
])

ARBITER_DTYPE = np.dtype([
    ("shape1", "<u4"),
    ("shape2", "<u4"),
    ("body1", "<i4"),
    ("body2", "<i4"),
    ("usage_mask", "<u4"),
    ("mode", "<u4"),
    ("friction", "<f4"),
    ("restitution", "<f4"),
    ("speculative_relaxation_factor", "<f4"),
    ("reserved", "<u4"),
    ("contacts", CONTACT_DTYPE, 4),
])

MANIFOLD_DTYPE = np.dtype([
    ("shape1", "<u4"),
    ("shape2", "<u4"),
    ("count", "<i4"),
    ("flags", "<i4"),
    ("normal", "<f4", 3),
    ("penetration", "<f4"),
    ("point_a", "<f4", 3),
    ("point_b", "<f4", 3),
    ("manifold_a", "<f4", (4, 3)),
    ("manifold_b", "<f4", (4, 3)),
])

STATS_DTYPE = np.dtype([
    ("bodies", "<i8"), ("shapes", "<i8"), ("pairs", "<i8"), ("colliding_pairs", "<i8"),
    ("arbiters", "<i8"), ("contacts", "<i8"), ("colours", "<i8"), ("epa_calls", "<i8"),
    ("steps", "<i8"), ("pair_overflow", "<i8"), ("arbiters_solved", "<i8"),
    ("ms_broadphase", "<f8"), ("ms_narrowphase", "<f8"), ("ms_contact_cache", "<f8"),
    ("ms_colouring", "<f8"), ("ms_solve", "<f8"), ("ms_update", "<f8"), ("ms_step", "<f8"), ("ms_epa", "<f8"),
    ("solver", "<i8"), ("max_degree", "<i8"), ("ms_solve_kernel", "<f8"), ("bodies_asleep", "<i8"),
])

TUNING_DTYPE = np.dtype([
    ("solver", "<i4"), ("solve_blocks", "<i4"), ("colour_blocks", "<i4"), ("candidates", "<i4"), ("steps", "<i4"),
    ("traversal", "<i4"),
    ("ms_solve_best", "<f8"), ("ms_solve_default", "<f8"), ("ms_colour_best", "<f8"), ("ms_colour_default", "<f8"),
    ("ms_broadphase_best", "<f8"), ("ms_broadphase_default", "<f8"),
])

CONSTRAINT_BALL_SOCKET, CONSTRAINT_HINGE_ANGLE, CONSTRAINT_FIXED_ANGLE, CONSTRAINT_DISTANCE_LIMIT = 0, 1, 2, 3
CONSTRAINT_TWIST_ANGLE, CONSTRAINT_CONE_LIMIT, CONSTRAINT_ANGULAR_MOTOR, CONSTRAINT_LINEAR_MOTOR = 4, 5, 6, 7
CONSTRAINT_POINT_ON_LINE, CONSTRAINT_POINT_ON_PLANE = 8, 9
CONSTRAINT_ANGULAR_SERVO, CONSTRAINT_LINEAR_SERVO, CONSTRAINT_LINEAR_AXIS_SERVO, CONSTRAINT_TWIST_SERVO = 10, 11, 12, 13
CONSTRAINT_ELLIPTICAL_CONE_LIMIT = 14

CONSTRAINT_DTYPE = np.dtype([
    ("type", "<i4"), ("body1", "<i4"), ("body2", "<i4"), ("enabled", "<i4"),
    ("params", "<f4", 16), ("state", "<f4", 8),
])

STRUCT_SIZES = {
    0: CONFIG_DTYPE.itemsize, 1: BODY_DTYPE.itemsize, 2: SHAPE_DTYPE.itemsize,
    3: CONTACT_DTYPE.itemsize, 4: ARBITER_DTYPE.itemsize, 5: MANIFOLD_DTYPE.itemsize,
    6: STATS_DTYPE.itemsize, 7: PROBE_DTYPE.itemsize, 8: TUNING_DTYPE.itemsize, 9: CONSTRAINT_DTYPE.itemsize,
}

PKG_DIR = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(PKG_DIR, "libjitterhip.so")


def default_config() -> np.ndarray:
    """World defaults (World.cs:306-412, World.Detect.cs:88-122)."""
    c = np.zeros(1, CONFIG_DTYPE)
    c["gravity"] = (0.0, -9.81, 0.0)
    c["substep_count"] = 1
    c["solver_iterations"] = 6
    c["relaxation_iterations"] = 4
    c["speculative_relaxation_factor"] = 0.9
    c["enable_auxiliary_contact_points"] = 1
    c["persistent_contact_manifold"] = 1
    c["pair_capacity"] = 0
    c["speculative_velocity_threshold"] = 10.0
    c["allow_deactivation"] = 1
    return c


def ptr(a: np.ndarray | None):
    """ctypes void pointer to a contiguous numpy array (or NULL).  The pointer object holds a
    reference to the array, so temporaries stay alive for the duration of the call."""
    if a is None:
        return None
    assert a.flags["C_CONTIGUOUS"], "array must be C-contiguous"
    return a.ctypes.data_as(ctypes.c_void_p)


_lib = None


def load_library(path: str | None = None) -> ctypes.CDLL:
    """Load libjitterhip.so (the HIP implementation).  Raises if it has not been built:
    there is no CPU fallback for the product path."""
    global _lib
    if _lib is not None and path is None:
        return _lib
    p = path or LIB_PATH
    if not os.path.exists(p):
        raise RuntimeError(
            f"{p} is missing: build the HIP library first (python -c 'import __graft_entry__ as g; g.build()')")
    lib = ctypes.CDLL(p)
    i32, vp = ctypes.c_int32, ctypes.c_void_p
    pi32 = ctypes.POINTER(ctypes.c_int32)
    sigs = {
        "jh_last_error": ([], ctypes.c_char_p),
        "jh_version": ([], i32),
        "jh_struct_size": ([i32], i32),
        "jh_world_create": ([vp, ctypes.POINTER(vp)], i32),
        "jh_world_destroy": ([vp], i32),
        "jh_world_set_config": ([vp, vp], i32),
        "jh_config_defaults": ([vp], i32),
        "jh_world_set_bodies": ([vp, vp, i32], i32),
        "jh_world_get_bodies": ([vp, vp, i32], i32),
        "jh_world_set_shapes": ([vp, vp, i32, vp, i32], i32),
        "jh_world_get_rigidbodydata": ([vp, vp, i32, i32], i32),
        "jh_world_set_rigidbodydata": ([vp, vp, i32, i32], i32),
        "jh_world_step": ([vp, ctypes.c_float], i32),
        "jh_world_set_spin_limit": ([vp, ctypes.c_uint32], i32),
        "jh_world_stabilize": ([vp, ctypes.c_float, i32, i32], i32),
        "jh_world_update_bodies": ([vp, vp, vp, i32], i32),
        "jh_world_get_activation": ([vp, vp, vp, i32], i32),
        "jh_world_set_activation_state": ([vp, i32, i32], i32),
        "jh_world_cancel": ([vp], i32),
        "jh_world_compute_pairs": ([vp, vp, i32, pi32], i32),
        "jh_world_get_manifolds": ([vp, vp, i32, pi32], i32),
        "jh_world_get_arbiters": ([vp, vp, i32, pi32], i32),
        "jh_world_set_arbiters": ([vp, vp, i32], i32),
        "jh_world_get_solve_order": ([vp, vp, i32, vp, i32, pi32, pi32], i32),
        "jh_world_get_collision_events": ([vp, vp, i32, pi32, vp, i32, pi32], i32),
        "jh_world_set_constraints": ([vp, vp, i32], i32),
        "jh_world_get_constraints": ([vp, vp, i32, pi32], i32),
        "jh_world_get_constraint_order": ([vp, vp, i32, vp, i32, pi32, pi32], i32),
        "jh_world_get_stats": ([vp, vp], i32),
        "jh_lws_create": ([ctypes.c_uint64, ctypes.POINTER(vp)], i32),
        "jh_lws_destroy": ([vp], i32),
        "jh_lws_set_params": ([vp, vp, i32], i32),
        "jh_lws_get_params": ([vp, vp, i32], i32),
        "jh_lws_set_normalisation": ([vp, vp, vp, vp], i32),
        "jh_lws_set_data": ([vp, vp, vp, ctypes.c_int64], i32),
        "jh_lws_train": ([vp, ctypes.c_int64, ctypes.c_int64, i32, i32, ctypes.c_float, vp], i32),
        "jh_lws_evaluate": ([vp, ctypes.c_int64, ctypes.c_int64, vp], i32),
        "jh_lws_gradient": ([vp, ctypes.c_int64, ctypes.c_int64, vp, vp], i32),
        "jh_lws_predict": ([vp, vp, ctypes.c_int64, vp], i32),
        "jh_lws_predict_device": ([vp, vp, ctypes.c_int64, vp], i32),
        "jh_world_set_warm_start_model": ([vp, vp], i32),
        "jh_world_step_probe": ([vp, ctypes.c_float, vp], i32),
        "jh_world_autotune": ([vp, ctypes.c_float, i32, vp], i32),
    }
    for name, (args, res) in sigs.items():
        f = getattr(lib, name)
        f.argtypes = args
        f.restype = res
    for k, size in STRUCT_SIZES.items():
        got = lib.jh_struct_size(k)
        if got != size:
            raise RuntimeError(f"struct {k}: library size {got} != numpy layout {size}")
    if path is None:
        _lib = lib
    return lib


def check(lib, status: int, what: str = "") -> None:
    if status != JH_OK:
        msg = lib.jh_last_error().decode(errors="replace")
        raise RuntimeError(f"{what}: {ERRORS.get(status, status)}: {msg}")
