"""GpuWorld: thin Python handle over the C-ABI of libjitterhip.so (include/jitterhip.h).

All compute runs in the HIP library; this module only moves arrays across the boundary.  It
raises if the library is missing — there is no CPU fallback on the product path.
"""
from __future__ import annotations

import ctypes

import numpy as np

from ._abi import (ARBITER_DTYPE, BODY_DTYPE, CONFIG_DTYPE, CONSTRAINT_DTYPE, LWS_FEATURES, MANIFOLD_DTYPE, PROBE_DTYPE,
                   STATS_DTYPE, TUNING_DTYPE, check, default_config, load_library, ptr)


class GpuWorld:
    """One jh_world: device-resident bodies, shapes and contact cache, stepped by World.Step
    semantics (src/Jitter2/World.Step.cs:141-296)."""

    def __init__(self, bodies: np.ndarray, shapes: np.ndarray, verts: np.ndarray | None = None,
                 config: np.ndarray | None = None):
        self.lib = load_library()
        self.handle = ctypes.c_void_p()
        cfg = default_config() if config is None else np.ascontiguousarray(config, CONFIG_DTYPE)
        check(self.lib, self.lib.jh_world_create(ptr(cfg), ctypes.byref(self.handle)), "jh_world_create")
        self.n_bodies = 0
        self.set_bodies(bodies)
        self.set_shapes(shapes, verts)

    def close(self) -> None:
        if getattr(self, "handle", None) and self.handle.value:
            self.lib.jh_world_destroy(self.handle)
            self.handle = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_config(self, config: np.ndarray) -> None:
        c = np.ascontiguousarray(config, CONFIG_DTYPE)
        check(self.lib, self.lib.jh_world_set_config(self.handle, ptr(c)), "jh_world_set_config")

    def set_bodies(self, bodies: np.ndarray) -> None:
        b = np.ascontiguousarray(bodies, BODY_DTYPE)
        check(self.lib, self.lib.jh_world_set_bodies(self.handle, ptr(b), len(b)), "jh_world_set_bodies")
        self.n_bodies = len(b)

    def update_bodies(self, index, records) -> None:
        """Replace the records of the listed bodies (jh_world_update_bodies): RigidBody property
        setters on those bodies, each woken for the next step."""
        idx = np.ascontiguousarray(index, np.int32).reshape(-1)
        rec = np.ascontiguousarray(records, BODY_DTYPE).reshape(-1)
        check(self.lib, self.lib.jh_world_update_bodies(self.handle, ptr(idx), ptr(rec), len(idx)),
              "jh_world_update_bodies")

    def activation(self) -> tuple[np.ndarray, np.ndarray]:
        """(RigidBody.IsActive as bool, sleep timer as float32) per body (jh_world_get_activation)."""
        act = np.zeros(self.n_bodies, np.uint8)
        sl = np.zeros(self.n_bodies, np.float32)
        check(self.lib, self.lib.jh_world_get_activation(self.handle, ptr(act), ptr(sl), self.n_bodies),
              "jh_world_get_activation")
        return act.astype(bool), sl

    def set_activation_state(self, body: int, active: bool) -> None:
        """RigidBody.SetActivationState (jh_world_set_activation_state), effective at the next step."""
        check(self.lib, self.lib.jh_world_set_activation_state(self.handle, int(body), 1 if active else 0),
              "jh_world_set_activation_state")

    def set_shapes(self, shapes: np.ndarray, verts: np.ndarray | None = None) -> None:
        s = np.ascontiguousarray(shapes)
        v = None if verts is None else np.ascontiguousarray(verts, np.float32).reshape(-1)
        nv = 0 if v is None else len(v) // 3
        check(self.lib, self.lib.jh_world_set_shapes(self.handle, ptr(s), len(s), ptr(v), nv), "jh_world_set_shapes")

    def step(self, dt: float) -> None:
        check(self.lib, self.lib.jh_world_step(self.handle, ctypes.c_float(dt)), "jh_world_step")

    def step_status(self, dt: float) -> int:
        """jh_world_step's status code, without raising."""
        return int(self.lib.jh_world_step(self.handle, ctypes.c_float(dt)))

    def stabilize(self, dt: float, iterations: int, relaxations: int = 0) -> None:
        """World.Stabilize (jh_world_stabilize): solve the held contacts and constraints at the velocity
        level, without detection or integration."""
        check(self.lib, self.lib.jh_world_stabilize(self.handle, ctypes.c_float(dt), iterations, relaxations),
              "jh_world_stabilize")

    def set_spin_limit(self, polls: int) -> None:
        check(self.lib, self.lib.jh_world_set_spin_limit(self.handle, ctypes.c_uint32(polls)), "jh_world_set_spin_limit")

    def cancel(self) -> None:
        check(self.lib, self.lib.jh_world_cancel(self.handle), "jh_world_cancel")

    def bodies(self) -> np.ndarray:
        out = np.zeros(self.n_bodies, BODY_DTYPE)
        check(self.lib, self.lib.jh_world_get_bodies(self.handle, ptr(out), self.n_bodies), "jh_world_get_bodies")
        return out

    def rigidbodydata(self, stride: int = 128) -> np.ndarray:
        out = np.zeros(self.n_bodies * stride, np.uint8)
        check(self.lib, self.lib.jh_world_get_rigidbodydata(self.handle, ptr(out), self.n_bodies, stride),
              "jh_world_get_rigidbodydata")
        return out.reshape(self.n_bodies, stride)

    def set_rigidbodydata(self, records: np.ndarray) -> None:
        """Write RawData records (n, stride) uint8 back (jh_world_set_rigidbodydata)."""
        r = np.ascontiguousarray(records, np.uint8)
        check(self.lib, self.lib.jh_world_set_rigidbodydata(self.handle, ptr(r), r.shape[0], r.shape[1]),
              "jh_world_set_rigidbodydata")

    def compute_pairs(self) -> np.ndarray:
        n = ctypes.c_int32(0)
        self.lib.jh_world_compute_pairs(self.handle, None, 0, ctypes.byref(n))
        out = np.zeros((max(n.value, 1), 2), np.uint32)
        check(self.lib, self.lib.jh_world_compute_pairs(self.handle, ptr(out), n.value, ctypes.byref(n)),
              "jh_world_compute_pairs")
        return out[:n.value]

    def manifolds(self) -> np.ndarray:
        n = ctypes.c_int32(0)
        self.lib.jh_world_get_manifolds(self.handle, None, 0, ctypes.byref(n))
        out = np.zeros(max(n.value, 1), MANIFOLD_DTYPE)
        check(self.lib, self.lib.jh_world_get_manifolds(self.handle, ptr(out), n.value, ctypes.byref(n)),
              "jh_world_get_manifolds")
        return out[:n.value]

    def arbiters(self) -> np.ndarray:
        n = ctypes.c_int32(0)
        self.lib.jh_world_get_arbiters(self.handle, None, 0, ctypes.byref(n))
        out = np.zeros(max(n.value, 1), ARBITER_DTYPE)
        check(self.lib, self.lib.jh_world_get_arbiters(self.handle, ptr(out), n.value, ctypes.byref(n)),
              "jh_world_get_arbiters")
        return out[:n.value]

    def set_constraints(self, constraints: np.ndarray) -> None:
        c = np.ascontiguousarray(constraints, CONSTRAINT_DTYPE)
        check(self.lib, self.lib.jh_world_set_constraints(self.handle, ptr(c), len(c)), "jh_world_set_constraints")

    def constraints(self) -> np.ndarray:
        n = ctypes.c_int32(0)
        self.lib.jh_world_get_constraints(self.handle, None, 0, ctypes.byref(n))
        out = np.zeros(max(n.value, 1), CONSTRAINT_DTYPE)
        check(self.lib, self.lib.jh_world_get_constraints(self.handle, ptr(out), n.value, ctypes.byref(n)),
              "jh_world_get_constraints")
        return out[:n.value]

    def constraint_order(self) -> tuple[np.ndarray, np.ndarray]:
        """(constraint indices in solve order, batch offsets) of the enabled constraints."""
        n, nb = ctypes.c_int32(0), ctypes.c_int32(0)
        self.lib.jh_world_get_constraint_order(self.handle, None, 0, None, 0, ctypes.byref(n), ctypes.byref(nb))
        order = np.zeros(max(n.value, 1), np.int32)
        offs = np.zeros(nb.value + 1, np.int32)
        check(self.lib, self.lib.jh_world_get_constraint_order(self.handle, ptr(order), n.value, ptr(offs),
                                                               nb.value + 1, ctypes.byref(n), ctypes.byref(nb)),
              "jh_world_get_constraint_order")
        return order[:n.value], offs

    def collision_events(self) -> tuple[np.ndarray, np.ndarray]:
        """(BeginCollide pairs, EndCollide pairs) of the last step, (K, 2) uint32 shape indices in key
        order (RigidBody.BeginCollide / EndCollide with the reference's arbiter lifetime)."""
        nb, ne = ctypes.c_int32(0), ctypes.c_int32(0)
        self.lib.jh_world_get_collision_events(self.handle, None, 0, ctypes.byref(nb), None, 0, ctypes.byref(ne))
        b = np.zeros((max(nb.value, 1), 2), np.uint32)
        e = np.zeros((max(ne.value, 1), 2), np.uint32)
        check(self.lib, self.lib.jh_world_get_collision_events(self.handle, ptr(b), nb.value, ctypes.byref(nb), ptr(e),
                                                               ne.value, ctypes.byref(ne)),
              "jh_world_get_collision_events")
        return b[:nb.value], e[:ne.value]

    def set_arbiters(self, arbiters: np.ndarray) -> None:
        a = np.ascontiguousarray(arbiters, ARBITER_DTYPE)
        check(self.lib, self.lib.jh_world_set_arbiters(self.handle, ptr(a), len(a)), "jh_world_set_arbiters")

    def solve_order(self) -> tuple[np.ndarray, np.ndarray]:
        """(keys (K, 2) uint32 in solve order, batch offsets (B+1,)) of the last step."""
        nk, nb = ctypes.c_int32(0), ctypes.c_int32(0)
        self.lib.jh_world_get_solve_order(self.handle, None, 0, None, 0, ctypes.byref(nk), ctypes.byref(nb))
        keys = np.zeros((max(nk.value, 1), 2), np.uint32)
        offs = np.zeros(nb.value + 1, np.int32)
        check(self.lib, self.lib.jh_world_get_solve_order(self.handle, ptr(keys), nk.value, ptr(offs), nb.value + 1,
                                                          ctypes.byref(nk), ctypes.byref(nb)), "jh_world_get_solve_order")
        return keys[:nk.value], offs

    def stats(self) -> dict:
        s = np.zeros(1, STATS_DTYPE)
        check(self.lib, self.lib.jh_world_get_stats(self.handle, ptr(s)), "jh_world_get_stats")
        return {k: s[k][0].item() for k in STATS_DTYPE.names}

    def set_warm_start_model(self, model) -> None:
        """Copy a LearnedWarmStart model into the world (used when config.warm_start = 1)."""
        check(self.lib, self.lib.jh_world_set_warm_start_model(self.handle, model.handle),
              "jh_world_set_warm_start_model")

    def step_probe(self, dt: float, iterations: int, modes: int, capacity: int = 0) -> dict:
        """One step with residual probes of its first substep (jh_world_step_probe).
        Returns residual (3, iterations + 1) sums of squares per mode (previous, learned, cold),
        points, and with capacity > 0 the trace rows: features (rows, 24), targets (rows, 3) and, with
        the learned mode, the network's warm-start impulses (rows, 3)."""
        pr = np.zeros(1, PROBE_DTYPE)
        pr["iterations"] = iterations
        pr["modes"] = modes
        pr["capacity"] = capacity
        feats = np.zeros((max(capacity, 1), LWS_FEATURES), np.float32)
        targs = np.zeros((max(capacity, 1), 3), np.float32)
        learned = np.zeros((max(capacity, 1), 3), np.float32)
        if capacity > 0:
            pr["features"] = feats.ctypes.data
            pr["targets"] = targs.ctypes.data
            pr["learned"] = learned.ctypes.data
        check(self.lib, self.lib.jh_world_step_probe(self.handle, ctypes.c_float(dt), ptr(pr)), "jh_world_step_probe")
        rows = int(pr["rows"][0])
        return {"residual": pr["residual"][0][:, :iterations + 1].copy(), "points": int(pr["points"][0]),
                "features": feats[:rows], "targets": targs[:rows], "learned": learned[:rows]}

    def autotune(self, dt: float, rounds: int = 2) -> dict:
        """Pick the solver and colouring workgroup counts by timing candidates on real steps
        (jh_world_autotune; advances the simulation)."""
        t = np.zeros(1, TUNING_DTYPE)
        check(self.lib, self.lib.jh_world_autotune(self.handle, ctypes.c_float(dt), rounds, ptr(t)), "jh_world_autotune")
        return {k: t[k][0].item() for k in TUNING_DTYPE.names}
