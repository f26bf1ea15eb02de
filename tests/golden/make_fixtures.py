"""Writes the golden fixtures of this directory with the oracle (oracle/liborc.so, the CPU restatement).

    python tests/golden/make_fixtures.py

step0_<scene>.npz — for scenes whose first step already has contacts: the broadphase pair set at the
    initial state (brute force over world AABBs, World.DefaultDynamicTreeFilter semantics) and the
    contact manifolds of the first step (normal, depth, points; EPA, the triangle edge filter and
    speculative contacts included).  Both are fixed by the initial state alone, whatever order the solver later uses, so
    the GPU must reproduce them bit for bit (tests/test_golden.py).
oracle_trajectory.npz — FNV-1a hashes of all body positions (FP32 bits, as the reference's
    ReproducibilityTest hashes them, Robustness/ReproducibilityTest.cs:107-147) while the oracle steps
    a 6-row pyramid in its own arbiter-key order: a regression anchor for the oracle's time evolution.
Each fixture records a hash of the scene arrays it was made from, so a changed scene builder is
reported as such.
"""
from __future__ import annotations

import hashlib
import os
import sys

import numpy as np
import numpy.lib.recfunctions as rfn

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

from jitterphysics2_amd import scenes  # noqa: E402


# Body fields outside the hash: the sleep parameters, which do not act in scenes that run without
# deactivation (the config's allow_deactivation is hashed), and the motion locks, which must be zero.
SLEEP_FIELDS = ("deactivation_threshold_linear_sq", "deactivation_threshold_angular_sq", "deactivation_time",
                "locked_axes")


def scene_hash(scene) -> str:
    bodies, shapes, verts, cfg, dt = scene
    h = hashlib.sha256()
    v = np.zeros(0, np.float32) if verts is None else np.asarray(verts, np.float32)
    b = np.ascontiguousarray(bodies)
    assert not np.any(b["locked_axes"]), "fixture scenes have no motion locks"
    b = rfn.repack_fields(b[[n for n in b.dtype.names if n not in SLEEP_FIELDS]])
    for a in (b, shapes, v, cfg, np.float32(dt)):
        h.update(np.ascontiguousarray(a).tobytes())
    return h.hexdigest()


def fnv1a_positions(bodies) -> int:
    """FNV-1a 64 over the FP32 bytes of every body position, in body order."""
    h = 0xCBF29CE484222325
    for byte in np.ascontiguousarray(bodies["position"], np.float32).tobytes():
        h ^= byte
        h = (h * 0x100000001B3) & 0xFFFFFFFFFFFFFFFF
    return h


def overlapping_heap():
    """400 mixed shapes on a 1.0 m grid: neighbours interpenetrate at the start (deep pairs, EPA)."""
    return scenes.heap(400, seed=0x4A2D0011, pitch=1.0)


def sunk_heightfield():
    """The full form at small scale (4 sub-worlds, 32x32-cell Perlin tiles, hulls), every dynamic body
    lowered so that the first layer starts inside the terrain (triangle pairs, edge filter, EPA)."""
    bodies, shapes, verts, cfg, dt = scenes.full_form(n_bodies=240, cells=32, worlds=4, feature=8.0, pitch=2.4)
    dyn = bodies["motion_type"] == scenes.MOTION_DYNAMIC
    bodies["position"][dyn, 1] -= np.float32(2.0)
    return bodies, shapes, verts, cfg, dt


def bullets_mesh():
    """Projectiles at 15-60 m/s over a triangle-mesh floor with the edge filter, half of them with
    speculative contacts: the first step's swept broadphase boxes and speculative manifolds."""
    return scenes.bullets(300, seed=0x4A2D0017, speculative=0.5, mesh=True)


STEP0_SCENES = {
    "pyramid20": lambda: scenes.pyramid(rows=20),
    "overlap_heap400": overlapping_heap,
    "sunk_heightfield240": sunk_heightfield,
    "bullets_mesh300": bullets_mesh,
}


def step0(scene):
    from tests.orc import OracleWorld

    bodies, shapes, verts, cfg, dt = scene
    ow = OracleWorld(bodies, shapes, verts, cfg)
    pairs = ow.pairs(brute=True)
    ow.step(dt)
    return pairs, ow.manifolds(), ow.epa_calls()


def trajectory(steps=(10, 25, 50, 75)):
    from tests.orc import OracleWorld

    bodies, shapes, verts, cfg, dt = scenes.pyramid(rows=6)
    ow = OracleWorld(bodies, shapes, verts, cfg)
    out = []
    for s in range(1, max(steps) + 1):
        ow.step(dt)
        if s in steps:
            out.append((s, fnv1a_positions(ow.bodies())))
    return out


def main() -> None:
    for name, make in STEP0_SCENES.items():
        scene = make()
        pairs, man, epa = step0(scene)
        np.savez_compressed(os.path.join(HERE, f"step0_{name}.npz"), pairs=pairs, manifolds=man,
                            epa_calls=np.int64(epa), scene_sha256=np.bytes_(scene_hash(scene)))
        print(f"step0_{name}: {len(pairs)} pairs, {len(man)} manifolds, {epa} EPA calls")
    traj = trajectory()
    np.savez_compressed(os.path.join(HERE, "oracle_trajectory.npz"), steps=np.array([s for s, _ in traj], np.int64),
                        fnv1a=np.array([h for _, h in traj], np.uint64),
                        scene_sha256=np.bytes_(scene_hash(scenes.pyramid(rows=6))))
    print("oracle_trajectory:", [(s, hex(h)) for s, h in traj])


if __name__ == "__main__":
    main()
