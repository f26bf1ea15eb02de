"""Golden fixtures (tests/golden/, written by tests/golden/make_fixtures.py with the oracle).

CPU: the oracle still reproduces every fixture bit for bit (a regression anchor for the restatement:
pair sets, first-step manifolds including EPA and the triangle edge filter, and the FNV-1a position
hashes of a pyramid stepped in the oracle's own order).
GPU: the first step of each fixture scene — pairs and manifolds depend only on the initial state, not
on the solve order — is bit-identical to the fixture.
"""
import os

import numpy as np
import pytest

from jitterphysics2_amd import scenes
from tests.golden import make_fixtures as mf
from tests.orc import OracleWorld

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
MANIFOLD_FIELDS = ("shape1", "shape2", "count", "flags", "normal", "penetration", "point_a", "point_b")


def load(name):
    return np.load(os.path.join(GOLDEN, name), allow_pickle=False)


def assert_manifolds(got, want, what):
    assert len(got) == len(want), f"{what}: {len(got)} manifolds, fixture has {len(want)}"
    for f in MANIFOLD_FIELDS:
        assert np.array_equal(got[f], want[f]), f"{what}: manifold field {f} differs from the fixture"
    for i in range(4):
        m = want["count"] > i
        assert np.array_equal(got["manifold_a"][m, i], want["manifold_a"][m, i]), f"{what}: manifold_a[{i}]"
        assert np.array_equal(got["manifold_b"][m, i], want["manifold_b"][m, i]), f"{what}: manifold_b[{i}]"


@pytest.mark.parametrize("name", sorted(mf.STEP0_SCENES))
def test_oracle_reproduces_step0_fixture(name):
    fx = load(f"step0_{name}.npz")
    scene = mf.STEP0_SCENES[name]()
    assert mf.scene_hash(scene) == fx["scene_sha256"].item().decode(), "scene builder changed: regenerate"
    bodies, shapes, verts, cfg, dt = scene
    ow = OracleWorld(bodies, shapes, verts, cfg)
    assert np.array_equal(ow.pairs(brute=True), fx["pairs"])
    assert np.array_equal(ow.pairs(brute=False), fx["pairs"])  # sort-and-sweep equals brute force
    ow.step(dt)
    assert_manifolds(ow.manifolds(), fx["manifolds"], name)
    assert ow.epa_calls() == int(fx["epa_calls"])


def test_fixtures_cover_epa_and_triangle_pairs():
    assert int(load("step0_overlap_heap400.npz")["epa_calls"]) > 100
    hf = load("step0_sunk_heightfield240.npz")
    assert int(hf["epa_calls"]) > 0 and len(hf["manifolds"]) > 0


def test_oracle_reproduces_trajectory_hashes():
    fx = load("oracle_trajectory.npz")
    assert mf.scene_hash(scenes.pyramid(rows=6)) == fx["scene_sha256"].item().decode()
    got = mf.trajectory(tuple(int(s) for s in fx["steps"]))
    assert [h for _, h in got] == [int(h) for h in fx["fnv1a"]]


@pytest.mark.gpu
@pytest.mark.parametrize("name", sorted(mf.STEP0_SCENES))
def test_gpu_first_step_matches_fixture(name):
    from jitterphysics2_amd.gpu import GpuWorld

    fx = load(f"step0_{name}.npz")
    bodies, shapes, verts, cfg, dt = mf.STEP0_SCENES[name]()
    gw = GpuWorld(bodies, shapes, verts, cfg)
    assert np.array_equal(gw.compute_pairs(), fx["pairs"]), f"{name}: pair set differs from the fixture"
    gw.step(dt)
    assert_manifolds(gw.manifolds(), fx["manifolds"], name)
