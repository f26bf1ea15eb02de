"""GPU parity of the whole step against the oracle.

The GPU solves contacts in graph-colour order; the oracle replays the same order (downloaded with
jh_world_get_solve_order), so bodies, contacts and impulses must agree bit for bit in FP32
(tolerance 0: the device code keeps the reference's operation order, FP contraction is off and
division/sqrt are correctly rounded).  Pair sets are compared with the oracle's brute-force pair
set.  FP64 error (oracle in double ≙ Jitter2.Double) is reported alongside.
"""
import numpy as np
import pytest

from jitterphysics2_amd import scenes
from jitterphysics2_amd._abi import MOTION_STATIC, default_config
from jitterphysics2_amd.gpu import GpuWorld
from tests.orc import OracleWorld

pytestmark = pytest.mark.gpu

STATE = ("position", "orientation", "velocity", "angular_velocity", "delta_velocity", "delta_angular_velocity")


def assert_bodies_equal(g, o, what):
    for f in STATE:
        if not np.array_equal(g[f], o[f]):
            bad = np.argwhere(np.any(g[f] != o[f], axis=-1)).ravel()
            d = np.abs(g[f].astype(np.float64) - o[f]).max()
            raise AssertionError(f"{what}: {f} differs in {len(bad)} bodies (first {bad[:5]}), max abs {d:.3e}")


def assert_arbiters_equal(ga, oa, what):
    assert len(ga) == len(oa), f"{what}: arbiter count {len(ga)} != {len(oa)}"
    for f in ("shape1", "shape2", "body1", "body2", "usage_mask", "mode", "friction", "restitution"):
        assert np.array_equal(ga[f], oa[f]), f"{what}: arbiter field {f} differs"
    gc, oc = ga["contacts"], oa["contacts"]
    used = ((ga["usage_mask"][:, None] >> np.arange(4)[None, :]) & 1).astype(bool)
    for f in gc.dtype.names:
        a, b = gc[f][used], oc[f][used]
        assert np.array_equal(a, b), f"{what}: contact field {f} differs"


def assert_manifolds_equal(gm, om, what):
    assert len(gm) == len(om), f"{what}: manifold count {len(gm)} != {len(om)}"
    for f in ("shape1", "shape2", "count", "flags", "normal", "penetration", "point_a", "point_b"):
        assert np.array_equal(gm[f], om[f]), f"{what}: manifold field {f} differs"
    for i in range(4):
        m = gm["count"] > i
        assert np.array_equal(gm["manifold_a"][m, i], om["manifold_a"][m, i]), f"{what}: manifold_a[{i}]"
        assert np.array_equal(gm["manifold_b"][m, i], om["manifold_b"][m, i]), f"{what}: manifold_b[{i}]"


def run_replay(scene, steps, check_every=1, brute_pairs=False):
    bodies, shapes, verts, cfg, dt = scene
    gw = GpuWorld(bodies, shapes, verts, cfg)
    ow = OracleWorld(bodies, shapes, verts, cfg)
    for s in range(steps):
        pg = gw.compute_pairs()
        po = ow.pairs(brute=brute_pairs)
        assert np.array_equal(pg, po), f"step {s}: pair set differs ({len(pg)} vs {len(po)})"
        gw.step(dt)
        keys, offs = gw.solve_order()
        ow.step(dt, keys)
        if s % check_every == 0 or s == steps - 1:
            assert_manifolds_equal(gw.manifolds(), ow.manifolds(), f"step {s}")
            assert_bodies_equal(gw.bodies(), ow.bodies(), f"step {s}")
            assert_arbiters_equal(gw.arbiters(), ow.arbiters(), f"step {s}")
    return gw, ow


def test_pyramid_small_bitexact():
    run_replay(scenes.pyramid(rows=6), steps=20, brute_pairs=True)


@pytest.mark.parametrize("solver", ["default", "flow"])
def test_pyramid_20_rows_bitexact(monkeypatch, solver):
    # default: a world this small runs the world-local solver; flow: the device-wide dataflow solver
    if solver != "default":
        monkeypatch.setenv("JH_SOLVER", solver)
    gw, ow = run_replay(scenes.pyramid(rows=20), steps=30, check_every=10)
    st = gw.stats()
    assert st["arbiters"] > 300 and st["colours"] >= 2
    assert st["solver"] == (2 if solver == "default" else 1)


def test_heap_mixed_shapes_bitexact(monkeypatch):
    # spheres, boxes and capsules with random orientations; deep impacts exercise EPA
    monkeypatch.setenv("JH_SOLVER", "flow")
    gw, ow = run_replay(scenes.heap(1500, seed=7), steps=120, check_every=20)
    assert gw.stats()["arbiters"] > 0 and gw.stats()["solver"] == 1


def test_heap_bitexact_barrier_solver(monkeypatch):
    # the colour-batch solver with a grid barrier per batch must agree with the dataflow solver
    monkeypatch.setenv("JH_SOLVER", "barrier")
    run_replay(scenes.heap(900, seed=5), steps=60, check_every=20)


@pytest.mark.parametrize("blocks", ["1", "7", "64", "256"])
def test_barrier_solver_launch_shapes(monkeypatch, blocks):
    # the barrier solver's shared body state is device-coherent and its grid barrier relaxed: exact for
    # any number of workgroups, including several per XCD and a single one
    monkeypatch.setenv("JH_SOLVER", "barrier")
    monkeypatch.setenv("JH_SOLVE_BLOCKS", blocks)
    run_replay(scenes.heap(700, seed=12), steps=40, check_every=13)


def test_dataflow_solver_launch_shapes(monkeypatch):
    # the dataflow solver is exact for any number of resident waves and any chunk ownership (taken
    # afresh from counters in every pass, taken in each phase's first pass and kept, round-robin)
    monkeypatch.setenv("JH_SOLVER", "flow")
    for queue in ("2", "1", "0"):
        monkeypatch.setenv("JH_FLOW_QUEUE", queue)
        for blocks in ("1", "7", "64"):
            monkeypatch.setenv("JH_SOLVE_BLOCKS", blocks)
            run_replay(scenes.heap(700, seed=11), steps=40, check_every=39)


def hub_scene(n_side=18):
    """A dynamic platform carrying n_side^2 spheres: the platform takes part in more arbiters than
    there are colours, so the last batch is the serial overflow colour, whose arbiters all share the
    platform and follow each other in the dataflow solver's chunks."""
    sb = scenes.SceneBuilder()
    g = sb.body((0.0, -0.5, 0.0), motion=MOTION_STATIC)
    sb.box(g, (60.0, 1.0, 60.0))
    p = sb.body((0.0, 0.25, 0.0))
    sb.box(p, (30.0, 0.5, 30.0))
    for i in range(n_side):
        for j in range(n_side):
            b = sb.body((1.5 * (i - n_side / 2), 0.9, 1.5 * (j - n_side / 2)))
            sb.sphere(b, 0.4)
    bodies, shapes, verts = sb.arrays()
    cfg = default_config()
    cfg["allow_deactivation"] = 0
    return bodies, shapes, verts, cfg, 0.01


@pytest.mark.parametrize("queue", ["0", "1", "2"])
def test_dataflow_solver_serial_colour(monkeypatch, queue):
    monkeypatch.setenv("JH_SOLVER", "flow")
    monkeypatch.setenv("JH_FLOW_QUEUE", queue)
    gw, ow = run_replay(hub_scene(), steps=12, check_every=3)
    keys, offs = gw.solve_order()
    assert len(offs) - 1 > gw.stats()["colours"], "the scene should overflow into the serial colour"
    assert offs[-1] - offs[-2] > 64  # serial arbiters fill whole chunks


@pytest.mark.parametrize("colour_blocks", ["default", "1", "16", "128"])
def test_solve_order_is_colouring(monkeypatch, colour_blocks):
    # checked after every step: the colouring kernel's shared lists, masks and colours are
    # device-coherent between its relaxed grid barriers, for any number of workgroups
    if colour_blocks != "default":
        monkeypatch.setenv("JH_COLOUR_BLOCKS", colour_blocks)
    bodies, shapes, verts, cfg, dt = scenes.heap(1200, seed=3)
    gw = GpuWorld(bodies, shapes, verts, cfg)
    body_of = shapes["body"]
    dyn = bodies["motion_type"] == 0
    for step in range(60):
        gw.step(dt)
        keys, offs = gw.solve_order()
        assert len(keys) >= len(gw.arbiters())  # arbiters were compacted after the solve
        for b in range(len(offs) - 1):
            k = keys[offs[b]:offs[b + 1]]
            bl = np.concatenate([body_of[k[:, 0]], body_of[k[:, 1]]])
            bl = bl[dyn[bl]]
            serial = b == len(offs) - 2 and len(offs) - 1 > gw.stats()["colours"]
            if not serial:
                assert len(np.unique(bl)) == len(bl), f"step {step}: batch {b} shares a dynamic body"


def test_step_argument_errors():
    bodies, shapes, verts, cfg, dt = scenes.pyramid(rows=3)
    gw = GpuWorld(bodies, shapes, verts, cfg)
    with pytest.raises(RuntimeError, match="invalid argument"):
        gw.step(-0.01)
    with pytest.raises(RuntimeError, match="non-finite"):
        gw.step(float("nan"))
    b0 = gw.bodies()
    gw.step(0.0)  # dt < epsilon: nothing happens (World.Step.cs:163-164)
    assert np.array_equal(gw.bodies()["position"], b0["position"])


def test_batched_worlds_bitexact():
    # independent worlds sharing coordinates: pairs never cross world ids; every world is solved in
    # the same persistent launch (the oracle filters pairs by world id by brute force)
    gw, ow = run_replay(scenes.batched_worlds(n_worlds=24, min_bodies=16, max_bodies=160, seed=9), steps=40,
                        check_every=13, brute_pairs=True)
    st = gw.stats()
    assert st["arbiters"] > 0


def test_autotune_keeps_results_exact():
    bodies, shapes, verts, cfg, dt = scenes.heap(1200, seed=21)
    gw = GpuWorld(bodies, shapes, verts, cfg)
    for _ in range(60):
        gw.step(dt)
    t = gw.autotune(dt, rounds=1)
    assert t["solve_blocks"] >= 1 and t["colour_blocks"] >= 1 and t["steps"] == t["candidates"]
    assert gw.stats()["solver"] == t["solver"] and t["solver"] in (1, 2, 3, 4)
    assert t["ms_solve_best"] <= t["ms_solve_default"] + 1e-9
    # continue from the tuned world's state: the oracle replay stays bit-exact with the new shapes
    ow = OracleWorld(gw.bodies(), shapes, verts, cfg)
    ow.set_arbiters(gw.arbiters())
    for s in range(20):
        gw.step(dt)
        ow.step(dt, gw.solve_order()[0])
    assert_bodies_equal(gw.bodies(), ow.bodies(), "after autotune")
    assert_arbiters_equal(gw.arbiters(), ow.arbiters(), "after autotune")


def _seam_scene(pos):
    sb = scenes.SceneBuilder()
    g = sb.body((0.0, 0.0, 0.0), motion=scenes.MOTION_STATIC, friction=0.5)
    sb.bodies[g]["inverse_mass"] = 0.0
    sb.bodies[g]["inverse_inertia"] = (0, 0, 0, 0, 0, 0)
    v = np.array([[-2, 0, -2], [2, 0, -2], [-2, 0, 2], [2, 0, 2]], np.float32)
    sb.triangle_mesh(g, v, np.array([[0, 2, 1], [1, 2, 3]]))
    b = sb.body(pos, velocity=(0.5, 0.0, -0.3))
    sb.sphere(b, 0.5)
    c = sb.body((1.0, 0.6, 1.0), orientation=(0.1, 0.2, 0.0, 0.9747), velocity=(-0.4, 0.0, -0.4))
    sb.box(c, (0.8, 0.8, 0.8))
    bodies, shapes, verts = sb.arrays()
    cfg = scenes.default_config()
    cfg["narrowphase_filter"] = 1
    cfg["substep_count"] = 2
    return bodies, shapes, verts, cfg, 0.01


def test_triangle_edge_filter_seam_bitexact():
    # contacts on the shared edge of a two-triangle mesh go through the filter's EPA re-run
    gw, ow = run_replay(_seam_scene((0.0, 0.45, 0.0)), steps=40, brute_pairs=True)
    assert gw.stats()["epa_calls"] >= 0


def test_heightfield_hulls_filter_bitexact():
    # the full form at small scale: 4 sub-worlds, each a Perlin heightfield tile (TriangleMesh with
    # the edge filter) under spheres, boxes, capsules and 32-point hulls
    scene = scenes.full_form(n_bodies=240, cells=32, worlds=4, feature=8.0, pitch=2.4)
    gw, ow = run_replay(scene, steps=90, check_every=15, brute_pairs=True)
    assert gw.stats()["arbiters"] > 50


def test_fp32_error_against_fp64_oracle():
    # the FP64 oracle (Jitter2.Double) replays the GPU's colour order; the FP32 error is reported and
    # bounded (tolerance: 1e-4 m in position, 1e-3 m/s in velocity after 30 steps)
    bodies, shapes, verts, cfg, dt = scenes.pyramid(rows=20)
    gw = GpuWorld(bodies, shapes, verts, cfg)
    o64 = OracleWorld(bodies, shapes, verts, cfg, fp64=True)
    for _ in range(30):
        gw.step(dt)
        o64.step(dt, gw.solve_order()[0])
    g = gw.bodies()
    s64 = o64.state64()
    dp = np.abs(g["position"].astype(np.float64) - s64[:, 0:3]).max()
    dv = np.abs(g["velocity"].astype(np.float64) - s64[:, 7:10]).max()
    print(f"FP32 GPU vs FP64 oracle after 30 steps: max |dp| {dp:.3e} m, max |dv| {dv:.3e} m/s")
    assert dp < 1e-4 and dv < 1e-3


def test_lane_solver_bitexact(monkeypatch):
    # one dataflow sequence per lane (positions L, L + lanes, ...): exact for any grid size
    monkeypatch.setenv("JH_SOLVER", "lanes")
    for blocks in ("1", "3", "64"):
        monkeypatch.setenv("JH_SOLVE_BLOCKS", blocks)
        gw, _ = run_replay(scenes.heap(900, seed=19), steps=50, check_every=49)
        assert gw.stats()["solver"] == 4
    monkeypatch.delenv("JH_SOLVE_BLOCKS")
    run_replay(scenes.pyramid(rows=12), steps=30, check_every=10)
    gw, _ = run_replay(scenes.batched_worlds(n_worlds=24, min_bodies=16, max_bodies=160, seed=23), steps=40,
                       check_every=13)
    assert gw.stats()["solver"] == 4


def test_tile_solver_bitexact(monkeypatch):
    # one workgroup per Morton range of bodies: private bodies in LDS, ordered by workgroup barriers;
    # bodies touched by other tiles' arbiters through the dataflow counters.  Every body still sees its
    # visits in colour order, so the replay stays bit-exact for any tile count.
    monkeypatch.setenv("JH_SOLVER", "tiles")
    for tiles in ("1", "5", "64"):
        monkeypatch.setenv("JH_TILES", tiles)
        gw, _ = run_replay(scenes.heap(900, seed=17), steps=50, check_every=49)
        assert gw.stats()["solver"] == 3
    monkeypatch.setenv("JH_TILES", "16")
    run_replay(scenes.pyramid(rows=12), steps=30, check_every=10)
    gw, _ = run_replay(scenes.full_form(n_bodies=200, cells=32, worlds=4, feature=8.0, pitch=2.4), steps=60,
                       check_every=20)
    assert gw.stats()["solver"] == 3


def test_world_solver_bitexact(monkeypatch):
    # one workgroup per world, colour by colour: same per-body order as the colour batches
    monkeypatch.setenv("JH_SOLVER", "world")
    run_replay(scenes.batched_worlds(n_worlds=24, min_bodies=16, max_bodies=160, seed=13), steps=40, check_every=13)
    run_replay(scenes.full_form(n_bodies=200, cells=32, worlds=4, feature=8.0, pitch=2.4), steps=60, check_every=20)
