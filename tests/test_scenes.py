"""Scene files for the native driver jh_cli (scenes.save_scene): version 1 without constraints, version 2
with the jh_constraint records appended; the layout is the one jh_cli.cpp reads."""
import numpy as np

from jitterphysics2_amd import scenes
from jitterphysics2_amd._abi import BODY_DTYPE, CONFIG_DTYPE, CONSTRAINT_DTYPE, SHAPE_DTYPE


def read_scene(path):
    raw = open(path, "rb").read()
    assert raw[:4] == b"JHSC"
    ver, nb, ns, nv = np.frombuffer(raw, np.int32, 4, 4)
    o = 20
    dt = np.frombuffer(raw, np.float32, 1, o)[0]
    o += 4
    cfg = np.frombuffer(raw, CONFIG_DTYPE, 1, o)[0]
    o += CONFIG_DTYPE.itemsize
    bodies = np.frombuffer(raw, BODY_DTYPE, nb, o)
    o += nb * BODY_DTYPE.itemsize
    shapes = np.frombuffer(raw, SHAPE_DTYPE, ns, o)
    o += ns * SHAPE_DTYPE.itemsize
    verts = np.frombuffer(raw, np.float32, 3 * nv, o).reshape(-1, 3)
    o += 12 * nv
    cons = None
    if ver == 2:
        nc = int(np.frombuffer(raw, np.int32, 1, o)[0])
        o += 4
        cons = np.frombuffer(raw, CONSTRAINT_DTYPE, nc, o)
        o += nc * CONSTRAINT_DTYPE.itemsize
    assert o == len(raw)
    return int(ver), dt, cfg, bodies, shapes, verts, cons


def test_scene_file_round_trip(tmp_path):
    scene = scenes.heap(200, seed=5)
    scenes.save_scene(str(tmp_path / "h.bin"), scene)
    ver, dt, cfg, bodies, shapes, verts, cons = read_scene(tmp_path / "h.bin")
    assert ver == 1 and cons is None and dt == np.float32(scene[4])
    assert bodies.tobytes() == scene[0].tobytes() and shapes.tobytes() == scene[1].tobytes()
    assert cfg.tobytes() == np.asarray(scene[3], CONFIG_DTYPE).tobytes()


def test_scene_file_with_constraints(tmp_path):
    scene = scenes.joints(9, seed=3)
    scenes.save_scene(str(tmp_path / "j.bin"), scene)
    ver, dt, cfg, bodies, shapes, verts, cons = read_scene(tmp_path / "j.bin")
    assert ver == 2 and len(cons) == len(scene[5]) > 0
    assert cons.tobytes() == scene[5].tobytes()
    assert bodies.tobytes() == scene[0].tobytes()
    # the scene holds every constraint kind
    assert set(np.unique(cons["type"]).tolist()) == set(range(15))
