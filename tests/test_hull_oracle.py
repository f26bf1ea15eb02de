"""ConvexHullShape (Collision/Shapes/ConvexHullShape.cs) in the oracle and the scene builder, pinned by
what the geometry fixes and by the reference's tests: the hill-climb support (:378-450) returns a
vertex of maximal extent (brute force over the vertices) for random directions on random hulls and on
hulls full of plateaus (cube, octahedron, prism faces split in triangles), and agrees with the
builder's own FP32 restatement; the cached-box AABB (:334-349) at identity is the vertex box and,
rotated, contains every rotated vertex while staying within the reference's volume bound
(Api/BoundingBoxTests.cs); the mass properties (:262-326) equal the closed-form tetrahedron moments
(Api/InertiaTests.cs ConvexHullInertia); a flat hull is rejected
(Robustness/ArgumentValidationTests.cs).  CPU only."""
import numpy as np
import pytest
from scipy.spatial import ConvexHull

from jitterphysics2_amd import scenes
from jitterphysics2_amd._abi import ptr
from tests.orc import lib


def hull_triangles(points):
    pts = np.asarray(points, np.float32)
    h = ConvexHull(pts.astype(np.float64))
    return pts[h.simplices]


def hull_shape(triangles, shift=(0.0, 0.0, 0.0)):
    sb = scenes.SceneBuilder()
    b = sb.body((0.0, 0.0, 0.0))
    sb.convex_hull(b, triangles, shift)
    bodies, shapes, verts = sb.arrays()
    return shapes[0:1].copy(), verts, bodies


def support(s, verts, d, fp64=0):
    out = np.zeros(3, np.float32)
    lib().orc_support(fp64, ptr(s), ptr(verts), ptr(np.asarray(d, np.float32)), ptr(out))
    return out


def aabb(s, verts, q, p):
    out = np.zeros(6, np.float32)
    lib().orc_aabb(0, ptr(s), ptr(verts), ptr(np.asarray(q, np.float32)), ptr(np.asarray(p, np.float32)), ptr(out))
    return out[:3], out[3:]


def cube_triangles(h=(0.5, 0.7, 0.9)):
    x, y, z = h
    c = np.array([[sx * x, sy * y, sz * z] for sx in (-1, 1) for sy in (-1, 1) for sz in (-1, 1)], np.float32)
    quads = [(0, 1, 3, 2), (4, 6, 7, 5), (0, 4, 5, 1), (2, 3, 7, 6), (0, 2, 6, 4), (1, 5, 7, 3)]
    tris = []
    for a, b, cc, d in quads:
        tris += [[c[a], c[b], c[cc]], [c[a], c[cc], c[d]]]
    return np.array(tris, np.float32)


def octahedron_triangles():
    e = np.eye(3, dtype=np.float32)
    p = [e[0], -e[0], e[1], -e[1], e[2], -e[2]]
    tris = [[p[i], p[j], p[k]] for i in (0, 1) for j in (2, 3) for k in (4, 5)]
    return np.array(tris, np.float32)


def prism_triangles(n=7, r=0.6, h=0.8):
    ang = 2 * np.pi * np.arange(n) / n
    top = np.stack([r * np.cos(ang), np.full(n, h), r * np.sin(ang)], -1).astype(np.float32)
    bot = top * np.array([1, -1, 1], np.float32)
    return hull_triangles(np.concatenate([top, bot]))


def random_hull(rng, n=32):
    p = rng.normal(size=(n, 3))
    p /= np.linalg.norm(p, axis=1, keepdims=True)
    return hull_triangles((p * rng.uniform(0.3, 1.0)).astype(np.float32))


HULLS = {"cube": cube_triangles, "octahedron": octahedron_triangles, "prism": prism_triangles,
         "sphere32": lambda: random_hull(np.random.default_rng(5)), "sphere200": lambda: random_hull(np.random.default_rng(6), 200)}


@pytest.mark.parametrize("name", sorted(HULLS))
@pytest.mark.parametrize("shift", [(0.0, 0.0, 0.0), (0.25, -0.5, 1.0)])
def test_hill_climb_finds_a_vertex_of_maximal_extent(name, shift):
    tris = HULLS[name]()
    s, verts, _ = hull_shape(tris, shift)
    V = int(s["vertex_count"][0])
    pts = verts[:3 * V].reshape(V, 3).astype(np.float64) + np.asarray(shift, np.float64)
    gv, _, nbrs = scenes.hull_graph(tris)
    rng = np.random.default_rng(7)
    dirs = np.concatenate([rng.normal(size=(300, 3)), np.eye(3), -np.eye(3), [[1, 1, 0], [1, -1, 1], [0, 1, 1]]])
    for d in dirs.astype(np.float32):
        r = support(s, verts, d)
        best = (pts @ d.astype(np.float64)).max()
        assert float(r.astype(np.float64) @ d) >= best - 1e-6 * max(1.0, abs(best)), (name, d)
        assert np.min(np.abs(pts - r).sum(axis=1)) < 1e-6  # a vertex (plus the shift)
        assert np.array_equal(r, scenes.hull_support(gv, nbrs, d, shift))  # the two restatements agree
        assert np.array_equal(support(s, verts, d, fp64=1).astype(np.float32), r) or name.startswith("sphere")


def test_detached_hull_bounding_box():  # BoundingBoxTests.cs:8-43
    a, b, c, d = (np.array(v, np.float32) for v in ((10, 10, 10), (11, 10, 10), (10, 11, 10), (10, 10, 11)))
    tris = [[a, b, c], [a, b, d], [a, c, d], [b, c, d]]
    s, verts, _ = hull_shape(tris)
    mn, mx = aabb(s, verts, (0, 0, 0, 1), (0, 0, 0))
    assert np.array_equal(mn, a) and np.array_equal(mx, [11, 11, 11])
    s, verts, _ = hull_shape(tris, shift=(3.0, 0.0, 0.0))
    assert aabb(s, verts, (0, 0, 0, 1), (0, 0, 0))[0][0] == 13


REFERENCE_TETRA = np.array([(1, 2, 3), (2, 2, 3), (1, -2, 4), (3, 3, 3)], np.float32)  # BoundingBoxTests.cs:70-84
HULLS["reference_tetra"] = lambda: REFERENCE_TETRA[[[0, 1, 2], [0, 1, 3], [0, 3, 2], [3, 1, 2]]]


@pytest.mark.parametrize("name", sorted(HULLS))
def test_rotated_hull_box_contains_the_hull(name):  # BoundingBoxTests.cs:45-58
    """The cached-box AABB holds every rotated vertex; for the reference's hull it also stays within
    the reference's volume bound (tight box volume / box volume > 0.2)."""
    s, verts, _ = hull_shape(HULLS[name]())
    V = int(s["vertex_count"][0])
    q = np.array([1, 2, 3, 4], np.float64)
    q /= np.linalg.norm(q)
    pos = np.array([12, -11, 17], np.float64)
    mn, mx = aabb(s, verts, q.astype(np.float32), pos.astype(np.float32))
    x, y, z, w = q
    R = np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
                  [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
                  [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)]])
    world = verts[:3 * V].reshape(V, 3).astype(np.float64) @ R.T + pos
    tight_mn, tight_mx = world.min(axis=0), world.max(axis=0)
    assert np.all(mn <= tight_mn + 1e-4) and np.all(mx >= tight_mx - 1e-4)
    fraction = np.prod(tight_mx - tight_mn) / np.prod(mx.astype(np.float64) - mn)
    assert fraction - 1e-4 < 1.0
    if name == "reference_tetra":
        assert fraction > 0.2


def tetra_moments(p):
    """Mass, centre and inertia about the origin of the tetrahedron p (4, 3) in FP64: the second
    moment is V/20 (sum p_k p_k^T + (sum p_k)(sum p_k)^T)."""
    p = np.asarray(p, np.float64)
    vol = abs(np.linalg.det(p[1:] - p[0])) / 6
    s = p.sum(axis=0)
    S = vol / 20 * (sum(np.outer(v, v) for v in p) + np.outer(s, s))
    return vol, s / 4, np.trace(S) * np.eye(3) - S


@pytest.mark.parametrize("shift", [(0.0, 0.0, 0.0), (1.0, -2.0, 0.5)])
def test_hull_mass_properties_equal_the_tetrahedron_closed_form(shift):  # InertiaTests.cs:144-162
    a, b, c, d = (np.array(v, np.float32) for v in ((0.234, 1.23, 3.54), (7.788, 0.23, 8.14), (2.234, 8.23, 8.14),
                                                      (6.234, 3.23, 9.04)))
    verts, faces, _ = scenes.hull_graph([[a, b, c], [a, b, d], [b, c, d], [a, c, d]])
    mass, inertia, com = scenes.hull_mass(verts, faces, shift)
    m, cm, I = tetra_moments(np.stack([a, b, c, d]) + np.asarray(shift, np.float32))
    assert abs(mass - m) < 1e-5 * m
    assert np.abs(com - cm).max() < 1e-4
    i11, i12, i13, i22, i23, i33 = (float(x) for x in inertia)
    got = np.array([[i11, i12, i13], [i12, i22, i23], [i13, i23, i33]])
    assert np.abs(got - I).max() < 1e-5 * np.abs(I).max()


def test_cube_hull_mass_properties():
    h = (0.5, 0.7, 0.9)
    verts, faces, _ = scenes.hull_graph(cube_triangles(h))
    mass, inertia, com = scenes.hull_mass(verts, faces)
    sx, sy, sz = (2 * v for v in h)
    m = sx * sy * sz
    assert abs(mass - m) < 1e-5 and np.abs(com).max() < 1e-6
    want = np.array([m * (sy * sy + sz * sz), 0, 0, m * (sx * sx + sz * sz), 0, m * (sx * sx + sy * sy)]) / 12
    assert np.abs(np.array(inertia, np.float64) - want).max() < 1e-5


def test_flat_hull_is_rejected():  # ArgumentValidationTests.cs:42-48
    with pytest.raises(ValueError):
        scenes.hull_graph([])
    verts, faces, _ = scenes.hull_graph([[(0, 0, 0), (1, 0, 0), (0, 1, 0)]])
    with pytest.raises(ValueError):
        scenes.hull_mass(verts, faces)


def test_hull_graph_is_the_reference_vertex_graph():
    tris = cube_triangles()
    verts, faces, nbrs = scenes.hull_graph(tris)
    assert len(verts) == 8 and len(faces) == 12
    for k, n in enumerate(nbrs):
        assert n == sorted(set(n)) and k not in n
        for j in n:
            assert k in nbrs[j]  # the graph is undirected
    # first-occurrence order of the corners
    seen = []
    for t in tris.reshape(-1, 3):
        if not any(np.array_equal(t, s) for s in seen):
            seen.append(t)
    assert np.array_equal(np.array(seen), verts)
