"""Seeded scene builders for the workloads of this path, producing jh_body / jh_shape arrays.

Mass properties follow the reference's shapes and RigidBody.SetMassInertia in FP32:
BoxShape.CalculateMassInertia (src/Jitter2/Collision/Shapes/BoxShape.cs:188-198),
SphereShape (SphereShape.cs:98-111), CapsuleShape (CapsuleShape.cs:109-129),
JSymmetricMatrix.Inverse (src/Jitter2/LinearMath/JSymmetricMatrix.cs:119-178),
RigidBody.SetMassInertia (src/Jitter2/Dynamics/RigidBody.cs:1575-1608).
Body defaults are those of a new RigidBody (RigidBody.cs:265-266, 398-409).

Random streams use SplitMix64 so that every scene is reproducible from its seed.
"""
from __future__ import annotations

import math

import numpy as np

from ._abi import (BODY_DTYPE, CONFIG_DTYPE, CONSTRAINT_DTYPE, FILTER_TRIANGLE_EDGE, MOTION_DYNAMIC, MOTION_STATIC,
                   SHAPE_BOX, SHAPE_CAPSULE, SHAPE_CONVEXHULL, SHAPE_DTYPE, SHAPE_POINTCLOUD, SHAPE_SPHERE,
                   SHAPE_TRIANGLE,
                   default_config)

f32 = np.float32


class SplitMix64:
    """SplitMix64 generator; uniform() returns doubles in [0, 1) from the top 53 bits."""

    def __init__(self, seed: int):
        self.state = seed & 0xFFFFFFFFFFFFFFFF

    def next_u64(self) -> int:
        self.state = (self.state + 0x9E3779B97F4A7C15) & 0xFFFFFFFFFFFFFFFF
        z = self.state
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & 0xFFFFFFFFFFFFFFFF
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & 0xFFFFFFFFFFFFFFFF
        return z ^ (z >> 31)

    def uniform(self, a: float = 0.0, b: float = 1.0) -> float:
        return a + (b - a) * ((self.next_u64() >> 11) * (1.0 / 9007199254740992.0))

    def normal(self) -> float:
        u1 = max(self.uniform(), 1e-300)
        u2 = self.uniform()
        return math.sqrt(-2.0 * math.log(u1)) * math.cos(2.0 * math.pi * u2)

    def quaternion(self) -> tuple[float, float, float, float]:
        """Uniform unit quaternion (Shoemake), as x, y, z, w."""
        u1, u2, u3 = self.uniform(), self.uniform(), self.uniform()
        a, b = math.sqrt(1.0 - u1), math.sqrt(u1)
        return (a * math.sin(2 * math.pi * u2), a * math.cos(2 * math.pi * u2),
                b * math.sin(2 * math.pi * u3), b * math.cos(2 * math.pi * u3))


def _sym_inverse(m11, m12, m13, m22, m23, m33):
    """JSymmetricMatrix.TryInverseUnscaled in FP32 (JSymmetricMatrix.cs:157-178)."""
    m11, m12, m13, m22, m23, m33 = (f32(v) for v in (m11, m12, m13, m22, m23, m33))
    det = m11 * (m22 * m33 - m23 * m23) + m12 * (m23 * m13 - m12 * m33) + m13 * (m12 * m23 - m22 * m13)
    idet = f32(1.0) / f32(det)
    a11 = m22 * m33 - m23 * m23
    a12 = m13 * m23 - m12 * m33
    a13 = m12 * m23 - m22 * m13
    a22 = m11 * m33 - m13 * m13
    a23 = m13 * m12 - m23 * m11
    a33 = m11 * m22 - m12 * m12
    return [f32(a11 * idet), f32(a12 * idet), f32(a13 * idet), f32(a22 * idet), f32(a23 * idet), f32(a33 * idet)]


def box_mass(size):
    """BoxShape.CalculateMassInertia (BoxShape.cs:188-198) for full extents `size`."""
    sx, sy, sz = (f32(v) for v in size)
    mass = sx * sy * sz
    k = f32(1.0 / 12.0)
    i11 = k * mass * (sy * sy + sz * sz)
    i22 = k * mass * (sx * sx + sz * sz)
    i33 = k * mass * (sx * sx + sy * sy)
    return f32(mass), (i11, f32(0), f32(0), i22, f32(0), i33)


def sphere_mass(r):
    """SphereShape.CalculateMassInertia."""
    r = f32(r)
    mass = f32(4.0 / 3.0) * f32(math.pi) * r * r * r
    i = f32(2.0 / 5.0) * mass * r * r
    return f32(mass), (i, f32(0), f32(0), i, f32(0), i)


def capsule_mass(r, length):
    """CapsuleShape.CalculateMassInertia."""
    r, length = f32(r), f32(length)
    pi = f32(math.pi)
    ms = f32(4.0 / 3.0) * pi * r * r * r
    mc = pi * r * r * length
    i11 = mc * (f32(1.0 / 12.0) * length * length + f32(1.0 / 4.0) * r * r) + ms * (
        f32(2.0 / 5.0) * r * r + f32(1.0 / 4.0) * length * length + f32(3.0 / 8.0) * length * r)
    i22 = f32(1.0 / 2.0) * mc * r * r + f32(2.0 / 5.0) * ms * r * r
    return f32(mc + ms), (i11, f32(0), f32(0), i22, f32(0), i11)


def _dot(a, b):
    return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]


def hull_graph(triangles):
    """ConvexHullShape's vertex graph (ConvexHullShape.cs:89-168): the distinct triangle corners in
    first-occurrence order (exact FP32 equality, -0 = +0), the triangles as vertex indices, and each
    vertex's neighbours ascending without repeats."""
    tri = np.asarray(triangles, np.float32).reshape(-1, 3, 3)
    if len(tri) == 0:
        raise ValueError("Triangle set must contain at least one triangle.")
    index, verts, faces = {}, [], []
    for t in tri:
        f = []
        for corner in t:
            key = tuple(float(x) + 0.0 for x in corner)
            if key not in index:
                if len(verts) >= 65535:
                    raise ValueError("The convex hull consists of too many vertices (>65535).")
                index[key] = len(verts)
                verts.append(corner.copy())
            f.append(index[key])
        faces.append(f)
    nbrs = [[] for _ in verts]
    for a, b, c in faces:
        nbrs[a] += [b, c]
        nbrs[b] += [a, c]
        nbrs[c] += [a, b]
    return np.array(verts, np.float32), np.array(faces, np.int64), [sorted(set(n)) for n in nbrs]


def hull_support(verts, nbrs, d, shift=(0.0, 0.0, 0.0)):
    """ConvexHullShape.InternalSupportMap (ConvexHullShape.cs:378-450) in FP32: the hill climb from
    vertex 0 with the plateau pass, plus the shift."""
    d = np.asarray(d, np.float32)
    eps = f32(1e-12)
    v = verts
    cur = 0
    dp = _dot(v[0], d)
    while True:
        verify, moved = False, False
        vdir = np.ones(3, np.float32)
        for nb in nbrs[cur]:
            npd = _dot(v[nb], d)
            if abs(npd - dp) < eps:
                vdir = v[nb] - v[cur]
                verify = True
            if npd > dp + eps:
                dp, cur, moved = npd, nb, True
                break
        if moved:
            continue
        if not verify:
            break
        d0 = _dot(vdir, v[cur])
        back = False
        while True:
            step = False
            for nb in nbrs[cur]:
                npd = _dot(v[nb], d)
                if npd > dp + eps:
                    dp, cur, back = npd, nb, True
                    break
                if abs(npd - dp) < eps:
                    d1 = _dot(vdir, v[nb])
                    if d1 > d0 + eps:
                        d0, cur, step = d1, nb, True
                        break
            if back or not step:
                break
        if not back:
            break
    return v[cur] + np.asarray(shift, np.float32)


def hull_mass(verts, faces, shift=(0.0, 0.0, 0.0)):
    """ConvexHullShape.CalculateMassInertia (ConvexHullShape.cs:262-326) in FP32: signed tetrahedra from
    the vertex mean to each face (winding fixed against that point), the canonical tetrahedron's second
    moments mapped by each face's corner matrix; returns (mass, inertia about the origin (6), centre of
    mass).  Density 1."""
    sh = np.asarray(shift, np.float32)
    a, b = f32(1.0 / 60.0), f32(1.0 / 120.0)
    can = (a, b, b, a, b, a)
    within = np.zeros(3, np.float32)
    for p in verts:
        within = within + p
    within = within * (f32(1.0) / f32(len(verts))) + sh
    inertia = [f32(0)] * 6
    center = np.zeros(3, np.float32)
    mass = f32(0)
    for ia, ib, ic in faces:
        c0, c1, c2 = verts[ia] + sh, verts[ib] + sh, verts[ic] + sh
        n = np.cross(c1 - c0, c2 - c0).astype(np.float32)
        if _dot(n, c0 - within) < f32(0):
            c0, c1 = c1, c0
        M = np.array([[c0[0], c1[0], c2[0]], [c0[1], c1[1], c2[1]], [c0[2], c1[2], c2[2]]], np.float32)
        det = (M[0, 0] * M[1, 1] * M[2, 2] + M[0, 1] * M[1, 2] * M[2, 0] + M[0, 2] * M[1, 0] * M[2, 1]
               - M[2, 0] * M[1, 1] * M[0, 2] - M[2, 1] * M[1, 2] * M[0, 0] - M[2, 2] * M[1, 0] * M[0, 1])
        m11, m12, m13, m22, m23, m33 = can
        prod = []
        for r in M:  # JVector.Transform(row, canonical)
            prod.append(np.array([m11 * r[0] + m12 * r[1] + m13 * r[2], m12 * r[0] + m22 * r[1] + m23 * r[2],
                                  m13 * r[0] + m23 * r[1] + m33 * r[2]], np.float32))
        t = [_dot(prod[0], M[0]), _dot(prod[0], M[1]), _dot(prod[0], M[2]), _dot(prod[1], M[1]),
             _dot(prod[1], M[2]), _dot(prod[2], M[2])]
        inertia = [inertia[k] + t[k] * det for k in range(6)]
        tmass = f32(1.0 / 6.0) * det
        center = center + tmass * (f32(0.25) * (c0 + c1 + c2))
        mass = mass + tmass
    if not np.isfinite(mass) or mass <= f32(1e-12):
        raise ValueError("Convex hull must define a non-degenerate volume.")
    tr = inertia[0] + inertia[3] + inertia[5]
    inertia = [tr - inertia[0], -inertia[1], -inertia[2], tr - inertia[3], -inertia[4], tr - inertia[5]]
    return mass, tuple(f32(x) for x in inertia), (center * (f32(1.0) / mass)).astype(np.float32)


def hull_block(verts, nbrs):
    """The vertex-array block of a JH_SHAPE_CONVEXHULL shape: vertices, (first, end) per vertex,
    neighbour list, padded to whole vertices."""
    ranges, flat = [], []
    for n in nbrs:
        ranges += [len(flat), len(flat) + len(n)]
        flat += n
    block = np.concatenate([np.asarray(verts, np.float32).ravel(), np.array(ranges + flat, np.float32)])
    pad = (-len(block)) % 3
    return np.concatenate([block, np.zeros(pad, np.float32)])


class SceneBuilder:
    """Accumulates bodies and shapes in creation order (shape ids ascend with creation)."""

    def __init__(self):
        self.bodies: list[np.void] = []
        self.shapes: list[np.void] = []
        self.vertices: list[float] = []
        self._next_id = 1
        self._shape_blocks: list[np.ndarray] = []   # earlier shapes, in order (bulk additions)
        self._vert_blocks: list[np.ndarray] = []
        self._n_vert_blocks = 0                     # vertices held in _vert_blocks

    def _flush(self):
        if self.shapes:
            self._shape_blocks.append(np.array(self.shapes, dtype=SHAPE_DTYPE))
            self.shapes = []
        if self.vertices:
            v = np.array(self.vertices, np.float32)
            self._vert_blocks.append(v)
            self._n_vert_blocks += len(v) // 3
            self.vertices = []

    def n_shapes(self) -> int:
        return sum(len(b) for b in self._shape_blocks) + len(self.shapes)

    def n_vertices(self) -> int:
        return self._n_vert_blocks + len(self.vertices) // 3

    def body(self, position, orientation=(0.0, 0.0, 0.0, 1.0), velocity=(0, 0, 0), angular_velocity=(0, 0, 0),
             motion=MOTION_DYNAMIC, friction=0.2, restitution=0.0, world=0) -> int:
        b = np.zeros(1, BODY_DTYPE)[0]
        b["position"] = position
        b["orientation"] = orientation
        b["velocity"] = velocity
        b["angular_velocity"] = angular_velocity
        b["inverse_inertia"] = (1, 0, 0, 1, 0, 1)
        b["inverse_mass"] = 1.0
        b["friction"] = friction
        b["restitution"] = restitution
        b["linear_damping_multiplier"] = f32(0.998)
        b["angular_damping_multiplier"] = f32(0.995)
        b["motion_type"] = motion
        b["affected_by_gravity"] = 1
        b["world_id"] = world
        b["deactivation_threshold_linear_sq"] = f32(0.1)  # RigidBody.cs:398-400 defaults
        b["deactivation_threshold_angular_sq"] = f32(0.1)
        b["deactivation_time"] = f32(1.0)
        self.bodies.append(b)
        return len(self.bodies) - 1

    def _shape(self, body: int, kind: int, params, mass_inertia, verts=None) -> int:
        s = np.zeros(1, SHAPE_DTYPE)[0]
        s["shape_id"] = self._next_id
        self._next_id += 1
        s["body"] = body
        s["type"] = kind
        p = np.zeros(12, np.float32)
        p[:len(params)] = params
        s["params"] = p
        if verts is not None:
            s["vertex_offset"] = self.n_vertices()
            s["vertex_count"] = len(verts)
            self.vertices.extend(np.asarray(verts, np.float32).ravel().tolist())
        self.shapes.append(s)
        # RigidBody.SetMassInertia with a single shape (RigidBody.cs:1575-1608)
        mass, inertia = mass_inertia
        b = self.bodies[body]
        b["inverse_inertia"] = _sym_inverse(*inertia)
        b["inverse_mass"] = f32(1.0) / f32(mass)
        return self.n_shapes() - 1

    def box(self, body: int, size) -> int:
        half = [f32(0.5) * f32(v) for v in size]
        return self._shape(body, SHAPE_BOX, half, box_mass(size))

    def sphere(self, body: int, radius) -> int:
        return self._shape(body, SHAPE_SPHERE, [radius], sphere_mass(radius))

    def capsule(self, body: int, radius, length) -> int:
        return self._shape(body, SHAPE_CAPSULE, [radius, f32(0.5) * f32(length)], capsule_mass(radius, length))

    def hull(self, body: int, points) -> int:
        """PointCloudShape over `points` (VertexSupportMap, PointCloudShape.cs:45-49).  GetCenter is
        the vertex mean (an interior point; the reference uses the hull's centre of mass) and the
        cached local box is the support box along +-x/y/z (CalcInitBox).  Mass properties: a solid
        sphere of the mean vertex radius (density 1) — body inputs, as with SetMassInertia."""
        pts = np.asarray(points, np.float32).reshape(-1, 3)
        c = pts.mean(axis=0, dtype=np.float64).astype(np.float32)
        lo, hi = pts.min(axis=0), pts.max(axis=0)
        r = float(np.mean(np.linalg.norm(pts.astype(np.float64) - c, axis=1)))
        return self._shape(body, SHAPE_POINTCLOUD, [*c, 0.0, 0.0, 0.0, *lo, *hi], sphere_mass(r), verts=pts)

    def convex_hull(self, body: int, triangles, shift=(0.0, 0.0, 0.0)) -> int:
        """ConvexHullShape from surface triangles (n, 3, 3) (ConvexHullShape.cs:89-168): params = centre
        of mass (GetCenter), shift, cached local box (CalcInitBox: the support along +-x/y/z,
        :352-376); the vertex block holds the vertices and their adjacency.  Mass properties:
        CalculateMassInertia, body inputs as SetMassInertia sets them (RigidBody.cs:1575-1608)."""
        verts, faces, nbrs = hull_graph(triangles)
        mass, inertia, com = hull_mass(verts, faces, shift)
        e = np.eye(3, dtype=np.float32)
        hi = [hull_support(verts, nbrs, e[k], shift)[k] for k in range(3)]
        lo = [hull_support(verts, nbrs, -e[k], shift)[k] for k in range(3)]
        block = hull_block(verts, nbrs)
        s = np.zeros(1, SHAPE_DTYPE)[0]
        s["shape_id"] = self._next_id
        self._next_id += 1
        s["body"] = body
        s["type"] = SHAPE_CONVEXHULL
        p = np.zeros(12, np.float32)
        p[0:3], p[3:6], p[6:9], p[9:12] = com, shift, lo, hi
        s["params"] = p
        s["vertex_offset"] = self.n_vertices()
        s["vertex_count"] = len(verts)
        self.vertices.extend(block.tolist())
        self.shapes.append(s)
        b = self.bodies[body]
        b["inverse_inertia"] = _sym_inverse(*inertia)
        b["inverse_mass"] = f32(1.0) / f32(mass)
        return self.n_shapes() - 1

    def triangle_mesh(self, body: int, vertices, triangles) -> np.ndarray:
        """TriangleShapes of a TriangleMesh (TriangleMesh.cs:117-319) on a static body: one shape per
        triangle, vertices A, B, C copied per triangle, params = triangle normal and the normals of
        NeighborA / B / C (AssignNeighbors: the triangle across the directed edge C->B, A->C, B->A;
        zero vector at a boundary).  Normals: normalize((B - A) x (C - A)) in float32.  Returns the
        shape indices."""
        self._flush()
        v = np.asarray(vertices, np.float32).reshape(-1, 3)
        t = np.asarray(triangles, np.int64).reshape(-1, 3)
        A, B, C = v[t[:, 0]], v[t[:, 1]], v[t[:, 2]]
        n = np.cross(B - A, C - A).astype(np.float32)
        n = (n * (np.float32(1.0) / np.sqrt((n * n).sum(axis=1, dtype=np.float32)))[:, None]).astype(np.float32)
        first = {}
        for i, (a, b, c) in enumerate(t.tolist()):
            for e in ((a, b), (b, c), (c, a)):
                first.setdefault(e, i)
        nb = np.full((len(t), 3), -1, np.int64)
        for i, (a, b, c) in enumerate(t.tolist()):
            nb[i, 0] = first.get((c, b), -1)
            nb[i, 1] = first.get((a, c), -1)
            nb[i, 2] = first.get((b, a), -1)
        params = np.zeros((len(t), 12), np.float32)
        params[:, 0:3] = n
        for k in range(3):
            ok = nb[:, k] >= 0
            params[ok, 3 + 3 * k:6 + 3 * k] = n[nb[ok, k]]
        out = np.zeros(len(t), SHAPE_DTYPE)
        out["shape_id"] = self._next_id + np.arange(len(t), dtype=np.uint64)
        self._next_id += len(t)
        out["body"] = body
        out["type"] = SHAPE_TRIANGLE
        out["params"] = params
        out["vertex_offset"] = self.n_vertices() + 3 * np.arange(len(t))
        out["vertex_count"] = 3
        start = self.n_shapes()
        self._shape_blocks.append(out)
        tri_verts = np.stack([A, B, C], axis=1).reshape(-1)
        self._vert_blocks.append(tri_verts.astype(np.float32))
        self._n_vert_blocks += len(tri_verts) // 3
        return np.arange(start, start + len(t))

    def arrays(self):
        self._flush()
        bodies = np.array(self.bodies, dtype=BODY_DTYPE) if self.bodies else np.zeros(0, BODY_DTYPE)
        shapes = np.concatenate(self._shape_blocks) if self._shape_blocks else np.zeros(0, SHAPE_DTYPE)
        verts = np.concatenate(self._vert_blocks) if self._vert_blocks else None
        # keep later builder calls consistent with what was returned
        self._shape_blocks = [shapes] if len(shapes) else []
        self._vert_blocks = [verts] if verts is not None else []
        return bodies, shapes, verts


def pyramid(rows: int = 20, seed: int = 0x4A2D0001, jitter: float = 0.01):
    """20-row pyramid of unit boxes on a static 200x1x200 ground box (mass 1, friction 0.6,
    restitution 0); layout of Helper.BuildPyramidBox (src/JitterTests/Helper.cs:63-79), all rows
    dynamic, lateral jitter U(-jitter, jitter).  Returns (bodies, shapes, verts, config, dt)."""
    rng = SplitMix64(seed)
    sb = SceneBuilder()
    g = sb.body((0.0, -0.5, 0.0), motion=MOTION_STATIC, friction=0.6)
    sb.box(g, (200.0, 1.0, 200.0))
    sb.bodies[g]["inverse_mass"] = 0.0
    sb.bodies[g]["inverse_inertia"] = (0, 0, 0, 0, 0, 0)
    for i in range(rows):
        for e in range(i, rows):
            x = (e - i * 0.5) * 1.01 + rng.uniform(-jitter, jitter)
            z = rng.uniform(-jitter, jitter)
            b = sb.body((x, 0.5 + i, z), friction=0.6)
            sb.box(b, (1.0, 1.0, 1.0))
    cfg = default_config()
    cfg["allow_deactivation"] = 0  # the configurations run without sleeping
    cfg["substep_count"] = 4
    cfg["solver_iterations"] = 8
    cfg["relaxation_iterations"] = 4
    return (*sb.arrays(), cfg, 1.0 / 100.0)


def _static_box(sb: SceneBuilder, pos, size, world=0, friction=0.2):
    b = sb.body(pos, motion=MOTION_STATIC, friction=friction, world=world)
    sb.box(b, size)
    sb.bodies[b]["inverse_mass"] = 0.0
    sb.bodies[b]["inverse_inertia"] = (0, 0, 0, 0, 0, 0)
    return b


def _random_shape(sb: SceneBuilder, rng: SplitMix64, body: int, kinds):
    k = kinds[int(rng.uniform() * len(kinds)) % len(kinds)]
    if k == SHAPE_SPHERE:
        sb.sphere(body, rng.uniform(0.3, 0.7))
    elif k == SHAPE_BOX:
        sb.box(body, [2.0 * rng.uniform(0.2, 0.8) for _ in range(3)])
    else:
        sb.capsule(body, rng.uniform(0.2, 0.4), rng.uniform(0.5, 1.5))


def heap(n_bodies: int = 32768, seed: int = 0x4A2D0002, interior: float = 60.0, height: float = 200.0,
         pitch: float = 2.7, kinds=(SHAPE_SPHERE, SHAPE_BOX, SHAPE_CAPSULE)):
    """Dynamic bodies dropped into a static box container (floor + 4 walls, interior
    `interior` x `interior`, height `height`, wall thickness 1): shape uniform over sphere
    r~U(0.3,0.7), box half extents ~U(0.2,0.8), capsule r~U(0.2,0.4) length~U(0.5,1.5); positions on a
    jittered grid (pitch 2.7 m, jitter U(-0.05,0.05), bottom layer at y=1.5), uniform random
    orientations, velocities N(0,1).  dt = 1/60, 4 substeps, 10 iterations.  The container widens
    with n so the grid keeps at most 68 layers for larger heaps."""
    rng = SplitMix64(seed)
    sb = SceneBuilder()
    per_row = max(22, int(math.ceil(math.sqrt(n_bodies / 68.0))))
    interior = max(interior, per_row * pitch + 0.6)
    h = interior / 2.0
    _static_box(sb, (0.0, -0.5, 0.0), (interior + 2.0, 1.0, interior + 2.0))
    _static_box(sb, (-(h + 0.5), height / 2.0, 0.0), (1.0, height, interior + 2.0))
    _static_box(sb, ((h + 0.5), height / 2.0, 0.0), (1.0, height, interior + 2.0))
    _static_box(sb, (0.0, height / 2.0, -(h + 0.5)), (interior, height, 1.0))
    _static_box(sb, (0.0, height / 2.0, (h + 0.5)), (interior, height, 1.0))
    x0 = -(per_row - 1) * pitch / 2.0
    count = 0
    layer = 0
    while count < n_bodies:
        for iz in range(per_row):
            for ix in range(per_row):
                if count >= n_bodies:
                    break
                pos = (x0 + ix * pitch + rng.uniform(-0.05, 0.05), 1.5 + layer * pitch + rng.uniform(-0.05, 0.05),
                       x0 + iz * pitch + rng.uniform(-0.05, 0.05))
                q = rng.quaternion()
                v = (rng.normal(), rng.normal(), rng.normal())
                b = sb.body(pos, orientation=q, velocity=v)
                _random_shape(sb, rng, b, kinds)
                count += 1
        layer += 1
    cfg = default_config()
    cfg["allow_deactivation"] = 0  # the configurations run without sleeping
    cfg["substep_count"] = 4
    cfg["solver_iterations"] = 10
    cfg["relaxation_iterations"] = 4
    return (*sb.arrays(), cfg, 1.0 / 60.0)


def batched_worlds(n_worlds: int = 4096, seed: int = 0x4A2D0003, min_bodies: int = 16, max_bodies: int = 2048,
                   spacing: float = 0.0):
    """Independent worlds with log-uniform body counts in [min_bodies, max_bodies]; each world is a
    random heap of boxes and spheres over its own static ground box.  Worlds share coordinates
    (they never interact: bodies carry a world id).  dt = 1/60, 4 substeps, 10 iterations."""
    rng = SplitMix64(seed)
    sb = SceneBuilder()
    pitch = 1.9
    for w in range(n_worlds):
        n = int(math.exp(rng.uniform(math.log(min_bodies), math.log(max_bodies))))
        side = max(4, int(math.ceil(math.sqrt(n / 8.0))))
        ox = w * spacing
        _static_box(sb, (ox, -0.5, 0.0), (side * pitch + 20.0, 1.0, side * pitch + 20.0), world=w)
        x0 = ox - (side - 1) * pitch / 2.0
        for i in range(n):
            layer, rem = divmod(i, side * side)
            iz, ix = divmod(rem, side)
            pos = (x0 + ix * pitch + rng.uniform(-0.05, 0.05), 1.0 + layer * pitch,
                   -(side - 1) * pitch / 2.0 + iz * pitch + rng.uniform(-0.05, 0.05))
            b = sb.body(pos, orientation=rng.quaternion(), velocity=(rng.normal(), rng.normal(), rng.normal()),
                        world=w)
            if rng.uniform() < 0.5:
                sb.sphere(b, rng.uniform(0.3, 0.7))
            else:
                sb.box(b, [2.0 * rng.uniform(0.2, 0.8) for _ in range(3)])
    cfg = default_config()
    cfg["allow_deactivation"] = 0  # the configurations run without sleeping
    cfg["substep_count"] = 4
    cfg["solver_iterations"] = 10
    cfg["relaxation_iterations"] = 4
    return (*sb.arrays(), cfg, 1.0 / 60.0)


def perlin2(x: np.ndarray, z: np.ndarray, seed: int) -> np.ndarray:
    """Perlin's improved gradient noise in 2-D (quintic fade, 8 gradient directions, permutation of
    0..255 shuffled by the SplitMix64 stream of `seed`); values in about [-1, 1]."""
    rng = SplitMix64(seed)
    perm = list(range(256))
    for i in range(255, 0, -1):
        j = int(rng.next_u64() % (i + 1))
        perm[i], perm[j] = perm[j], perm[i]
    P = np.array(perm + perm, np.int64)
    g = np.array([[1, 0], [-1, 0], [0, 1], [0, -1], [0.7071, 0.7071], [-0.7071, 0.7071], [0.7071, -0.7071],
                  [-0.7071, -0.7071]])
    xi, zi = np.floor(x).astype(np.int64), np.floor(z).astype(np.int64)
    xf, zf = x - xi, z - zi
    xi &= 255
    zi &= 255

    def grad(h, dx, dz):
        gg = g[h & 7]
        return gg[..., 0] * dx + gg[..., 1] * dz

    fade = lambda t: t * t * t * (t * (t * 6 - 15) + 10)  # noqa: E731
    u, w = fade(xf), fade(zf)
    aa, ab = P[P[xi] + zi], P[P[xi] + zi + 1]
    ba, bb = P[P[xi + 1] + zi], P[P[xi + 1] + zi + 1]
    x1 = grad(aa, xf, zf) * (1 - u) + grad(ba, xf - 1, zf) * u
    x2 = grad(ab, xf, zf - 1) * (1 - u) + grad(bb, xf - 1, zf - 1) * u
    return x1 * (1 - w) + x2 * w


def heightfield(cells: int, cell_size: float, amplitude: float, seed: int, x0: float = 0.0, z0: float = 0.0,
                feature: float = 48.0):
    """(cells + 1)^2 grid vertices with heights amplitude * (perlin(x / feature) + 0.5 perlin(x /
    (feature / 2))) / 1.5 and 2 cells^2 triangles, wound so their normals point up (+y)."""
    i = np.arange(cells + 1)
    X, Z = np.meshgrid(x0 + i * cell_size, z0 + i * cell_size, indexing="ij")
    H = (perlin2(X / feature, Z / feature, seed) + 0.5 * perlin2(X / (feature / 2), Z / (feature / 2), seed ^ 0x5A5A))
    H = amplitude * H / 1.5
    verts = np.stack([X, H, Z], axis=-1).reshape(-1, 3).astype(np.float32)
    k = np.arange(cells)
    a, b = np.meshgrid(k, k, indexing="ij")
    v00 = a * (cells + 1) + b          # (x0, z0)
    v10 = (a + 1) * (cells + 1) + b    # (x1, z0)
    v01 = a * (cells + 1) + b + 1      # (x0, z1)
    v11 = (a + 1) * (cells + 1) + b + 1
    t1 = np.stack([v00, v01, v10], axis=-1).reshape(-1, 3)
    t2 = np.stack([v10, v01, v11], axis=-1).reshape(-1, 3)
    tris = np.stack([t1, t2], axis=1).reshape(-1, 3)
    return verts, tris, H


def full_form(n_bodies: int = 262144, cells: int = 512, worlds: int = 64, seed: int = 0x4A2D0005,
              cell_size: float = 1.0, amplitude: float = 5.0, pitch: float = 2.2, feature: float = 48.0):
    """Bodies over a static triangle-mesh heightfield: `cells` x `cells` cells (Perlin heights,
    amplitude 5, main feature size `feature` metres) split into `worlds` square tiles, each tile the static ground of its own sub-world
    (a TriangleMesh body) with an uneven, log-uniform share of the bodies dropped on a jittered grid
    above it.  Shapes uniform over sphere, box, capsule (as in the heap) and 32-point hulls (points
    uniform on the unit sphere, scaled U(0.3, 1.0), PointCloudShape).  TriangleEdgeCollisionFilter
    on; dt = 1/60, 4 substeps, 10 iterations."""
    rng = SplitMix64(seed)
    tiles = int(round(math.sqrt(worlds)))
    assert tiles * tiles == worlds and cells % tiles == 0
    tc = cells // tiles
    wts = np.array([math.exp(rng.uniform(0.0, math.log(16.0))) for _ in range(worlds)])
    counts = np.floor(wts / wts.sum() * n_bodies).astype(int)
    counts[: n_bodies - counts.sum()] += 1
    sb = SceneBuilder()
    side = tc * cell_size
    per_row = max(1, int((side - 2.0) // pitch))
    for w in range(worlds):
        tx, tz = divmod(w, tiles)
        x0, z0 = tx * side, tz * side
        verts, tris, H = heightfield(tc, cell_size, amplitude, seed, x0, z0, feature)
        g = sb.body((0.0, 0.0, 0.0), motion=MOTION_STATIC, friction=0.5, world=w)
        sb.bodies[g]["inverse_mass"] = 0.0
        sb.bodies[g]["inverse_inertia"] = (0, 0, 0, 0, 0, 0)
        sb.triangle_mesh(g, verts, tris)
        top = float(H.max()) + 2.0
        for i in range(int(counts[w])):
            layer, rem = divmod(i, per_row * per_row)
            iz, ix = divmod(rem, per_row)
            pos = (x0 + 1.0 + (ix + 0.5) * pitch + rng.uniform(-0.05, 0.05), top + layer * pitch,
                   z0 + 1.0 + (iz + 0.5) * pitch + rng.uniform(-0.05, 0.05))
            b = sb.body(pos, orientation=rng.quaternion(), velocity=(rng.normal(), rng.normal(), rng.normal()),
                        world=w)
            k = int(rng.uniform() * 4) % 4
            if k == 3:
                pts = []
                scale = rng.uniform(0.3, 1.0)
                for _ in range(32):
                    d = np.array([rng.normal(), rng.normal(), rng.normal()])
                    pts.append(d / np.linalg.norm(d) * scale)
                sb.hull(b, pts)
            else:
                _random_shape(sb, rng, b, (SHAPE_SPHERE, SHAPE_BOX, SHAPE_CAPSULE)[k:k + 1])
    cfg = default_config()
    cfg["allow_deactivation"] = 0  # the configurations run without sleeping
    cfg["substep_count"] = 4
    cfg["solver_iterations"] = 10
    cfg["relaxation_iterations"] = 4
    cfg["narrowphase_filter"] = FILTER_TRIANGLE_EDGE
    return (*sb.arrays(), cfg, 1.0 / 60.0)


def bullets(n_bodies: int = 512, seed: int = 0x4A2D0006, speed=(15.0, 60.0), speculative: float = 1.0,
            mesh: bool = False, arena: float = 30.0, hulls: bool = True):
    """Fast projectiles against thin static slabs: a floor (a flat box, or with mesh=True a wavy
    triangle-mesh heightfield and the TriangleEdgeCollisionFilter), four walls and a middle slab
    0.2 thick around an `arena` x `arena` area, and dynamic spheres, boxes, capsules (and 16-point
    hulls) at heights 1..20 with speeds U(speed) in random directions biased downwards.  A fraction
    `speculative` of the bodies has speculative contacts on (RigidBody.EnableSpeculativeContacts).
    At these speeds a body moves further than a slab is thick in one step: without speculative
    contacts it tunnels.  dt = 1/60, 4 substeps, 10 iterations, speculative velocity threshold 10."""
    rng = SplitMix64(seed)
    sb = SceneBuilder()
    h = arena / 2.0
    t = 0.2
    if mesh:
        cells = 16
        verts, tris, _ = heightfield(cells, arena / cells, 1.0, seed, -h, -h, feature=8.0)
        g = sb.body((0.0, 0.0, 0.0), motion=MOTION_STATIC, friction=0.5)
        sb.bodies[g]["inverse_mass"] = 0.0
        sb.bodies[g]["inverse_inertia"] = (0, 0, 0, 0, 0, 0)
        sb.triangle_mesh(g, verts, tris)
    else:
        _static_box(sb, (0.0, -t / 2.0, 0.0), (arena + 2.0 * t, t, arena + 2.0 * t))
    _static_box(sb, (-(h + t / 2.0), 10.0, 0.0), (t, 20.0, arena + 2.0 * t))
    _static_box(sb, ((h + t / 2.0), 10.0, 0.0), (t, 20.0, arena + 2.0 * t))
    _static_box(sb, (0.0, 10.0, -(h + t / 2.0)), (arena, 20.0, t))
    _static_box(sb, (0.0, 10.0, (h + t / 2.0)), (arena, 20.0, t))
    _static_box(sb, (0.0, 6.0, 0.0), (t, 12.0, arena * 0.6))
    kinds = 4 if hulls else 3
    for _ in range(n_bodies):
        pos = (rng.uniform(-h + 1.0, h - 1.0), rng.uniform(1.5, 20.0), rng.uniform(-h + 1.0, h - 1.0))
        d = np.array([rng.normal(), rng.normal() - 0.7, rng.normal()])
        d = d / np.linalg.norm(d) * rng.uniform(*speed)
        w = (rng.normal(), rng.normal(), rng.normal())
        b = sb.body(pos, orientation=rng.quaternion(), velocity=tuple(d), angular_velocity=w)
        k = int(rng.uniform() * kinds) % kinds
        if k == 3:
            pts = []
            scale = rng.uniform(0.3, 0.8)
            for _ in range(16):
                q = np.array([rng.normal(), rng.normal(), rng.normal()])
                pts.append(q / np.linalg.norm(q) * scale)
            sb.hull(b, pts)
        else:
            _random_shape(sb, rng, b, (SHAPE_SPHERE, SHAPE_BOX, SHAPE_CAPSULE)[k:k + 1])
        sb.bodies[b]["enable_speculative_contacts"] = 1 if rng.uniform() < speculative else 0
    cfg = default_config()
    cfg["allow_deactivation"] = 0  # the configurations run without sleeping
    cfg["substep_count"] = 4
    cfg["solver_iterations"] = 10
    cfg["relaxation_iterations"] = 4
    if mesh:
        cfg["narrowphase_filter"] = FILTER_TRIANGLE_EDGE
    return (*sb.arrays(), cfg, 1.0 / 60.0)


def joints(n_groups: int = 24, seed: int = 0x4A2D0007, links: int = 8):
    """Jointed bodies over a static floor, for the constraint solver: per group one of (cycling)
    a hinge chain of boxes hanging from a static post (HingeJoint, limits +-60 degrees), a rope of
    spheres (BallSocket chain), welded box pairs (WeldJoint) dropped on the floor, capsules tied to a
    post by distance limits with a slack range, a ragdoll-like limb chain (BallSocket + ConeLimit +
    TwistAngle), motor-driven wheels on hinges (AngularMotor) with a block pulled by a LinearMotor,
    prismatic sliders (PointOnLine + FixedAngle or HingeAngle, limits, motors), universal-joint chains
    with a body held on a plane (PointOnPlane), and servos (TwistServo on a hinged arm, AngularServo on a
    ball-socketed box, LinearServo and LinearAxisServo on free blocks, an EllipticalConeLimit on a swinging
    capsule); free boxes fall between them.  Returns (bodies, shapes,
    verts, cfg, dt, constraints).  dt = 1/60, 4 substeps, 10 iterations."""
    from . import constraints as C

    rng = SplitMix64(seed)
    sb = SceneBuilder()
    side = int(math.ceil(math.sqrt(n_groups)))
    floor = max(200.0, 8.0 * side + 40.0)
    _static_box(sb, (0.0, -0.5, 0.0), (floor, 1.0, floor))
    recs: list = []
    plan = []  # deferred: constraints need the final body array
    for g in range(n_groups):
        gx, gz = (g % side - side / 2.0) * 8.0, (g // side - side / 2.0) * 8.0
        kind = g % 9
        top = 14.0
        if kind != 2:
            post = _static_box(sb, (gx, top + 0.5, gz), (0.4, 1.0, 0.4))
        if kind == 0:  # hinge chain
            prev, anchor_y = post, top
            for k in range(links):
                b = sb.body((gx, anchor_y - 0.55, gz), velocity=(rng.normal(), 0.0, rng.normal()))
                sb.box(b, (0.3, 1.0, 0.3))
                plan.append(("hinge", prev, b, (gx, anchor_y, gz)))
                prev, anchor_y = b, anchor_y - 1.1
        elif kind == 1:  # rope of spheres
            prev, anchor_y = post, top
            for k in range(links):
                b = sb.body((gx + 0.6 * (k + 1), anchor_y, gz))
                sb.sphere(b, 0.25)
                ax = gx + 0.6 * k
                plan.append(("ball", prev, b, (ax + 0.3 if k else gx, anchor_y, gz)))
                prev = b
        elif kind == 2:  # welded pairs dropping onto the floor
            for k in range(links // 2):
                y = 2.0 + 2.2 * k
                a = sb.body((gx - 0.5, y, gz), orientation=rng.quaternion())
                sb.box(a, (1.0, 0.5, 0.5))
                qa = sb.bodies[a]["orientation"].copy()
                b = sb.body((gx + 0.5, y, gz), orientation=qa)
                sb.box(b, (1.0, 0.5, 0.5))
                plan.append(("weld", a, b, (gx, y, gz)))
        elif kind == 4:  # limb chain: ball sockets with cone and twist limits
            prev, anchor_y = post, top
            for k in range(links):
                b = sb.body((gx, anchor_y - 0.6, gz), velocity=(rng.normal(), rng.normal(), rng.normal()),
                            angular_velocity=(rng.normal(), 3.0 * rng.normal(), rng.normal()))
                sb.capsule(b, 0.2, 0.8)
                plan.append(("limb", prev, b, (gx, anchor_y, gz)))
                prev, anchor_y = b, anchor_y - 1.2
        elif kind == 5:  # motor-driven wheels and a block on a linear motor
            for k in range(links // 2):
                b = sb.body((gx + 1.2 * k - 2.0, top - 1.0, gz + 0.6))
                sb.box(b, (1.0, 1.0, 0.2))
                plan.append(("wheel", post, b, (gx + 1.2 * k - 2.0, top - 1.0, gz + 0.6)))
            b = sb.body((gx + 2.5, top - 3.0, gz))
            sb.box(b, (0.5, 0.5, 0.5))
            plan.append(("slider", post, b, None))
        elif kind == 6:  # prismatic sliders
            for k in range(links // 2):
                b = sb.body((gx - 2.0 + 1.3 * k, top - 1.5, gz), velocity=(rng.normal(), rng.normal(), rng.normal()))
                sb.box(b, (0.5, 0.5, 0.5))
                plan.append(("slider_joint", post, b, (gx - 2.0 + 1.3 * k, top - 1.5, gz), k))
        elif kind == 7:  # universal-joint chain, and a sphere held on a plane
            prev, anchor_y = post, top
            for k in range(links // 2):
                b = sb.body((gx, anchor_y - 0.55, gz), angular_velocity=(rng.normal(), 2.0, rng.normal()))
                sb.box(b, (0.3, 1.0, 0.3))
                plan.append(("universal", prev, b, (gx, anchor_y, gz)))
                prev, anchor_y = b, anchor_y - 1.1
            b = sb.body((gx + 2.0, top - 2.0, gz + 1.0), velocity=(1.0, 0.0, -1.0))
            sb.sphere(b, 0.3)
            plan.append(("plane", post, b, (gx + 2.0, top - 2.0, gz + 1.0)))
        elif kind == 8:  # servos: a twist-driven arm, an oriented box, a block pulled to a point and along an axis
            arm = sb.body((gx, top - 0.55, gz), angular_velocity=(0.0, rng.normal(), 0.0))
            sb.box(arm, (0.3, 1.0, 0.3))
            plan.append(("twist_servo", post, arm, (gx, top, gz)))
            b = sb.body((gx + 2.0, top - 1.0, gz), orientation=rng.quaternion(),
                        angular_velocity=(rng.normal(), rng.normal(), rng.normal()))
            sb.box(b, (0.6, 0.6, 0.6))
            plan.append(("angular_servo", post, b, (gx + 2.0, top - 1.0, gz)))
            b = sb.body((gx - 2.0, top - 2.0, gz + 1.0), velocity=(rng.normal(), 0.0, rng.normal()))
            sb.box(b, (0.5, 0.5, 0.5))
            plan.append(("linear_servo", post, b, (gx - 2.0, top - 2.0, gz + 1.0)))
            b = sb.body((gx, top - 3.0, gz - 2.0))
            sb.sphere(b, 0.3)
            plan.append(("axis_servo", post, b, (gx, top - 3.0, gz - 2.0)))
            b = sb.body((gx + 2.0, top - 0.6, gz - 2.0), angular_velocity=(3.0 * rng.normal(), rng.normal(), 3.0))
            sb.capsule(b, 0.2, 0.8)
            plan.append(("ellipse", post, b, (gx + 2.0, top, gz - 2.0)))
        else:  # capsules on slack distance limits
            for k in range(links // 2):
                b = sb.body((gx + rng.uniform(-2, 2), top - 2.0 - 1.5 * k, gz + rng.uniform(-2, 2)),
                            orientation=rng.quaternion())
                sb.capsule(b, 0.25, 0.8)
                plan.append(("dist", post, b, None))
        for k in range(3):  # free boxes
            b = sb.body((gx + rng.uniform(-3, 3), 4.0 + 3.0 * k, gz + rng.uniform(-3, 3)), orientation=rng.quaternion())
            sb.box(b, [2.0 * rng.uniform(0.2, 0.5) for _ in range(3)])
    bodies, shapes, verts = sb.arrays()
    for item in plan:
        kind, a, b, anchor = item[:4]
        if kind == "slider_joint":
            k = item[4]
            ax = ((1.0, 0.0, 0.0), (0.0, 1.0, 0.0), (1.0, 1.0, 0.0), (0.0, 0.0, 1.0))[k % 4]
            recs += C.prismatic_joint(bodies, a, b, anchor, ax, limit=(-0.6, 0.8), pinned=k % 2 == 0,
                                      motor=(0.5, 10.0) if k % 3 == 0 else None)
        elif kind == "universal":
            recs += C.universal_joint(bodies, a, b, anchor, (1.0, 0.0, 0.0), (0.0, 0.0, 1.0))
        elif kind == "plane":
            recs.append(C.point_on_plane(bodies, a, b, (0.0, 1.0, 0.0), bodies["position"][a], anchor))
        elif kind == "twist_servo":  # hinge about y driven to +1 rad with a torque limit
            recs += C.hinge_joint(bodies, a, b, anchor, (0.0, 1.0, 0.0), servo=(1.0, 40.0))
        elif kind == "angular_servo":  # ball socket; servo to the post's orientation, unlimited torque
            recs.append(C.ball_socket(bodies, a, b, anchor))
            recs.append(C.angular_servo(bodies, a, b, target=(0.0, 0.0, 0.0, 1.0), max_torque=math.inf))
        elif kind == "linear_servo":  # pulled 1 m toward the post's side, limited force
            recs.append(C.linear_servo(bodies, a, b, anchor, target=(1.0, 0.0, -1.0), max_force=60.0))
        elif kind == "axis_servo":  # held 2 m below the post along y (the other directions are free)
            recs.append(C.linear_axis_servo(bodies, a, b, (0.0, 1.0, 0.0), bodies["position"][a], anchor,
                                            target=-4.0, max_force=math.inf))
        elif kind == "ellipse":  # ball socket with an elliptical swing cone (0.8 rad along x, 0.3 along z)
            recs.append(C.ball_socket(bodies, a, b, anchor))
            recs.append(C.elliptical_cone_limit(bodies, a, b, (0.0, -1.0, 0.0), (0.0, -1.0, 0.0), (1.0, 0.0, 0.0),
                                                (0.0, 0.8), (0.0, 0.3)))
        elif kind == "hinge":
            recs += C.hinge_joint(bodies, a, b, anchor, (1.0, 0.0, 0.0), (-math.pi / 3, math.pi / 3))
        elif kind == "ball":
            recs.append(C.ball_socket(bodies, a, b, anchor))
        elif kind == "weld":
            recs += C.weld_joint(bodies, a, b, anchor)
        elif kind == "limb":
            recs += C.ragdoll_joint(bodies, a, b, anchor, (0.0, -1.0, 0.0), math.pi / 5, (-math.pi / 8, math.pi / 8))
        elif kind == "wheel":
            recs += C.hinge_joint(bodies, a, b, anchor, (0.0, 0.0, 1.0), motor=(4.0, 5.0))
        elif kind == "slider":
            recs.append(C.linear_motor(bodies, a, b, (0.0, 1.0, 0.0), (0.0, 1.0, 0.0), 1.0, 20.0))
        else:
            pa = bodies["position"][a].astype(np.float32)
            pb = bodies["position"][b].astype(np.float32)
            recs.append(C.distance_limit(bodies, a, b, pa, pb, (-1.0, 0.5)))
    cfg = default_config()
    cfg["allow_deactivation"] = 0  # the configurations run without sleeping
    cfg["substep_count"] = 4
    cfg["solver_iterations"] = 10
    cfg["relaxation_iterations"] = 4
    return bodies, shapes, verts, cfg, 1.0 / 60.0, C.as_array(recs)


def two_bodies(shape_a, shape_b, pos_a, pos_b, q_a=(0, 0, 0, 1), q_b=(0, 0, 0, 1), static_a=False):
    """Tiny two-body scene for unit tests; shape_x = ("box", size) | ("sphere", r) | ("capsule", r, len)."""
    sb = SceneBuilder()
    for spec, pos, q, static in ((shape_a, pos_a, q_a, static_a), (shape_b, pos_b, q_b, False)):
        b = sb.body(pos, orientation=q, motion=MOTION_STATIC if static else MOTION_DYNAMIC)
        if spec[0] == "box":
            sb.box(b, spec[1])
        elif spec[0] == "sphere":
            sb.sphere(b, spec[1])
        else:
            sb.capsule(b, spec[1], spec[2])
        if static:
            sb.bodies[b]["inverse_mass"] = 0.0
            sb.bodies[b]["inverse_inertia"] = (0, 0, 0, 0, 0, 0)
    return sb.arrays()


def save_scene(path: str, scene) -> None:
    """Write (bodies, shapes, verts, config, dt[, constraints]) in the binary form read by the native
    driver jh_cli: "JHSC", int32 version (1, or 2 with constraints), int32 body/shape/vertex counts,
    float32 dt, the 64-byte jh_config, then the jh_body, jh_shape and float32[3*nv] arrays; version 2
    appends int32 nc and the jh_constraint records."""
    bodies, shapes, verts, cfg, dt = scene[:5]
    cons = scene[5] if len(scene) > 5 else None
    v = np.zeros((0, 3), np.float32) if verts is None else np.ascontiguousarray(verts, np.float32).reshape(-1, 3)
    cfg = np.asarray(cfg, dtype=CONFIG_DTYPE).reshape(())
    with open(path, "wb") as f:
        f.write(b"JHSC")
        f.write(np.array([1 if cons is None else 2, len(bodies), len(shapes), len(v)], np.int32).tobytes())
        f.write(np.float32(dt).tobytes())
        f.write(cfg.tobytes())
        f.write(np.ascontiguousarray(bodies, BODY_DTYPE).tobytes())
        f.write(np.ascontiguousarray(shapes, SHAPE_DTYPE).tobytes())
        f.write(v.tobytes())
        if cons is not None:
            f.write(np.int32(len(cons)).tobytes())
            f.write(np.ascontiguousarray(cons, CONSTRAINT_DTYPE).tobytes())
