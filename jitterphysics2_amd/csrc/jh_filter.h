// TriangleEdgeCollisionFilter (Collision/CollisionFilter/TriangleEdgeCollisionFilter.cs:22-324): for a
// contact between a static/kinematic triangle and another shape, discard back-face hits and, near a
// shared edge, snap the normal to the triangle's or its neighbour's normal.  Defaults: EdgeThreshold
// 0.01 (:40), cos(AngleThreshold) 0.999 (:42), ProjectionThreshold 0.5 (:61), FilterDynamicBodies
// false (:48).
//
// Triangle shapes carry their mesh data in the shape parameters: p[0..2] = the triangle's normal
// (TriangleMesh.Triangle.Normal), p[3..5] / p[6..8] / p[9..11] = the normals of NeighborA / B / C
// (the triangles across the edges opposite vertices A, B, C; zero = no neighbour), all in the body
// frame (TriangleMesh.cs:56-63, 300-319).
#pragma once
#include "jh_collide.h"
#include "jh_math.h"

namespace jh {
namespace tef {

constexpr float kEdgeThreshold = 0.01f;
constexpr float kCosAngle = 0.999f;
constexpr float kProjectionThreshold = 0.5f;
enum { Keep = 0, Discard = 1, Edge = 2 };

// World-space narrowphase answer (MprEpa output, NarrowPhase.cs:902-927).
struct WRes {
    vec3 pa, pb, n;
    float pen;
};

__device__ __forceinline__ vec3 param3(const ShapeRec& t, int k) { return {t.p[3 * k], t.p[3 * k + 1], t.p[3 * k + 2]}; }

// EvaluateTriangleContact (:86-147).  cp = the contact point on the triangle's side.
__device__ __forceinline__ int evaluate(const ShapeRec& t, const float* verts, quat q, vec3 pos, bool tri_is_b,
                                        vec3 cp, vec3 normal, vec3& tri_n, vec3& nb_n) {
    tri_n = rot(param3(t, 0), q);
    nb_n = {0.0f, 0.0f, 0.0f};
    if (tri_is_b) tri_n = -tri_n;
    if (dot(normal, tri_n) < -kCosAngle) return Discard;
    // TriangleShape.GetWorldVertices (TriangleShape.cs:89-108)
    const float* v = verts + 3 * t.voff;
    vec3 a = rot(vec3{v[0], v[1], v[2]}, q), b = rot(vec3{v[3], v[4], v[5]}, q), c = rot(vec3{v[6], v[7], v[8]}, q);
    a = a + pos;
    b = b + pos;
    c = c + pos;
    // ProjectPointOnPlane (:74-84)
    vec3 nrm = cross(b - a, c - a);
    float inv = 1.0f / sqrtf(len_sq(nrm));
    nrm = {nrm.x * inv, nrm.y * inv, nrm.z * inv};
    float dist = dot(cp - a, nrm);
    vec3 pp = cp - dist * nrm;
    vec3 n = b - a;
    vec3 pma = pp - a;
    float d0 = len_sq(pma - (dot(pma, n) * n) * (1.0f / len_sq(n)));
    n = c - a;
    pma = pp - a;
    float d1 = len_sq(pma - (dot(pma, n) * n) * (1.0f / len_sq(n)));
    n = c - b;
    pma = pp - b;
    float d2 = len_sq(pma - (dot(pma, n) * n) * (1.0f / len_sq(n)));
    if (fminf(fminf(d0, d1), d2) > kEdgeThreshold * kEdgeThreshold) return Keep;
    vec3 nn;
    if (d0 < d1 && d0 < d2) nn = param3(t, 3);       // NeighborC
    else if (d1 < d2) nn = param3(t, 2);             // NeighborB
    else nn = param3(t, 1);                          // NeighborA
    if (nn.x == 0.0f && nn.y == 0.0f && nn.z == 0.0f) return Keep;  // boundary edge
    nb_n = rot(nn, q);
    if (tri_is_b) nb_n = -nb_n;
    return Edge;
}

// The normal snapping of Filter (:213-323) for an Edge evaluation; false = discard.
__device__ __forceinline__ bool adjust(WRes& r, vec3 tri_n, vec3 nb_n) {
    bool spec = r.pen < 0.0f;
    vec3 mid = 0.5f * (r.pa + r.pb);
    if (dot(tri_n, nb_n) > kCosAngle) {
        float f5 = dot(r.n, nb_n);
        float f6 = dot(r.n, tri_n);
        if (!spec) {
            r.pen = 0.0f;
            r.pa = mid;
            r.pb = mid;
        }
        r.n = f5 > f6 ? nb_n : tri_n;
        return true;
    }
    vec3 cr = cross(nb_n, tri_n);
    float cl = len_sq(cr);
    vec3 proj = r.n - (dot(cr, r.n) / cl) * cr;
    if (len_sq(proj) < kProjectionThreshold * kProjectionThreshold) return false;
    float f1 = dot(cross(proj, nb_n), cr);
    float f2 = dot(cross(proj, tri_n), cr);
    bool between = f1 <= 0.0f && f2 >= 0.0f;
    if (!between) {
        float f3 = dot(r.n, nb_n);
        float f4 = dot(r.n, tri_n);
        if (!spec) {
            r.pen = 0.0f;
            r.pa = mid;
            r.pb = mid;
        }
        r.n = f3 > f4 ? nb_n : tri_n;
    }
    return true;
}

// Which shape of a pair the filter applies to: 1 = shape A is the triangle, 2 = shape B, 0 = none
// (both or neither are triangles, or the triangle's body is dynamic, :155-177).
__device__ __forceinline__ int applies(const ShapeRec& sa, const ShapeRec& sb, const int* motion) {
    bool c1 = sa.type == JH_SHAPE_TRIANGLE, c2 = sb.type == JH_SHAPE_TRIANGLE;
    if (c1 == c2) return 0;
    const ShapeRec& t = c1 ? sa : sb;
    if (motion[t.body] == JH_MOTION_DYNAMIC) return 0;
    return c1 ? 1 : 2;
}

struct Ctx {
    const ShapeRec* tri;
    const float* verts;
    quat q;
    vec3 pos;
    bool tri_is_b;
};

__device__ __forceinline__ int eval_res(const Ctx& c, const WRes& r, vec3& tri_n, vec3& nb_n) {
    return evaluate(*c.tri, c.verts, c.q, c.pos, c.tri_is_b, c.tri_is_b ? r.pb : r.pa, r.n, tri_n, nb_n);
}

// Second half of Filter (:186-210 then the snapping) on the answer of the re-run with EPA threshold 0.
__device__ __forceinline__ bool after_rerun(const Ctx& c, WRes& r) {
    vec3 tn, nn;
    int e = eval_res(c, r, tn, nn);
    if (e == Discard) return false;
    if (e == Keep) return true;
    return adjust(r, tn, nn);
}

// The whole Filter when the re-run returns the same answer (the first call already ran EPA, or the
// answer is speculative).
__device__ __forceinline__ bool filter_same(const Ctx& c, WRes& r) {
    vec3 tn, nn;
    int e = eval_res(c, r, tn, nn);
    if (e == Discard) return false;
    if (e == Keep) return true;
    if (r.pen < 0.0f) return adjust(r, tn, nn);
    return after_rerun(c, r);
}

}  // namespace tef
}  // namespace jh
