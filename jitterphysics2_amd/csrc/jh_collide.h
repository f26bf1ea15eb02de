// Narrowphase device code: shape support maps and AABBs, XenoCollide MPR (FP32, registers), the
// one-shot contact manifold, and the EPA refinement (FP64 with filtered exact predicates) that runs
// out of a per-thread workspace in HBM.  Each function cites the Jitter2 code it restates.
#pragma once
#include "jh_math.h"
#include "../../include/jitterhip.h"

namespace jh {

// Device shape record: 64 bytes, read-only during a step.
struct ShapeRec {
    int type;
    int body;
    float p[12];
    int voff;
    int vcnt;
};

// Hill climb over the hull's vertex graph from vertex 0 (ConvexHullShape.InternalSupportMap): move to
// the first neighbour that is better by more than 1e-12; at a local maximum with a neighbour level
// within 1e-12 (a plateau), climb along that neighbour's direction over equally good vertices.  The
// dot products strictly increase (or, on one plateau, the secondary product does), so no vertex is
// visited twice; the walk is capped at V + 1 moves, which a valid hull never reaches.
__device__ __forceinline__ vec3 hull_support(const float* v, int V, vec3 d) {
    const float* adj = v + 3 * V;
    const float* nbl = adj + 2 * V;
    auto vx = [&](int i) { return vec3{v[3 * i], v[3 * i + 1], v[3 * i + 2]}; };
    const float eps = 1e-12f;
    int cur = 0;
    float dp = dot(vx(0), d);
    for (int moves = 0; moves <= V;) {
        bool verify = false, moved = false;
        vec3 vdir{1.0f, 1.0f, 1.0f};  // JVector.Arbitrary
        int lo = int(adj[2 * cur]), hi = int(adj[2 * cur + 1]);
        for (int i = lo; i < hi; ++i) {
            int nb = int(nbl[i]);
            float np = dot(vx(nb), d);
            if (fabsf(np - dp) < eps) {
                vdir = vx(nb) - vx(cur);
                verify = true;
            }
            if (np > dp + eps) {
                dp = np;
                cur = nb;
                moved = true;
                break;
            }
        }
        if (moved) { ++moves; continue; }
        if (!verify) break;
        float d0 = dot(vdir, vx(cur));
        bool back_to_main = false;
        while (moves <= V) {
            bool step = false;
            lo = int(adj[2 * cur]);
            hi = int(adj[2 * cur + 1]);
            for (int i = lo; i < hi; ++i) {
                int nb = int(nbl[i]);
                float np = dot(vx(nb), d);
                if (np > dp + eps) {
                    dp = np;
                    cur = nb;
                    back_to_main = true;
                    break;
                }
                if (fabsf(np - dp) < eps) {
                    float d1 = dot(vdir, vx(nb));
                    if (d1 > d0 + eps) {
                        d0 = d1;
                        cur = nb;
                        step = true;
                        break;
                    }
                }
            }
            ++moves;
            if (back_to_main || !step) break;
        }
        if (!back_to_main) break;
    }
    return vx(cur);
}

struct ShapeView {
    const ShapeRec* s;
    const float* verts;           // all vertices (x, y, z) of the world
    const float* lv = nullptr;    // a point cloud's own vertices staged in LDS (k_narrow_hulls), else null

    // ISupportMappable.SupportMap, local frame, direction not normalised.
    __device__ __forceinline__ vec3 support(vec3 d) const {
        const float* p = s->p;
        switch (s->type) {
            case JH_SHAPE_BOX:  // BoxShape.cs:86-91
                return {signbit_f(d.x) * p[0], signbit_f(d.y) * p[1], signbit_f(d.z) * p[2]};
            case JH_SHAPE_SPHERE:  // SphereShape.cs:54-58
                return normalize(d) * p[0];
            case JH_SHAPE_CAPSULE: {  // CapsuleShape.cs:74-89
                vec3 r = normalize(d) * p[0];
                r.y += sign_f(d.y) * p[1];
                return r;
            }
            case JH_SHAPE_POINTCLOUD: {  // VertexSupportMap.cs:64-109 (last maximum wins) + shift
                const float* v = lv ? lv : verts + 3 * s->voff;
                float best = -3.402823466e+38f;
                int idx = 0;
                for (int i = 0; i < s->vcnt; ++i) {
                    float dx = v[3 * i] * d.x, dy = v[3 * i + 1] * d.y, dz = v[3 * i + 2] * d.z;
                    float dp = dx + (dy + dz);
                    if (dp >= best) { best = dp; idx = i; }
                }
                return vec3{v[3 * idx], v[3 * idx + 1], v[3 * idx + 2]} + vec3{p[3], p[4], p[5]};
            }
            case JH_SHAPE_CONVEXHULL:  // ConvexHullShape.InternalSupportMap (ConvexHullShape.cs:378-450) + shift
                return hull_support(verts + 3 * s->voff, s->vcnt, d) + vec3{p[3], p[4], p[5]};
            default: {  // TriangleShape.cs:167-191
                const float* v = verts + 3 * s->voff;
                vec3 a{v[0], v[1], v[2]}, b{v[3], v[4], v[5]}, c{v[6], v[7], v[8]};
                float mn = dot(a, d);
                float dp = dot(b, d);
                vec3 r = a;
                if (dp > mn) { mn = dp; r = b; }
                dp = dot(c, d);
                if (dp > mn) r = c;
                return r;
            }
        }
    }
    // ISupportMappable.GetCenter
    __device__ __forceinline__ vec3 center() const {
        if (s->type == JH_SHAPE_POINTCLOUD || s->type == JH_SHAPE_CONVEXHULL) return {s->p[0], s->p[1], s->p[2]};
        if (s->type == JH_SHAPE_TRIANGLE) {
            const float* v = verts + 3 * s->voff;
            vec3 a{v[0], v[1], v[2]}, b{v[3], v[4], v[5]}, c{v[6], v[7], v[8]};
            return float(1.0 / 3.0) * (a + b + c);
        }
        return {0.0f, 0.0f, 0.0f};
    }
};

// Shape world AABB (RigidBodyShape.UpdateWorldBoundingBox, RigidBodyShape.cs:37-58).
__device__ __forceinline__ void shape_aabb(const ShapeRec& s, const float* verts, quat q, vec3 pos, vec3& mn,
                                           vec3& mx) {
    const float* p = s.p;
    switch (s.type) {
        case JH_SHAPE_BOX: {  // BoxShape.cs:180-186
            vec3 ths = abs_mat_mul(vec3{p[0], p[1], p[2]}, mat_from_quat(q));
            mn = pos - ths;
            mx = pos + ths;
            return;
        }
        case JH_SHAPE_SPHERE:  // SphereShape.cs:65-71
            mn = vec3{-p[0], -p[0], -p[0]} + pos;
            mx = vec3{p[0], p[0], p[0]} + pos;
            return;
        case JH_SHAPE_CAPSULE: {  // CapsuleShape.cs:97-107
            vec3 dl = p[1] * basis_y(q);
            mx = {p[0] + fabsf(dl.x), p[0] + fabsf(dl.y), p[0] + fabsf(dl.z)};
            mn = -mx;
            mn = mn + pos;
            mx = mx + pos;
            return;
        }
        case JH_SHAPE_POINTCLOUD:    // PointCloudShape.cs:140-154
        case JH_SHAPE_CONVEXHULL: {  // ConvexHullShape.cs:334-349: the same cached-box formula
            vec3 bmin{p[6], p[7], p[8]}, bmax{p[9], p[10], p[11]};
            vec3 half = 0.5f * (bmax - bmin);
            vec3 cen = 0.5f * (bmax + bmin);
            vec3 t = abs_mat_mul(half, mat_from_quat(q));
            vec3 t2 = rot(cen, q);
            mx = t;
            mn = -t;
            mn = mn + (pos + t2);
            mx = mx + (pos + t2);
            return;
        }
        default: {  // TriangleShape.cs:111-134
            const float* v = verts + 3 * s.voff;
            vec3 a = rot(vec3{v[0], v[1], v[2]}, q), b = rot(vec3{v[3], v[4], v[5]}, q),
                 c = rot(vec3{v[6], v[7], v[8]}, q);
            vec3 extra{0.01f, 0.01f, 0.01f};
            mn = {fminf(fminf(a.x, b.x), c.x), fminf(fminf(a.y, b.y), c.y), fminf(fminf(a.z, b.z), c.z)};
            mx = {fmaxf(fmaxf(a.x, b.x), c.x), fmaxf(fmaxf(a.y, b.y), c.y), fmaxf(fmaxf(a.z, b.z), c.z)};
            mn = mn + (pos - extra);
            mx = mx + (pos + extra);
            return;
        }
    }
}

// MinkowskiDifference.Vertex (MinkowskiDifference.cs:20-42)
struct MVertex {
    vec3 V, A, B;
};

// MinkowskiDifference.Support / GetCenter (MinkowskiDifference.cs:57-90), B expressed in A's frame.
struct MinkowskiPair {
    ShapeView a, b;
    quat qb;
    vec3 pb;
    __device__ __forceinline__ MVertex support(vec3 d) const {
        MVertex v;
        vec3 tmp = -d;
        v.A = a.support(d);
        vec3 tmp2 = rot_inv(tmp, qb);
        v.B = b.support(tmp2);
        v.B = rot(v.B, qb);
        v.B = v.B + pb;
        v.V = v.A - v.B;
        return v;
    }
    __device__ __forceinline__ MVertex center() const {
        MVertex c;
        c.A = a.center();
        c.B = b.center();
        c.B = rot(c.B, qb);
        c.B = pb + c.B;
        c.V = c.A - c.B;
        return c;
    }
};

// Output of MPR for one pair, in A's frame until mapped back to world space.
struct MprOut {
    int hit;       // MPR overlap
    int deep;      // penetration above the EPA threshold: portal stored for the EPA kernel
    vec3 pa, pb, normal;
    float pen;
    MVertex portal[4];  // v0..v3 when deep
};

// MprEpaSolver.SolveMpr (NarrowPhase.cs:98-347).  When the penetration exceeds epa_threshold the
// portal is returned for EPA; the MPR answer (normal, depth, barycentric points) is also returned
// because it is kept whenever EPA fails (NarrowPhase.cs:287-310).
__device__ __forceinline__ void solve_mpr(const MinkowskiPair& mk, float epa_threshold, MprOut& out) {
    const float eps = 1e-5f;
    const int maxIter = 34;
    const float NumericEpsilon = 1e-16f;
    MVertex v0, v1, v2, v3, v4;
    vec3 t1{0, 0, 0}, t2{0, 0, 0}, t3{0, 0, 0};
    vec3 normal;
    float pen = 0.0f;
    out.hit = 0;
    out.deep = 0;
    v0 = mk.center();
    if (fabsf(v0.V.x) < NumericEpsilon && fabsf(v0.V.y) < NumericEpsilon && fabsf(v0.V.z) < NumericEpsilon)
        v0.V.x = 1e-05f;
    normal = -v0.V;
    v1 = mk.support(normal);
    out.pa = v1.A;
    out.pb = v1.B;
    out.normal = normal;
    out.pen = 0.0f;
    if (dot(v1.V, normal) <= 0.0f) return;
    normal = cross(v1.V, v0.V);
    if (len_sq(normal) < 1e-12f) {  // origin, v0 and v1 collinear (NarrowPhase.cs:160-174)
        normal = v1.V - v0.V;
        normal = normalize(normal);
        t1 = v1.A - v1.B;
        out.pen = dot(t1, normal);
        out.normal = normal;
        out.hit = 1;
        return;
    }
    v2 = mk.support(normal);
    if (dot(v2.V, normal) <= 0.0f) return;
    t1 = v1.V - v0.V;
    t2 = v2.V - v0.V;
    normal = cross(t1, t2);
    float dist = dot(normal, v0.V);
    if (dist > 0.0f) {
        MVertex t = v1;
        v1 = v2;
        v2 = t;
        normal = -normal;
    }
    int phase1 = 0, phase2 = 0;
    bool hit = false;
    while (true) {  // phase one: find a portal (:201-237)
        if (phase1 > maxIter) return;
        phase1++;
        v3 = mk.support(normal);
        if (dot(v3.V, normal) <= 0.0f) return;
        t1 = cross(v1.V, v3.V);
        if (dot(t1, v0.V) < 0.0f) {
            v2 = v3;
            t1 = v1.V - v0.V;
            t2 = v3.V - v0.V;
            normal = cross(t1, t2);
            continue;
        }
        t1 = cross(v3.V, v2.V);
        if (dot(t1, v0.V) < 0.0f) {
            v1 = v3;
            t1 = v3.V - v0.V;
            t2 = v2.V - v0.V;
            normal = cross(t1, t2);
            continue;
        }
        break;
    }
    while (true) {  // phase two: refine the portal (:241-346)
        phase2++;
        t1 = v2.V - v1.V;
        t2 = v3.V - v1.V;
        normal = cross(t1, t2);
        float nsq = len_sq(normal);
        if (nsq < NumericEpsilon) return;
        if (!hit) {
            float d = dot(normal, v1.V);
            hit = d >= 0;
        }
        v4 = mk.support(normal);
        t3 = v4.V - v3.V;
        float delta = dot(t3, normal);
        pen = dot(v4.V, normal);
        if (delta * delta <= eps * eps * nsq || pen <= 0.0f || phase2 > maxIter) {
            if (hit) {
                float invn = 1.0f / sqrtf(nsq);
                pen *= invn;
                if (pen > epa_threshold) {
                    out.deep = 1;
                    out.portal[0] = v0;
                    out.portal[1] = v1;
                    out.portal[2] = v2;
                    out.portal[3] = v3;
                }
                normal = normal * invn;
                t3 = cross(v1.V, t1);
                float gamma = dot(t3, normal) * invn;
                t3 = cross(t2, v1.V);
                float beta = dot(t3, normal) * invn;
                float alpha = 1.0f - gamma - beta;
                out.pa = alpha * v1.A + beta * v2.A + gamma * v3.A;
                out.pb = alpha * v1.B + beta * v2.B + gamma * v3.B;
                out.normal = normal;
                out.pen = pen;
                out.hit = 1;
            }
            return;
        }
        t1 = cross(v4.V, v0.V);
        float d = dot(t1, v1.V);
        if (d >= 0.0f) {
            d = dot(t1, v2.V);
            if (d >= 0.0f) v1 = v4;
            else v3 = v4;
        } else {
            d = dot(t1, v3.V);
            if (d >= 0.0f) v2 = v4;
            else v1 = v4;
        }
    }
}

// ---- One-shot manifold (CollisionManifold.cs) --------------------------------------------------
struct ManifoldBuf {
    vec3 mA[12], mB[12];
    int count;
};

namespace mf {
constexpr float DupSq = 0.001f;
__device__ __forceinline__ void push(vec3* arr, int& n, vec3 v) {  // PushLeft/PushRight (:58-88)
    if (n > 0 && len_sq(arr[0] - v) < DupSq) return;
    if (n > 1 && len_sq(arr[n - 1] - v) < DupSq) return;
    arr[n++] = v;
}
__device__ __forceinline__ bool polygon_sign(const vec3* poly, int n, vec3 normal, float& sign) {  // :92-113
    float area = 0.0f;
    vec3 prev = poly[n - 1];
    for (int i = 0; i < n; ++i) {
        vec3 cur = poly[i];
        area += dot(cross(prev, cur), normal);
        prev = cur;
    }
    if (fabsf(area) <= 1e-6f) { sign = 0.0f; return false; }
    sign = area < 0.0f ? -1.0f : 1.0f;
    return true;
}
__device__ __forceinline__ bool contains(const vec3* poly, int n, vec3 pt, vec3 normal, float sign) {  // :116-133
    vec3 prev = poly[n - 1];
    for (int i = 0; i < n; ++i) {
        vec3 cur = poly[i];
        float s = sign * dot(cross(cur - prev, pt - prev), normal);
        if (s < -1e-3f) return false;
        prev = cur;
    }
    return true;
}
__device__ __forceinline__ int clip_segment(vec3 start, vec3 end, const vec3* poly, int n, vec3 normal, float sign,
                                            vec3* out) {  // :137-185
    const float e = 1e-3f;
    float enter = 0.0f, exit = 1.0f;
    vec3 delta = end - start;
    for (int i = 0; i < n; ++i) {
        vec3 es = poly[i], ee = poly[(i + 1) % n];
        float ss = sign * dot(cross(ee - es, start - es), normal);
        float se = sign * dot(cross(ee - es, end - es), normal);
        bool si = ss >= -e, ei = se >= -e;
        if (!si && !ei) return 0;
        if (si && ei) continue;
        float den = ss - se;
        if (fabsf(den) <= e) return 0;
        float t = clampr(ss / den, 0.0f, 1.0f);
        if (!si) enter = fmaxr(enter, t);
        else exit = fminr(exit, t);
        if (exit < enter) return 0;
    }
    out[0] = start + enter * delta;
    if (len_sq(start + exit * delta - out[0]) < e) return 1;
    out[1] = start + exit * delta;
    return 2;
}
__device__ __forceinline__ float cross_n(vec3 l, vec3 r, vec3 n) { return dot(cross(l, r), n); }
__device__ __forceinline__ int intersect(vec3 sA, vec3 eA, vec3 sB, vec3 eB, vec3 normal, vec3* cA, vec3* cB) {  // :196-253
    const float e = 1e-3f;
    vec3 dA = eA - sA, dB = eB - sB, off = sB - sA;
    float cr = cross_n(dA, dB, normal);
    if (fabsf(cr) <= e) {
        if (fabsf(cross_n(off, dA, normal)) > e) return 0;
        float la = len_sq(dA), lb = len_sq(dB);
        if (la <= e || lb <= e) return 0;
        float pB0 = dot(sB - sA, dA) / la;
        float pB1 = dot(eB - sA, dA) / la;
        float enter = fmaxr(0.0f, fminr(pB0, pB1));
        float exit = fminr(1.0f, fmaxr(pB0, pB1));
        if (exit < enter - e) return 0;
        cA[0] = sA + enter * dA;
        float pA0 = clampr(dot(cA[0] - sB, dB) / lb, 0.0f, 1.0f);
        cB[0] = sB + pA0 * dB;
        cA[1] = sA + exit * dA;
        if (len_sq(cA[1] - cA[0]) <= e) return 1;
        float pA1 = clampr(dot(cA[1] - sB, dB) / lb, 0.0f, 1.0f);
        cB[1] = sB + pA1 * dB;
        return 2;
    }
    float t = cross_n(off, dB, normal) / cr;
    float u = cross_n(off, dA, normal) / cr;
    if (t < -e || t > 1.0f + e || u < -e || u > 1.0f + e) return 0;
    t = clampr(t, 0.0f, 1.0f);
    u = clampr(u, 0.0f, 1.0f);
    cA[0] = sA + t * dA;
    cB[0] = sB + u * dB;
    return 1;
}
__device__ __forceinline__ bool is_dup(const vec3* m, int n, vec3 p) {  // :256-264
    for (int i = 0; i < n; ++i)
        if (len_sq(m[i] - p) < DupSq) return true;
    return false;
}
__device__ __forceinline__ void add_on_a(ManifoldBuf& m, vec3 pt, vec3 pB, vec3 normal) {  // :268-278
    if (m.count == 12) return;
    float diff = dot(pt - pB, normal);
    if (diff < 0.0f) return;
    if (is_dup(m.mA, m.count, pt)) return;
    m.mA[m.count] = pt;
    m.mB[m.count++] = pt - diff * normal;
}
__device__ __forceinline__ void add_on_b(ManifoldBuf& m, vec3 pt, vec3 pA, vec3 normal) {  // :281-291
    if (m.count == 12) return;
    float diff = dot(pt - pA, normal);
    if (diff > 0.0f) return;
    if (is_dup(m.mB, m.count, pt)) return;
    m.mB[m.count] = pt;
    m.mA[m.count++] = pt - diff * normal;
}
__device__ __forceinline__ void add_pair(ManifoldBuf& m, vec3 pa, vec3 pb, vec3 normal) {  // :294-304
    if (m.count == 12) return;
    if (dot(pa - pb, normal) < 0.0f) return;
    if (is_dup(m.mA, m.count, pa)) return;
    if (is_dup(m.mB, m.count, pb)) return;
    m.mA[m.count] = pa;
    m.mB[m.count++] = pb;
}
// ReduceManifold (:321-375)
__device__ __forceinline__ void reduce(ManifoldBuf& m, vec3 normal) {
    if (m.count <= 4) return;
    vec3 c{0.0f, 0.0f, 0.0f};
    for (int i = 0; i < m.count; ++i) c = c + m.mA[i];
    c = c * float(1.0 / m.count);
    vec3 t1 = create_orthonormal(normal);
    vec3 t2 = cross(normal, t1);
    float ang[12];
    int order[12];
    for (int i = 0; i < m.count; ++i) {
        vec3 d = m.mA[i] - c;
        ang[i] = stable::atan2(dot(d, t2), dot(d, t1));
        order[i] = i;
    }
    for (int i = 1; i < m.count; ++i) {  // insertion sort (stable)
        int cur = order[i];
        float ca = ang[cur];
        int j = i - 1;
        while (j >= 0 && ang[order[j]] > ca) {
            order[j + 1] = order[j];
            j--;
        }
        order[j + 1] = cur;
    }
    vec3 rA[4], rB[4];
    for (int i = 0; i < 4; ++i) {
        int sel = order[((2 * i + 1) * m.count) / 8];
        rA[i] = m.mA[sel];
        rB[i] = m.mB[sel];
    }
    for (int i = 0; i < 4; ++i) {
        m.mA[i] = rA[i];
        m.mB[i] = rB[i];
    }
    m.count = 4;
}
}  // namespace mf

// CollisionManifold.BuildManifold (CollisionManifold.cs:393-566); results in world space.
__device__ __noinline__ void build_manifold(const ShapeView& sa, const ShapeView& sb, quat qa, quat qb, vec3 posA,
                                            vec3 posB, vec3 pA, vec3 pB, vec3 normal, ManifoldBuf& m) {
    const float hex[12] = {1.0f, 0.0f, 0.5f, 0.86602540378f, -0.5f, 0.86602540378f,
                           -1.0f, 0.0f, -0.5f, -0.86602540378f, 0.5f, -0.86602540378f};
    const float pert = 0.01f;
    int lc = 0, rc = 0;
    m.count = 0;
    vec3 cv1 = create_orthonormal(normal);
    vec3 cv2 = cross(normal, cv1);
    vec3 relB = posB - posA, pAl = pA - posA, pBl = pB - posA;
    vec3 left[6], right[6];
    for (int e = 0; e < 6; ++e) {
        vec3 ptn = normal + hex[2 * e] * pert * cv1 + hex[2 * e + 1] * pert * cv2;
        vec3 tmp = rot_inv(ptn, qa);
        vec3 np1 = sa.support(tmp);
        np1 = rot(np1, qa);
        mf::push(left, lc, np1);
        ptn = -ptn;
        tmp = rot_inv(ptn, qb);
        vec3 np2 = sb.support(tmp);
        np2 = rot(np2, qb);
        np2 = np2 + relB;
        mf::push(right, rc, np2);
    }
    float ls = 0.0f, rs = 0.0f;
    bool lp = lc > 2 && mf::polygon_sign(left, lc, normal, ls);
    bool rp = rc > 2 && mf::polygon_sign(right, rc, normal, rs);
    if (lp) {
        for (int e = 0; e < rc; ++e) {
            vec3 p = right[e];
            if (!mf::contains(left, lc, p, normal, ls)) continue;
            mf::add_on_b(m, p, pAl, normal);
            if (m.count == 12) goto finalize;
        }
        if (rc == 2) {
            vec3 cl[2];
            int n = mf::clip_segment(right[0], right[1], left, lc, normal, ls, cl);
            for (int i = 0; i < n; ++i) {
                mf::add_on_b(m, cl[i], pAl, normal);
                if (m.count == 12) goto finalize;
            }
        }
    }
    if (rp) {
        for (int e = 0; e < lc; ++e) {
            vec3 p = left[e];
            if (!mf::contains(right, rc, p, normal, rs)) continue;
            mf::add_on_a(m, p, pBl, normal);
            if (m.count == 12) goto finalize;
        }
        if (lc == 2) {
            vec3 cl[2];
            int n = mf::clip_segment(left[0], left[1], right, rc, normal, rs, cl);
            for (int i = 0; i < n; ++i) {
                mf::add_on_a(m, cl[i], pBl, normal);
                if (m.count == 12) goto finalize;
            }
        }
    }
    if (lc == 2 && rc == 2) {
        vec3 cA[2], cB[2];
        int n = mf::intersect(left[0], left[1], right[0], right[1], normal, cA, cB);
        for (int i = 0; i < n; ++i) {
            mf::add_pair(m, cA[i], cB[i], normal);
            if (m.count == 12) goto finalize;
        }
    }
    if (lp && rp) {
        vec3 cA[2], cB[2];
        for (int i = 0; i < lc; ++i) {
            vec3 sA = left[i], eA = left[(i + 1) % lc];
            for (int j = 0; j < rc; ++j) {
                vec3 sB = right[j], eB = right[(j + 1) % rc];
                int n = mf::intersect(sA, eA, sB, eB, normal, cA, cB);
                for (int k = 0; k < n; ++k) {
                    mf::add_pair(m, cA[k], cB[k], normal);
                    if (m.count == 12) goto finalize;
                }
            }
        }
    }
    if (m.count <= 1) {
        m.mA[0] = pAl;
        m.mB[0] = pBl;
        m.count = 1;
    } else if (m.count < 4) {
        m.mA[m.count] = pAl;
        m.mB[m.count] = pBl;
        m.count++;
    }
finalize:
    mf::reduce(m, normal);
    for (int i = 0; i < m.count; ++i) {
        m.mA[i] = m.mA[i] + posA;
        m.mB[i] = m.mB[i] + posA;
    }
}

}  // namespace jh
