// ORACLE — test infrastructure only (see orc_math.hpp).  Shapes (support maps, centres, world
// AABBs), the Minkowski difference, XenoCollide MPR, the FP64 EPA with filtered exact predicates,
// and the one-shot contact manifold, restated from Jitter2 in the cited operation order.
#pragma once
#include <algorithm>
#include <vector>

#include "../include/jitterhip.h"
#include "orc_math.hpp"

namespace orc {

template <class R> struct Shape {
    int type = JH_SHAPE_BOX;
    int body = 0;
    uint64_t id = 0;
    R p[12] = {};
    const float* verts = nullptr;  // point-cloud / triangle vertices (FP32 as uploaded)
    int nverts = 0;

    // ConvexHullShape.InternalSupportMap: hill climb from vertex 0 over the hull's vertex graph
    // (adjacency after the vertices: per vertex (first, end) into the neighbour list, then the list),
    // with the plateau pass; capped at nverts + 1 moves, never reached by a valid hull.
    V3<R> hull_support(V3<R> d) const {
        const float* adj = verts + 3 * nverts;
        const float* nbl = adj + 2 * nverts;
        auto vx = [&](int i) { return V3<R>{R(verts[3 * i]), R(verts[3 * i + 1]), R(verts[3 * i + 2])}; };
        const R eps = R(1e-12);
        int cur = 0;
        R dp = dot(vx(0), d);
        int moves = 0;
    main_pass:
        if (moves > nverts) return vx(cur);
        {
            bool verify = false;
            V3<R> vdir{R(1), R(1), R(1)};  // JVector.Arbitrary
            for (int i = int(adj[2 * cur]); i < int(adj[2 * cur + 1]); ++i) {
                int nb = int(nbl[i]);
                R np = dot(vx(nb), d);
                if (std::fabs(np - dp) < eps) {
                    vdir = vx(nb) - vx(cur);
                    verify = true;
                }
                if (np > dp + eps) {
                    dp = np;
                    cur = nb;
                    ++moves;
                    goto main_pass;
                }
            }
            if (!verify) return vx(cur);
            R d0 = dot(vdir, vx(cur));
        secondary:
            if (moves > nverts) return vx(cur);
            ++moves;
            for (int i = int(adj[2 * cur]); i < int(adj[2 * cur + 1]); ++i) {
                int nb = int(nbl[i]);
                R np = dot(vx(nb), d);
                if (np > dp + eps) {
                    dp = np;
                    cur = nb;
                    goto main_pass;
                }
                if (std::fabs(np - dp) < eps) {
                    R d1 = dot(vdir, vx(nb));
                    if (d1 > d0 + eps) {
                        d0 = d1;
                        cur = nb;
                        goto secondary;
                    }
                }
            }
        }
        return vx(cur);
    }

    // ISupportMappable.SupportMap (local frame, direction not normalised).
    V3<R> support(V3<R> d) const {
        switch (type) {
            case JH_SHAPE_BOX:  // BoxShape.SupportMap (BoxShape.cs:86-91)
                return {R(signbit_i(d.x)) * p[0], R(signbit_i(d.y)) * p[1], R(signbit_i(d.z)) * p[2]};
            case JH_SHAPE_SPHERE:  // SphereShape.SupportMap (SphereShape.cs:54-58)
                return normalize(d) * p[0];
            case JH_SHAPE_CAPSULE: {  // CapsuleShape.SupportMap (CapsuleShape.cs:74-89)
                V3<R> r = normalize(d) * p[0];
                r.y += R(sign_i(d.y)) * p[1];
                return r;
            }
            case JH_SHAPE_POINTCLOUD: {  // PointCloudShape.SupportMap -> VertexSupportMap (:64-109) + shift
                R best = -std::numeric_limits<R>::max();
                int idx = 0;
                for (int i = 0; i < nverts; ++i) {
                    R vx = R(verts[3 * i]), vy = R(verts[3 * i + 1]), vz = R(verts[3 * i + 2]);
                    R dx = vx * d.x, dy = vy * d.y, dz = vz * d.z;
                    R dp = dx + (dy + dz);
                    if (dp >= best) { best = dp; idx = i; }
                }
                V3<R> r{R(verts[3 * idx]), R(verts[3 * idx + 1]), R(verts[3 * idx + 2])};
                return r + V3<R>{p[3], p[4], p[5]};
            }
            case JH_SHAPE_CONVEXHULL:  // ConvexHullShape.InternalSupportMap (ConvexHullShape.cs:378-450) + shift
                return hull_support(d) + V3<R>{p[3], p[4], p[5]};
            default: {  // TriangleShape.SupportMap (TriangleShape.cs:167-191)
                V3<R> a{R(verts[0]), R(verts[1]), R(verts[2])};
                V3<R> b{R(verts[3]), R(verts[4]), R(verts[5])};
                V3<R> c{R(verts[6]), R(verts[7]), R(verts[8])};
                R mn = dot(a, d);
                R dp = dot(b, d);
                V3<R> r = a;
                if (dp > mn) { mn = dp; r = b; }
                dp = dot(c, d);
                if (dp > mn) r = c;
                return r;
            }
        }
    }
    // ISupportMappable.GetCenter.
    V3<R> center() const {
        if (type == JH_SHAPE_POINTCLOUD || type == JH_SHAPE_CONVEXHULL) return {p[0], p[1], p[2]};
        if (type == JH_SHAPE_TRIANGLE) {
            V3<R> a{R(verts[0]), R(verts[1]), R(verts[2])};
            V3<R> b{R(verts[3]), R(verts[4]), R(verts[5])};
            V3<R> c{R(verts[6]), R(verts[7]), R(verts[8])};
            return R(1.0 / 3.0) * (a + b + c);
        }
        return {R(0), R(0), R(0)};
    }
    // CalculateBoundingBox of each shape from the body pose.
    void aabb(Q4<R> q, V3<R> pos, V3<R>& mn, V3<R>& mx) const {
        switch (type) {
            case JH_SHAPE_BOX: {  // BoxShape.cs:180-186
                M3<R> absm = mat_abs(mat_from_quat(q));
                V3<R> ths = mat_mul(V3<R>{p[0], p[1], p[2]}, absm);
                mn = pos - ths;
                mx = pos + ths;
                return;
            }
            case JH_SHAPE_SPHERE: {  // SphereShape.cs:65-71
                mn = V3<R>{-p[0], -p[0], -p[0]} + pos;
                mx = V3<R>{p[0], p[0], p[0]} + pos;
                return;
            }
            case JH_SHAPE_CAPSULE: {  // CapsuleShape.cs:97-107
                V3<R> delta = p[1] * basis_y(q);
                mx = {p[0] + std::fabs(delta.x), p[0] + std::fabs(delta.y), p[0] + std::fabs(delta.z)};
                mn = -mx;
                mn = mn + pos;
                mx = mx + pos;
                return;
            }
            case JH_SHAPE_POINTCLOUD:    // PointCloudShape.cs:140-154
            case JH_SHAPE_CONVEXHULL: {  // ConvexHullShape.cs:334-349 (the same formula)
                V3<R> bmin{p[6], p[7], p[8]}, bmax{p[9], p[10], p[11]};
                V3<R> half = R(0.5) * (bmax - bmin);
                V3<R> cen = R(0.5) * (bmax + bmin);
                M3<R> absm = mat_abs(mat_from_quat(q));
                V3<R> t = mat_mul(half, absm);
                V3<R> t2 = rotate(cen, q);
                mx = t;
                mn = -t;
                mn = mn + (pos + t2);
                mx = mx + (pos + t2);
                return;
            }
            default: {  // TriangleShape.cs:111-134: rotated corners, then position -/+ 0.01 margin
                V3<R> a = rotate(V3<R>{R(verts[0]), R(verts[1]), R(verts[2])}, q);
                V3<R> b = rotate(V3<R>{R(verts[3]), R(verts[4]), R(verts[5])}, q);
                V3<R> c = rotate(V3<R>{R(verts[6]), R(verts[7]), R(verts[8])}, q);
                const V3<R> extra{R(0.01), R(0.01), R(0.01)};
                mn = {std::min(std::min(a.x, b.x), c.x), std::min(std::min(a.y, b.y), c.y),
                      std::min(std::min(a.z, b.z), c.z)};
                mx = {std::max(std::max(a.x, b.x), c.x), std::max(std::max(a.y, b.y), c.y),
                      std::max(std::max(a.z, b.z), c.z)};
                mn = mn + (pos - extra);
                mx = mx + (pos + extra);
                return;
            }
        }
    }
};

// JBoundingBox.Disjoint (JBoundingBox.cs:314-318): strict comparisons, touching boxes overlap.
template <class R> inline bool disjoint(V3<R> amin, V3<R> amax, V3<R> bmin, V3<R> bmax) {
    return amax.x < bmin.x || amin.x > bmax.x || amax.y < bmin.y || amin.y > bmax.y || amax.z < bmin.z ||
           amin.z > bmax.z;
}

// MinkowskiDifference.Vertex (MinkowskiDifference.cs:20-42).
template <class R> struct Vertex { V3<R> V, A, B; };

// MinkowskiDifference.Support / GetCenter (MinkowskiDifference.cs:57-90), B in A's frame.
template <class R> struct MinkowskiPair {
    const Shape<R>* a;
    const Shape<R>* b;
    Q4<R> qb;
    V3<R> pb;
    Vertex<R> support(V3<R> d) const {
        Vertex<R> v;
        V3<R> tmp = -d;
        v.A = a->support(d);
        V3<R> tmp2 = rotate_inv(tmp, qb);
        v.B = b->support(tmp2);
        v.B = rotate(v.B, qb);
        v.B = v.B + pb;
        v.V = v.A - v.B;
        return v;
    }
    Vertex<R> center() const {
        Vertex<R> c;
        c.A = a->center();
        c.B = b->center();
        c.B = rotate(c.B, qb);
        c.B = pb + c.B;
        c.V = c.A - c.B;
        return c;
    }
};

// ---- Exact arithmetic (ExactArithmetic/ExpansionArithmetic.cs) ------------------------------
namespace exact {
inline int difference(double a, double b, double* r) {  // :106-116
    double value = a - b;
    double vb = a - value;
    double residual = (a - (value + vb)) + (vb - b);
    int n = 0;
    if (residual != 0) r[n++] = residual;
    if (value != 0) r[n++] = value;
    return n;
}
inline int grow(double* e, int len, double value) {  // :88-103
    int count = 0;
    for (int i = 0; i < len; ++i) {
        double part = e[i];
        double sum = value + part;
        double vp = sum - value;
        double residual = (value - (sum - vp)) + (part - vp);
        if (residual != 0) e[count++] = residual;
        value = sum;
    }
    if (value != 0) e[count++] = value;
    return count;
}
inline int add_product(double* res, int len, const double* a, int na, const double* b, int nb, bool neg) {  // :71-86
    for (int i = 0; i < na; ++i)
        for (int j = 0; j < nb; ++j) {
            double x = a[i], y = b[j];
            double product = x * y;
            double residual = std::fma(x, y, -product);
            if (neg) { product = -product; residual = -residual; }
            if (residual != 0) len = grow(res, len, residual);
            if (product != 0) len = grow(res, len, product);
        }
    return len;
}
inline int cross(const double* a, int na, const double* b, int nb, const double* c, int nc, const double* d, int nd,
                 double* res) {  // :63-69
    int len = add_product(res, 0, a, na, b, nb, false);
    return add_product(res, len, c, nc, d, nd, true);
}
inline double estimate(const double* e, int n) {
    double v = 0;
    for (int i = 0; i < n; ++i) v += e[i];
    return v;
}
// ExpansionArithmetic.Side (:21-41): sign of dot(cross(a-b, a-c), p-a), exactly.
template <class R> inline int side(V3<R> a, V3<R> b, V3<R> c, V3<R> p) {
    double ux[2], uy[2], uz[2], vx[2], vy[2], vz[2], dx[2], dy[2], dz[2];
    int nx = difference(a.x, b.x, ux), ny = difference(a.y, b.y, uy), nz = difference(a.z, b.z, uz);
    int mx = difference(a.x, c.x, vx), my = difference(a.y, c.y, vy), mz = difference(a.z, c.z, vz);
    int px = difference(p.x, a.x, dx), py = difference(p.y, a.y, dy), pz = difference(p.z, a.z, dz);
    double cr[16], det[192];
    int len = cross(uy, ny, vz, mz, uz, nz, vy, my, cr);
    int count = add_product(det, 0, cr, len, dx, px, false);
    len = cross(uz, nz, vx, mx, ux, nx, vz, mz, cr);
    count = add_product(det, count, cr, len, dy, py, false);
    len = cross(ux, nx, vy, my, uy, ny, vx, mx, cr);
    count = add_product(det, count, cr, len, dz, pz, false);
    if (count == 0) return 0;
    double top = det[count - 1];
    return top > 0 ? 1 : (top < 0 ? -1 : 0);
}
// ExpansionArithmetic.Normal (:43-55).
template <class R> inline void normal(V3<R> a, V3<R> b, V3<R> c, double& x, double& y, double& z) {
    double ux[2], uy[2], uz[2], vx[2], vy[2], vz[2];
    int nx = difference(a.x, b.x, ux), ny = difference(a.y, b.y, uy), nz = difference(a.z, b.z, uz);
    int mx = difference(a.x, c.x, vx), my = difference(a.y, c.y, vy), mz = difference(a.z, c.z, vz);
    double cr[16];
    x = estimate(cr, cross(uy, ny, vz, mz, uz, nz, vy, my, cr));
    y = estimate(cr, cross(uz, nz, vx, mx, ux, nx, vz, mz, cr));
    z = estimate(cr, cross(ux, nx, vy, my, uy, ny, vx, mx, cr));
}
}  // namespace exact

// PlaneFilter (ExactArithmetic/PlaneFilter.cs:17-79).
struct PlaneFilter {
    static constexpr double UnitRoundoff = 1.1102230246251565e-16;
    static constexpr double ErrorFactor = 16 * UnitRoundoff;
    double ex = 0, ey = 0, ez = 0;
    double nx = 0, ny = 0, nz = 0;
    template <class R> void init(V3<R> a, V3<R> b, V3<R> c) {
        double ux = (double)a.x - b.x, uy = (double)a.y - b.y, uz = (double)a.z - b.z;
        double vx = (double)a.x - c.x, vy = (double)a.y - c.y, vz = (double)a.z - c.z;
        double yz = uy * vz, zy = uz * vy, zx = uz * vx, xz = ux * vz, xy = ux * vy, yx = uy * vx;
        nx = yz - zy;
        ny = zx - xz;
        nz = xy - yx;
        ex = ErrorFactor * (std::fabs(yz) + std::fabs(zy));
        ey = ErrorFactor * (std::fabs(zx) + std::fabs(xz));
        ez = ErrorFactor * (std::fabs(xy) + std::fabs(yx));
    }
    bool reliable_normal() const {
        return std::max(std::fabs(nx), std::max(std::fabs(ny), std::fabs(nz))) > 1048576 * std::max(ex, std::max(ey, ez));
    }
    void negate() { nx = -nx; ny = -ny; nz = -nz; }
    template <class R> bool try_side(V3<R> anchor, V3<R> point, int& sign) const {
        sign = 0;
        double dx = (double)point.x - anchor.x, dy = (double)point.y - anchor.y, dz = (double)point.z - anchor.z;
        double x = nx * dx, y = ny * dy, z = nz * dz;
        double value = (x + y) + z;
        double roundoff = ErrorFactor * ((std::fabs(x) + std::fabs(y)) + std::fabs(z));
        double coef = (ex * std::fabs(dx) + ey * std::fabs(dy)) + ez * std::fabs(dz);
        double bound = roundoff + 2 * coef;
        if (std::fabs(value) <= bound) return false;
        sign = value > 0 ? 1 : -1;
        return true;
    }
    static bool safe_coord(double v) {
        return v == 0 || (std::fabs(v) >= 6.223015277861142e-61 && std::fabs(v) <= 1.6069380442589903e60);
    }
    template <class R> static bool is_safe(V3<R> p) { return safe_coord(p.x) && safe_coord(p.y) && safe_coord(p.z); }
};

struct D3 {
    double x, y, z;
};
inline D3 operator+(D3 a, D3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
inline D3 operator-(D3 a, D3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline D3 operator*(D3 a, double s) { return {a.x * s, a.y * s, a.z * s}; }
inline double ddot(D3 a, D3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline D3 dcross(D3 a, D3 b) { return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
inline bool dfinite(D3 a) { return std::isfinite(a.x) && std::isfinite(a.y) && std::isfinite(a.z); }

// ConvexPolytope (NarrowPhase/ConvexPolytope.cs:27-398).
template <class R> struct Polytope {
    enum Result { Added, NoHorizon, Degenerate, CapacityExceeded, NumericalFailure };
    static constexpr int MaxVertices = 128, MaxTriangles = 256;
    struct Triangle {
        int16_t a, b, c;
        bool facing_origin;
        V3<R> normal, closest;
        R normal_sq, closest_sq;
    };
    struct FaceData {
        PlaneFilter filter;
        D3 bary;
        double dist_sq;
    };
    Vertex<R> vertices[MaxVertices];
    Triangle tris[MaxTriangles];
    FaceData face[MaxTriangles];
    int vcount = 0, tcount = 0;
    bool origin_enclosed = false;

    static D3 fromv(V3<R> v) { return {double(v.x), double(v.y), double(v.z)}; }
    static V3<R> tov(D3 d) { return {R(d.x), R(d.y), R(d.z)}; }

    // InitTetrahedron() (:113-131).
    bool init_tetrahedron() {
        vcount = 4;
        tcount = 0;
        origin_enclosed = false;
        for (int i = 0; i < 4; ++i)
            if (!PlaneFilter::is_safe(vertices[i].V)) return false;
        static const int16_t idx[12] = {0, 2, 1, 0, 1, 3, 0, 3, 2, 1, 2, 3};
        for (int i = 0; i < 4; ++i)
            if (create_triangle(idx[3 * i], idx[3 * i + 1], idx[3 * i + 2], tris[i], face[i]) != Added) return false;
        tcount = 4;
        return true;
    }
    // GetClosestTriangle (:147-166).
    Triangle closest_triangle() {
        int closest = 0;
        double minimum = std::numeric_limits<double>::infinity();
        bool skip = origin_enclosed;
        origin_enclosed = true;
        for (int i = 0; i < tcount; ++i) {
            if (face[i].dist_sq < minimum) { minimum = face[i].dist_sq; closest = i; }
            if (!tris[i].facing_origin) origin_enclosed = skip;
        }
        return tris[closest];
    }
    // CalculatePoints (:169-186).
    bool calculate_points(const Triangle& t, V3<R>& pa, V3<R>& pb) const {
        for (int i = 0; i < tcount; ++i) {
            if (tris[i].a != t.a || tris[i].b != t.b || tris[i].c != t.c) continue;
            D3 bc = face[i].bary;
            pa = tov(fromv(vertices[t.a].A) * bc.x + fromv(vertices[t.b].A) * bc.y + fromv(vertices[t.c].A) * bc.z);
            pb = tov(fromv(vertices[t.a].B) * bc.x + fromv(vertices[t.b].B) * bc.y + fromv(vertices[t.c].B) * bc.z);
            return true;
        }
        return false;
    }
    // Side (:320-326).
    int side(const FaceData& d, int a, int b, int c, int pi) const {
        V3<R> pt = pi < 0 ? V3<R>{R(0), R(0), R(0)} : vertices[pi].V;
        if (veq(pt, vertices[a].V) || veq(pt, vertices[b].V) || veq(pt, vertices[c].V)) return 0;
        int s;
        if (d.filter.try_side(vertices[a].V, pt, s)) return s;
        return exact::side(vertices[a].V, vertices[b].V, vertices[c].V, pt);
    }
    // CalcBarycentric (:328-395).
    static bool calc_bary(D3 a, D3 b, D3 c, D3 n, double nscale, double nsq, D3& out) {
        D3 u = a - b, v = a - c;
        double gamma = (ddot(dcross(u, a), n) / nsq) / nscale;
        double beta = (ddot(dcross(a, v), n) / nsq) / nscale;
        double alpha = 1 - gamma - beta;
        bool clamped = false;
        if (alpha >= 0 && beta < 0) {
            double t = ddot(a, u);
            if (gamma < 0 && t > 0) {
                beta = std::min(1.0, t / ddot(u, u));
                alpha = 1 - beta;
                gamma = 0;
            } else {
                gamma = rclamp(ddot(a, v) / ddot(v, v), 0.0, 1.0);
                alpha = 1 - gamma;
                beta = 0;
            }
            clamped = true;
        } else if (beta >= 0 && gamma < 0) {
            D3 w = b - c;
            double t = ddot(b, w);
            if (alpha < 0 && t > 0) {
                gamma = std::min(1.0, t / ddot(w, w));
                beta = 1 - gamma;
                alpha = 0;
            } else {
                alpha = rclamp(-ddot(b, u) / ddot(u, u), 0.0, 1.0);
                beta = 1 - alpha;
                gamma = 0;
            }
            clamped = true;
        } else if (gamma >= 0 && alpha < 0) {
            D3 w = b - c;
            double t = -ddot(c, v);
            if (beta < 0 && t > 0) {
                alpha = std::min(1.0, t / ddot(v, v));
                gamma = 1 - alpha;
                beta = 0;
            } else {
                beta = rclamp(-ddot(c, w) / ddot(w, w), 0.0, 1.0);
                gamma = 1 - beta;
                alpha = 0;
            }
            clamped = true;
        }
        out = {alpha, beta, gamma};
        return clamped;
    }
    // CreateTriangle (:267-318).
    Result create_triangle(int16_t a, int16_t b, int16_t c, Triangle& tri, FaceData& data) {
        tri = Triangle{};
        data = FaceData{};
        data.filter.init(vertices[a].V, vertices[b].V, vertices[c].V);
        int orientation = 0;
        for (int16_t i = 0; i < 4 && orientation == 0; ++i) {
            if (i == a || i == b || i == c) continue;
            orientation = -side(data, a, b, c, i);
        }
        if (orientation == 0) return Degenerate;
        if (orientation < 0) {
            std::swap(a, b);
            data.filter.negate();
        }
        bool facing = side(data, a, b, c, -1) <= 0;
        D3 raw{data.filter.nx, data.filter.ny, data.filter.nz};
        if (!data.filter.reliable_normal()) {
            double x, y, z;
            exact::normal(vertices[a].V, vertices[b].V, vertices[c].V, x, y, z);
            raw = {x, y, z};
        }
        double nscale = std::max(std::fabs(raw.x), std::max(std::fabs(raw.y), std::fabs(raw.z)));
        if (!(nscale > 0) || !std::isfinite(nscale)) return NumericalFailure;
        D3 sn{raw.x / nscale, raw.y / nscale, raw.z / nscale};
        double nsq = ddot(sn, sn);
        D3 va = fromv(vertices[a].V), vb = fromv(vertices[b].V), vc = fromv(vertices[c].V);
        bool clamped = calc_bary(va, vb, vc, sn, nscale, nsq, data.bary);
        D3 closest = clamped ? va * data.bary.x + vb * data.bary.y + vc * data.bary.z : sn * (ddot(sn, va) / nsq);
        data.dist_sq = ddot(closest, closest);
        if (!dfinite(closest) || !dfinite(data.bary) || !std::isfinite(data.dist_sq)) return NumericalFailure;
        tri.a = a;
        tri.b = b;
        tri.c = c;
        tri.facing_origin = facing;
        tri.normal = tov(sn);
        tri.normal_sq = len_sq(tri.normal);
        tri.closest = tov(closest);
        tri.closest_sq = R(data.dist_sq);
        bool fin = std::isfinite(tri.closest.x) && std::isfinite(tri.closest.y) && std::isfinite(tri.closest.z) &&
                   std::isfinite(tri.closest_sq);
        return fin ? Added : NumericalFailure;
    }
    // AddVertexDetailed (:199-265).
    Result add_vertex(const Vertex<R>& vx) {
        if (tcount == 0 || !PlaneFilter::is_safe(vx.V)) return NumericalFailure;
        if (vcount == MaxVertices) return CapacityExceeded;
        vertices[vcount] = vx;
        bool lit[MaxTriangles];
        int16_t ea[3 * MaxTriangles], eb[3 * MaxTriangles];
        int ecount = 0, litcount = 0;
        for (int index = tcount; index-- > 0;) {
            const Triangle& f = tris[index];
            lit[index] = side(face[index], f.a, f.b, f.c, vcount) > 0;
            if (!lit[index]) continue;
            litcount++;
            int16_t corners[3] = {f.a, f.b, f.c};
            for (int k = 0; k < 3; ++k) {
                int16_t A = corners[k], B = corners[(k + 1) % 3];
                bool added = true;
                for (int e = ecount; e-- > 0;) {
                    bool m = (ea[e] == A && eb[e] == B) || (ea[e] == B && eb[e] == A);
                    if (!m) continue;
                    --ecount;
                    ea[e] = ea[ecount];
                    eb[e] = eb[ecount];
                    added = false;
                    break;
                }
                if (added) { ea[ecount] = A; eb[ecount] = B; ecount++; }
            }
        }
        if (ecount == 0) return NoHorizon;
        if (tcount - litcount + ecount > MaxTriangles) return CapacityExceeded;
        int surv = 0;
        for (int i = 0; i < tcount; ++i) {
            if (lit[i]) continue;
            if (surv != i) { tris[surv] = tris[i]; face[surv] = face[i]; }
            surv++;
        }
        for (int i = 0; i < ecount; ++i) {
            int index = surv + i;
            Result r = create_triangle(ea[i], eb[i], int16_t(vcount), tris[index], face[index]);
            if (r != Added) {
                tcount = 0;
                origin_enclosed = false;
                return r;
            }
        }
        tcount = surv + ecount;
        vcount++;
        return Added;
    }
};

template <class R> struct MprResult {
    bool hit = false;
    bool epa = false;
    V3<R> pa{}, pb{}, normal{};
    R penetration = 0;
};

// MprEpaSolver.SolveMprEpa (NarrowPhase.cs:39-95).
template <class R>
bool solve_epa(Polytope<R>& poly, const MinkowskiPair<R>& mk, V3<R>& p1, V3<R>& p2, V3<R>& normal, R& pen) {
    const R eps = R(1e-5);
    const int maxIter = 85;
    const R NumericEpsilon = R(1e-16);
    if (!poly.init_tetrahedron()) return false;
    int iter = 0;
    typename Polytope<R>::Triangle ctri{};
    bool converged = false;
    while (++iter < maxIter) {
        ctri = poly.closest_triangle();
        V3<R> dir = ctri.closest;
        R dirsq = ctri.closest_sq;
        if (!poly.origin_enclosed) dir = -dir;
        if (ctri.closest_sq < NumericEpsilon) { dir = ctri.normal; dirsq = ctri.normal_sq; }
        Vertex<R> v = mk.support(dir);
        R delta = dot(ctri.closest - v.V, dir);
        if (delta * delta <= eps * eps * dirsq) { converged = true; break; }
        auto r = poly.add_vertex(v);
        if (r == Polytope<R>::NoHorizon) { converged = true; break; }
        if (r != Polytope<R>::Added) return false;
    }
    if (!converged) return false;
    if (!poly.origin_enclosed) return false;
    if (!poly.calculate_points(ctri, p1, p2)) return false;
    normal = ctri.normal * (R(1.0) / std::sqrt(ctri.normal_sq));
    pen = std::sqrt(ctri.closest_sq);
    return true;
}

// MprEpaSolver.SolveMpr (NarrowPhase.cs:98-347) in A's frame.
template <class R> MprResult<R> solve_mpr(const MinkowskiPair<R>& mk, R epaThreshold, Polytope<R>* poly) {
    const R eps = R(1e-5);
    const int maxIter = 34;
    const R NumericEpsilon = R(1e-16);
    MprResult<R> out;
    Vertex<R> v0, v1, v2, v3, v4;
    V3<R> t1{}, t2{}, t3{};
    V3<R>& normal = out.normal;
    R& pen = out.penetration;
    pen = R(0);
    v0 = mk.center();
    if (std::fabs(v0.V.x) < NumericEpsilon && std::fabs(v0.V.y) < NumericEpsilon && std::fabs(v0.V.z) < NumericEpsilon)
        v0.V.x = R(1e-05);
    normal = -v0.V;
    v1 = mk.support(normal);
    out.pa = v1.A;
    out.pb = v1.B;
    if (dot(v1.V, normal) <= R(0.0)) return out;
    normal = cross(v1.V, v0.V);
    if (len_sq(normal) < R(1e-12)) {
        normal = v1.V - v0.V;
        normal = normalize(normal);
        t1 = v1.A - v1.B;
        pen = dot(t1, normal);
        out.hit = true;
        return out;
    }
    v2 = mk.support(normal);
    if (dot(v2.V, normal) <= R(0.0)) return out;
    t1 = v1.V - v0.V;
    t2 = v2.V - v0.V;
    normal = cross(t1, t2);
    R dist = dot(normal, v0.V);
    if (dist > R(0.0)) {
        std::swap(v1.V, v2.V);
        std::swap(v1.A, v2.A);
        std::swap(v1.B, v2.B);
        normal = -normal;
    }
    int phase2 = 0, phase1 = 0;
    bool hit = false;
    while (true) {
        if (phase1 > maxIter) return out;
        phase1++;
        v3 = mk.support(normal);
        if (dot(v3.V, normal) <= R(0.0)) return out;
        t1 = cross(v1.V, v3.V);
        if (dot(t1, v0.V) < R(0.0)) {
            v2 = v3;
            t1 = v1.V - v0.V;
            t2 = v3.V - v0.V;
            normal = cross(t1, t2);
            continue;
        }
        t1 = cross(v3.V, v2.V);
        if (dot(t1, v0.V) < R(0.0)) {
            v1 = v3;
            t1 = v3.V - v0.V;
            t2 = v2.V - v0.V;
            normal = cross(t1, t2);
            continue;
        }
        break;
    }
    while (true) {
        phase2++;
        t1 = v2.V - v1.V;
        t2 = v3.V - v1.V;
        normal = cross(t1, t2);
        R nsq = len_sq(normal);
        if (nsq < NumericEpsilon) return out;
        if (!hit) {
            R d = dot(normal, v1.V);
            hit = d >= 0;
        }
        v4 = mk.support(normal);
        t3 = v4.V - v3.V;
        R delta = dot(t3, normal);
        pen = dot(v4.V, normal);
        if (delta * delta <= eps * eps * nsq || pen <= R(0.0) || phase2 > maxIter) {
            if (hit) {
                R invn = R(1.0) / std::sqrt(nsq);
                pen *= invn;
                if (pen > epaThreshold && poly) {
                    poly->vertices[0] = v0;
                    poly->vertices[1] = v1;
                    poly->vertices[2] = v2;
                    poly->vertices[3] = v3;
                    if (solve_epa(*poly, mk, out.pa, out.pb, out.normal, out.penetration)) {
                        out.hit = true;
                        out.epa = true;
                        return out;
                    }
                }
                normal = normal * invn;
                t3 = cross(v1.V, t1);
                R gamma = dot(t3, normal) * invn;
                t3 = cross(t2, v1.V);
                R beta = dot(t3, normal) * invn;
                R alpha = R(1.0) - gamma - beta;
                out.pa = alpha * v1.A + beta * v2.A + gamma * v3.A;
                out.pb = alpha * v1.B + beta * v2.B + gamma * v3.B;
            }
            out.hit = hit;
            return out;
        }
        t1 = cross(v4.V, v0.V);
        R d = dot(t1, v1.V);
        if (d >= R(0.0)) {
            d = dot(t1, v2.V);
            if (d >= R(0.0)) v1 = v4;
            else v3 = v4;
        } else {
            d = dot(t1, v3.V);
            if (d >= R(0.0)) v2 = v4;
            else v1 = v4;
        }
    }
}

// NarrowPhase.MprEpa (NarrowPhase.cs:902-927): world-space wrapper.
template <class R>
MprResult<R> mpr_epa(const Shape<R>& sa, const Shape<R>& sb, Q4<R> qa, Q4<R> qb, V3<R> pa, V3<R> pb,
                     Polytope<R>* poly, R epaThreshold = R(0.02)) {
    MinkowskiPair<R> mk;
    mk.a = &sa;
    mk.b = &sb;
    mk.qb = qconjmul(qa, qb);
    V3<R> pos = pb - pa;
    mk.pb = rotate_inv(pos, qa);
    MprResult<R> r = solve_mpr(mk, epaThreshold, poly);
    r.pa = rotate(r.pa, qa);
    r.pa = r.pa + pa;
    r.pb = rotate(r.pb, qa);
    r.pb = r.pb + pa;
    r.normal = rotate(r.normal, qa);
    return r;
}

// CollisionManifold (NarrowPhase/CollisionManifold.cs).
template <class R> struct Manifold {
    static constexpr int MaxSupportPoints = 6, MaxManifoldPoints = 12, SolverContactLimit = 4;
    V3<R> mA[MaxManifoldPoints], mB[MaxManifoldPoints];
    int count = 0;
    static constexpr R DupSq = R(0.001);

    static void push(V3<R>* arr, int& n, V3<R> v) {  // PushLeft / PushRight (:58-88)
        if (n > 0 && len_sq(arr[0] - v) < DupSq) return;
        if (n > 1 && len_sq(arr[n - 1] - v) < DupSq) return;
        arr[n++] = v;
    }
    static bool polygon_sign(const V3<R>* poly, int n, V3<R> normal, R& sign) {  // :92-113
        const R e = R(1e-6);
        R area = R(0.0);
        V3<R> prev = poly[n - 1];
        for (int i = 0; i < n; ++i) {
            V3<R> cur = poly[i];
            area += dot(cross(prev, cur), normal);
            prev = cur;
        }
        if (std::fabs(area) <= e) { sign = R(0.0); return false; }
        sign = area < R(0.0) ? -R(1.0) : R(1.0);
        return true;
    }
    static bool contains(const V3<R>* poly, int n, V3<R> pt, V3<R> normal, R sign) {  // :116-133
        const R e = R(1e-3);
        V3<R> prev = poly[n - 1];
        for (int i = 0; i < n; ++i) {
            V3<R> cur = poly[i];
            R s = sign * dot(cross(cur - prev, pt - prev), normal);
            if (s < -e) return false;
            prev = cur;
        }
        return true;
    }
    static int clip_segment(V3<R> start, V3<R> end, const V3<R>* poly, int n, V3<R> normal, R sign, V3<R>* out) {  // :137-185
        const R e = R(1e-3);
        R enter = R(0.0), exit = R(1.0);
        V3<R> delta = end - start;
        for (int i = 0; i < n; ++i) {
            V3<R> es = poly[i], ee = poly[(i + 1) % n];
            R ss = sign * dot(cross(ee - es, start - es), normal);
            R se = sign * dot(cross(ee - es, end - es), normal);
            bool si = ss >= -e, ei = se >= -e;
            if (!si && !ei) return 0;
            if (si && ei) continue;
            R den = ss - se;
            if (std::fabs(den) <= e) return 0;
            R t = rclamp(ss / den, R(0.0), R(1.0));
            if (!si) enter = rmax(enter, t);
            else exit = rmin(exit, t);
            if (exit < enter) return 0;
        }
        out[0] = start + enter * delta;
        if (len_sq(start + exit * delta - out[0]) < e) return 1;
        out[1] = start + exit * delta;
        return 2;
    }
    static R cross_n(V3<R> l, V3<R> r, V3<R> n) { return dot(cross(l, r), n); }
    static int intersect(V3<R> sA, V3<R> eA, V3<R> sB, V3<R> eB, V3<R> normal, V3<R>* cA, V3<R>* cB) {  // :196-253
        const R e = R(1e-3);
        V3<R> dA = eA - sA, dB = eB - sB, off = sB - sA;
        R cr = cross_n(dA, dB, normal);
        if (std::fabs(cr) <= e) {
            if (std::fabs(cross_n(off, dA, normal)) > e) return 0;
            R la = len_sq(dA), lb = len_sq(dB);
            if (la <= e || lb <= e) return 0;
            R pB0 = dot(sB - sA, dA) / la;
            R pB1 = dot(eB - sA, dA) / la;
            R enter = rmax(R(0.0), rmin(pB0, pB1));
            R exit = rmin(R(1.0), rmax(pB0, pB1));
            if (exit < enter - e) return 0;
            cA[0] = sA + enter * dA;
            R pA0 = rclamp(dot(cA[0] - sB, dB) / lb, R(0.0), R(1.0));
            cB[0] = sB + pA0 * dB;
            cA[1] = sA + exit * dA;
            if (len_sq(cA[1] - cA[0]) <= e) return 1;
            R pA1 = rclamp(dot(cA[1] - sB, dB) / lb, R(0.0), R(1.0));
            cB[1] = sB + pA1 * dB;
            return 2;
        }
        R t = cross_n(off, dB, normal) / cr;
        R u = cross_n(off, dA, normal) / cr;
        if (t < -e || t > R(1.0) + e || u < -e || u > R(1.0) + e) return 0;
        t = rclamp(t, R(0.0), R(1.0));
        u = rclamp(u, R(0.0), R(1.0));
        cA[0] = sA + t * dA;
        cB[0] = sB + u * dB;
        return 1;
    }
    static bool is_dup(const V3<R>* m, int n, V3<R> p) {  // :256-264
        for (int i = 0; i < n; ++i)
            if (len_sq(m[i] - p) < DupSq) return true;
        return false;
    }
    void add_on_a(V3<R> pt, V3<R> pB, V3<R> normal) {  // :268-278
        if (count == MaxManifoldPoints) return;
        R diff = dot(pt - pB, normal);
        if (diff < R(0.0)) return;
        if (is_dup(mA, count, pt)) return;
        mA[count] = pt;
        mB[count++] = pt - diff * normal;
    }
    void add_on_b(V3<R> pt, V3<R> pA, V3<R> normal) {  // :281-291
        if (count == MaxManifoldPoints) return;
        R diff = dot(pt - pA, normal);
        if (diff > R(0.0)) return;
        if (is_dup(mB, count, pt)) return;
        mB[count] = pt;
        mA[count++] = pt - diff * normal;
    }
    void add_pair(V3<R> pa, V3<R> pb, V3<R> normal) {  // :294-304
        if (count == MaxManifoldPoints) return;
        if (dot(pa - pb, normal) < R(0.0)) return;
        if (is_dup(mA, count, pa)) return;
        if (is_dup(mB, count, pb)) return;
        mA[count] = pa;
        mB[count++] = pb;
    }
    void reduce(V3<R> normal) {  // ReduceManifold (:321-375)
        if (count <= SolverContactLimit) return;
        V3<R> c{R(0), R(0), R(0)};
        for (int i = 0; i < count; ++i) c = c + mA[i];
        c = c * R(1.0 / count);
        V3<R> t1 = create_orthonormal(normal);
        V3<R> t2 = cross(normal, t1);
        double ang[MaxManifoldPoints];
        int order[MaxManifoldPoints];
        for (int i = 0; i < count; ++i) {
            V3<R> d = mA[i] - c;
            ang[i] = double(StableMath<R>::atan2(dot(d, t2), dot(d, t1)));
            order[i] = i;
        }
        for (int i = 1; i < count; ++i) {
            int cur = order[i];
            double ca = ang[cur];
            int j = i - 1;
            while (j >= 0 && ang[order[j]] > ca) { order[j + 1] = order[j]; j--; }
            order[j + 1] = cur;
        }
        V3<R> rA[SolverContactLimit], rB[SolverContactLimit];
        for (int i = 0; i < SolverContactLimit; ++i) {
            int sel = order[((2 * i + 1) * count) / (2 * SolverContactLimit)];
            rA[i] = mA[sel];
            rB[i] = mB[sel];
        }
        for (int i = 0; i < SolverContactLimit; ++i) { mA[i] = rA[i]; mB[i] = rB[i]; }
        count = SolverContactLimit;
    }
    // BuildManifold (:393-566).
    void build(const Shape<R>& sa, const Shape<R>& sb, Q4<R> qa, Q4<R> qb, V3<R> posA, V3<R> posB, V3<R> pA,
               V3<R> pB, V3<R> normal) {
        static const R hex[12] = {R(1.0), R(0.0), R(0.5), R(0.86602540378), -R(0.5), R(0.86602540378),
                                  -R(1.0), R(0.0), -R(0.5), -R(0.86602540378), R(0.5), -R(0.86602540378)};
        const R pert = R(0.01);
        int lc = 0, rc = 0;
        count = 0;
        V3<R> cv1 = create_orthonormal(normal);
        V3<R> cv2 = cross(normal, cv1);
        V3<R> relB = posB - posA, pAl = pA - posA, pBl = pB - posA;
        V3<R> left[MaxSupportPoints], right[MaxSupportPoints];
        for (int e = 0; e < MaxSupportPoints; ++e) {
            V3<R> ptn = normal + hex[2 * e] * pert * cv1 + hex[2 * e + 1] * pert * cv2;
            V3<R> tmp = rotate_inv(ptn, qa);
            V3<R> np1 = sa.support(tmp);
            np1 = rotate(np1, qa);
            push(left, lc, np1);
            ptn = -ptn;
            tmp = rotate_inv(ptn, qb);
            V3<R> np2 = sb.support(tmp);
            np2 = rotate(np2, qb);
            np2 = np2 + relB;
            push(right, rc, np2);
        }
        R ls = R(0.0), rs = R(0.0);
        bool lp = lc > 2 && polygon_sign(left, lc, normal, ls);
        bool rp = rc > 2 && polygon_sign(right, rc, normal, rs);
        if (lp) {
            for (int e = 0; e < rc; ++e) {
                V3<R> p = right[e];
                if (!contains(left, lc, p, normal, ls)) continue;
                add_on_b(p, pAl, normal);
                if (count == MaxManifoldPoints) goto finalize;
            }
            if (rc == 2) {
                V3<R> cl[2];
                int n = clip_segment(right[0], right[1], left, lc, normal, ls, cl);
                for (int i = 0; i < n; ++i) {
                    add_on_b(cl[i], pAl, normal);
                    if (count == MaxManifoldPoints) goto finalize;
                }
            }
        }
        if (rp) {
            for (int e = 0; e < lc; ++e) {
                V3<R> p = left[e];
                if (!contains(right, rc, p, normal, rs)) continue;
                add_on_a(p, pBl, normal);
                if (count == MaxManifoldPoints) goto finalize;
            }
            if (lc == 2) {
                V3<R> cl[2];
                int n = clip_segment(left[0], left[1], right, rc, normal, rs, cl);
                for (int i = 0; i < n; ++i) {
                    add_on_a(cl[i], pBl, normal);
                    if (count == MaxManifoldPoints) goto finalize;
                }
            }
        }
        if (lc == 2 && rc == 2) {
            V3<R> cA[2], cB[2];
            int n = intersect(left[0], left[1], right[0], right[1], normal, cA, cB);
            for (int i = 0; i < n; ++i) {
                add_pair(cA[i], cB[i], normal);
                if (count == MaxManifoldPoints) goto finalize;
            }
        }
        if (lp && rp) {
            V3<R> cA[2], cB[2];
            for (int i = 0; i < lc; ++i) {
                V3<R> sA = left[i], eA = left[(i + 1) % lc];
                for (int j = 0; j < rc; ++j) {
                    V3<R> sB = right[j], eB = right[(j + 1) % rc];
                    int n = intersect(sA, eA, sB, eB, normal, cA, cB);
                    for (int k = 0; k < n; ++k) {
                        add_pair(cA[k], cB[k], normal);
                        if (count == MaxManifoldPoints) goto finalize;
                    }
                }
            }
        }
        if (count <= 1) {
            mA[0] = pAl;
            mB[0] = pBl;
            count = 1;
        } else if (count < SolverContactLimit) {
            mA[count] = pAl;
            mB[count] = pBl;
            count++;
        }
    finalize:
        reduce(normal);
        for (int i = 0; i < count; ++i) {
            mA[i] = mA[i] + posA;
            mB[i] = mB[i] + posA;
        }
    }
};

// ---- Speculative contacts: linear sweep ---------------------------------------------------------
// SimplexSolverAB (Collision/NarrowPhase/SimplexSolverAB.cs): the point of a simplex of Minkowski
// vertices closest to the origin, with barycentric weights kept for the witness points.  Weights are
// stored per slot; only the slots of `mask` are meaningful (the reference copies whole weight sets).
template <class R> struct SimplexAB {
    static constexpr R Epsilon = R(1e-8);
    Vertex<R> v[4];
    R bc[4] = {R(0), R(0), R(0), R(0)};
    uint32_t mask = 0;

    // ClosestSegment (:62-104)
    V3<R> closest_segment(int i0, int i1, R* b, uint32_t& m) const {
        V3<R> a = v[i0].V, bb = v[i1].V;
        V3<R> d = bb - a;
        R vsq = len_sq(d);
        if (vsq < Epsilon) {
            m = 1u << i1;
            b[i0] = R(0.0);
            b[i1] = R(1.0);
            return bb;
        }
        R t = -dot(a, d) / vsq;
        if (t > R(1.0)) {
            m = 1u << i1;
            b[i0] = R(0.0);
            b[i1] = R(1.0);
            return bb;
        }
        if (t < R(0.0)) {
            m = 1u << i0;
            b[i0] = R(1.0);
            b[i1] = R(0.0);
            return a;
        }
        b[i0] = R(1.0) - t;
        b[i1] = t;
        m = (1u << i0) | (1u << i1);
        return cross(d, cross(a, d)) * (R(1.0) / vsq);
    }

    // ClosestTriangle (:106-185)
    V3<R> closest_triangle(int i0, int i1, int i2, R* b, uint32_t& m) const {
        m = 0;
        V3<R> a = v[i0].V, bb = v[i1].V, c = v[i2].V;
        V3<R> u = a - bb;
        V3<R> w = a - c;
        V3<R> normal = cross(u, w);
        R t = len_sq(normal);
        R it = R(1.0) / t;
        bool degenerate = t < Epsilon;
        V3<R> c1 = cross(u, a);
        V3<R> c2 = cross(a, w);
        R lambda2 = dot(c1, normal) * it;
        R lambda1 = dot(c2, normal) * it;
        R lambda0 = R(1.0) - lambda2 - lambda1;
        R best = std::numeric_limits<R>::max();
        V3<R> closest{};
        R b0[4];
        if (lambda0 < R(0.0) || degenerate) {
            uint32_t mm;
            V3<R> cl = closest_segment(i1, i2, b0, mm);
            R dist = len_sq(cl);
            if (dist < best) { std::memcpy(b, b0, sizeof(b0)); m = mm; best = dist; closest = cl; }
        }
        if (lambda1 < R(0.0) || degenerate) {
            uint32_t mm;
            V3<R> cl = closest_segment(i0, i2, b0, mm);
            R dist = len_sq(cl);
            if (dist < best) { std::memcpy(b, b0, sizeof(b0)); m = mm; best = dist; closest = cl; }
        }
        if (lambda2 < R(0.0) || degenerate) {
            uint32_t mm;
            V3<R> cl = closest_segment(i0, i1, b0, mm);
            R dist = len_sq(cl);
            if (dist < best) { std::memcpy(b, b0, sizeof(b0)); m = mm; closest = cl; }
        }
        if (m != 0) return closest;
        b[i0] = lambda0;
        b[i1] = lambda1;
        b[i2] = lambda2;
        m = (1u << i0) | (1u << i1) | (1u << i2);
        return normal * (dot(normal, a) * it);
    }

    // Determinant of ClosestTetrahedron: (b - a) . ((c - a) x (d - a))
    static R det(V3<R> a, V3<R> b, V3<R> c, V3<R> d) { return dot(b - a, cross(c - a, d - a)); }

    // ClosestTetrahedron (:187-279)
    V3<R> closest_tetrahedron(R* b, uint32_t& m) const {
        const V3<R> z{R(0), R(0), R(0)};
        R detT = det(v[0].V, v[1].V, v[2].V, v[3].V);
        R inv = R(1.0) / detT;
        bool degenerate = detT * detT < Epsilon;
        R lambda0 = det(z, v[1].V, v[2].V, v[3].V) * inv;
        R lambda1 = det(v[0].V, z, v[2].V, v[3].V) * inv;
        R lambda2 = det(v[0].V, v[1].V, z, v[3].V) * inv;
        R lambda3 = R(1.0) - lambda0 - lambda1 - lambda2;
        R best = std::numeric_limits<R>::max();
        V3<R> closest{};
        R b0[4];
        m = 0;
        static const int face[4][3] = {{1, 2, 3}, {0, 2, 3}, {0, 1, 3}, {0, 1, 2}};
        const R lam[4] = {lambda0, lambda1, lambda2, lambda3};
        for (int f = 0; f < 4; ++f) {
            if (!(lam[f] < R(0.0) || degenerate)) continue;
            uint32_t mm;
            V3<R> cl = closest_triangle(face[f][0], face[f][1], face[f][2], b0, mm);
            R dist = len_sq(cl);
            if (dist < best) {
                std::memcpy(b, b0, sizeof(b0));
                m = mm;
                best = dist;
                closest = cl;
            }
        }
        if (m != 0) return closest;
        b[0] = lambda0;
        b[1] = lambda1;
        b[2] = lambda2;
        b[3] = lambda3;
        m = 15u;
        return z;
    }

    // GetClosest (:281-298): witness points from the weights of the used slots, in slot order.
    void get_closest(V3<R>& pa, V3<R>& pb) const {
        pa = V3<R>{R(0), R(0), R(0)};
        pb = V3<R>{R(0), R(0), R(0)};
        for (int i = 0; i < 4; ++i) {
            if ((mask & (1u << i)) == 0) continue;
            pa = pa + bc[i] * v[i].A;
            pb = pb + bc[i] * v[i].B;
        }
    }

    // AddVertex (:309-357): the new vertex takes the last free slot; used slots keep their order.
    bool add_vertex(const Vertex<R>& x, V3<R>& closest) {
        int ix[4];
        int use = 0, free_slot = 0;
        for (int i = 0; i < 4; ++i) {
            if (mask & (1u << i)) ix[use++] = i;
            else free_slot = i;
        }
        ix[use++] = free_slot;
        v[free_slot] = x;
        switch (use) {
            case 1:
                closest = v[ix[0]].V;
                mask = 1u << ix[0];
                bc[ix[0]] = R(1.0);
                return true;
            case 2: closest = closest_segment(ix[0], ix[1], bc, mask); return true;
            case 3: closest = closest_triangle(ix[0], ix[1], ix[2], bc, mask); return true;
            default: closest = closest_tetrahedron(bc, mask); return mask != 15u;
        }
    }
};

// JVector.NormalizeSafe (LinearMath/JVector.cs:561-567).
template <class R> inline V3<R> normalize_safe(V3<R> v, R eps_sq) {
    R len2 = v.x * v.x + v.y * v.y + v.z * v.z;
    if (len2 < eps_sq) return V3<R>{R(0), R(0), R(0)};
    return (R(1.0) / std::sqrt(len2)) * v;
}

// NarrowPhase.Sweep with A fixed at the origin (NarrowPhase.cs:1155-1221): conservative advancement
// of B along `sweep` against a SimplexSolverAB distance solver.  mk.pb is B's position at time 0.
template <class R>
bool sweep_local(MinkowskiPair<R> mk, V3<R> sweep, V3<R>& pointA, V3<R>& pointB, V3<R>& normal, R& lambda) {
    const R collideEpsilon = R(1e-4);
    const int maxIter = 34;
    SimplexAB<R> s;
    Vertex<R> center = mk.center();
    const V3<R> positionB = mk.pb;
    V3<R> posB = positionB;
    lambda = R(0.0);
    pointA = pointB = V3<R>{R(0), R(0), R(0)};
    V3<R> r = sweep;
    V3<R> v = -center.V;
    normal = V3<R>{R(0), R(0), R(0)};
    int iter = maxIter;
    R distSq = len_sq(v);
    while ((distSq > collideEpsilon * collideEpsilon) && (iter-- != 0)) {
        mk.pb = posB;
        Vertex<R> vx = mk.support(v);
        V3<R> w = vx.V;
        R vDotW = -dot(v, w);
        if (vDotW > R(0.0)) {
            R vDotR = dot(v, r);
            if (vDotR >= -R(1e-12)) {
                lambda = std::numeric_limits<R>::infinity();
                return false;
            }
            lambda = lambda - vDotW / vDotR;
            posB = positionB + lambda * r;
            normal = v;
        }
        if (!s.add_vertex(vx, v)) break;
        v = -v;
        distSq = len_sq(v);
    }
    s.get_closest(pointA, pointB);
    normal = normalize_safe(normal, R(1e-16));
    return true;
}

// NarrowPhase.Sweep for two moving shapes, linear velocities (NarrowPhase.cs:1100-1137): the relative
// sweep in A's frame, then the witness points back in world space at time 0.
template <class R>
bool sweep(const Shape<R>& sa, const Shape<R>& sb, Q4<R> qa, Q4<R> qb, V3<R> pa, V3<R> pb, V3<R> va, V3<R> vb,
           V3<R>& pointA, V3<R>& pointB, V3<R>& normal, R& lambda) {
    MinkowskiPair<R> mk;
    mk.a = &sa;
    mk.b = &sb;
    mk.qb = qconjmul(qa, qb);
    V3<R> position = pb - pa;
    mk.pb = rotate_inv(position, qa);
    V3<R> sw = vb - va;
    sw = rotate_inv(sw, qa);
    if (!sweep_local(mk, sw, pointA, pointB, normal, lambda)) return false;
    pointA = rotate(pointA, qa);
    pointA = pointA + pa;
    pointB = rotate(pointB, qa);
    pointB = pointB + pa;
    normal = rotate(normal, qa);
    pointB = pointB + lambda * (va - vb);
    return true;
}

}  // namespace orc
