// EPA refinement on the GPU: the FP64 polytope with filtered exact orientation predicates of
// ConvexPolytope.cs / PlaneFilter.cs / ExpansionArithmetic.cs.  One thread owns one deep pair and a
// 38 KB polytope (far too large for registers) in an HBM workspace interleaved across threads; the
// EPA kernel runs over the compacted queue of deep pairs.
#pragma once
#include "jh_collide.h"

namespace jh {
namespace epa {

constexpr int MaxVertices = 128, MaxTriangles = 256;

struct Tri {
    short a, b, c;
    short facing;
    vec3 normal, closest;
    float normal_sq, closest_sq;
};
struct Plane {  // PlaneFilter (PlaneFilter.cs:17-79)
    double ex, ey, ez, nx, ny, nz;
};
struct Face {
    Plane f;
    double bx, by, bz;  // barycentric
    double dist_sq;
};

// Per-thread polytope storage in HBM, interleaved across threads: element e of field f of thread t
// lives at [(f * Elems + e) * T + t], so the lanes of a wave that touch the same element together
// (face scans, vertex reads) issue coalesced accesses.  Scalars stay in registers.
constexpr size_t kVertBytes = size_t(9) * MaxVertices * 4;
constexpr size_t kTriIdxBytes = size_t(MaxTriangles) * 8;
constexpr size_t kTriFBytes = size_t(8) * MaxTriangles * 4;
constexpr size_t kFaceBytes = size_t(10) * MaxTriangles * 8;
constexpr size_t kEdgeBytes = size_t(3) * MaxTriangles * 4;
constexpr size_t kLitBytes = size_t(MaxTriangles);
constexpr size_t kBytesPerThread = kVertBytes + kTriIdxBytes + kTriFBytes + kFaceBytes + kEdgeBytes + kLitBytes;

struct Work {
    float* vf;            // 9 x MaxVertices
    unsigned long long* ti;  // MaxTriangles: packed a, b, c, facing
    float* tf;            // 8 x MaxTriangles: normal, closest, normal_sq, closest_sq
    double* fd;           // 10 x MaxTriangles: plane ex ey ez nx ny nz, bary x y z, dist_sq
    unsigned* ed;         // 3 x MaxTriangles edges (a | b << 16)
    unsigned char* lt;    // MaxTriangles
    int T, tid;
    int vcount, tcount, enclosed;

    __device__ void bind(char* base, int threads, int t) {
        T = threads;
        tid = t;
        size_t TT = size_t(threads);
        vf = (float*)base;
        ti = (unsigned long long*)(base + kVertBytes * TT);
        tf = (float*)(base + (kVertBytes + kTriIdxBytes) * TT);
        fd = (double*)(base + (kVertBytes + kTriIdxBytes + kTriFBytes) * TT);
        ed = (unsigned*)(base + (kVertBytes + kTriIdxBytes + kTriFBytes + kFaceBytes) * TT);
        lt = (unsigned char*)(base + (kVertBytes + kTriIdxBytes + kTriFBytes + kFaceBytes + kEdgeBytes) * TT);
    }
    __device__ __forceinline__ size_t at(int field, int elems, int e) const {
        return (size_t(field) * elems + e) * size_t(T) + tid;
    }
    __device__ __forceinline__ vec3 V(int i) const {
        return {vf[at(0, MaxVertices, i)], vf[at(1, MaxVertices, i)], vf[at(2, MaxVertices, i)]};
    }
    __device__ __forceinline__ MVertex vertex(int i) const {
        MVertex m;
        m.V = V(i);
        m.A = {vf[at(3, MaxVertices, i)], vf[at(4, MaxVertices, i)], vf[at(5, MaxVertices, i)]};
        m.B = {vf[at(6, MaxVertices, i)], vf[at(7, MaxVertices, i)], vf[at(8, MaxVertices, i)]};
        return m;
    }
    __device__ __forceinline__ void set_vertex(int i, const MVertex& m) {
        vf[at(0, MaxVertices, i)] = m.V.x; vf[at(1, MaxVertices, i)] = m.V.y; vf[at(2, MaxVertices, i)] = m.V.z;
        vf[at(3, MaxVertices, i)] = m.A.x; vf[at(4, MaxVertices, i)] = m.A.y; vf[at(5, MaxVertices, i)] = m.A.z;
        vf[at(6, MaxVertices, i)] = m.B.x; vf[at(7, MaxVertices, i)] = m.B.y; vf[at(8, MaxVertices, i)] = m.B.z;
    }
    __device__ __forceinline__ void tri_idx(int i, short& a, short& b, short& c, short& facing) const {
        unsigned long long p = ti[at(0, MaxTriangles, i)];
        a = short(p & 0xffff); b = short((p >> 16) & 0xffff); c = short((p >> 32) & 0xffff); facing = short(p >> 48);
    }
    __device__ __forceinline__ Tri tri(int i) const {
        Tri t;
        tri_idx(i, t.a, t.b, t.c, t.facing);
        t.normal = {tf[at(0, MaxTriangles, i)], tf[at(1, MaxTriangles, i)], tf[at(2, MaxTriangles, i)]};
        t.closest = {tf[at(3, MaxTriangles, i)], tf[at(4, MaxTriangles, i)], tf[at(5, MaxTriangles, i)]};
        t.normal_sq = tf[at(6, MaxTriangles, i)];
        t.closest_sq = tf[at(7, MaxTriangles, i)];
        return t;
    }
    __device__ __forceinline__ void set_tri(int i, const Tri& t) {
        ti[at(0, MaxTriangles, i)] = (unsigned long long)(unsigned short)t.a | ((unsigned long long)(unsigned short)t.b << 16) |
                                     ((unsigned long long)(unsigned short)t.c << 32) |
                                     ((unsigned long long)(unsigned short)t.facing << 48);
        tf[at(0, MaxTriangles, i)] = t.normal.x; tf[at(1, MaxTriangles, i)] = t.normal.y; tf[at(2, MaxTriangles, i)] = t.normal.z;
        tf[at(3, MaxTriangles, i)] = t.closest.x; tf[at(4, MaxTriangles, i)] = t.closest.y; tf[at(5, MaxTriangles, i)] = t.closest.z;
        tf[at(6, MaxTriangles, i)] = t.normal_sq;
        tf[at(7, MaxTriangles, i)] = t.closest_sq;
    }
    __device__ __forceinline__ Plane plane(int i) const {
        return {fd[at(0, MaxTriangles, i)], fd[at(1, MaxTriangles, i)], fd[at(2, MaxTriangles, i)],
                fd[at(3, MaxTriangles, i)], fd[at(4, MaxTriangles, i)], fd[at(5, MaxTriangles, i)]};
    }
    __device__ __forceinline__ double dist_sq(int i) const { return fd[at(9, MaxTriangles, i)]; }
    __device__ __forceinline__ Face face(int i) const {
        Face f;
        f.f = plane(i);
        f.bx = fd[at(6, MaxTriangles, i)];
        f.by = fd[at(7, MaxTriangles, i)];
        f.bz = fd[at(8, MaxTriangles, i)];
        f.dist_sq = fd[at(9, MaxTriangles, i)];
        return f;
    }
    __device__ __forceinline__ void set_face(int i, const Face& f) {
        fd[at(0, MaxTriangles, i)] = f.f.ex; fd[at(1, MaxTriangles, i)] = f.f.ey; fd[at(2, MaxTriangles, i)] = f.f.ez;
        fd[at(3, MaxTriangles, i)] = f.f.nx; fd[at(4, MaxTriangles, i)] = f.f.ny; fd[at(5, MaxTriangles, i)] = f.f.nz;
        fd[at(6, MaxTriangles, i)] = f.bx; fd[at(7, MaxTriangles, i)] = f.by; fd[at(8, MaxTriangles, i)] = f.bz;
        fd[at(9, MaxTriangles, i)] = f.dist_sq;
    }
    __device__ __forceinline__ void edge(int e, short& a, short& b) const {
        unsigned p = ed[at(0, 3 * MaxTriangles, e)];
        a = short(p & 0xffff);
        b = short(p >> 16);
    }
    __device__ __forceinline__ void set_edge(int e, short a, short b) {
        ed[at(0, 3 * MaxTriangles, e)] = (unsigned)(unsigned short)a | ((unsigned)(unsigned short)b << 16);
    }
    __device__ __forceinline__ bool lit(int i) const { return lt[at(0, MaxTriangles, i)] != 0; }
    __device__ __forceinline__ void set_lit(int i, bool v) { lt[at(0, MaxTriangles, i)] = v ? 1 : 0; }
};

constexpr double ErrorFactor = 16 * 1.1102230246251565e-16;

__device__ __forceinline__ bool safe_coord(double v) {
    return v == 0 || (fabs(v) >= 6.223015277861142e-61 && fabs(v) <= 1.6069380442589903e60);
}
__device__ __forceinline__ bool is_safe(vec3 p) { return safe_coord(p.x) && safe_coord(p.y) && safe_coord(p.z); }

__device__ __forceinline__ void plane_init(Plane& pf, vec3 a, vec3 b, vec3 c) {
    double ux = (double)a.x - b.x, uy = (double)a.y - b.y, uz = (double)a.z - b.z;
    double vx = (double)a.x - c.x, vy = (double)a.y - c.y, vz = (double)a.z - c.z;
    double yz = uy * vz, zy = uz * vy, zx = uz * vx, xz = ux * vz, xy = ux * vy, yx = uy * vx;
    pf.nx = yz - zy;
    pf.ny = zx - xz;
    pf.nz = xy - yx;
    pf.ex = ErrorFactor * (fabs(yz) + fabs(zy));
    pf.ey = ErrorFactor * (fabs(zx) + fabs(xz));
    pf.ez = ErrorFactor * (fabs(xy) + fabs(yx));
}
__device__ __forceinline__ bool plane_reliable(const Plane& pf) {
    return fmax(fabs(pf.nx), fmax(fabs(pf.ny), fabs(pf.nz))) > 1048576 * fmax(pf.ex, fmax(pf.ey, pf.ez));
}
__device__ __forceinline__ bool plane_try_side(const Plane& pf, vec3 anchor, vec3 point, int& sign) {
    double dx = (double)point.x - anchor.x, dy = (double)point.y - anchor.y, dz = (double)point.z - anchor.z;
    double x = pf.nx * dx, y = pf.ny * dy, z = pf.nz * dz;
    double value = (x + y) + z;
    double roundoff = ErrorFactor * ((fabs(x) + fabs(y)) + fabs(z));
    double coef = (pf.ex * fabs(dx) + pf.ey * fabs(dy)) + pf.ez * fabs(dz);
    double bound = roundoff + 2 * coef;
    if (fabs(value) <= bound) return false;
    sign = value > 0 ? 1 : -1;
    return true;
}

// ExpansionArithmetic (ExpansionArithmetic.cs:71-116): exact sums of non-overlapping doubles.
__device__ __forceinline__ int xdiff(double a, double b, double* r) {
    double value = a - b;
    double vb = a - value;
    double residual = (a - (value + vb)) + (vb - b);
    int n = 0;
    if (residual != 0) r[n++] = residual;
    if (value != 0) r[n++] = value;
    return n;
}
__device__ __forceinline__ int xgrow(double* e, int len, double value) {
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
__device__ __forceinline__ int xaddprod(double* res, int len, const double* a, int na, const double* b, int nb,
                                        bool neg) {
    for (int i = 0; i < na; ++i)
        for (int j = 0; j < nb; ++j) {
            double x = a[i], y = b[j];
            double product = x * y;
            double residual = fma(x, y, -product);
            if (neg) {
                product = -product;
                residual = -residual;
            }
            if (residual != 0) len = xgrow(res, len, residual);
            if (product != 0) len = xgrow(res, len, product);
        }
    return len;
}
__device__ __forceinline__ int xcross(const double* a, int na, const double* b, int nb, const double* c, int nc,
                                      const double* d, int nd, double* res) {
    int len = xaddprod(res, 0, a, na, b, nb, false);
    return xaddprod(res, len, c, nc, d, nd, true);
}
// ExpansionArithmetic.Side (:21-41)
__device__ __noinline__ int exact_side(vec3 a, vec3 b, vec3 c, vec3 p) {
    double ux[2], uy[2], uz[2], vx[2], vy[2], vz[2], dx[2], dy[2], dz[2];
    int nx = xdiff(a.x, b.x, ux), ny = xdiff(a.y, b.y, uy), nz = xdiff(a.z, b.z, uz);
    int mx = xdiff(a.x, c.x, vx), my = xdiff(a.y, c.y, vy), mz = xdiff(a.z, c.z, vz);
    int px = xdiff(p.x, a.x, dx), py = xdiff(p.y, a.y, dy), pz = xdiff(p.z, a.z, dz);
    double cr[16], det[192];
    int len = xcross(uy, ny, vz, mz, uz, nz, vy, my, cr);
    int count = xaddprod(det, 0, cr, len, dx, px, false);
    len = xcross(uz, nz, vx, mx, ux, nx, vz, mz, cr);
    count = xaddprod(det, count, cr, len, dy, py, false);
    len = xcross(ux, nx, vy, my, uy, ny, vx, mx, cr);
    count = xaddprod(det, count, cr, len, dz, pz, false);
    if (count == 0) return 0;
    double top = det[count - 1];
    return top > 0 ? 1 : (top < 0 ? -1 : 0);
}
// ExpansionArithmetic.Normal (:43-55)
__device__ __noinline__ void exact_normal(vec3 a, vec3 b, vec3 c, double& x, double& y, double& z) {
    double ux[2], uy[2], uz[2], vx[2], vy[2], vz[2];
    int nx = xdiff(a.x, b.x, ux), ny = xdiff(a.y, b.y, uy), nz = xdiff(a.z, b.z, uz);
    int mx = xdiff(a.x, c.x, vx), my = xdiff(a.y, c.y, vy), mz = xdiff(a.z, c.z, vz);
    double cr[16];
    int n = xcross(uy, ny, vz, mz, uz, nz, vy, my, cr);
    x = 0;
    for (int i = 0; i < n; ++i) x += cr[i];
    n = xcross(uz, nz, vx, mx, ux, nx, vz, mz, cr);
    y = 0;
    for (int i = 0; i < n; ++i) y += cr[i];
    n = xcross(ux, nx, vy, my, uy, ny, vx, mx, cr);
    z = 0;
    for (int i = 0; i < n; ++i) z += cr[i];
}

enum { Added = 0, NoHorizon, Degenerate, CapacityExceeded, NumericalFailure };

// ConvexPolytope.Side (:320-326)
__device__ __forceinline__ int side(const Work& w, const Plane& pf, int a, int b, int c, int pi) {
    vec3 pt = pi < 0 ? vec3{0.0f, 0.0f, 0.0f} : w.V(pi);
    vec3 va = w.V(a), vb = w.V(b), vc = w.V(c);
    if (veq(pt, va) || veq(pt, vb) || veq(pt, vc)) return 0;
    int s;
    if (plane_try_side(pf, va, pt, s)) return s;
    return exact_side(va, vb, vc, pt);
}

__device__ __forceinline__ double ddot(double ax, double ay, double az, double bx, double by, double bz) {
    return ax * bx + ay * by + az * bz;
}

// ConvexPolytope.CreateTriangle (:267-318) with CalcBarycentric (:328-395).
__device__ __noinline__ int create_triangle(Work& w, short a, short b, short c, int slot) {
    Tri tri{};
    Face fd{};
    plane_init(fd.f, w.V(a), w.V(b), w.V(c));
    int orientation = 0;
    for (short i = 0; i < 4 && orientation == 0; ++i) {
        if (i == a || i == b || i == c) continue;
        orientation = -side(w, fd.f, a, b, c, i);
    }
    if (orientation == 0) return Degenerate;
    if (orientation < 0) {
        short t = a;
        a = b;
        b = t;
        fd.f.nx = -fd.f.nx;
        fd.f.ny = -fd.f.ny;
        fd.f.nz = -fd.f.nz;
    }
    bool facing = side(w, fd.f, a, b, c, -1) <= 0;
    double rx = fd.f.nx, ry = fd.f.ny, rz = fd.f.nz;
    if (!plane_reliable(fd.f)) exact_normal(w.V(a), w.V(b), w.V(c), rx, ry, rz);
    double ns = fmax(fabs(rx), fmax(fabs(ry), fabs(rz)));
    if (!(ns > 0) || !isfinite(ns)) return NumericalFailure;
    double nx = rx / ns, ny = ry / ns, nz = rz / ns;
    double nsq = ddot(nx, ny, nz, nx, ny, nz);
    vec3 Va = w.V(a), Vb = w.V(b), Vc = w.V(c);
    double ax = Va.x, ay = Va.y, az = Va.z;
    double bx = Vb.x, by = Vb.y, bz = Vb.z;
    double cx = Vc.x, cy = Vc.y, cz = Vc.z;
    // CalcBarycentric
    double ux = ax - bx, uy = ay - by, uz = az - bz;
    double vx = ax - cx, vy = ay - cy, vz = az - cz;
    double c1x = uy * az - uz * ay, c1y = uz * ax - ux * az, c1z = ux * ay - uy * ax;  // cross(u, a)
    double c2x = ay * vz - az * vy, c2y = az * vx - ax * vz, c2z = ax * vy - ay * vx;  // cross(a, v)
    double gamma = (ddot(c1x, c1y, c1z, nx, ny, nz) / nsq) / ns;
    double beta = (ddot(c2x, c2y, c2z, nx, ny, nz) / nsq) / ns;
    double alpha = 1 - gamma - beta;
    bool clamped = false;
    if (alpha >= 0 && beta < 0) {
        double t = ddot(ax, ay, az, ux, uy, uz);
        if (gamma < 0 && t > 0) {
            beta = fmin(1.0, t / ddot(ux, uy, uz, ux, uy, uz));
            alpha = 1 - beta;
            gamma = 0;
        } else {
            double g = ddot(ax, ay, az, vx, vy, vz) / ddot(vx, vy, vz, vx, vy, vz);
            gamma = g < 0 ? 0 : (g > 1 ? 1 : g);
            alpha = 1 - gamma;
            beta = 0;
        }
        clamped = true;
    } else if (beta >= 0 && gamma < 0) {
        double wx = bx - cx, wy = by - cy, wz = bz - cz;
        double t = ddot(bx, by, bz, wx, wy, wz);
        if (alpha < 0 && t > 0) {
            gamma = fmin(1.0, t / ddot(wx, wy, wz, wx, wy, wz));
            beta = 1 - gamma;
            alpha = 0;
        } else {
            double g = -ddot(bx, by, bz, ux, uy, uz) / ddot(ux, uy, uz, ux, uy, uz);
            alpha = g < 0 ? 0 : (g > 1 ? 1 : g);
            beta = 1 - alpha;
            gamma = 0;
        }
        clamped = true;
    } else if (gamma >= 0 && alpha < 0) {
        double wx = bx - cx, wy = by - cy, wz = bz - cz;
        double t = -ddot(cx, cy, cz, vx, vy, vz);
        if (beta < 0 && t > 0) {
            alpha = fmin(1.0, t / ddot(vx, vy, vz, vx, vy, vz));
            gamma = 1 - alpha;
            beta = 0;
        } else {
            double g = -ddot(cx, cy, cz, wx, wy, wz) / ddot(wx, wy, wz, wx, wy, wz);
            beta = g < 0 ? 0 : (g > 1 ? 1 : g);
            gamma = 1 - beta;
            alpha = 0;
        }
        clamped = true;
    }
    fd.bx = alpha;
    fd.by = beta;
    fd.bz = gamma;
    double clx, cly, clz;
    if (clamped) {
        clx = ax * alpha + bx * beta + cx * gamma;
        cly = ay * alpha + by * beta + cy * gamma;
        clz = az * alpha + bz * beta + cz * gamma;
    } else {
        double s = ddot(nx, ny, nz, ax, ay, az) / nsq;
        clx = nx * s;
        cly = ny * s;
        clz = nz * s;
    }
    fd.dist_sq = ddot(clx, cly, clz, clx, cly, clz);
    if (!isfinite(clx) || !isfinite(cly) || !isfinite(clz) || !isfinite(alpha) || !isfinite(beta) ||
        !isfinite(gamma) || !isfinite(fd.dist_sq))
        return NumericalFailure;
    tri.a = a;
    tri.b = b;
    tri.c = c;
    tri.facing = facing ? 1 : 0;
    tri.normal = vec3{float(nx), float(ny), float(nz)};
    tri.normal_sq = len_sq(tri.normal);
    tri.closest = vec3{float(clx), float(cly), float(clz)};
    tri.closest_sq = float(fd.dist_sq);
    w.set_tri(slot, tri);
    w.set_face(slot, fd);
    bool fin = isfinite(tri.closest.x) && isfinite(tri.closest.y) && isfinite(tri.closest.z) &&
               isfinite(tri.closest_sq);
    return fin ? Added : NumericalFailure;
}

// ConvexPolytope.InitTetrahedron (:113-131)
__device__ __forceinline__ bool init_tetrahedron(Work& w) {
    w.vcount = 4;
    w.tcount = 0;
    w.enclosed = 0;
    for (int i = 0; i < 4; ++i)
        if (!is_safe(w.V(i))) return false;
    const short idx[12] = {0, 2, 1, 0, 1, 3, 0, 3, 2, 1, 2, 3};
    for (int i = 0; i < 4; ++i)
        if (create_triangle(w, idx[3 * i], idx[3 * i + 1], idx[3 * i + 2], i) != Added) return false;
    w.tcount = 4;
    return true;
}

// ConvexPolytope.GetClosestTriangle (:147-166)
__device__ __forceinline__ Tri closest_triangle(Work& w) {
    int closest = 0;
    double minimum = __longlong_as_double(0x7ff0000000000000LL);
    int skip = w.enclosed;
    int enclosed = 1;
    for (int i = 0; i < w.tcount; ++i) {
        double d = w.dist_sq(i);
        if (d < minimum) {
            minimum = d;
            closest = i;
        }
        short a, b, c, facing;
        w.tri_idx(i, a, b, c, facing);
        if (!facing) enclosed = skip;
    }
    w.enclosed = enclosed;
    return w.tri(closest);
}

// ConvexPolytope.AddVertexDetailed (:199-265)
__device__ __noinline__ int add_vertex(Work& w, const MVertex& vx) {
    if (w.tcount == 0 || !is_safe(vx.V)) return NumericalFailure;
    if (w.vcount == MaxVertices) return CapacityExceeded;
    w.set_vertex(w.vcount, vx);
    int ecount = 0, litcount = 0;
    for (int index = w.tcount; index-- > 0;) {
        short fa, fb, fc, ffacing;
        w.tri_idx(index, fa, fb, fc, ffacing);
        bool lit = side(w, w.plane(index), fa, fb, fc, w.vcount) > 0;
        w.set_lit(index, lit);
        if (!lit) continue;
        litcount++;
        short corners[3] = {fa, fb, fc};
        for (int k = 0; k < 3; ++k) {
            short A = corners[k], B = corners[(k + 1) % 3];
            bool added = true;
            for (int e = ecount; e-- > 0;) {
                short ea, eb;
                w.edge(e, ea, eb);
                bool m = (ea == A && eb == B) || (ea == B && eb == A);
                if (!m) continue;
                --ecount;
                short la, lb;
                w.edge(ecount, la, lb);
                w.set_edge(e, la, lb);
                added = false;
                break;
            }
            if (added) {
                w.set_edge(ecount, A, B);
                ecount++;
            }
        }
    }
    if (ecount == 0) return NoHorizon;
    if (w.tcount - litcount + ecount > MaxTriangles) return CapacityExceeded;
    int surv = 0;
    for (int i = 0; i < w.tcount; ++i) {
        if (w.lit(i)) continue;
        if (surv != i) {
            w.set_tri(surv, w.tri(i));
            w.set_face(surv, w.face(i));
        }
        surv++;
    }
    for (int i = 0; i < ecount; ++i) {
        short ea, eb;
        w.edge(i, ea, eb);
        int r = create_triangle(w, ea, eb, short(w.vcount), surv + i);
        if (r != Added) {
            w.tcount = 0;
            w.enclosed = 0;
            return r;
        }
    }
    w.tcount = surv + ecount;
    w.vcount++;
    return Added;
}

// MprEpaSolver.SolveMprEpa (NarrowPhase.cs:39-95); vertices 0..3 hold the MPR portal.
__device__ __noinline__ bool solve(Work& w, const MinkowskiPair& mk, vec3& p1, vec3& p2, vec3& normal, float& pen) {
    const float eps = 1e-5f;
    const float NumericEpsilon = 1e-16f;
    if (!init_tetrahedron(w)) return false;
    int iter = 0;
    Tri ctri{};
    bool converged = false;
    while (++iter < 85) {
        ctri = closest_triangle(w);
        vec3 dir = ctri.closest;
        float dirsq = ctri.closest_sq;
        if (!w.enclosed) dir = -dir;
        if (ctri.closest_sq < NumericEpsilon) {
            dir = ctri.normal;
            dirsq = ctri.normal_sq;
        }
        MVertex v = mk.support(dir);
        float delta = dot(ctri.closest - v.V, dir);
        if (delta * delta <= eps * eps * dirsq) {
            converged = true;
            break;
        }
        int r = add_vertex(w, v);
        if (r == NoHorizon) {
            converged = true;
            break;
        }
        if (r != Added) return false;
    }
    if (!converged || !w.enclosed) return false;
    // CalculatePoints (:169-186)
    for (int i = 0; i < w.tcount; ++i) {
        short ta, tb, tc, tfacing;
        w.tri_idx(i, ta, tb, tc, tfacing);
        if (ta != ctri.a || tb != ctri.b || tc != ctri.c) continue;
        Face fc = w.face(i);
        double bx = fc.bx, by = fc.by, bz = fc.bz;
        const MVertex A = w.vertex(ctri.a), B = w.vertex(ctri.b), C = w.vertex(ctri.c);
        p1 = vec3{float((double)A.A.x * bx + (double)B.A.x * by + (double)C.A.x * bz),
                  float((double)A.A.y * bx + (double)B.A.y * by + (double)C.A.y * bz),
                  float((double)A.A.z * bx + (double)B.A.z * by + (double)C.A.z * bz)};
        p2 = vec3{float((double)A.B.x * bx + (double)B.B.x * by + (double)C.B.x * bz),
                  float((double)A.B.y * bx + (double)B.B.y * by + (double)C.B.y * bz),
                  float((double)A.B.z * bx + (double)B.B.z * by + (double)C.B.z * bz)};
        normal = ctri.normal * (1.0f / sqrtf(ctri.normal_sq));
        pen = sqrtf(ctri.closest_sq);
        return true;
    }
    return false;
}

}  // namespace epa
}  // namespace jh
