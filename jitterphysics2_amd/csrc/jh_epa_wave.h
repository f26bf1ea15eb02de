// Wave-cooperative EPA: one wavefront per deep pair, the polytope's hot state in LDS.
//
// Same algorithm, operation for operation, as the one-thread-per-pair form in jh_epa.h (and the
// reference's ConvexPolytope.cs / MprEpaSolver.SolveMprEpa, NarrowPhase.cs:39-95); only the
// bookkeeping is parallel:
//  * triangles live in fixed slots; their logical order (the order of the reference's triangle
//    array, which decides the closest-triangle tie-break and the horizon walk) is a byte array `ord`,
//    so removing lit triangles moves bytes, not triangle records;
//  * the closest-triangle search is a wave reduction over (distance, logical index);
//  * the lit test of every triangle runs on all lanes;
//  * the horizon edge list is built in exactly the sequential order of insertions and removals, each
//    step one wave-wide match search (ballot) over the list in LDS;
//  * the new triangles of the horizon are created one per lane.
// LDS per wave: vertices (positions), packed triangle indices, squared distances, order / free-slot /
// lit bytes: kLdsBytes (6.3 KB).  The rest (plane filters, triangle normals and closest points,
// barycentric coordinates, vertex support points, the edge list of the rare long horizon) is written
// once per triangle / vertex and read at most once per iteration: it lives in a per-wave HBM area.
#pragma once
#include "jh_epa.h"

namespace jh {
namespace epw {

using epa::MaxTriangles;
using epa::MaxVertices;
constexpr int kMaxEdges = 3 * MaxTriangles;

struct Lds {
    float vx[MaxVertices], vy[MaxVertices], vz[MaxVertices];
    unsigned long long tri[MaxTriangles];  // a | b << 16 | c << 32 | facing << 48
    double dist[MaxTriangles];
    unsigned char ord[MaxTriangles];       // logical position -> slot
    unsigned char freel[MaxTriangles];     // free slots
    unsigned char lit[MaxTriangles];       // by logical position
};
constexpr size_t kLdsBytes = sizeof(Lds);

// Per-wave HBM area: pl[6][256] doubles (plane filter ex ey ez nx ny nz), bary[3][256] doubles,
// tf[8][256] floats (normal xyz, closest xyz, normal_sq, closest_sq), ab[6][128] floats (support points
// on A and B), edge[768] (a | b << 16).
constexpr size_t kPlBytes = size_t(6) * MaxTriangles * 8;
constexpr size_t kBaryBytes = size_t(3) * MaxTriangles * 8;
constexpr size_t kTfBytes = size_t(8) * MaxTriangles * 4;
constexpr size_t kAbBytes = size_t(6) * MaxVertices * 4;
constexpr size_t kEdgeBytes = size_t(kMaxEdges) * 4;
constexpr size_t kSideBytes = kPlBytes + kBaryBytes + kTfBytes + kAbBytes + kEdgeBytes;

struct Poly {
    Lds* L;
    double* pl;
    double* bary;
    float* tf;
    float* ab;
    unsigned* edge;
    int vcount, tcount, enclosed, nfree;

    __device__ __forceinline__ void bind(char* base) {
        pl = (double*)base;
        bary = (double*)(base + kPlBytes);
        tf = (float*)(base + kPlBytes + kBaryBytes);
        ab = (float*)(base + kPlBytes + kBaryBytes + kTfBytes);
        edge = (unsigned*)(base + kPlBytes + kBaryBytes + kTfBytes + kAbBytes);
    }
    __device__ __forceinline__ vec3 V(int i) const { return {L->vx[i], L->vy[i], L->vz[i]}; }
    __device__ __forceinline__ epa::Plane plane(int s) const {
        return {pl[0 * MaxTriangles + s], pl[1 * MaxTriangles + s], pl[2 * MaxTriangles + s],
                pl[3 * MaxTriangles + s], pl[4 * MaxTriangles + s], pl[5 * MaxTriangles + s]};
    }
};

__device__ __forceinline__ void unpack(unsigned long long p, short& a, short& b, short& c, short& facing) {
    a = short(p & 0xffff);
    b = short((p >> 16) & 0xffff);
    c = short((p >> 32) & 0xffff);
    facing = short(p >> 48);
}

// ConvexPolytope.Side (:320-326)
__device__ __forceinline__ int side(const Poly& w, const epa::Plane& pf, int a, int b, int c, int pi) {
    vec3 pt = pi < 0 ? vec3{0.0f, 0.0f, 0.0f} : w.V(pi);
    vec3 va = w.V(a), vb = w.V(b), vc = w.V(c);
    if (veq(pt, va) || veq(pt, vb) || veq(pt, vc)) return 0;
    int s;
    if (epa::plane_try_side(pf, va, pt, s)) return s;
    return epa::exact_side(va, vb, vc, pt);
}

// ConvexPolytope.CreateTriangle (:267-318) with CalcBarycentric (:328-395) into `slot`; the arithmetic
// of epa::create_triangle.
__device__ __noinline__ int create_triangle(Poly& w, short a, short b, short c, int slot) {
    using epa::ddot;
    epa::Plane pf;
    epa::plane_init(pf, w.V(a), w.V(b), w.V(c));
    int orientation = 0;
    for (short i = 0; i < 4 && orientation == 0; ++i) {
        if (i == a || i == b || i == c) continue;
        orientation = -side(w, pf, a, b, c, i);
    }
    if (orientation == 0) return epa::Degenerate;
    if (orientation < 0) {
        short t = a;
        a = b;
        b = t;
        pf.nx = -pf.nx;
        pf.ny = -pf.ny;
        pf.nz = -pf.nz;
    }
    bool facing = side(w, pf, a, b, c, -1) <= 0;
    double rx = pf.nx, ry = pf.ny, rz = pf.nz;
    if (!epa::plane_reliable(pf)) epa::exact_normal(w.V(a), w.V(b), w.V(c), rx, ry, rz);
    double ns = fmax(fabs(rx), fmax(fabs(ry), fabs(rz)));
    if (!(ns > 0) || !isfinite(ns)) return epa::NumericalFailure;
    double nx = rx / ns, ny = ry / ns, nz = rz / ns;
    double nsq = ddot(nx, ny, nz, nx, ny, nz);
    vec3 Va = w.V(a), Vb = w.V(b), Vc = w.V(c);
    double ax = Va.x, ay = Va.y, az = Va.z;
    double bx = Vb.x, by = Vb.y, bz = Vb.z;
    double cx = Vc.x, cy = Vc.y, cz = Vc.z;
    double ux = ax - bx, uy = ay - by, uz = az - bz;
    double vx = ax - cx, vy = ay - cy, vz = az - cz;
    double c1x = uy * az - uz * ay, c1y = uz * ax - ux * az, c1z = ux * ay - uy * ax;
    double c2x = ay * vz - az * vy, c2y = az * vx - ax * vz, c2z = ax * vy - ay * vx;
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
    double dsq = ddot(clx, cly, clz, clx, cly, clz);
    if (!isfinite(clx) || !isfinite(cly) || !isfinite(clz) || !isfinite(alpha) || !isfinite(beta) ||
        !isfinite(gamma) || !isfinite(dsq))
        return epa::NumericalFailure;
    vec3 nf{float(nx), float(ny), float(nz)};
    vec3 cf{float(clx), float(cly), float(clz)};
    float nfsq = len_sq(nf);
    float cfsq = float(dsq);
    w.L->tri[slot] = (unsigned long long)(unsigned short)a | ((unsigned long long)(unsigned short)b << 16) |
                     ((unsigned long long)(unsigned short)c << 32) |
                     ((unsigned long long)(unsigned short)(facing ? 1 : 0) << 48);
    w.L->dist[slot] = dsq;
    double* pl = w.pl;
    pl[0 * MaxTriangles + slot] = pf.ex; pl[1 * MaxTriangles + slot] = pf.ey; pl[2 * MaxTriangles + slot] = pf.ez;
    pl[3 * MaxTriangles + slot] = pf.nx; pl[4 * MaxTriangles + slot] = pf.ny; pl[5 * MaxTriangles + slot] = pf.nz;
    float* tf = w.tf;
    tf[0 * MaxTriangles + slot] = nf.x; tf[1 * MaxTriangles + slot] = nf.y; tf[2 * MaxTriangles + slot] = nf.z;
    tf[3 * MaxTriangles + slot] = cf.x; tf[4 * MaxTriangles + slot] = cf.y; tf[5 * MaxTriangles + slot] = cf.z;
    tf[6 * MaxTriangles + slot] = nfsq;
    tf[7 * MaxTriangles + slot] = cfsq;
    w.bary[0 * MaxTriangles + slot] = alpha;
    w.bary[1 * MaxTriangles + slot] = beta;
    w.bary[2 * MaxTriangles + slot] = gamma;
    bool fin = isfinite(cf.x) && isfinite(cf.y) && isfinite(cf.z) && isfinite(cfsq);
    return fin ? epa::Added : epa::NumericalFailure;
}

__device__ __forceinline__ void wave_sync() { __syncthreads(); }  // the workgroup is one wave

__device__ __forceinline__ void set_vertex(Poly& w, int i, const MVertex& m, int lane) {
    if (lane == 0) {
        w.L->vx[i] = m.V.x;
        w.L->vy[i] = m.V.y;
        w.L->vz[i] = m.V.z;
        float* ab = w.ab;
        ab[0 * MaxVertices + i] = m.A.x; ab[1 * MaxVertices + i] = m.A.y; ab[2 * MaxVertices + i] = m.A.z;
        ab[3 * MaxVertices + i] = m.B.x; ab[4 * MaxVertices + i] = m.B.y; ab[5 * MaxVertices + i] = m.B.z;
    }
    wave_sync();
}

// Lowest failing code of the lanes (0 = all added).
__device__ __forceinline__ int wave_any_fail(int r) {
    int worst = r;
    for (int o = 1; o < 64; o <<= 1) worst = max(worst, __shfl_xor(worst, o, 64));
    return worst;
}

// ConvexPolytope.InitTetrahedron (:113-131); vertices 0..3 hold the portal.
__device__ __forceinline__ bool init_tetrahedron(Poly& w, int lane) {
    w.vcount = 4;
    w.tcount = 0;
    w.enclosed = 0;
    for (int i = 0; i < 4; ++i)
        if (!epa::is_safe(w.V(i))) return false;
    const short idx[12] = {0, 2, 1, 0, 1, 3, 0, 3, 2, 1, 2, 3};
    int r = epa::Added;
    if (lane < 4) r = create_triangle(w, idx[3 * lane], idx[3 * lane + 1], idx[3 * lane + 2], lane);
    for (int i = lane; i < MaxTriangles; i += 64) {
        if (i < 4) w.L->ord[i] = (unsigned char)i;
        if (i < MaxTriangles - 4) w.L->freel[i] = (unsigned char)(MaxTriangles - 1 - i);
    }
    w.nfree = MaxTriangles - 4;
    wave_sync();
    if (wave_any_fail(r) != epa::Added) return false;
    w.tcount = 4;
    return true;
}

// Wave minimum of a 64-bit key with DPP row shifts and row broadcasts (VALU only, no LDS traffic):
// prefix minima within each 16-lane row, then rows folded into lane 63, which is read back.
template <int CTRL, int ROWS>
__device__ __forceinline__ unsigned long long dpp_u64(unsigned long long x) {
    int lo = int(uint32_t(x)), hi = int(uint32_t(x >> 32));
    lo = __builtin_amdgcn_update_dpp(lo, lo, CTRL, ROWS, 0xf, false);
    hi = __builtin_amdgcn_update_dpp(hi, hi, CTRL, ROWS, 0xf, false);
    return (unsigned long long)uint32_t(lo) | ((unsigned long long)uint32_t(hi) << 32);
}
__device__ __forceinline__ unsigned long long wave_min_u64(unsigned long long x) {
    unsigned long long y;
    y = dpp_u64<0x111, 0xf>(x); x = y < x ? y : x;  // row_shr:1
    y = dpp_u64<0x112, 0xf>(x); x = y < x ? y : x;  // row_shr:2
    y = dpp_u64<0x114, 0xf>(x); x = y < x ? y : x;  // row_shr:4
    y = dpp_u64<0x118, 0xf>(x); x = y < x ? y : x;  // row_shr:8
    y = dpp_u64<0x142, 0xa>(x); x = y < x ? y : x;  // row_bcast:15 into rows 1, 3
    y = dpp_u64<0x143, 0xc>(x); x = y < x ? y : x;  // row_bcast:31 into rows 2, 3
    uint32_t lo = __builtin_amdgcn_readlane(int(uint32_t(x)), 63);
    uint32_t hi = __builtin_amdgcn_readlane(int(uint32_t(x >> 32)), 63);
    return (unsigned long long)lo | ((unsigned long long)hi << 32);
}

// ConvexPolytope.GetClosestTriangle (:147-166): the first triangle (logical order) of minimum
// distance; returns its slot.  Distances are finite and non-negative, so their bit patterns order
// like the values: the minimum is a wave minimum of the bits, and the first logical index holding it
// comes from ballots (lane l holds indices l + 64 j, its own first minimum in j).
__device__ __forceinline__ int closest_triangle(Poly& w, int lane) {
    unsigned long long best = 0x7ff0000000000000ull;  // +inf
    int bj = 4;
    bool all_facing = true;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        int i = j * 64 + lane;
        if (i >= w.tcount) continue;
        int s = w.L->ord[i];
        unsigned long long d = (unsigned long long)__double_as_longlong(w.L->dist[s]);
        if (d < best) {
            best = d;
            bj = j;
        }
        if (!(w.L->tri[s] >> 48)) all_facing = false;
    }
    unsigned long long m = wave_min_u64(best);
    int bi = 0;
#pragma unroll
    for (int j = 3; j >= 0; --j) {
        unsigned long long mm = __ballot(bj == j && best == m);
        if (mm) bi = j * 64 + __ffsll((long long)mm) - 1;
    }
    bool every = __all(all_facing);
    w.enclosed = every ? 1 : w.enclosed;
    return w.L->ord[bi];
}

// Horizon of the lit triangles (the sequential walk of ConvexPolytope.AddVertexDetailed :217-246): lit
// triangles in descending logical order, edges (a,b), (b,c), (c,a); a matching edge (either direction)
// already in the list cancels (the last edge moves into its place), otherwise the edge is appended.
// This form keeps the list in the HBM side area (any length); horizon_regs below is the fast path.
__device__ __noinline__ int horizon_lds(Poly& w, int lane, int& litcount) {
    int ecount = 0;
    litcount = 0;
    for (int i = w.tcount; i-- > 0;) {
        if (!w.L->lit[i]) continue;
        litcount++;
        short fa, fb, fc, ff;
        unpack(w.L->tri[w.L->ord[i]], fa, fb, fc, ff);
        short corners[3] = {fa, fb, fc};
        for (int k = 0; k < 3; ++k) {
            unsigned A = (unsigned short)corners[k], B = (unsigned short)corners[(k + 1) % 3];
            unsigned fwd = A | (B << 16), rev = B | (A << 16);
            int found = -1;
            for (int base = ((ecount - 1) / 64) * 64; base >= 0 && found < 0 && ecount > 0; base -= 64) {
                int e = base + lane;
                unsigned v = e < ecount ? w.edge[e] : 0xffffffffu;
                unsigned long long m = __ballot(e < ecount && (v == fwd || v == rev));
                if (m) found = base + 63 - __clzll(m);
            }
            wave_sync();
            if (found >= 0) {
                --ecount;
                if (lane == 0) w.edge[found] = w.edge[ecount];
            } else {
                if (lane == 0) w.edge[ecount] = fwd;
                ++ecount;
            }
            wave_sync();
        }
    }
    return ecount;
}

// The same walk with the list in registers: edge e is held by lane e % 64 in register e / 64, so a
// match search is one compare and ballot per register, a removal one readlane, and the lit triangles
// come from the ballot masks of the lit test instead of a scan of the lit flags.  Returns -1 when the
// list would outgrow the registers (the caller then runs horizon_lds).
constexpr int kEdgeRegs = 4;
__device__ __forceinline__ unsigned ereg_get(const unsigned (&r)[kEdgeRegs], int q) {
    return q == 0 ? r[0] : q == 1 ? r[1] : q == 2 ? r[2] : r[3];
}
__device__ __forceinline__ void ereg_set(unsigned (&r)[kEdgeRegs], int q, unsigned v) {
    if (q == 0) r[0] = v;
    else if (q == 1) r[1] = v;
    else if (q == 2) r[2] = v;
    else r[3] = v;
}
__device__ __forceinline__ int horizon_regs(Poly& w, int lane, const unsigned long long (&litm)[4], int& litcount,
                                            unsigned (&ereg)[kEdgeRegs]) {
    int ecount = 0;
    litcount = 0;
    for (int j = 3; j >= 0; --j) {
        unsigned long long m = litm[j];
        while (m) {
            int hi = 63 - __clzll(m);
            m &= ~(1ull << hi);
            litcount++;
            short fa, fb, fc, ff;
            unpack(w.L->tri[w.L->ord[j * 64 + hi]], fa, fb, fc, ff);
            short corners[3] = {fa, fb, fc};
            for (int k = 0; k < 3; ++k) {
                unsigned A = (unsigned short)corners[k], B = (unsigned short)corners[(k + 1) % 3];
                unsigned fwd = A | (B << 16), rev = B | (A << 16);
                int found = -1;
#pragma unroll
                for (int q = kEdgeRegs - 1; q >= 0; --q) {
                    if (found >= 0 || q * 64 >= ecount) continue;
                    unsigned v = ereg[q];
                    unsigned long long mm = __ballot(q * 64 + lane < ecount && (v == fwd || v == rev));
                    if (mm) found = q * 64 + 63 - __clzll(mm);
                }
                if (found >= 0) {
                    --ecount;
                    unsigned last = __builtin_amdgcn_readlane(ereg_get(ereg, ecount >> 6), ecount & 63);
                    if (lane == (found & 63)) ereg_set(ereg, found >> 6, last);
                } else {
                    if (ecount == kEdgeRegs * 64) return -1;
                    if (lane == (ecount & 63)) ereg_set(ereg, ecount >> 6, fwd);
                    ++ecount;
                }
            }
        }
    }
    return ecount;
}

// ConvexPolytope.AddVertexDetailed (:199-265).
__device__ __forceinline__ int add_vertex(Poly& w, const MVertex& vx, int lane) {
    if (w.tcount == 0 || !epa::is_safe(vx.V)) return epa::NumericalFailure;
    if (w.vcount == MaxVertices) return epa::CapacityExceeded;
    set_vertex(w, w.vcount, vx, lane);
    // lit test of every triangle (MaxTriangles = 4 x 64: lane + 64 j covers them)
    static_assert(MaxTriangles <= 256, "lit masks hold 256 triangles");
    unsigned long long litm[4];
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        int i = j * 64 + lane;
        bool lit = false;
        if (i < w.tcount) {
            int s = w.L->ord[i];
            short fa, fb, fc, ff;
            unpack(w.L->tri[s], fa, fb, fc, ff);
            lit = side(w, w.plane(s), fa, fb, fc, w.vcount) > 0;
            w.L->lit[i] = lit ? 1 : 0;
        }
        litm[j] = __ballot(lit);
    }
    wave_sync();
    int litcount = 0;
    unsigned ereg[kEdgeRegs] = {0u, 0u, 0u, 0u};
    int ecount = horizon_regs(w, lane, litm, litcount, ereg);
    bool in_regs = ecount >= 0;
    if (!in_regs) ecount = horizon_lds(w, lane, litcount);
    if (ecount == 0) return epa::NoHorizon;
    if (w.tcount - litcount + ecount > MaxTriangles) return epa::CapacityExceeded;
    // survivors keep their order; lit slots return to the free list (all reads before the writes:
    // the compaction is in place)
    int surv_before = 0, lit_before = 0;
    unsigned long long below = lane ? (~0ull >> (64 - lane)) : 0ull;
    unsigned char slot_j[4];
    int pos_j[4];
    bool valid_j[4], lit_j[4];
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        int i = j * 64 + lane;
        bool valid = i < w.tcount;
        unsigned long long vm = __ballot(valid);
        unsigned long long lm = litm[j] & vm, sm = vm & ~litm[j];
        bool lit = ((litm[j] >> lane) & 1ull) != 0;
        valid_j[j] = valid;
        lit_j[j] = lit;
        slot_j[j] = valid ? w.L->ord[i] : (unsigned char)0;
        pos_j[j] = lit ? lit_before + __popcll(lm & below) : surv_before + __popcll(sm & below);
        surv_before += __popcll(sm);
        lit_before += __popcll(lm);
    }
    wave_sync();
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        if (!valid_j[j]) continue;
        if (lit_j[j]) w.L->freel[w.nfree + pos_j[j]] = slot_j[j];
        else w.L->ord[pos_j[j]] = slot_j[j];
    }
    int surv = surv_before;
    w.nfree += lit_before;
    wave_sync();
    // one new triangle per horizon edge, in edge order
    int r = epa::Added;
    for (int e = lane; e < ecount; e += 64) {
        unsigned ev = in_regs ? ereg_get(ereg, e >> 6) : w.edge[e];
        int slot = w.L->freel[w.nfree - 1 - e];
        w.L->ord[surv + e] = (unsigned char)slot;
        int rr = create_triangle(w, short(ev & 0xffff), short(ev >> 16), short(w.vcount), slot);
        if (rr != epa::Added) r = rr;
    }
    wave_sync();
    if (wave_any_fail(r) != epa::Added) {
        w.tcount = 0;
        w.enclosed = 0;
        return epa::NumericalFailure;
    }
    w.nfree -= ecount;
    w.tcount = surv + ecount;
    w.vcount++;
    return epa::Added;
}

// MprEpaSolver.SolveMprEpa (NarrowPhase.cs:39-95); the portal must be in vertices 0..3.
__device__ __forceinline__ bool solve(Poly& w, const MinkowskiPair& mk, int lane, vec3& p1, vec3& p2, vec3& normal,
                                      float& pen) {
    const float eps = 1e-5f;
    const float NumericEpsilon = 1e-16f;
    if (!init_tetrahedron(w, lane)) return false;
    int iter = 0;
    int cslot = 0;
    bool converged = false;
    vec3 cn{}, cc{};
    float cnsq = 0.0f, ccsq = 0.0f;
    while (++iter < 85) {
        cslot = closest_triangle(w, lane);
        const float* tf = w.tf;
        cn = {tf[0 * MaxTriangles + cslot], tf[1 * MaxTriangles + cslot], tf[2 * MaxTriangles + cslot]};
        cc = {tf[3 * MaxTriangles + cslot], tf[4 * MaxTriangles + cslot], tf[5 * MaxTriangles + cslot]};
        cnsq = tf[6 * MaxTriangles + cslot];
        ccsq = tf[7 * MaxTriangles + cslot];
        vec3 dir = cc;
        float dirsq = ccsq;
        if (!w.enclosed) dir = -dir;
        if (ccsq < NumericEpsilon) {
            dir = cn;
            dirsq = cnsq;
        }
        MVertex v = mk.support(dir);
        float delta = dot(cc - v.V, dir);
        if (delta * delta <= eps * eps * dirsq) {
            converged = true;
            break;
        }
        int r = add_vertex(w, v, lane);
        if (r == epa::NoHorizon) {
            converged = true;
            break;
        }
        if (r != epa::Added) return false;
    }
    if (!converged || !w.enclosed) return false;
    // CalculatePoints (:169-186): the first triangle with the closest triangle's corners
    short ca, cb, ccn, cf;
    unpack(w.L->tri[cslot], ca, cb, ccn, cf);
    int pos = 0x7fffffff;
    for (int i = lane; i < w.tcount; i += 64) {
        short ta, tb, tc, tfc;
        unpack(w.L->tri[w.L->ord[i]], ta, tb, tc, tfc);
        if (ta == ca && tb == cb && tc == ccn && i < pos) pos = i;
    }
    for (int o = 1; o < 64; o <<= 1) pos = min(pos, __shfl_xor(pos, o, 64));
    if (pos == 0x7fffffff) return false;
    int s = w.L->ord[pos];
    double bx = w.bary[0 * MaxTriangles + s], by = w.bary[1 * MaxTriangles + s], bz = w.bary[2 * MaxTriangles + s];
    const float* ab = w.ab;
    auto A = [&](int comp, int vi) { return double(ab[comp * MaxVertices + vi]); };
    p1 = vec3{float(A(0, ca) * bx + A(0, cb) * by + A(0, ccn) * bz),
              float(A(1, ca) * bx + A(1, cb) * by + A(1, ccn) * bz),
              float(A(2, ca) * bx + A(2, cb) * by + A(2, ccn) * bz)};
    p2 = vec3{float(A(3, ca) * bx + A(3, cb) * by + A(3, ccn) * bz),
              float(A(4, ca) * bx + A(4, cb) * by + A(4, ccn) * bz),
              float(A(5, ca) * bx + A(5, cb) * by + A(5, ccn) * bz)};
    normal = cn * (1.0f / sqrtf(cnsq));
    pen = sqrtf(ccsq);
    return true;
}

}  // namespace epw
}  // namespace jh
