// Speculative contacts on the device: the SimplexSolverAB distance sub-solver and the linear
// conservative-advancement sweep of NarrowPhase.Sweep, FP32 in the reference's operation order.
// Only pairs queued by the narrowphase (MPR found them separated, one body has speculative contacts
// on, relative speed above the threshold) reach this code, one thread per pair.
#pragma once
#include "jh_collide.h"

namespace jh {

// SimplexSolverAB (Collision/NarrowPhase/SimplexSolverAB.cs).  Weights are kept per slot; only the
// slots of `mask` are meaningful (the reference copies whole weight sets between its temporaries).
struct SimplexAB {
    MVertex v[4];
    float bc[4];
    uint32_t mask;

    // ClosestSegment (:62-104)
    __device__ vec3 closest_segment(int i0, int i1, float* b, uint32_t& m) const {
        vec3 a = v[i0].V, bb = v[i1].V;
        vec3 d = bb - a;
        float vsq = len_sq(d);
        if (vsq < 1e-8f) {
            m = 1u << i1;
            b[i0] = 0.0f;
            b[i1] = 1.0f;
            return bb;
        }
        float t = -dot(a, d) / vsq;
        if (t > 1.0f) {
            m = 1u << i1;
            b[i0] = 0.0f;
            b[i1] = 1.0f;
            return bb;
        }
        if (t < 0.0f) {
            m = 1u << i0;
            b[i0] = 1.0f;
            b[i1] = 0.0f;
            return a;
        }
        b[i0] = 1.0f - t;
        b[i1] = t;
        m = (1u << i0) | (1u << i1);
        return cross(d, cross(a, d)) * (1.0f / vsq);
    }

    // ClosestTriangle (:106-185)
    __device__ vec3 closest_triangle(int i0, int i1, int i2, float* b, uint32_t& m) const {
        m = 0;
        vec3 a = v[i0].V, bb = v[i1].V, c = v[i2].V;
        vec3 u = a - bb;
        vec3 w = a - c;
        vec3 normal = cross(u, w);
        float t = len_sq(normal);
        float it = 1.0f / t;
        bool degenerate = t < 1e-8f;
        vec3 c1 = cross(u, a);
        vec3 c2 = cross(a, w);
        float lambda2 = dot(c1, normal) * it;
        float lambda1 = dot(c2, normal) * it;
        float lambda0 = 1.0f - lambda2 - lambda1;
        float best = 3.402823466e+38f;
        vec3 closest{0.0f, 0.0f, 0.0f};
        float b0[4];
        const int e0[3] = {i1, i0, i0}, e1[3] = {i2, i2, i1};
        const float lam[3] = {lambda0, lambda1, lambda2};
        for (int e = 0; e < 3; ++e) {
            if (!(lam[e] < 0.0f || degenerate)) continue;
            uint32_t mm;
            vec3 cl = closest_segment(e0[e], e1[e], b0, mm);
            float dist = len_sq(cl);
            if (dist < best) {
                for (int k = 0; k < 4; ++k) b[k] = b0[k];
                m = mm;
                best = dist;
                closest = cl;
            }
        }
        if (m != 0) return closest;
        b[i0] = lambda0;
        b[i1] = lambda1;
        b[i2] = lambda2;
        m = (1u << i0) | (1u << i1) | (1u << i2);
        return normal * (dot(normal, a) * it);
    }

    __device__ static float det(vec3 a, vec3 b, vec3 c, vec3 d) { return dot(b - a, cross(c - a, d - a)); }

    // ClosestTetrahedron (:187-279)
    __device__ vec3 closest_tetrahedron(float* b, uint32_t& m) const {
        const vec3 z{0.0f, 0.0f, 0.0f};
        float detT = det(v[0].V, v[1].V, v[2].V, v[3].V);
        float inv = 1.0f / detT;
        bool degenerate = detT * detT < 1e-8f;
        float lambda0 = det(z, v[1].V, v[2].V, v[3].V) * inv;
        float lambda1 = det(v[0].V, z, v[2].V, v[3].V) * inv;
        float lambda2 = det(v[0].V, v[1].V, z, v[3].V) * inv;
        float lambda3 = 1.0f - lambda0 - lambda1 - lambda2;
        float best = 3.402823466e+38f;
        vec3 closest = z;
        float b0[4];
        m = 0;
        const int f0[4] = {1, 0, 0, 0}, f1[4] = {2, 2, 1, 1}, f2[4] = {3, 3, 3, 2};
        const float lam[4] = {lambda0, lambda1, lambda2, lambda3};
        for (int f = 0; f < 4; ++f) {
            if (!(lam[f] < 0.0f || degenerate)) continue;
            uint32_t mm;
            vec3 cl = closest_triangle(f0[f], f1[f], f2[f], b0, mm);
            float dist = len_sq(cl);
            if (dist < best) {
                for (int k = 0; k < 4; ++k) b[k] = b0[k];
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

    // GetClosest (:281-298)
    __device__ void get_closest(vec3& pa, vec3& pb) const {
        pa = vec3{0.0f, 0.0f, 0.0f};
        pb = vec3{0.0f, 0.0f, 0.0f};
        for (int i = 0; i < 4; ++i) {
            if ((mask & (1u << i)) == 0) continue;
            pa = pa + bc[i] * v[i].A;
            pb = pb + bc[i] * v[i].B;
        }
    }

    // AddVertex (:309-357): the new vertex takes the last free slot; used slots keep their order.
    __device__ bool add_vertex(const MVertex& x, vec3& closest) {
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
                bc[ix[0]] = 1.0f;
                return true;
            case 2: closest = closest_segment(ix[0], ix[1], bc, mask); return true;
            case 3: closest = closest_triangle(ix[0], ix[1], ix[2], bc, mask); return true;
            default: closest = closest_tetrahedron(bc, mask); return mask != 15u;
        }
    }
};

// JVector.NormalizeSafe (LinearMath/JVector.cs:561-567)
__device__ __forceinline__ vec3 normalize_safe(vec3 v, float eps_sq) {
    float len2 = v.x * v.x + v.y * v.y + v.z * v.z;
    if (len2 < eps_sq) return vec3{0.0f, 0.0f, 0.0f};
    return (1.0f / sqrtf(len2)) * v;
}

// NarrowPhase.Sweep with A fixed at the origin (NarrowPhase.cs:1155-1221).  mk.pb is B's position in
// A's frame at time 0; `sweep` is B's velocity relative to A in A's frame.
__device__ bool sweep_local(MinkowskiPair mk, vec3 sweep, vec3& pointA, vec3& pointB, vec3& normal, float& lambda) {
    const float collideEpsilon = 1e-4f;
    SimplexAB s;
    s.mask = 0;
    for (int k = 0; k < 4; ++k) s.bc[k] = 0.0f;
    MVertex center = mk.center();
    const vec3 positionB = mk.pb;
    vec3 posB = positionB;
    lambda = 0.0f;
    pointA = pointB = vec3{0.0f, 0.0f, 0.0f};
    vec3 r = sweep;
    vec3 v = -center.V;
    normal = vec3{0.0f, 0.0f, 0.0f};
    int iter = 34;
    float distSq = len_sq(v);
    while ((distSq > collideEpsilon * collideEpsilon) && (iter-- != 0)) {
        mk.pb = posB;
        MVertex vx = mk.support(v);
        vec3 w = vx.V;
        float vDotW = -dot(v, w);
        if (vDotW > 0.0f) {
            float vDotR = dot(v, r);
            if (vDotR >= -1e-12f) {
                lambda = __builtin_inff();
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
    normal = normalize_safe(normal, 1e-16f);
    return true;
}

// NarrowPhase.Sweep for two moving shapes with linear velocities (NarrowPhase.cs:1100-1137).  mk holds
// B in A's frame (as set up for MPR); the witness points come back in world space at time 0.
__device__ bool sweep_pair(const MinkowskiPair& mk, quat qa, vec3 pa, vec3 va, vec3 vb, vec3& pointA, vec3& pointB,
                           vec3& normal, float& lambda) {
    vec3 sw = vb - va;
    sw = rot_inv(sw, qa);
    if (!sweep_local(mk, sw, pointA, pointB, normal, lambda)) return false;
    pointA = rot(pointA, qa);
    pointA = pointA + pa;
    pointB = rot(pointB, qa);
    pointB = pointB + pa;
    normal = rot(normal, qa);
    pointB = pointB + lambda * (va - vb);
    return true;
}

}  // namespace jh
