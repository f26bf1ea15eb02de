// Constraints on the device: BallSocket, HingeAngle, FixedAngle, DistanceLimit, TwistAngle, ConeLimit,
// AngularMotor, LinearMotor, PointOnLine, PointOnPlane, AngularServo, LinearServo, LinearAxisServo,
// TwistServo, EllipticalConeLimit (Dynamics/Constraints/*.cs) with the JMatrix / JSymmetricMatrix / MathHelper / AngularConstraintMath
// pieces they use, FP32 in the reference's operation order.  One thread solves one constraint; the
// solver visits constraints in batches with disjoint dynamic bodies.
#pragma once
#include "jh_math.h"
#include "../../include/jitterhip.h"

namespace jh {
namespace con {

// Leading entries of jh_constraint.state that hold the accumulated impulse, which ResetWarmStart of
// each constraint class zeroes (the hemisphere and clamp records after it stay).
__host__ __device__ __forceinline__ int impulse_count(int type) {
    switch (type) {
        case JH_CONSTRAINT_BALL_SOCKET:
        case JH_CONSTRAINT_HINGE_ANGLE:
        case JH_CONSTRAINT_FIXED_ANGLE:
        case JH_CONSTRAINT_POINT_ON_LINE:
        case JH_CONSTRAINT_ANGULAR_SERVO:
        case JH_CONSTRAINT_LINEAR_SERVO: return 3;
        default: return 1;
    }
}

// ---- JMatrix (LinearMath/JMatrix.cs), row-major fields m11 .. m33 -----------------------------
__device__ __forceinline__ mat3 mzero() { return {0, 0, 0, 0, 0, 0, 0, 0, 0}; }
__device__ __forceinline__ mat3 mmul(const mat3& a, const mat3& b) {  // operator * (:620-633)
    mat3 r;
    r.m11 = a.m11 * b.m11 + a.m12 * b.m21 + a.m13 * b.m31;
    r.m12 = a.m11 * b.m12 + a.m12 * b.m22 + a.m13 * b.m32;
    r.m13 = a.m11 * b.m13 + a.m12 * b.m23 + a.m13 * b.m33;
    r.m21 = a.m21 * b.m11 + a.m22 * b.m21 + a.m23 * b.m31;
    r.m22 = a.m21 * b.m12 + a.m22 * b.m22 + a.m23 * b.m32;
    r.m23 = a.m21 * b.m13 + a.m22 * b.m23 + a.m23 * b.m33;
    r.m31 = a.m31 * b.m11 + a.m32 * b.m21 + a.m33 * b.m31;
    r.m32 = a.m31 * b.m12 + a.m32 * b.m22 + a.m33 * b.m32;
    r.m33 = a.m31 * b.m13 + a.m32 * b.m23 + a.m33 * b.m33;
    return r;
}
__device__ __forceinline__ mat3 madd(const mat3& a, const mat3& b) {  // Add (:353-364)
    return {a.m11 + b.m11, a.m12 + b.m12, a.m13 + b.m13, a.m21 + b.m21, a.m22 + b.m22,
            a.m23 + b.m23, a.m31 + b.m31, a.m32 + b.m32, a.m33 + b.m33};
}
__device__ __forceinline__ mat3 msub(const mat3& a, const mat3& b) {  // Subtract (:373-384)
    return {a.m11 - b.m11, a.m12 - b.m12, a.m13 - b.m13, a.m21 - b.m21, a.m22 - b.m22,
            a.m23 - b.m23, a.m31 - b.m31, a.m32 - b.m32, a.m33 - b.m33};
}
__device__ __forceinline__ mat3 mscale(const mat3& a, float s) {  // Multiply(matrix, scale) (:503-514)
    return {a.m11 * s, a.m12 * s, a.m13 * s, a.m21 * s, a.m22 * s, a.m23 * s, a.m31 * s, a.m32 * s, a.m33 * s};
}
__device__ __forceinline__ mat3 mtranspose(const mat3& m) {  // Transpose (:603-614)
    return {m.m11, m.m21, m.m31, m.m12, m.m22, m.m32, m.m13, m.m23, m.m33};
}
__device__ __forceinline__ float mdet(const mat3& m) {  // Determinant (:391-395)
    return m.m11 * m.m22 * m.m33 + m.m12 * m.m23 * m.m31 + m.m13 * m.m21 * m.m32 - m.m31 * m.m22 * m.m13 -
           m.m32 * m.m23 * m.m11 - m.m33 * m.m21 * m.m12;
}
__device__ __forceinline__ mat3 cross_matrix(vec3 v) {  // CreateCrossProduct (:597-600)
    return {0.0f, -v.z, v.y, v.z, 0.0f, -v.x, -v.y, v.x, 0.0f};
}
__device__ __forceinline__ bool fin(float v) { return isfinite(v); }
__device__ __forceinline__ bool mfinite(const mat3& m) {
    return fin(m.m11) && fin(m.m12) && fin(m.m13) && fin(m.m21) && fin(m.m22) && fin(m.m23) && fin(m.m31) &&
           fin(m.m32) && fin(m.m33);
}
// Real.IsNormal: finite, non-zero, not subnormal
__device__ __forceinline__ bool is_normal(float v) {
    float a = fabsf(v);
    return a >= 1.17549435e-38f && a <= 3.40282347e+38f;
}
// TryInverseUnscaled (:455-481)
__device__ __forceinline__ bool try_inverse_unscaled(const mat3& m, mat3& r) {
    float idet = 1.0f / mdet(m);
    if (!is_normal(idet)) {
        r = mzero();
        return false;
    }
    float n11 = m.m22 * m.m33 - m.m23 * m.m32;
    float n12 = m.m13 * m.m32 - m.m12 * m.m33;
    float n13 = m.m12 * m.m23 - m.m22 * m.m13;
    float n21 = m.m23 * m.m31 - m.m33 * m.m21;
    float n22 = m.m11 * m.m33 - m.m31 * m.m13;
    float n23 = m.m13 * m.m21 - m.m23 * m.m11;
    float n31 = m.m21 * m.m32 - m.m31 * m.m22;
    float n32 = m.m12 * m.m31 - m.m32 * m.m11;
    float n33 = m.m11 * m.m22 - m.m21 * m.m12;
    r = {n11 * idet, n12 * idet, n13 * idet, n21 * idet, n22 * idet, n23 * idet, n31 * idet, n32 * idet, n33 * idet};
    return mfinite(r);
}
// MathR.Max with NaN propagation
__device__ __forceinline__ float max_nan(float a, float b) {
    if (a != a) return a;
    if (b != b) return b;
    return a > b ? a : b;
}
// Inverse (:404-444)
__device__ __forceinline__ bool minverse(const mat3& m, mat3& out) {
    mat3 inv;
    if (try_inverse_unscaled(m, inv)) {
        out = inv;
        return true;
    }
    float scale = max_nan(fabsf(m.m11), max_nan(fabsf(m.m12), fabsf(m.m13)));
    scale = max_nan(scale, max_nan(fabsf(m.m21), max_nan(fabsf(m.m22), fabsf(m.m23))));
    scale = max_nan(scale, max_nan(fabsf(m.m31), max_nan(fabsf(m.m32), fabsf(m.m33))));
    if (scale > 0.0f && fin(scale)) {
        mat3 n{m.m11 / scale, m.m12 / scale, m.m13 / scale, m.m21 / scale, m.m22 / scale,
               m.m23 / scale, m.m31 / scale, m.m32 / scale, m.m33 / scale};
        if (try_inverse_unscaled(n, inv)) {
            inv = {inv.m11 / scale, inv.m12 / scale, inv.m13 / scale, inv.m21 / scale, inv.m22 / scale,
                   inv.m23 / scale, inv.m31 / scale, inv.m32 / scale, inv.m33 / scale};
            if (mfinite(inv)) {
                out = inv;
                return true;
            }
        }
    }
    out = mzero();
    return false;
}
// JVector.Transform(v, JMatrix) (JVector.cs:290-298): M v
__device__ __forceinline__ vec3 mat_mul(vec3 v, const mat3& m) {
    return {v.x * m.m11 + v.y * m.m12 + v.z * m.m13, v.x * m.m21 + v.y * m.m22 + v.z * m.m23,
            v.x * m.m31 + v.y * m.m32 + v.z * m.m33};
}
// JVector.TransposedTransform(v, JMatrix) (JVector.cs:308-317): M^T v
__device__ __forceinline__ vec3 tmat_mul(vec3 v, const mat3& m) {
    return {v.x * m.m11 + v.y * m.m21 + v.z * m.m31, v.x * m.m12 + v.y * m.m22 + v.z * m.m32,
            v.x * m.m13 + v.y * m.m23 + v.z * m.m33};
}

// ---- JSymmetricMatrix (LinearMath/JSymmetricMatrix.cs) ---------------------------------------
__device__ __forceinline__ sym3 sym_add(const sym3& a, const sym3& b) {  // Add (:190-196)
    return {a.m11 + b.m11, a.m12 + b.m12, a.m13 + b.m13, a.m22 + b.m22, a.m23 + b.m23, a.m33 + b.m33};
}
__device__ __forceinline__ sym3 sym_scale_vec(vec3 s) { return {s.x, 0.0f, 0.0f, s.y, 0.0f, s.z}; }  // CreateScale
__device__ __forceinline__ sym3 sym_outer(vec3 v) {  // Outer (:89-93)
    return {v.x * v.x, v.x * v.y, v.x * v.z, v.y * v.y, v.y * v.z, v.z * v.z};
}
__device__ __forceinline__ mat3 sym_to_matrix(const sym3& s) {  // ToMatrix (:68-71)
    return {s.m11, s.m12, s.m13, s.m12, s.m22, s.m23, s.m13, s.m23, s.m33};
}
__device__ __forceinline__ float sym_det(const sym3& s) {  // Determinant (:104-109)
    return s.m11 * (s.m22 * s.m33 - s.m23 * s.m23) + s.m12 * (s.m23 * s.m13 - s.m12 * s.m33) +
           s.m13 * (s.m12 * s.m23 - s.m22 * s.m13);
}

// ---- MathHelper (LinearMath/MathHelper.cs) ---------------------------------------------------
constexpr float kTolerance = 1e-6f;  // single precision

// JacobiRotate (:374-407) on a[9] / basis[9] held in registers (p < q, both compile-time)
template <int p, int q>
__device__ __forceinline__ bool jacobi_rotate(float* a, float* basis) {
    float off = a[3 * p + q];
    if (fabsf(off) <= kTolerance) return false;
    float tau = (a[3 * q + q] - a[3 * p + p]) / (2.0f * off);
    float tangent = copysignf(1.0f, tau) / (fabsf(tau) + sqrtf(1.0f + tau * tau));
    float cosine = 1.0f / sqrtf(1.0f + tangent * tangent);
    float sine = tangent * cosine;
    a[3 * p + p] -= tangent * off;
    a[3 * q + q] += tangent * off;
    a[3 * p + q] = a[3 * q + p] = 0.0f;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        if (k != p && k != q) {
            float akp = a[3 * k + p], akq = a[3 * k + q];
            a[3 * k + p] = a[3 * p + k] = cosine * akp - sine * akq;
            a[3 * k + q] = a[3 * q + k] = sine * akp + cosine * akq;
        }
        float vkp = basis[3 * k + p], vkq = basis[3 * k + q];
        basis[3 * k + p] = cosine * vkp - sine * vkq;
        basis[3 * k + q] = sine * vkp + cosine * vkq;
    }
    return true;
}

// PseudoInverseSymmetric (:282-325)
__device__ __forceinline__ mat3 pseudo_inverse_symmetric(const mat3& m) {
    float scale = max_nan(fabsf(m.m11), max_nan(fabsf(m.m22), fabsf(m.m33)));
    if (!(scale > 0.0f) || !fin(scale)) return mzero();
    sym3 n{m.m11 / scale, m.m12 / scale, m.m13 / scale, m.m22 / scale, m.m23 / scale, m.m33 / scale};
    mat3 inv;
    if (fabsf(sym_det(n)) > kTolerance && minverse(m, inv)) return inv;
    float a[9] = {n.m11, n.m12, n.m13, n.m12, n.m22, n.m23, n.m13, n.m23, n.m33};
    float basis[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1};
    for (int sweep = 0; sweep < 8; ++sweep) {
        bool changed = jacobi_rotate<0, 1>(a, basis);
        changed |= jacobi_rotate<0, 2>(a, basis);
        changed |= jacobi_rotate<1, 2>(a, basis);
        if (!changed) break;
    }
    sym3 result{0, 0, 0, 0, 0, 0};
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        float eig = a[3 * i + i];
        if (!(eig > kTolerance)) continue;
        float weight = (1.0f / eig) / scale;
        vec3 ax{basis[i], basis[3 + i], basis[6 + i]};
        sym3 o{ax.x * ax.x, ax.x * ax.y, ax.x * ax.z, ax.y * ax.y, ax.y * ax.z, ax.z * ax.z};
        sym3 w{o.m11 * weight, o.m12 * weight, o.m13 * weight, o.m22 * weight, o.m23 * weight, o.m33 * weight};
        result = sym_add(result, w);
    }
    return sym_to_matrix(result);
}

// InverseSymmetric (:328-364)
__device__ __forceinline__ mat3 inverse_symmetric(const mat3& m) {
    float diag = m.m11 * m.m22 * m.m33;
    mat3 inv;
    if (mdet(m) / diag > kTolerance && minverse(m, inv)) return inv;
    float x = m.m11 > 0.0f ? sqrtf(m.m11) : 0.0f;
    float y = m.m22 > 0.0f ? sqrtf(m.m22) : 0.0f;
    float z = m.m33 > 0.0f ? sqrtf(m.m33) : 0.0f;
    if (x == 0.0f && y == 0.0f && z == 0.0f) return mzero();
    float xy = x > 0.0f && y > 0.0f ? m.m12 / x / y : 0.0f;
    float xz = x > 0.0f && z > 0.0f ? m.m13 / x / z : 0.0f;
    float yz = y > 0.0f && z > 0.0f ? m.m23 / y / z : 0.0f;
    mat3 normalized{1.0f, xy, xz, xy, 1.0f, yz, xz, yz, 1.0f};
    mat3 resp = pseudo_inverse_symmetric(normalized);
    xy = x > 0.0f && y > 0.0f ? resp.m12 / x / y : 0.0f;
    xz = x > 0.0f && z > 0.0f ? resp.m13 / x / z : 0.0f;
    yz = y > 0.0f && z > 0.0f ? resp.m23 / y / z : 0.0f;
    mat3 r{x > 0.0f ? resp.m11 / x / x : 0.0f, xy, xz, xy, y > 0.0f ? resp.m22 / y / y : 0.0f, yz, xz, yz,
           z > 0.0f ? resp.m33 / z / z : 0.0f};
    bool ok = fin(r.m11) && fin(r.m12) && fin(r.m13) && fin(r.m22) && fin(r.m23) && fin(r.m33);
    return ok ? r : mzero();
}

// InverseBilateralBlock (:367-372)
__device__ __forceinline__ vec3 inverse_bilateral_block(const mat3& m) {
    mat3 block{m.m11, m.m12, 0.0f, m.m12, m.m22, 0.0f, 0.0f, 0.0f, 0.0f};
    mat3 inv = inverse_symmetric(block);
    return {inv.m11, inv.m12, inv.m22};
}

// ---- AngularConstraintMath (Dynamics/Constraints/Internal/AngularConstraintMath.cs) ------------
__device__ __forceinline__ quat qconj(quat q) { return {-q.x, -q.y, -q.z, q.w}; }
__device__ __forceinline__ quat qscale(quat q, float s) { return {q.x * s, q.y * s, q.z * s, q.w * s}; }

// ShortestAngle (:13-21)
__device__ __forceinline__ float shortest_angle(float angle, float half) {
    if (angle >= -half && angle <= half) return angle;
    float period = 2.0f * half;
    angle = fmodf(angle, period);
    if (angle > half) angle -= period;
    else if (angle < -half) angle += period;
    return angle;
}

// CalculateJacobian (:63-94)
__device__ __forceinline__ mat3 calc_jacobian(quat left, quat right, vec3& error, int& hemisphere) {
    quat q = qmul(left, right);
    hemisphere = q.w < 0.0f ? -1 : 1;
    if (q.w < 0.0f) q = qscale(q, -1.0f);
    vec3 vec{q.x, q.y, q.z};
    float sine = sqrtf(len_sq(vec));
    mat3 derivative;
    if (sine > 0.0f) {
        float half_angle = stable::atan2(sine, q.w);
        vec3 axis = vec * (1.0f / sine);
        error = axis * half_angle;
        float diagonal = half_angle * q.w / sine;
        mat3 ident{1, 0, 0, 0, 1, 0, 0, 0, 1};
        mat3 t1 = mscale(ident, diagonal);
        mat3 t2 = mscale(sym_to_matrix(sym_outer(axis)), 1.0f - diagonal);
        derivative = msub(madd(t1, t2), cross_matrix(error));
    } else {
        error = vec3{0.0f, 0.0f, 0.0f};
        derivative = mat3{1, 0, 0, 0, 1, 0, 0, 0, 1};
    }
    return mmul(mscale(derivative, -0.5f), mat_from_quat(left));
}

// QMatrix.ProjectMultiplyLeftRight (Dynamics/Constraints/Internal/QMatrix.cs:109-124)
__device__ __forceinline__ mat3 project_multiply_left_right(quat l, quat r) {
    mat3 m;
    m.m11 = -l.x * r.x + l.w * r.w + l.z * r.z + l.y * r.y;
    m.m12 = -l.x * r.y + l.w * r.z - l.z * r.w - l.y * r.x;
    m.m13 = -l.x * r.z - l.w * r.y - l.z * r.x + l.y * r.w;
    m.m21 = -l.y * r.x + l.z * r.w - l.w * r.z - l.x * r.y;
    m.m22 = -l.y * r.y + l.z * r.z + l.w * r.w + l.x * r.x;
    m.m23 = -l.y * r.z - l.z * r.y + l.w * r.x - l.x * r.w;
    m.m31 = -l.z * r.x - l.y * r.w - l.x * r.z + l.w * r.y;
    m.m32 = -l.z * r.y - l.y * r.z + l.x * r.w - l.w * r.x;
    m.m33 = -l.z * r.z + l.y * r.y + l.x * r.x + l.w * r.w;
    return m;
}

// CalculateTwist (AngularConstraintMath.cs:31-61)
__device__ __forceinline__ float calc_twist(quat left, quat right, vec3 axis, vec3& jacobian, int& hemisphere) {
    mat3 m = mscale(project_multiply_left_right(left, right), -float(1.0 / 2.0));
    quat q = qmul(left, right);
    jacobian = tmat_mul(axis, m);
    float projection = dot(axis, vec3{q.x, q.y, q.z});
    hemisphere = q.w < 0.0f ? -1 : 1;
    if (q.w < 0.0f) {
        q = qscale(q, -1.0f);
        projection *= -1.0f;
        jacobian = jacobian * -1.0f;
    }
    float dsq = len_sq(jacobian);
    if (dsq > 0.0f) jacobian = jacobian * (1.0f / sqrtf(dsq));
    else jacobian = -rot(axis, right);
    vec3 perpendicular = vec3{q.x, q.y, q.z} - axis * projection;
    float cosine = sqrtf(q.w * q.w + len_sq(perpendicular));
    return 2.0f * stable::atan2(projection, cosine);
}

// Math.Clamp(value, min, max) for min <= max
__device__ __forceinline__ float clamp_cs(float v, float lo, float hi) {
    if (v < lo) return lo;
    if (v > hi) return hi;
    return v;
}

// std::min / std::max forms (the oracle's)
__device__ __forceinline__ float cmin(float a, float b) { return b < a ? b : a; }
__device__ __forceinline__ float cmax(float a, float b) { return a < b ? b : a; }

// ---- Constraints ----------------------------------------------------------------------------
// What PrepareForIteration leaves for Iterate, per constraint in device memory (48 floats).
struct Work {
    vec3 r1, r2, bias, bilateral;  // 12
    mat3 em, jac;                  // 18
    float em_s, bias_s, cx, cy;    // 4
    vec3 j0, j1, j2, j3;           // 12
    int clamp;                     // 1
    float pad;
};
static_assert(sizeof(Work) == 48 * sizeof(float), "constraint work record is 48 floats");

struct CBody {
    vec3 v, w, pos, imv;
    quat q;
    sym3 I;
    bool dyn;
};

__device__ __forceinline__ vec3 p3(const float* p) { return {p[0], p[1], p[2]}; }
__device__ __forceinline__ void set3(float* p, vec3 v) { p[0] = v.x; p[1] = v.y; p[2] = v.z; }

// PrepareForIterationBallSocket (BallSocket.cs:140-183)
__device__ __forceinline__ void prepare_ball_socket(jh_constraint& c, Work& k, CBody& A, CBody& B, float inv_dt,
                                                    float inv_sub) {
    k.r1 = rot(p3(c.params), A.q);
    k.r2 = rot(p3(c.params + 3), B.q);
    vec3 p1 = A.pos + k.r1;
    vec3 p2 = B.pos + k.r2;
    mat3 cr1 = cross_matrix(k.r1), cr2 = cross_matrix(k.r2);
    sym3 s = sym_add(sym_add(sym_add(sym_scale_vec(A.imv), sym_rotate(A.I, cr1)), sym_scale_vec(B.imv)),
                     sym_rotate(B.I, cr2));
    k.em = sym_to_matrix(s);
    float softness = c.params[7] * inv_sub;
    k.em.m11 += softness;
    k.em.m22 += softness;
    k.em.m33 += softness;
    k.em = inverse_symmetric(k.em);
    k.bias = (p2 - p1) * c.params[6] * inv_dt;
    vec3 acc = p3(c.state);
    if (A.dyn) {
        A.v = A.v - cmul(acc, A.imv);
        A.w = A.w - sym_mul(mat_mul(acc, cr1), A.I);
    }
    if (B.dyn) {
        B.v = B.v + cmul(acc, B.imv);
        B.w = B.w + sym_mul(mat_mul(acc, cr2), B.I);
    }
}
// IterateBallSocket (BallSocket.cs:223-252)
__device__ __forceinline__ void iterate_ball_socket(jh_constraint& c, Work& k, CBody& A, CBody& B, float inv_sub) {
    mat3 cr1 = cross_matrix(k.r1), cr2 = cross_matrix(k.r2);
    vec3 acc = p3(c.state);
    vec3 soft = acc * c.params[7] * inv_sub;
    vec3 jv = -A.v + mat_mul(A.w, cr1) + B.v - mat_mul(B.w, cr2);
    vec3 lambda = -1.0f * mat_mul(jv + k.bias + soft, k.em);
    set3(c.state, acc + lambda);
    if (A.dyn) {
        A.v = A.v - cmul(lambda, A.imv);
        A.w = A.w - sym_mul(mat_mul(lambda, cr1), A.I);
    }
    if (B.dyn) {
        B.v = B.v + cmul(lambda, B.imv);
        B.w = B.w + sym_mul(mat_mul(lambda, cr2), B.I);
    }
}

// PrepareForIterationHingeAngle (HingeAngle.cs:133-220)
__device__ __forceinline__ void prepare_hinge_angle(jh_constraint& c, Work& k, CBody& A, CBody& B, float inv_dt,
                                                    float inv_sub) {
    const float* p = c.params;
    vec3 axis = p3(p + 7);
    quat Q0{p[10], p[11], p[12], p[13]};
    vec3 p0 = create_orthonormal(axis);
    vec3 p1 = cross(axis, p0);
    vec3 rotation;
    int hemisphere;
    mat3 m0 = calc_jacobian(qmul(Q0, qconj(A.q)), B.q, rotation, hemisphere);
    int old_h = int(c.state[3]);
    if (old_h != 0 && old_h != hemisphere) c.state[0] = c.state[1] = c.state[2] = 0.0f;
    c.state[3] = float(hemisphere);
    vec3 error{dot(p0, rotation), dot(p1, rotation), dot(axis, rotation)};
    k.clamp = 0;
    vec3 j0 = tmat_mul(p0, m0), j1 = tmat_mul(p1, m0), j2 = tmat_mul(axis, m0);  // Jacobian columns
    k.jac = {j0.x, j1.x, j2.x, j0.y, j1.y, j2.y, j0.z, j1.z, j2.z};
    k.em = sym_to_matrix(sym_rotate(sym_add(A.I, B.I), mtranspose(k.jac)));
    k.em.m11 += p[6] * inv_sub;
    k.em.m22 += p[6] * inv_sub;
    k.em.m33 += p[5] * inv_sub;
    float maxA = p[1], minA = p[0];
    bool full_range = minA <= -stable::Pi / 2.0f && maxA >= stable::Pi / 2.0f;
    if (minA >= maxA) {
        k.clamp = 3;
        error.z = shortest_angle(error.z - p[2], stable::Pi / 2.0f);
    } else if (!full_range && error.z >= maxA) {
        k.clamp = 1;
        error.z -= maxA;
    } else if (!full_range && error.z <= minA) {
        k.clamp = 2;
        error.z -= minA;
    } else {
        c.state[2] = 0.0f;
        k.em.m33 = 0.0f;
        k.em.m31 = k.em.m13 = 0.0f;
        k.em.m32 = k.em.m23 = 0.0f;
        k.jac.m13 = k.jac.m23 = k.jac.m33 = 0.0f;
    }
    if (k.clamp == 1 || k.clamp == 2) {
        c.state[2] = k.clamp == 1 ? cmin(c.state[2], 0.0f) : cmax(c.state[2], 0.0f);
        k.bilateral = inverse_bilateral_block(k.em);
        k.cx = k.em.m13;
        k.cy = k.em.m23;
    }
    k.em = inverse_symmetric(k.em);
    k.bias = error * inv_dt;
    k.bias.x *= p[3];
    k.bias.y *= p[3];
    k.bias.z *= p[4];
    vec3 acc = p3(c.state);
    if (A.dyn) A.w = A.w + sym_mul(mat_mul(acc, k.jac), A.I);
    if (B.dyn) B.w = B.w - sym_mul(mat_mul(acc, k.jac), B.I);
}
// IterateHingeAngle (HingeAngle.cs:313-364)
__device__ __forceinline__ void iterate_hinge_angle(jh_constraint& c, Work& k, CBody& A, CBody& B, float inv_sub) {
    const float* p = c.params;
    vec3 jv = tmat_mul(A.w - B.w, k.jac);
    vec3 acc = p3(c.state);
    vec3 soft = acc * inv_sub;
    soft.x *= p[6];
    soft.y *= p[6];
    soft.z *= p[5];
    vec3 residual = jv + k.bias + soft;
    vec3 lambda = -1.0f * mat_mul(residual, k.em);
    vec3 orig = acc;
    acc = acc + lambda;
    if ((k.clamp == 1 && acc.z > 0.0f) || (k.clamp == 2 && acc.z < 0.0f)) {
        lambda.z = -orig.z;
        float x = residual.x + k.cx * lambda.z;
        float y = residual.y + k.cy * lambda.z;
        lambda.x = -(k.bilateral.x * x + k.bilateral.y * y);
        lambda.y = -(k.bilateral.y * x + k.bilateral.z * y);
        acc.x = orig.x + lambda.x;
        acc.y = orig.y + lambda.y;
        acc.z = 0.0f;
    } else if (k.clamp == 0) {
        orig.z = 0.0f;
        acc.z = 0.0f;
    }
    set3(c.state, acc);
    lambda = acc - orig;
    if (A.dyn) A.w = A.w + sym_mul(mat_mul(lambda, k.jac), A.I);
    if (B.dyn) B.w = B.w - sym_mul(mat_mul(lambda, k.jac), B.I);
}

// PrepareForIterationFixedAngle (FixedAngle.cs:83-122)
__device__ __forceinline__ void prepare_fixed_angle(jh_constraint& c, Work& k, CBody& A, CBody& B, float inv_dt,
                                                    float inv_sub) {
    const float* p = c.params;
    quat Q0{p[2], p[3], p[4], p[5]};
    vec3 error;
    int hemisphere;
    k.clamp = 1024;
    k.jac = mscale(calc_jacobian(qmul(Q0, qconj(A.q)), B.q, error, hemisphere), -2.0f);
    int old_h = int(c.state[3]);
    if (old_h != 0 && old_h != hemisphere) c.state[0] = c.state[1] = c.state[2] = 0.0f;
    c.state[3] = float(hemisphere);
    k.em = sym_to_matrix(sym_rotate(sym_add(A.I, B.I), k.jac));
    k.em.m11 += p[1] * inv_sub;
    k.em.m22 += p[1] * inv_sub;
    k.em.m33 += p[1] * inv_sub;
    k.em = inverse_symmetric(k.em);
    k.bias = -error * p[0] * inv_dt;
    vec3 acc = p3(c.state);
    if (A.dyn) A.w = A.w + sym_mul(tmat_mul(acc, k.jac), A.I);
    if (B.dyn) B.w = B.w - sym_mul(tmat_mul(acc, k.jac), B.I);
}
// IterateFixedAngle (FixedAngle.cs:161-181)
__device__ __forceinline__ void iterate_fixed_angle(jh_constraint& c, Work& k, CBody& A, CBody& B, float inv_sub) {
    vec3 jv = mat_mul(A.w - B.w, k.jac);
    vec3 acc = p3(c.state);
    vec3 soft = acc * (c.params[1] * inv_sub);
    vec3 lambda = -1.0f * mat_mul(jv + k.bias + soft, k.em);
    set3(c.state, acc + lambda);
    if (A.dyn) A.w = A.w + sym_mul(tmat_mul(lambda, k.jac), A.I);
    if (B.dyn) B.w = B.w - sym_mul(tmat_mul(lambda, k.jac), B.I);
}

// PrepareForIterationDistanceLimit (DistanceLimit.cs:210-291)
__device__ __forceinline__ void prepare_distance_limit(jh_constraint& c, Work& k, CBody& A, CBody& B, float inv_dt,
                                                       float inv_sub) {
    const float* p = c.params;
    vec3 r1 = rot(p3(p), A.q);
    vec3 r2 = rot(p3(p + 3), B.q);
    vec3 p1 = A.pos + r1;
    vec3 p2 = B.pos + r2;
    vec3 dp = p2 - p1;
    float distance = sqrtf(len_sq(dp));
    float error = distance - p[8];
    k.clamp = 0;
    float lmin = p[9], lmax = p[10];
    if (lmin >= lmax) {
        error -= 0.5f * lmin + 0.5f * lmax;
        k.clamp = 3;
    } else if (error >= lmax) {
        k.clamp = 1;
        error -= lmax;
    } else if (error <= lmin) {
        k.clamp = 2;
        error -= lmin;
    } else {
        c.state[0] = 0.0f;
        return;
    }
    if (distance == 0.0f) {
        k.clamp = 0;
        c.state[0] = 0.0f;
        return;
    }
    vec3 n = dp * (1.0f / distance);
    k.j0 = -1.0f * n;
    k.j1 = -1.0f * cross(r1, n);
    k.j2 = 1.0f * n;
    k.j3 = cross(r2, n);
    k.em_s = dot(cmul(k.j0, A.imv), k.j0) + dot(cmul(k.j2, B.imv), k.j2) + dot(sym_mul(k.j1, A.I), k.j1) +
             dot(sym_mul(k.j3, B.I), k.j3);
    k.em_s += p[7] * inv_sub;
    k.em_s = k.em_s > 0.0f ? 1.0f / k.em_s : 0.0f;
    if (k.em_s == 0.0f) c.state[0] = 0.0f;
    k.bias_s = error * p[6] * inv_dt;
    float acc = c.state[0];
    if (A.dyn) {
        A.v = A.v + cmul(acc * k.j0, A.imv);
        A.w = A.w + sym_mul(acc * k.j1, A.I);
    }
    if (B.dyn) {
        B.v = B.v + cmul(acc * k.j2, B.imv);
        B.w = B.w + sym_mul(acc * k.j3, B.I);
    }
}
// IterateDistanceLimit (DistanceLimit.cs:330-376)
__device__ __forceinline__ void iterate_distance_limit(jh_constraint& c, Work& k, CBody& A, CBody& B, float inv_sub) {
    if (k.clamp == 0) return;
    float jv = dot(A.v, k.j0) + dot(A.w, k.j1) + dot(B.v, k.j2) + dot(B.w, k.j3);
    float soft = c.state[0] * c.params[7] * inv_sub;
    float lambda = -k.em_s * (jv + k.bias_s + soft);
    float old = c.state[0];
    float acc = old + lambda;
    if (k.clamp == 1) acc = cmin(acc, 0.0f);
    else if (k.clamp == 2) acc = cmax(acc, 0.0f);
    c.state[0] = acc;
    lambda = acc - old;
    if (A.dyn) {
        A.v = A.v + cmul(lambda * k.j0, A.imv);
        A.w = A.w + sym_mul(lambda * k.j1, A.I);
    }
    if (B.dyn) {
        B.v = B.v + cmul(lambda * k.j2, B.imv);
        B.w = B.w + sym_mul(lambda * k.j3, B.I);
    }
}

// PrepareForIterationTwistAngle (TwistAngle.cs:152-210)
__device__ __forceinline__ void prepare_twist_angle(jh_constraint& c, Work& k, CBody& A, CBody& B, float inv_dt,
                                                    float inv_sub) {
    const float* p = c.params;
    quat Q0{p[3], p[4], p[5], p[6]};
    int hemisphere;
    float error = calc_twist(qmul(Q0, qconj(A.q)), B.q, p3(p), k.j0, hemisphere);
    float lower = p[7], upper = p[8], target = p[9];
    int old_h = int(c.state[1]);
    if (old_h != 0 && old_h != hemisphere) c.state[0] = 0.0f;
    c.state[1] = float(hemisphere);
    k.clamp = 0;
    bool full_range = lower <= -stable::Pi && upper >= stable::Pi;
    if (lower >= upper) {
        k.clamp = 3;
        error = shortest_angle(error - target, stable::Pi);
    } else if (!full_range && error >= upper) {
        k.clamp = 1;
        error -= upper;
    } else if (!full_range && error <= lower) {
        k.clamp = 2;
        error -= lower;
    } else {
        c.state[0] = 0.0f;
        return;
    }
    if (k.clamp == 1) c.state[0] = cmin(c.state[0], 0.0f);
    else if (k.clamp == 2) c.state[0] = cmax(c.state[0], 0.0f);
    k.em_s = dot(sym_mul(k.j0, sym_add(A.I, B.I)), k.j0);
    k.em_s += p[11] * inv_sub;
    k.em_s = k.em_s > 0.0f ? 1.0f / k.em_s : 0.0f;
    if (k.em_s == 0.0f) c.state[0] = 0.0f;
    k.bias_s = error * p[10] * inv_dt;
    float acc = c.state[0];
    if (A.dyn) A.w = A.w + sym_mul(acc * k.j0, A.I);
    if (B.dyn) B.w = B.w - sym_mul(acc * k.j0, B.I);
}
// IterateTwistAngle (TwistAngle.cs:292-328)
__device__ __forceinline__ void iterate_twist_angle(jh_constraint& c, Work& k, CBody& A, CBody& B, float inv_sub) {
    if (k.clamp == 0) return;
    float jv = dot(A.w - B.w, k.j0);
    float soft = c.state[0] * (c.params[11] * inv_sub);
    float lambda = -k.em_s * (jv + k.bias_s + soft);
    float orig = c.state[0];
    float acc = orig + lambda;
    if (k.clamp == 1) acc = cmin(acc, 0.0f);
    else if (k.clamp == 2) acc = cmax(acc, 0.0f);
    c.state[0] = acc;
    lambda = acc - orig;
    if (A.dyn) A.w = A.w + sym_mul(lambda * k.j0, A.I);
    if (B.dyn) B.w = B.w - sym_mul(lambda * k.j0, B.I);
}

// PrepareForIterationConeLimit (ConeLimit.cs:256-347)
__device__ __forceinline__ void prepare_cone_limit(jh_constraint& c, Work& k, CBody& A, CBody& B, float inv_dt,
                                                   float inv_sub) {
    const float* p = c.params;
    vec3 a1 = rot(p3(p), A.q);
    vec3 a2 = rot(p3(p + 3), B.q);
    vec3 previous = p3(c.state + 1);
    vec3 j0 = cross(a2, a1);
    k.clamp = 0;
    float error = dot(a1, a2);
    float lower = -p[8], upper = -p[9];
    float sine = sqrtf(len_sq(j0));
    if (sine > 0.0f) {
        j0 = j0 * (1.0f / sine);
    } else {
        vec3 axis = create_orthonormal(a1);
        vec3 other = cross(a1, axis);
        sym3 ii = sym_add(A.I, B.I);
        j0 = dot(sym_mul(axis, ii), axis) >= dot(sym_mul(other, ii), other) ? axis : other;
    }
    vec3 j1 = -j0;
    set3(c.state + 1, j0);
    if (dot(previous, j0) < 0.0f) c.state[0] = 0.0f;
    error = -stable::atan2(sine, error);
    if (lower == upper) {
        k.clamp = 3;
        error -= upper;
    } else if (error <= upper && p[9] < stable::Pi) {
        k.clamp = 1;
        error -= upper;
    } else if (error >= lower && p[8] > 0.0f) {
        k.clamp = 2;
        error -= lower;
    } else {
        c.state[0] = 0.0f;
        return;
    }
    k.j0 = j0;
    k.j1 = j1;
    if (k.clamp == 1) c.state[0] = cmin(c.state[0], 0.0f);
    else if (k.clamp == 2) c.state[0] = cmax(c.state[0], 0.0f);
    k.em_s = dot(sym_mul(j0, A.I), j0) + dot(sym_mul(j1, B.I), j1);
    k.em_s += p[7] * inv_sub;
    k.em_s = k.em_s > 0.0f ? 1.0f / k.em_s : 0.0f;
    if (k.em_s == 0.0f) c.state[0] = 0.0f;
    k.bias_s = -error * p[6] * inv_dt;
    float acc = c.state[0];
    if (A.dyn) A.w = A.w + sym_mul(acc * j0, A.I);
    if (B.dyn) B.w = B.w + sym_mul(acc * j1, B.I);
}
// IterateConeLimit (ConeLimit.cs:386-427)
__device__ __forceinline__ void iterate_cone_limit(jh_constraint& c, Work& k, CBody& A, CBody& B, float inv_sub) {
    if (k.clamp == 0) return;
    float jv = dot(A.w, k.j0) + dot(B.w, k.j1);
    float soft = c.state[0] * c.params[7] * inv_sub;
    float lambda = -k.em_s * (jv + k.bias_s + soft);
    float old = c.state[0];
    float acc = old + lambda;
    if (k.clamp == 1) acc = cmin(acc, 0.0f);
    else if (k.clamp == 2) acc = cmax(acc, 0.0f);
    c.state[0] = acc;
    lambda = acc - old;
    if (A.dyn) A.w = A.w + sym_mul(lambda * k.j0, A.I);
    if (B.dyn) B.w = B.w + sym_mul(lambda * k.j1, B.I);
}

// PrepareForIterationAngularMotor (AngularMotor.cs:150-175) / PrepareForIterationLinearMotor
// (LinearMotor.cs:152-176); MaxLambda kept in bias_s
__device__ __forceinline__ void prepare_motor(jh_constraint& c, Work& k, CBody& A, CBody& B, float sub_dt,
                                              bool linear) {
    vec3 j1 = rot(p3(c.params), A.q);
    vec3 j2 = rot(p3(c.params + 3), B.q);
    if (linear) k.em_s = dot(cmul(j1, A.imv), j1) + dot(cmul(j2, B.imv), j2);
    else k.em_s = dot(sym_mul(j1, A.I), j1) + dot(sym_mul(j2, B.I), j2);
    k.em_s = k.em_s > 0.0f ? 1.0f / k.em_s : 0.0f;
    if (k.em_s == 0.0f) c.state[0] = 0.0f;
    k.bias_s = sub_dt * c.params[7];
    float acc = c.state[0];
    if (linear) {
        if (A.dyn) A.v = A.v - cmul(j1 * acc, A.imv);
        if (B.dyn) B.v = B.v + cmul(j2 * acc, B.imv);
    } else {
        if (A.dyn) A.w = A.w - sym_mul(j1 * acc, A.I);
        if (B.dyn) B.w = B.w + sym_mul(j2 * acc, B.I);
    }
}
// IterateAngularMotor (AngularMotor.cs:177-207) / IterateLinearMotor (LinearMotor.cs:203-232)
__device__ __forceinline__ void iterate_motor(jh_constraint& c, Work& k, CBody& A, CBody& B, bool linear) {
    vec3 j1 = rot(p3(c.params), A.q);
    vec3 j2 = rot(p3(c.params + 3), B.q);
    float jv = linear ? dot(-j1, A.v) + dot(j2, B.v) : dot(-j1, A.w) + dot(j2, B.w);
    float lambda = -(jv - c.params[6]) * k.em_s;
    float old = c.state[0];
    float acc = clamp_cs(old + lambda, -k.bias_s, k.bias_s);
    c.state[0] = acc;
    lambda = acc - old;
    if (linear) {
        if (A.dyn) A.v = A.v - cmul(j1 * lambda, A.imv);
        if (B.dyn) B.v = B.v + cmul(j2 * lambda, B.imv);
    } else {
        if (A.dyn) A.w = A.w - sym_mul(j1 * lambda, A.I);
        if (B.dyn) B.w = B.w + sym_mul(j2 * lambda, B.I);
    }
}

// Linear rows of the slider constraints: for direction d, Jacobian (-d, -((r1 + u) x d), d, r2 x d).
struct LinRow {
    vec3 j0, j1, j2, j3;
};
__device__ __forceinline__ LinRow lin_row(vec3 r1, vec3 r2, vec3 u, vec3 d) {
    return {-d, -cross(r1 + u, d), d, cross(r2, d)};
}
__device__ __forceinline__ float lin_mass(vec3 a, vec3 imv1, vec3 b, vec3 c, vec3 imv2, vec3 e) {
    return dot(cmul(a, imv1), b) + dot(cmul(c, imv2), e);
}
__device__ __forceinline__ float ang_mass(vec3 a, const sym3& i1, vec3 b, vec3 c, const sym3& i2, vec3 e) {
    return dot(sym_mul(a, i1), b) + dot(sym_mul(c, i2), e);
}

// PrepareForIterationPointOnPlane (PointOnPlane.cs:118-192)
__device__ __forceinline__ void prepare_point_on_plane(jh_constraint& c, Work& k, CBody& A, CBody& B, float inv_dt,
                                                       float inv_sub) {
    const float* p = c.params;
    vec3 axis = rot(p3(p), A.q);
    vec3 r1 = rot(p3(p + 3), A.q);
    vec3 r2 = rot(p3(p + 6), B.q);
    vec3 p1 = A.pos + r1;
    vec3 p2 = B.pos + r2;
    k.clamp = 0;
    vec3 u = p2 - p1;
    k.j0 = -axis;
    k.j1 = -cross(r1 + u, axis);
    k.j2 = axis;
    k.j3 = cross(r2, axis);
    float error = dot(u, axis);
    k.em_s = 1.0f;
    float mn = p[11], mx = p[12];
    if (mn >= mx) {
        error -= 0.5f * mn + 0.5f * mx;
        k.clamp = 3;
    } else if (error >= mx) {
        error -= mx;
        k.clamp = 1;
    } else if (error <= mn) {
        error -= mn;
        k.clamp = 2;
    } else {
        c.state[0] = 0.0f;
        return;
    }
    k.em_s = dot(cmul(k.j0, A.imv), k.j0) + dot(cmul(k.j2, B.imv), k.j2) + dot(sym_mul(k.j1, A.I), k.j1) +
             dot(sym_mul(k.j3, B.I), k.j3);
    k.em_s += (p[10] * inv_sub);
    k.em_s = k.em_s > 0.0f ? 1.0f / k.em_s : 0.0f;
    if (k.em_s == 0.0f) c.state[0] = 0.0f;
    k.bias_s = error * p[9] * inv_dt;
    float acc = c.state[0];
    if (A.dyn) {
        A.v = A.v + cmul(k.j0 * acc, A.imv);
        A.w = A.w + sym_mul(k.j1 * acc, A.I);
    }
    if (B.dyn) {
        B.v = B.v + cmul(k.j2 * acc, B.imv);
        B.w = B.w + sym_mul(k.j3 * acc, B.I);
    }
}
// IteratePointOnPlane (PointOnPlane.cs:231-274)
__device__ __forceinline__ void iterate_point_on_plane(jh_constraint& c, Work& k, CBody& A, CBody& B, float inv_sub) {
    if (k.clamp == 0) return;
    float jv = dot(k.j0, A.v) + dot(k.j1, A.w) + dot(k.j2, B.v) + dot(k.j3, B.w);
    float soft = c.state[0] * c.params[10] * inv_sub;
    float lambda = -1.0f * (jv + k.bias_s + soft) * k.em_s;
    float orig = c.state[0];
    float acc = orig + lambda;
    if (k.clamp == 1) acc = cmin(acc, 0.0f);
    else if (k.clamp == 2) acc = cmax(acc, 0.0f);
    c.state[0] = acc;
    lambda = acc - orig;
    if (A.dyn) {
        A.v = A.v + cmul(k.j0 * lambda, A.imv);
        A.w = A.w + sym_mul(k.j1 * lambda, A.I);
    }
    if (B.dyn) {
        B.v = B.v + cmul(k.j2 * lambda, B.imv);
        B.w = B.w + sym_mul(k.j3 * lambda, B.I);
    }
}

// PointOnLine: the three rows rebuilt from the poses, as in prepare and iterate of PointOnLine.cs
__device__ __forceinline__ void point_on_line_rows(const jh_constraint& c, const CBody& A, const CBody& B, LinRow* J,
                                                   vec3& err) {
    const float* p = c.params;
    vec3 aw = rot(p3(p), A.q);
    vec3 n1 = create_orthonormal(p3(p));
    n1 = rot(n1, A.q);
    vec3 n2 = cross(aw, n1);
    vec3 r1 = rot(p3(p + 3), A.q);
    vec3 r2 = rot(p3(p + 6), B.q);
    vec3 p1 = A.pos + r1;
    vec3 p2 = B.pos + r2;
    vec3 u = p2 - p1;
    J[0] = lin_row(r1, r2, u, n1);
    J[1] = lin_row(r1, r2, u, n2);
    J[2] = lin_row(r1, r2, u, aw);
    err = vec3{dot(u, n1), dot(u, n2), dot(u, aw)};
}
// PrepareForIterationPointOnLine (PointOnLine.cs:149-300)
__device__ __forceinline__ void prepare_point_on_line(jh_constraint& c, Work& k, CBody& A, CBody& B, float inv_dt,
                                                      float inv_sub) {
    const float* p = c.params;
    LinRow J[3];
    vec3 error;
    point_on_line_rows(c, A, B, J, error);
    k.clamp = 0;
    k.em = mat3{1, 0, 0, 0, 1, 0, 0, 0, 1};
    k.em.m11 = lin_mass(J[0].j0, A.imv, J[0].j0, J[0].j2, B.imv, J[0].j2) + dot(sym_mul(J[0].j1, A.I), J[0].j1) +
               dot(sym_mul(J[0].j3, B.I), J[0].j3);
    k.em.m22 = lin_mass(J[1].j0, A.imv, J[1].j0, J[1].j2, B.imv, J[1].j2) + dot(sym_mul(J[1].j1, A.I), J[1].j1) +
               dot(sym_mul(J[1].j3, B.I), J[1].j3);
    k.em.m12 = ang_mass(J[0].j1, A.I, J[1].j1, J[0].j3, B.I, J[1].j3);
    k.em.m21 = ang_mass(J[1].j1, A.I, J[0].j1, J[1].j3, B.I, J[0].j3);
    float coupling12 = lin_mass(J[0].j0, A.imv, J[1].j0, J[0].j2, B.imv, J[1].j2);
    k.em.m12 += coupling12;
    k.em.m21 += coupling12;
    float mn = p[13], mx = p[14];
    if (mn >= mx) {
        error.z -= 0.5f * mn + 0.5f * mx;
        k.clamp = 3;
    } else if (error.z >= mx) {
        error.z -= mx;
        k.clamp = 1;
    } else if (error.z <= mn) {
        error.z -= mn;
        k.clamp = 2;
    } else {
        c.state[2] = 0.0f;
    }
    if (k.clamp != 0) {
        k.em.m33 = lin_mass(J[2].j0, A.imv, J[2].j0, J[2].j2, B.imv, J[2].j2) + dot(sym_mul(J[2].j1, A.I), J[2].j1) +
                   dot(sym_mul(J[2].j3, B.I), J[2].j3);
        k.em.m13 = ang_mass(J[0].j1, A.I, J[2].j1, J[0].j3, B.I, J[2].j3);
        k.em.m31 = ang_mass(J[2].j1, A.I, J[0].j1, J[2].j3, B.I, J[0].j3);
        k.em.m23 = ang_mass(J[1].j1, A.I, J[2].j1, J[1].j3, B.I, J[2].j3);
        k.em.m32 = ang_mass(J[2].j1, A.I, J[1].j1, J[2].j3, B.I, J[1].j3);
        float coupling13 = lin_mass(J[0].j0, A.imv, J[2].j0, J[0].j2, B.imv, J[2].j2);
        float coupling23 = lin_mass(J[1].j0, A.imv, J[2].j0, J[1].j2, B.imv, J[2].j2);
        k.em.m13 += coupling13;
        k.em.m31 += coupling13;
        k.em.m23 += coupling23;
        k.em.m32 += coupling23;
    }
    k.em.m11 += p[11] * inv_sub;
    k.em.m22 += p[11] * inv_sub;
    k.em.m33 += p[12] * inv_sub;
    if (k.clamp == 1 || k.clamp == 2) {
        k.bilateral = inverse_bilateral_block(k.em);
        k.cx = k.em.m13;
        k.cy = k.em.m23;
    }
    if (k.clamp == 0) k.em.m33 = 0.0f;
    k.em = inverse_symmetric(k.em);
    k.bias = error * inv_dt;
    k.bias.x *= p[9];
    k.bias.y *= p[9];
    k.bias.z *= p[10];
    if (k.clamp == 1) c.state[2] = cmin(0.0f, c.state[2]);
    else if (k.clamp == 2) c.state[2] = cmax(0.0f, c.state[2]);
    vec3 acc = p3(c.state);
    if (A.dyn) {
        A.v = A.v + cmul(J[0].j0 * acc.x + J[1].j0 * acc.y + J[2].j0 * acc.z, A.imv);
        A.w = A.w + sym_mul(J[0].j1 * acc.x + J[1].j1 * acc.y + J[2].j1 * acc.z, A.I);
    }
    if (B.dyn) {
        B.v = B.v + cmul(J[0].j2 * acc.x + J[1].j2 * acc.y + J[2].j2 * acc.z, B.imv);
        B.w = B.w + sym_mul(J[0].j3 * acc.x + J[1].j3 * acc.y + J[2].j3 * acc.z, B.I);
    }
}
// IteratePointOnLine (PointOnLine.cs:372-465)
__device__ __forceinline__ void iterate_point_on_line(jh_constraint& c, Work& k, CBody& A, CBody& B, float inv_sub) {
    const float* p = c.params;
    LinRow J[3];
    vec3 unused;
    point_on_line_rows(c, A, B, J, unused);
    vec3 jv;
    jv.x = dot(J[0].j0, A.v) + dot(J[0].j1, A.w) + dot(J[0].j2, B.v) + dot(J[0].j3, B.w);
    jv.y = dot(J[1].j0, A.v) + dot(J[1].j1, A.w) + dot(J[1].j2, B.v) + dot(J[1].j3, B.w);
    jv.z = dot(J[2].j0, A.v) + dot(J[2].j1, A.w) + dot(J[2].j2, B.v) + dot(J[2].j3, B.w);
    vec3 acc = p3(c.state);
    vec3 soft = acc * inv_sub;
    soft.x *= p[11];
    soft.y *= p[11];
    soft.z *= p[12];
    vec3 residual = jv + k.bias + soft;
    vec3 lambda = -1.0f * mat_mul(residual, k.em);
    vec3 orig = acc;
    acc = acc + lambda;
    if ((k.clamp == 1 && acc.z > 0.0f) || (k.clamp == 2 && acc.z < 0.0f)) {
        lambda.z = -orig.z;
        float x = residual.x + k.cx * lambda.z;
        float y = residual.y + k.cy * lambda.z;
        lambda.x = -(k.bilateral.x * x + k.bilateral.y * y);
        lambda.y = -(k.bilateral.y * x + k.bilateral.z * y);
        acc.x = orig.x + lambda.x;
        acc.y = orig.y + lambda.y;
        acc.z = 0.0f;
    } else if (k.clamp == 0) {
        acc.z = 0.0f;
        orig.z = 0.0f;
    }
    set3(c.state, acc);
    lambda = acc - orig;
    if (A.dyn) {
        A.v = A.v + cmul(J[0].j0 * lambda.x + J[1].j0 * lambda.y + J[2].j0 * lambda.z, A.imv);
        A.w = A.w + sym_mul(J[0].j1 * lambda.x + J[1].j1 * lambda.y + J[2].j1 * lambda.z, A.I);
    }
    if (B.dyn) {
        B.v = B.v + cmul(J[0].j2 * lambda.x + J[1].j2 * lambda.y + J[2].j2 * lambda.z, B.imv);
        B.w = B.w + sym_mul(J[0].j3 * lambda.x + J[1].j3 * lambda.y + J[2].j3 * lambda.z, B.I);
    }
}

// ---- Servos (Dynamics/Constraints/Internal/ServoMath.cs); the substep's maximum impulse is kept in cx --
// MaximumImpulse (:30)
__device__ __forceinline__ float servo_max_impulse(float maximum, float dt) { return dt > 0.0f ? maximum * dt : 0.0f; }
__device__ __forceinline__ bool pos_inf(float v) { return v == __builtin_inff(); }
// ClampMagnitude (:32-38)
__device__ __forceinline__ vec3 servo_clamp_magnitude(vec3 impulse, float maximum) {
    if (pos_inf(maximum)) return impulse;
    float l2 = len_sq(impulse);
    if (l2 > maximum * maximum) impulse = impulse * (maximum / sqrtf(l2));
    return impulse;
}
// ClampAngular (:40-48)
__device__ __forceinline__ vec3 servo_clamp_angular(vec3 impulse, const mat3& jac, float maximum) {
    if (pos_inf(maximum)) return impulse;
    float l2 = len_sq(tmat_mul(impulse, jac));
    if (l2 > maximum * maximum) impulse = impulse * (maximum / sqrtf(l2));
    return impulse;
}

// AngularServo.ApplyImpulse (AngularServo.cs:164-172)
__device__ __forceinline__ void angular_servo_apply(CBody& A, CBody& B, vec3 impulse) {
    if (A.dyn) A.w = A.w + sym_mul(impulse, A.I);
    if (B.dyn) B.w = B.w - sym_mul(impulse, B.I);
}
// PrepareForIterationAngularServo (AngularServo.cs:125-149)
__device__ __forceinline__ void prepare_angular_servo(jh_constraint& c, Work& k, CBody& A, CBody& B, float inv_dt,
                                                      float inv_sub, float sub_dt) {
    const float* p = c.params;
    quat target{p[0], p[1], p[2], p[3]};
    vec3 error;
    int hemisphere;
    k.jac = mscale(calc_jacobian(qmul(qconj(target), qconj(A.q)), B.q, error, hemisphere), -2.0f);
    int old_h = int(c.state[3]);
    if (old_h != 0 && old_h != hemisphere) c.state[0] = c.state[1] = c.state[2] = 0.0f;
    c.state[3] = float(hemisphere);
    k.em = sym_to_matrix(sym_rotate(sym_add(A.I, B.I), k.jac));
    float softness = p[5] * inv_sub;
    k.em.m11 += softness;
    k.em.m22 += softness;
    k.em.m33 += softness;
    k.em = inverse_symmetric(k.em);
    k.bias = -error * p[6] * inv_dt;
    k.cx = servo_max_impulse(p[4], sub_dt);
    vec3 acc = servo_clamp_angular(p3(c.state), k.jac, k.cx);
    set3(c.state, acc);
    angular_servo_apply(A, B, tmat_mul(acc, k.jac));
}
// IterateAngularServo (AngularServo.cs:151-162)
__device__ __forceinline__ void iterate_angular_servo(jh_constraint& c, Work& k, CBody& A, CBody& B, float inv_sub) {
    vec3 jv = mat_mul(A.w - B.w, k.jac);
    vec3 previous = p3(c.state);
    vec3 softness = previous * (c.params[5] * inv_sub);
    vec3 lambda = -mat_mul(jv + k.bias + softness, k.em);
    vec3 acc = servo_clamp_angular(previous + lambda, k.jac, k.cx);
    set3(c.state, acc);
    angular_servo_apply(A, B, tmat_mul(acc - previous, k.jac));
}

// LinearServo.ApplyImpulse (LinearServo.cs:189-203)
__device__ __forceinline__ void linear_servo_apply(Work& k, CBody& A, CBody& B, vec3 impulse) {
    if (A.dyn) {
        A.v = A.v - cmul(impulse, A.imv);
        A.w = A.w - sym_mul(cross(k.r1, impulse), A.I);
    }
    if (B.dyn) {
        B.v = B.v + cmul(impulse, B.imv);
        B.w = B.w + sym_mul(cross(k.r2, impulse), B.I);
    }
}
// PrepareForIterationLinearServo (LinearServo.cs:149-171)
__device__ __forceinline__ void prepare_linear_servo(jh_constraint& c, Work& k, CBody& A, CBody& B, float inv_dt,
                                                     float inv_sub, float sub_dt) {
    const float* p = c.params;
    k.r1 = rot(p3(p) + p3(p + 6), A.q);
    k.r2 = rot(p3(p + 3), B.q);
    mat3 cr1 = cross_matrix(k.r1), cr2 = cross_matrix(k.r2);
    sym3 s = sym_add(sym_add(sym_add(sym_scale_vec(A.imv), sym_rotate(A.I, cr1)), sym_scale_vec(B.imv)),
                     sym_rotate(B.I, cr2));
    k.em = sym_to_matrix(s);
    float softness = p[10] * inv_sub;
    k.em.m11 += softness;
    k.em.m22 += softness;
    k.em.m33 += softness;
    k.em = inverse_symmetric(k.em);
    k.bias = (B.pos + k.r2 - A.pos - k.r1) * (p[11] * inv_dt);
    k.cx = servo_max_impulse(p[9], sub_dt);
    vec3 acc = servo_clamp_magnitude(p3(c.state), k.cx);
    set3(c.state, acc);
    linear_servo_apply(k, A, B, acc);
}
// IterateLinearServo (LinearServo.cs:173-187)
__device__ __forceinline__ void iterate_linear_servo(jh_constraint& c, Work& k, CBody& A, CBody& B, float inv_sub) {
    vec3 jv = -A.v - cross(A.w, k.r1) + B.v + cross(B.w, k.r2);
    vec3 previous = p3(c.state);
    vec3 softness = previous * (c.params[10] * inv_sub);
    vec3 lambda = -mat_mul(jv + k.bias + softness, k.em);
    vec3 acc = servo_clamp_magnitude(previous + lambda, k.cx);
    set3(c.state, acc);
    linear_servo_apply(k, A, B, acc - previous);
}

// LinearAxisServo.ApplyImpulse (LinearAxisServo.cs:211-226)
__device__ __forceinline__ void linear_axis_servo_apply(Work& k, CBody& A, CBody& B, float impulse) {
    if (A.dyn) {
        A.v = A.v + cmul(k.j0 * impulse, A.imv);
        A.w = A.w + sym_mul(k.j1 * impulse, A.I);
    }
    if (B.dyn) {
        B.v = B.v + cmul(k.j2 * impulse, B.imv);
        B.w = B.w + sym_mul(k.j3 * impulse, B.I);
    }
}
// PrepareForIterationLinearAxisServo (LinearAxisServo.cs:169-194)
__device__ __forceinline__ void prepare_linear_axis_servo(jh_constraint& c, Work& k, CBody& A, CBody& B,
                                                          float inv_dt, float inv_sub, float sub_dt) {
    const float* p = c.params;
    vec3 axis = rot(p3(p), A.q);
    vec3 r1 = rot(p3(p + 3), A.q);
    vec3 r2 = rot(p3(p + 6), B.q);
    vec3 separation = B.pos + r2 - A.pos - r1;
    k.j0 = -axis;
    k.j1 = -cross(r1 + separation, axis);
    k.j2 = axis;
    k.j3 = cross(r2, axis);
    float mass = dot(cmul(k.j0, A.imv), k.j0) + dot(cmul(k.j2, B.imv), k.j2) + dot(sym_mul(k.j1, A.I), k.j1) +
                 dot(sym_mul(k.j3, B.I), k.j3) + p[11] * inv_sub;
    k.em_s = mass > 0.0f ? 1.0f / mass : 0.0f;
    if (k.em_s == 0.0f) c.state[0] = 0.0f;
    k.bias_s = (dot(separation, axis) - p[9]) * p[12] * inv_dt;
    k.cx = servo_max_impulse(p[10], sub_dt);
    c.state[0] = clamp_cs(c.state[0], -k.cx, k.cx);
    linear_axis_servo_apply(k, A, B, c.state[0]);
}
// IterateLinearAxisServo (LinearAxisServo.cs:196-209)
__device__ __forceinline__ void iterate_linear_axis_servo(jh_constraint& c, Work& k, CBody& A, CBody& B,
                                                          float inv_sub) {
    float jv = dot(k.j0, A.v) + dot(k.j1, A.w) + dot(k.j2, B.v) + dot(k.j3, B.w);
    float lambda = -(jv + k.bias_s + c.state[0] * c.params[11] * inv_sub) * k.em_s;
    float previous = c.state[0];
    c.state[0] = clamp_cs(previous + lambda, -k.cx, k.cx);
    linear_axis_servo_apply(k, A, B, c.state[0] - previous);
}

// TwistServo.ApplyImpulse (TwistServo.cs:184-192)
__device__ __forceinline__ void twist_servo_apply(Work& k, CBody& A, CBody& B, float impulse) {
    if (A.dyn) A.w = A.w + sym_mul(k.j0 * impulse, A.I);
    if (B.dyn) B.w = B.w - sym_mul(k.j0 * impulse, B.I);
}
// PrepareForIterationTwistServo (TwistServo.cs:150-170)
__device__ __forceinline__ void prepare_twist_servo(jh_constraint& c, Work& k, CBody& A, CBody& B, float inv_dt,
                                                    float inv_sub, float sub_dt) {
    const float* p = c.params;
    quat ref{p[3], p[4], p[5], p[6]};
    int hemisphere;
    float angle = calc_twist(qmul(ref, qconj(A.q)), B.q, p3(p), k.j0, hemisphere);
    int old_h = int(c.state[1]);
    if (old_h != 0 && old_h != hemisphere) c.state[0] = 0.0f;
    c.state[1] = float(hemisphere);
    float mass = dot(sym_mul(k.j0, sym_add(A.I, B.I)), k.j0) + p[9] * inv_sub;
    k.em_s = mass > 0.0f ? 1.0f / mass : 0.0f;
    if (k.em_s == 0.0f) c.state[0] = 0.0f;
    float error = angle - p[7];
    k.bias_s = error * p[10] * inv_dt;
    k.cx = servo_max_impulse(p[8], sub_dt);
    c.state[0] = clamp_cs(c.state[0], -k.cx, k.cx);
    twist_servo_apply(k, A, B, c.state[0]);
}
// IterateTwistServo (TwistServo.cs:172-182)
__device__ __forceinline__ void iterate_twist_servo(jh_constraint& c, Work& k, CBody& A, CBody& B, float inv_sub) {
    float jv = dot(A.w - B.w, k.j0);
    float lambda = -k.em_s * (jv + k.bias_s + c.state[0] * c.params[9] * inv_sub);
    float previous = c.state[0];
    c.state[0] = clamp_cs(previous + lambda, -k.cx, k.cx);
    twist_servo_apply(k, A, B, c.state[0] - previous);
}

// EllipticalConeLimit (EllipticalConeLimit.cs, layout as in jh_constraint): params LocalAxis1 (0-2), LocalAxis2 (3-5), LocalReference1
// (6-8), MinAngle1 (9), MaxAngle1 (10), MinAngle2 (11), MaxAngle2 (12), Softness (13), BiasFactor (14);
// state AccumulatedImpulse (0), Jacobian (1-3), Clamp (4) — the last two persist because the next
// prepare compares against them.
__device__ __forceinline__ float min_nan(float a, float b) {  // MathR.Min
    if (a != a) return a;
    if (b != b) return b;
    return a < b ? a : b;
}
// EllipseRadius (:258-268): radius of the ellipse with semiaxes a, b at the azimuth (cosine, sine) and
// the logarithmic azimuth derivative of that radius
__device__ __forceinline__ float ellipse_radius(float a, float b, float cosine, float sine, float& derivative) {
    float scale = min_nan(a, b);
    float u = scale / a;
    float v = scale / b;
    float squared = cosine * cosine * u * u + sine * sine * v * v;
    derivative = cosine * sine * (v * v - u * u) / squared;
    return scale / sqrtf(squared);
}
// PrepareForIterationEllipticalConeLimit (:270-389)
__device__ __forceinline__ void prepare_elliptical_cone(jh_constraint& c, Work& k, CBody& A, CBody& B, float inv_dt,
                                                        float inv_sub) {
    const float* p = c.params;
    vec3 axis = rot(p3(p), A.q);
    vec3 reference = rot(p3(p + 6), A.q);
    vec3 other = cross(axis, reference);
    vec3 target = rot(p3(p + 3), B.q);
    float x = dot(target, reference);
    float y = dot(target, other);
    float z = dot(target, axis);
    float sine = sqrtf(x * x + y * y);
    float angle = stable::atan2(sine, z);
    float ca, sa;
    if (sine > 0.0f) {
        ca = x / sine;
        sa = y / sine;
    } else {
        sym3 inertia = sym_add(A.I, B.I);
        bool use_reference = dot(sym_mul(other, inertia), other) >= dot(sym_mul(reference, inertia), reference);
        ca = use_reference ? 1.0f : 0.0f;
        sa = use_reference ? 0.0f : 1.0f;
    }
    float derivative;
    float outer = ellipse_radius(p[10], p[12], ca, sa, derivative);
    bool fixed_cone = p[9] == p[10] && p[11] == p[12];
    float boundary = outer;
    int clamp = 0;
    if (fixed_cone) {
        clamp = 3;
    } else if (angle >= outer) {
        clamp = 1;
    } else if (p[9] > 0.0f) {
        float inner_derivative;
        float inner = ellipse_radius(p[9], p[11], ca, sa, inner_derivative);
        if (angle <= inner) {
            clamp = 2;
            boundary = inner;
            derivative = inner_derivative;
        }
    }
    int previous = int(c.state[4]);
    c.state[4] = float(clamp);
    k.clamp = clamp;
    if (clamp == 0) {
        c.state[0] = 0.0f;
        return;
    }
    vec3 tilt = reference * ca + other * sa;
    vec3 azimuth_dir = cross(axis, tilt);
    vec3 angle_dir = tilt * z - axis * sine;
    float azimuth_factor = sine > 0.0f ? angle / sine * derivative : 0.0f;
    float scale = max_nan(1.0f, fabsf(azimuth_factor));
    vec3 jac = -azimuth_dir * (1.0f / scale) + angle_dir * (azimuth_factor / scale);
    float inv_len = 1.0f / sqrtf(len_sq(jac));
    jac = jac * inv_len;
    inv_len /= scale;
    if (previous != clamp || dot(p3(c.state + 1), jac) < 0.0f) c.state[0] = 0.0f;
    set3(c.state + 1, jac);
    if (clamp == 1) c.state[0] = min_nan(c.state[0], 0.0f);
    else if (clamp == 2) c.state[0] = max_nan(c.state[0], 0.0f);
    float mass = dot(sym_mul(jac, A.I), jac) + dot(sym_mul(jac, B.I), jac) + p[13] * inv_sub;
    k.em_s = mass > 0.0f ? 1.0f / mass : 0.0f;
    if (k.em_s == 0.0f) c.state[0] = 0.0f;
    k.bias_s = (angle - boundary) * inv_len * p[14] * inv_dt;
    float acc = c.state[0];
    if (A.dyn) A.w = A.w + sym_mul(jac * acc, A.I);
    if (B.dyn) B.w = B.w - sym_mul(jac * acc, B.I);
}
// IterateEllipticalConeLimit (:391-425)
__device__ __forceinline__ void iterate_elliptical_cone(jh_constraint& c, Work& k, CBody& A, CBody& B, float inv_sub) {
    if (k.clamp == 0) return;
    vec3 jac = p3(c.state + 1);
    float jv = dot(A.w - B.w, jac);
    float lambda = -k.em_s * (jv + k.bias_s + c.state[0] * c.params[13] * inv_sub);
    float old = c.state[0];
    float acc = old + lambda;
    if (k.clamp == 1) acc = min_nan(acc, 0.0f);
    else if (k.clamp == 2) acc = max_nan(acc, 0.0f);
    c.state[0] = acc;
    lambda = acc - old;
    if (A.dyn) A.w = A.w + sym_mul(jac * lambda, A.I);
    if (B.dyn) B.w = B.w - sym_mul(jac * lambda, B.I);
}

__device__ __forceinline__ void prepare(jh_constraint& c, Work& k, CBody& A, CBody& B, float inv_dt, float inv_sub,
                                        float sub_dt) {
    switch (c.type) {
        case JH_CONSTRAINT_BALL_SOCKET: prepare_ball_socket(c, k, A, B, inv_dt, inv_sub); break;
        case JH_CONSTRAINT_HINGE_ANGLE: prepare_hinge_angle(c, k, A, B, inv_dt, inv_sub); break;
        case JH_CONSTRAINT_FIXED_ANGLE: prepare_fixed_angle(c, k, A, B, inv_dt, inv_sub); break;
        case JH_CONSTRAINT_DISTANCE_LIMIT: prepare_distance_limit(c, k, A, B, inv_dt, inv_sub); break;
        case JH_CONSTRAINT_TWIST_ANGLE: prepare_twist_angle(c, k, A, B, inv_dt, inv_sub); break;
        case JH_CONSTRAINT_CONE_LIMIT: prepare_cone_limit(c, k, A, B, inv_dt, inv_sub); break;
        case JH_CONSTRAINT_ANGULAR_MOTOR: prepare_motor(c, k, A, B, sub_dt, false); break;
        case JH_CONSTRAINT_LINEAR_MOTOR: prepare_motor(c, k, A, B, sub_dt, true); break;
        case JH_CONSTRAINT_POINT_ON_LINE: prepare_point_on_line(c, k, A, B, inv_dt, inv_sub); break;
        case JH_CONSTRAINT_POINT_ON_PLANE: prepare_point_on_plane(c, k, A, B, inv_dt, inv_sub); break;
        case JH_CONSTRAINT_ANGULAR_SERVO: prepare_angular_servo(c, k, A, B, inv_dt, inv_sub, sub_dt); break;
        case JH_CONSTRAINT_LINEAR_SERVO: prepare_linear_servo(c, k, A, B, inv_dt, inv_sub, sub_dt); break;
        case JH_CONSTRAINT_LINEAR_AXIS_SERVO: prepare_linear_axis_servo(c, k, A, B, inv_dt, inv_sub, sub_dt); break;
        case JH_CONSTRAINT_TWIST_SERVO: prepare_twist_servo(c, k, A, B, inv_dt, inv_sub, sub_dt); break;
        default: prepare_elliptical_cone(c, k, A, B, inv_dt, inv_sub); break;
    }
}
__device__ __forceinline__ void iterate(jh_constraint& c, Work& k, CBody& A, CBody& B, float inv_sub) {
    switch (c.type) {
        case JH_CONSTRAINT_BALL_SOCKET: iterate_ball_socket(c, k, A, B, inv_sub); break;
        case JH_CONSTRAINT_HINGE_ANGLE: iterate_hinge_angle(c, k, A, B, inv_sub); break;
        case JH_CONSTRAINT_FIXED_ANGLE: iterate_fixed_angle(c, k, A, B, inv_sub); break;
        case JH_CONSTRAINT_DISTANCE_LIMIT: iterate_distance_limit(c, k, A, B, inv_sub); break;
        case JH_CONSTRAINT_TWIST_ANGLE: iterate_twist_angle(c, k, A, B, inv_sub); break;
        case JH_CONSTRAINT_CONE_LIMIT: iterate_cone_limit(c, k, A, B, inv_sub); break;
        case JH_CONSTRAINT_ANGULAR_MOTOR: iterate_motor(c, k, A, B, false); break;
        case JH_CONSTRAINT_LINEAR_MOTOR: iterate_motor(c, k, A, B, true); break;
        case JH_CONSTRAINT_POINT_ON_LINE: iterate_point_on_line(c, k, A, B, inv_sub); break;
        case JH_CONSTRAINT_POINT_ON_PLANE: iterate_point_on_plane(c, k, A, B, inv_sub); break;
        case JH_CONSTRAINT_ANGULAR_SERVO: iterate_angular_servo(c, k, A, B, inv_sub); break;
        case JH_CONSTRAINT_LINEAR_SERVO: iterate_linear_servo(c, k, A, B, inv_sub); break;
        case JH_CONSTRAINT_LINEAR_AXIS_SERVO: iterate_linear_axis_servo(c, k, A, B, inv_sub); break;
        case JH_CONSTRAINT_TWIST_SERVO: iterate_twist_servo(c, k, A, B, inv_sub); break;
        default: iterate_elliptical_cone(c, k, A, B, inv_sub); break;
    }
}

}  // namespace con
}  // namespace jh
