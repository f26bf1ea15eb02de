// ORACLE — test infrastructure only (see orc_math.hpp).  Constraints solved with the contacts —
// BallSocket, HingeAngle, FixedAngle, DistanceLimit, TwistAngle, ConeLimit, AngularMotor, LinearMotor,
// PointOnLine, PointOnPlane, AngularServo, LinearServo, LinearAxisServo, TwistServo, EllipticalConeLimit
// (Dynamics/Constraints/*.cs) — and the JMatrix,
// JSymmetricMatrix, MathHelper and AngularConstraintMath pieces they use, restated in the reference's
// operation order and templated on the precision R.
#pragma once
#include <cmath>

#include "orc_math.hpp"

namespace orc {

// ---- JMatrix (LinearMath/JMatrix.cs) ---------------------------------------------------------
template <class R> inline M3<R> mzero() { return M3<R>{R(0), R(0), R(0), R(0), R(0), R(0), R(0), R(0), R(0)}; }
template <class R> inline M3<R> midentity() {
    return M3<R>{R(1), R(0), R(0), R(0), R(1), R(0), R(0), R(0), R(1)};
}
// operator * (:620-633)
template <class R> inline M3<R> mmul(const M3<R>& a, const M3<R>& b) {
    M3<R> r;
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
// Add / Subtract (:353-385), Multiply(matrix, scale) (:503-514)
template <class R> inline M3<R> madd(const M3<R>& a, const M3<R>& b) {
    return {a.m11 + b.m11, a.m12 + b.m12, a.m13 + b.m13, a.m21 + b.m21, a.m22 + b.m22,
            a.m23 + b.m23, a.m31 + b.m31, a.m32 + b.m32, a.m33 + b.m33};
}
template <class R> inline M3<R> msub(const M3<R>& a, const M3<R>& b) {
    return {a.m11 - b.m11, a.m12 - b.m12, a.m13 - b.m13, a.m21 - b.m21, a.m22 - b.m22,
            a.m23 - b.m23, a.m31 - b.m31, a.m32 - b.m32, a.m33 - b.m33};
}
template <class R> inline M3<R> mscale(const M3<R>& a, R s) {
    return {a.m11 * s, a.m12 * s, a.m13 * s, a.m21 * s, a.m22 * s, a.m23 * s, a.m31 * s, a.m32 * s, a.m33 * s};
}
// Transpose (:603-614)
template <class R> inline M3<R> mtranspose(const M3<R>& m) {
    return {m.m11, m.m21, m.m31, m.m12, m.m22, m.m32, m.m13, m.m23, m.m33};
}
// Determinant (:391-395)
template <class R> inline R mdet(const M3<R>& m) {
    return m.m11 * m.m22 * m.m33 + m.m12 * m.m23 * m.m31 + m.m13 * m.m21 * m.m32 - m.m31 * m.m22 * m.m13 -
           m.m32 * m.m23 * m.m11 - m.m33 * m.m21 * m.m12;
}
// CreateCrossProduct (:597-600)
template <class R> inline M3<R> cross_matrix(V3<R> v) {
    return {R(0), -v.z, v.y, v.z, R(0), -v.x, -v.y, v.x, R(0)};
}
template <class R> inline bool finite_r(R v) { return std::isfinite(v); }
template <class R> inline bool mfinite(const M3<R>& m) {
    return finite_r(m.m11) && finite_r(m.m12) && finite_r(m.m13) && finite_r(m.m21) && finite_r(m.m22) &&
           finite_r(m.m23) && finite_r(m.m31) && finite_r(m.m32) && finite_r(m.m33);
}
// TryInverseUnscaled (:455-481)
template <class R> inline bool try_inverse_unscaled(const M3<R>& m, M3<R>& r) {
    R idet = R(1.0) / mdet(m);
    if (!std::isnormal(idet)) {
        r = mzero<R>();
        return false;
    }
    R n11 = m.m22 * m.m33 - m.m23 * m.m32;
    R n12 = m.m13 * m.m32 - m.m12 * m.m33;
    R n13 = m.m12 * m.m23 - m.m22 * m.m13;
    R n21 = m.m23 * m.m31 - m.m33 * m.m21;
    R n22 = m.m11 * m.m33 - m.m31 * m.m13;
    R n23 = m.m13 * m.m21 - m.m23 * m.m11;
    R n31 = m.m21 * m.m32 - m.m31 * m.m22;
    R n32 = m.m12 * m.m31 - m.m32 * m.m11;
    R n33 = m.m11 * m.m22 - m.m21 * m.m12;
    r = {n11 * idet, n12 * idet, n13 * idet, n21 * idet, n22 * idet, n23 * idet, n31 * idet, n32 * idet, n33 * idet};
    return mfinite(r);
}
// MathR.Max with NaN propagation (Math.Max)
template <class R> inline R max_nan(R a, R b) {
    if (a != a) return a;
    if (b != b) return b;
    return a > b ? a : b;
}
// Inverse (:404-444): the direct inverse, else the inverse of the matrix scaled by its largest entry.
template <class R> inline bool minverse(const M3<R>& m, M3<R>& out) {
    M3<R> inv;
    if (try_inverse_unscaled(m, inv)) {
        out = inv;
        return true;
    }
    R scale = max_nan(std::fabs(m.m11), max_nan(std::fabs(m.m12), std::fabs(m.m13)));
    scale = max_nan(scale, max_nan(std::fabs(m.m21), max_nan(std::fabs(m.m22), std::fabs(m.m23))));
    scale = max_nan(scale, max_nan(std::fabs(m.m31), max_nan(std::fabs(m.m32), std::fabs(m.m33))));
    if (scale > R(0) && finite_r(scale)) {
        M3<R> n{m.m11 / scale, m.m12 / scale, m.m13 / scale, m.m21 / scale, m.m22 / scale,
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
    out = mzero<R>();
    return false;
}
// JVector.TransposedTransform(v, M) (JVector.cs:308-317): M^T v
template <class R> inline V3<R> tmat_mul(V3<R> v, const M3<R>& m) {
    return {v.x * m.m11 + v.y * m.m21 + v.z * m.m31, v.x * m.m12 + v.y * m.m22 + v.z * m.m32,
            v.x * m.m13 + v.y * m.m23 + v.z * m.m33};
}

// ---- JSymmetricMatrix (LinearMath/JSymmetricMatrix.cs) ---------------------------------------
template <class R> inline Sym<R> sym_add(const Sym<R>& a, const Sym<R>& b) {  // Add (:190-196)
    return {a.m11 + b.m11, a.m12 + b.m12, a.m13 + b.m13, a.m22 + b.m22, a.m23 + b.m23, a.m33 + b.m33};
}
template <class R> inline Sym<R> sym_scale_vec(V3<R> s) { return {s.x, R(0), R(0), s.y, R(0), s.z}; }  // CreateScale
template <class R> inline Sym<R> sym_outer(V3<R> v) {  // Outer (:89-93)
    return {v.x * v.x, v.x * v.y, v.x * v.z, v.y * v.y, v.y * v.z, v.z * v.z};
}
template <class R> inline M3<R> sym_to_matrix(const Sym<R>& s) {  // ToMatrix (:68-71)
    return {s.m11, s.m12, s.m13, s.m12, s.m22, s.m23, s.m13, s.m23, s.m33};
}
template <class R> inline R sym_det(const Sym<R>& s) {  // Determinant (:104-109)
    return s.m11 * (s.m22 * s.m33 - s.m23 * s.m23) + s.m12 * (s.m23 * s.m13 - s.m12 * s.m33) +
           s.m13 * (s.m12 * s.m23 - s.m22 * s.m13);
}

// ---- MathHelper (LinearMath/MathHelper.cs) ---------------------------------------------------
template <class R> constexpr R solver_tolerance() { return sizeof(R) == 8 ? R(1e-12) : R(1e-6); }

// JacobiRotate (:374-407)
template <class R> inline bool jacobi_rotate(R* a, R* basis, int p, int q, R tolerance) {
    R off = a[3 * p + q];
    if (std::fabs(off) <= tolerance) return false;
    R tau = (a[3 * q + q] - a[3 * p + p]) / (R(2.0) * off);
    R tangent = std::copysign(R(1.0), tau) / (std::fabs(tau) + std::sqrt(R(1.0) + tau * tau));
    R cosine = R(1.0) / std::sqrt(R(1.0) + tangent * tangent);
    R sine = tangent * cosine;
    a[3 * p + p] -= tangent * off;
    a[3 * q + q] += tangent * off;
    a[3 * p + q] = a[3 * q + p] = R(0);
    for (int k = 0; k < 3; ++k) {
        if (k != p && k != q) {
            R akp = a[3 * k + p], akq = a[3 * k + q];
            a[3 * k + p] = a[3 * p + k] = cosine * akp - sine * akq;
            a[3 * k + q] = a[3 * q + k] = sine * akp + cosine * akq;
        }
        R vkp = basis[3 * k + p], vkq = basis[3 * k + q];
        basis[3 * k + p] = cosine * vkp - sine * vkq;
        basis[3 * k + q] = sine * vkp + cosine * vkq;
    }
    return true;
}

// PseudoInverseSymmetric (:282-325)
template <class R> inline M3<R> pseudo_inverse_symmetric(const M3<R>& m) {
    R scale = max_nan(std::fabs(m.m11), max_nan(std::fabs(m.m22), std::fabs(m.m33)));
    if (!(scale > R(0)) || !finite_r(scale)) return mzero<R>();
    Sym<R> n{m.m11 / scale, m.m12 / scale, m.m13 / scale, m.m22 / scale, m.m23 / scale, m.m33 / scale};
    const R tol = solver_tolerance<R>();
    M3<R> inv;
    if (std::fabs(sym_det(n)) > tol && minverse(m, inv)) return inv;
    R a[9] = {n.m11, n.m12, n.m13, n.m12, n.m22, n.m23, n.m13, n.m23, n.m33};
    R basis[9] = {R(1), R(0), R(0), R(0), R(1), R(0), R(0), R(0), R(1)};
    for (int sweep = 0; sweep < 8; ++sweep) {
        bool changed = jacobi_rotate(a, basis, 0, 1, tol);
        changed |= jacobi_rotate(a, basis, 0, 2, tol);
        changed |= jacobi_rotate(a, basis, 1, 2, tol);
        if (!changed) break;
    }
    Sym<R> result{R(0), R(0), R(0), R(0), R(0), R(0)};
    for (int i = 0; i < 3; ++i) {
        R eig = a[3 * i + i];
        if (!(eig > tol)) continue;
        R weight = (R(1.0) / eig) / scale;
        V3<R> ax{basis[i], basis[3 + i], basis[6 + i]};
        Sym<R> o{ax.x * ax.x, ax.x * ax.y, ax.x * ax.z, ax.y * ax.y, ax.y * ax.z, ax.z * ax.z};
        Sym<R> w{o.m11 * weight, o.m12 * weight, o.m13 * weight, o.m22 * weight, o.m23 * weight, o.m33 * weight};
        result = sym_add(result, w);
    }
    return sym_to_matrix(result);
}

// InverseSymmetric (:328-364)
template <class R> inline M3<R> inverse_symmetric(const M3<R>& m) {
    const R tol = solver_tolerance<R>();
    R diag = m.m11 * m.m22 * m.m33;
    M3<R> inv;
    if (mdet(m) / diag > tol && minverse(m, inv)) return inv;
    R x = m.m11 > R(0) ? std::sqrt(m.m11) : R(0);
    R y = m.m22 > R(0) ? std::sqrt(m.m22) : R(0);
    R z = m.m33 > R(0) ? std::sqrt(m.m33) : R(0);
    if (x == R(0) && y == R(0) && z == R(0)) return mzero<R>();
    R xy = x > R(0) && y > R(0) ? m.m12 / x / y : R(0);
    R xz = x > R(0) && z > R(0) ? m.m13 / x / z : R(0);
    R yz = y > R(0) && z > R(0) ? m.m23 / y / z : R(0);
    M3<R> normalized{R(1), xy, xz, xy, R(1), yz, xz, yz, R(1)};
    M3<R> resp = pseudo_inverse_symmetric(normalized);
    xy = x > R(0) && y > R(0) ? resp.m12 / x / y : R(0);
    xz = x > R(0) && z > R(0) ? resp.m13 / x / z : R(0);
    yz = y > R(0) && z > R(0) ? resp.m23 / y / z : R(0);
    M3<R> r{x > R(0) ? resp.m11 / x / x : R(0), xy, xz, xy, y > R(0) ? resp.m22 / y / y : R(0), yz, xz, yz,
            z > R(0) ? resp.m33 / z / z : R(0)};
    bool ok = finite_r(r.m11) && finite_r(r.m12) && finite_r(r.m13) && finite_r(r.m22) && finite_r(r.m23) &&
              finite_r(r.m33);
    return ok ? r : mzero<R>();
}

// InverseBilateralBlock (:367-372): the inverse of the upper-left 2x2 block as (M11, M12, M22).
template <class R> inline V3<R> inverse_bilateral_block(const M3<R>& m) {
    M3<R> block{m.m11, m.m12, R(0), m.m12, m.m22, R(0), R(0), R(0), R(0)};
    M3<R> inv = inverse_symmetric(block);
    return {inv.m11, inv.m12, inv.m22};
}

// ---- AngularConstraintMath (Dynamics/Constraints/Internal/AngularConstraintMath.cs) ------------
template <class R> inline Q4<R> qconj(Q4<R> q) { return {-q.x, -q.y, -q.z, q.w}; }
template <class R> inline Q4<R> qscale(Q4<R> q, R s) { return {q.x * s, q.y * s, q.z * s, q.w * s}; }

// ShortestAngle (:13-21)
template <class R> inline R shortest_angle(R angle, R half) {
    if (angle >= -half && angle <= half) return angle;
    R period = R(2) * half;
    angle = std::fmod(angle, period);
    if (angle > half) angle -= period;
    else if (angle < -half) angle += period;
    return angle;
}

// CalculateJacobian (:63-94): d(rotation coordinate)/d(angular velocity difference) and the error.
template <class R> inline M3<R> calc_jacobian(Q4<R> left, Q4<R> right, V3<R>& error, int& hemisphere) {
    Q4<R> q = qmul(left, right);
    hemisphere = q.w < R(0) ? -1 : 1;
    if (q.w < R(0)) q = qscale(q, R(-1));
    V3<R> vec{q.x, q.y, q.z};
    R sine = length(vec);
    M3<R> derivative;
    if (sine > R(0)) {
        R half_angle = StableMath<R>::atan2(sine, q.w);
        V3<R> axis = vec * (R(1) / sine);
        error = axis * half_angle;
        R diagonal = half_angle * q.w / sine;
        M3<R> t1 = mscale(midentity<R>(), diagonal);
        M3<R> t2 = mscale(sym_to_matrix(sym_outer(axis)), R(1) - diagonal);
        derivative = msub(madd(t1, t2), cross_matrix(error));
    } else {
        error = V3<R>{R(0), R(0), R(0)};
        derivative = midentity<R>();
    }
    return mmul(mscale(derivative, -R(0.5)), mat_from_quat(left));
}

// QMatrix.ProjectMultiplyLeftRight (Dynamics/Constraints/Internal/QMatrix.cs:109-124)
template <class R> inline M3<R> project_multiply_left_right(Q4<R> l, Q4<R> r) {
    M3<R> m;
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

// CalculateTwist (AngularConstraintMath.cs:31-61): twist angle about `axis` and its unit Jacobian.
template <class R> inline R calc_twist(Q4<R> left, Q4<R> right, V3<R> axis, V3<R>& jacobian, int& hemisphere) {
    M3<R> m = mscale(project_multiply_left_right(left, right), -R(1.0 / 2.0));
    Q4<R> q = qmul(left, right);
    jacobian = tmat_mul(axis, m);
    R projection = dot(axis, V3<R>{q.x, q.y, q.z});
    hemisphere = q.w < R(0) ? -1 : 1;
    if (q.w < R(0)) {
        q = qscale(q, R(-1));
        projection *= -R(1);
        jacobian = jacobian * R(-1);
    }
    R dsq = len_sq(jacobian);
    if (dsq > R(0)) jacobian = jacobian * (R(1) / std::sqrt(dsq));
    else jacobian = -rotate(axis, right);
    V3<R> perpendicular = V3<R>{q.x, q.y, q.z} - axis * projection;
    R cosine = std::sqrt(q.w * q.w + len_sq(perpendicular));
    return R(2) * StableMath<R>::atan2(projection, cosine);
}

// Math.Clamp(value, min, max) for min <= max
template <class R> inline R clamp_cs(R v, R lo, R hi) {
    if (v < lo) return lo;
    if (v > hi) return hi;
    return v;
}

// ---- Constraints ----------------------------------------------------------------------------
// One constraint: kind, bodies, the Data fields of the reference's constraint struct that persist
// (Initialize / properties: params; accumulated impulse and hemisphere: state) and the fields its
// PrepareForIteration computes for Iterate (work).  Layouts per kind as in jh_constraint.
template <class R> struct Constraint {
    int type = 0, b1 = 0, b2 = 0, enabled = 1;
    R p[16] = {};
    R s[8] = {};
    // work
    V3<R> r1{}, r2{}, bias{}, bilateral{};
    M3<R> em{}, jac{};
    R em_s = 0, bias_s = 0, cx = 0, cy = 0;
    V3<R> j0{}, j1{}, j2{}, j3{};
    int clamp = 0;
};

template <class R> struct ConstraintBody {
    V3<R>* vel;
    V3<R>* angvel;
    V3<R> pos, imv;
    Q4<R> q;
    Sym<R> iw;
    bool dynamic;
};

template <class R> inline V3<R> vecp(const R* p) { return {p[0], p[1], p[2]}; }
template <class R> inline void setv(R* p, V3<R> v) { p[0] = v.x; p[1] = v.y; p[2] = v.z; }

// BallSocket (BallSocket.cs): params LocalAnchor1 (0-2), LocalAnchor2 (3-5), BiasFactor (6), Softness (7);
// state AccumulatedImpulse (0-2).
// PrepareForIterationBallSocket (:140-183)
// Leading entries of the constraint state that hold the accumulated impulse (ResetWarmStart of each
// constraint class zeroes its AccumulatedImpulse; the hemisphere and clamp records after it stay).
inline int constraint_impulse_count(int type) {
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

template <class R> void prepare_ball_socket(Constraint<R>& c, ConstraintBody<R>& A, ConstraintBody<R>& B, R inv_dt,
                                            R inv_sub) {
    c.r1 = rotate(vecp(c.p), A.q);
    c.r2 = rotate(vecp(c.p + 3), B.q);
    V3<R> p1 = A.pos + c.r1;
    V3<R> p2 = B.pos + c.r2;
    M3<R> cr1 = cross_matrix(c.r1), cr2 = cross_matrix(c.r2);
    Sym<R> s = sym_add(sym_add(sym_add(sym_scale_vec(A.imv), sym_transform(A.iw, cr1)), sym_scale_vec(B.imv)),
                       sym_transform(B.iw, cr2));
    c.em = sym_to_matrix(s);
    R softness = c.p[7] * inv_sub;
    c.em.m11 += softness;
    c.em.m22 += softness;
    c.em.m33 += softness;
    c.em = inverse_symmetric(c.em);
    c.bias = (p2 - p1) * c.p[6] * inv_dt;
    V3<R> acc = vecp(c.s);
    if (A.dynamic) {
        *A.vel = *A.vel - cmul(acc, A.imv);
        *A.angvel = *A.angvel - sym_mul(mat_mul(acc, cr1), A.iw);
    }
    if (B.dynamic) {
        *B.vel = *B.vel + cmul(acc, B.imv);
        *B.angvel = *B.angvel + sym_mul(mat_mul(acc, cr2), B.iw);
    }
}
// IterateBallSocket (:223-252)
template <class R> void iterate_ball_socket(Constraint<R>& c, ConstraintBody<R>& A, ConstraintBody<R>& B, R inv_sub) {
    M3<R> cr1 = cross_matrix(c.r1), cr2 = cross_matrix(c.r2);
    V3<R> acc = vecp(c.s);
    V3<R> soft = acc * c.p[7] * inv_sub;
    V3<R> jv = -*A.vel + mat_mul(*A.angvel, cr1) + *B.vel - mat_mul(*B.angvel, cr2);
    V3<R> lambda = -R(1.0) * mat_mul(jv + c.bias + soft, c.em);
    setv(c.s, acc + lambda);
    if (A.dynamic) {
        *A.vel = *A.vel - cmul(lambda, A.imv);
        *A.angvel = *A.angvel - sym_mul(mat_mul(lambda, cr1), A.iw);
    }
    if (B.dynamic) {
        *B.vel = *B.vel + cmul(lambda, B.imv);
        *B.angvel = *B.angvel + sym_mul(mat_mul(lambda, cr2), B.iw);
    }
}

// HingeAngle (HingeAngle.cs): params MinAngle (0), MaxAngle (1), FixedAngle (2), BiasFactor (3),
// LimitBias (4), LimitSoftness (5), Softness (6), Axis (7-9), Q0 (10-13, x y z w); state
// AccumulatedImpulse (0-2), Hemisphere (3).
// PrepareForIterationHingeAngle (:133-220)
template <class R> void prepare_hinge_angle(Constraint<R>& c, ConstraintBody<R>& A, ConstraintBody<R>& B, R inv_dt,
                                            R inv_sub) {
    const R kPi = R(3.14159265358979323846);
    Q4<R> q1 = A.q, q2 = B.q;
    V3<R> axis = vecp(c.p + 7);
    Q4<R> Q0{c.p[10], c.p[11], c.p[12], c.p[13]};
    V3<R> p0 = create_orthonormal(axis);
    V3<R> p1 = cross(axis, p0);
    V3<R> rotation;
    int hemisphere;
    M3<R> m0 = calc_jacobian(qmul(Q0, qconj(q1)), q2, rotation, hemisphere);
    int old_h = int(c.s[3]);
    if (old_h != 0 && old_h != hemisphere) c.s[0] = c.s[1] = c.s[2] = R(0);
    c.s[3] = R(hemisphere);
    V3<R> error{dot(p0, rotation), dot(p1, rotation), dot(axis, rotation)};
    c.clamp = 0;
    // Jacobian.UnsafeGet(i) is column i (JMatrix stores its columns contiguously, JMatrix.cs:19-27, 75-79)
    V3<R> j0 = tmat_mul(p0, m0), j1 = tmat_mul(p1, m0), j2 = tmat_mul(axis, m0);
    c.jac = {j0.x, j1.x, j2.x, j0.y, j1.y, j2.y, j0.z, j1.z, j2.z};
    c.em = sym_to_matrix(sym_transform(sym_add(A.iw, B.iw), mtranspose(c.jac)));
    c.em.m11 += c.p[6] * inv_sub;
    c.em.m22 += c.p[6] * inv_sub;
    c.em.m33 += c.p[5] * inv_sub;
    R maxA = c.p[1], minA = c.p[0];
    bool full_range = minA <= -kPi / R(2) && maxA >= kPi / R(2);
    if (minA >= maxA) {
        c.clamp = 3;
        error.z = shortest_angle(error.z - c.p[2], kPi / R(2));
    } else if (!full_range && error.z >= maxA) {
        c.clamp = 1;
        error.z -= maxA;
    } else if (!full_range && error.z <= minA) {
        c.clamp = 2;
        error.z -= minA;
    } else {
        c.s[2] = R(0);
        c.em.m33 = R(0);
        c.em.m31 = c.em.m13 = R(0);
        c.em.m32 = c.em.m23 = R(0);
        c.jac.m13 = c.jac.m23 = c.jac.m33 = R(0);
    }
    if (c.clamp == 1 || c.clamp == 2) {
        c.s[2] = c.clamp == 1 ? std::min(c.s[2], R(0)) : std::max(c.s[2], R(0));
        c.bilateral = inverse_bilateral_block(c.em);
        c.cx = c.em.m13;
        c.cy = c.em.m23;
    }
    c.em = inverse_symmetric(c.em);
    c.bias = error * inv_dt;
    c.bias.x *= c.p[3];
    c.bias.y *= c.p[3];
    c.bias.z *= c.p[4];
    V3<R> acc = vecp(c.s);
    if (A.dynamic) *A.angvel = *A.angvel + sym_mul(mat_mul(acc, c.jac), A.iw);
    if (B.dynamic) *B.angvel = *B.angvel - sym_mul(mat_mul(acc, c.jac), B.iw);
}
// IterateHingeAngle (:313-364)
template <class R> void iterate_hinge_angle(Constraint<R>& c, ConstraintBody<R>& A, ConstraintBody<R>& B, R inv_sub) {
    V3<R> jv = tmat_mul(*A.angvel - *B.angvel, c.jac);
    V3<R> acc = vecp(c.s);
    V3<R> soft = acc * inv_sub;
    soft.x *= c.p[6];
    soft.y *= c.p[6];
    soft.z *= c.p[5];
    V3<R> residual = jv + c.bias + soft;
    V3<R> lambda = -R(1.0) * mat_mul(residual, c.em);
    V3<R> orig = acc;
    acc = acc + lambda;
    if ((c.clamp == 1 && acc.z > R(0)) || (c.clamp == 2 && acc.z < R(0))) {
        lambda.z = -orig.z;
        R x = residual.x + c.cx * lambda.z;
        R y = residual.y + c.cy * lambda.z;
        lambda.x = -(c.bilateral.x * x + c.bilateral.y * y);
        lambda.y = -(c.bilateral.y * x + c.bilateral.z * y);
        acc.x = orig.x + lambda.x;
        acc.y = orig.y + lambda.y;
        acc.z = R(0);
    } else if (c.clamp == 0) {
        orig.z = R(0);
        acc.z = R(0);
    }
    setv(c.s, acc);
    lambda = acc - orig;
    if (A.dynamic) *A.angvel = *A.angvel + sym_mul(mat_mul(lambda, c.jac), A.iw);
    if (B.dynamic) *B.angvel = *B.angvel - sym_mul(mat_mul(lambda, c.jac), B.iw);
}

// FixedAngle (FixedAngle.cs): params BiasFactor (0), Softness (1), Q0 (2-5, x y z w); state
// AccumulatedImpulse (0-2), Hemisphere (3).
// PrepareForIterationFixedAngle (:83-122)
template <class R> void prepare_fixed_angle(Constraint<R>& c, ConstraintBody<R>& A, ConstraintBody<R>& B, R inv_dt,
                                            R inv_sub) {
    Q4<R> Q0{c.p[2], c.p[3], c.p[4], c.p[5]};
    V3<R> error;
    int hemisphere;
    c.clamp = 1024;
    c.jac = mscale(calc_jacobian(qmul(Q0, qconj(A.q)), B.q, error, hemisphere), -R(2));
    int old_h = int(c.s[3]);
    if (old_h != 0 && old_h != hemisphere) c.s[0] = c.s[1] = c.s[2] = R(0);
    c.s[3] = R(hemisphere);
    c.em = sym_to_matrix(sym_transform(sym_add(A.iw, B.iw), c.jac));
    c.em.m11 += c.p[1] * inv_sub;
    c.em.m22 += c.p[1] * inv_sub;
    c.em.m33 += c.p[1] * inv_sub;
    c.em = inverse_symmetric(c.em);
    c.bias = -error * c.p[0] * inv_dt;
    V3<R> acc = vecp(c.s);
    if (A.dynamic) *A.angvel = *A.angvel + sym_mul(tmat_mul(acc, c.jac), A.iw);
    if (B.dynamic) *B.angvel = *B.angvel - sym_mul(tmat_mul(acc, c.jac), B.iw);
}
// IterateFixedAngle (:161-181)
template <class R> void iterate_fixed_angle(Constraint<R>& c, ConstraintBody<R>& A, ConstraintBody<R>& B, R inv_sub) {
    V3<R> jv = mat_mul(*A.angvel - *B.angvel, c.jac);
    V3<R> acc = vecp(c.s);
    V3<R> soft = acc * (c.p[1] * inv_sub);
    V3<R> lambda = -R(1.0) * mat_mul(jv + c.bias + soft, c.em);
    setv(c.s, acc + lambda);
    if (A.dynamic) *A.angvel = *A.angvel + sym_mul(tmat_mul(lambda, c.jac), A.iw);
    if (B.dynamic) *B.angvel = *B.angvel - sym_mul(tmat_mul(lambda, c.jac), B.iw);
}

// DistanceLimit (DistanceLimit.cs): params LocalAnchor1 (0-2), LocalAnchor2 (3-5), BiasFactor (6),
// Softness (7), Distance (8), LimitMin (9), LimitMax (10); state AccumulatedImpulse (0).
// PrepareForIterationDistanceLimit (:210-291)
template <class R> void prepare_distance_limit(Constraint<R>& c, ConstraintBody<R>& A, ConstraintBody<R>& B, R inv_dt,
                                               R inv_sub) {
    V3<R> r1 = rotate(vecp(c.p), A.q);
    V3<R> r2 = rotate(vecp(c.p + 3), B.q);
    V3<R> p1 = A.pos + r1;
    V3<R> p2 = B.pos + r2;
    V3<R> dp = p2 - p1;
    R distance = length(dp);
    R error = distance - c.p[8];
    c.clamp = 0;
    R lmin = c.p[9], lmax = c.p[10];
    if (lmin >= lmax) {
        error -= R(0.5) * lmin + R(0.5) * lmax;
        c.clamp = 3;
    } else if (error >= lmax) {
        c.clamp = 1;
        error -= lmax;
    } else if (error <= lmin) {
        c.clamp = 2;
        error -= lmin;
    } else {
        c.s[0] = R(0.0);
        return;
    }
    if (distance == R(0)) {
        c.clamp = 0;
        c.s[0] = R(0);
        return;
    }
    V3<R> n = dp * (R(1.0) / distance);
    c.j0 = -R(1.0) * n;
    c.j1 = -R(1.0) * cross(r1, n);
    c.j2 = R(1.0) * n;
    c.j3 = cross(r2, n);
    c.em_s = dot(cmul(c.j0, A.imv), c.j0) + dot(cmul(c.j2, B.imv), c.j2) + dot(sym_mul(c.j1, A.iw), c.j1) +
             dot(sym_mul(c.j3, B.iw), c.j3);
    c.em_s += c.p[7] * inv_sub;
    c.em_s = c.em_s > R(0) ? R(1.0) / c.em_s : R(0);
    if (c.em_s == R(0)) c.s[0] = R(0);
    c.bias_s = error * c.p[6] * inv_dt;
    R acc = c.s[0];
    if (A.dynamic) {
        *A.vel = *A.vel + cmul(acc * c.j0, A.imv);
        *A.angvel = *A.angvel + sym_mul(acc * c.j1, A.iw);
    }
    if (B.dynamic) {
        *B.vel = *B.vel + cmul(acc * c.j2, B.imv);
        *B.angvel = *B.angvel + sym_mul(acc * c.j3, B.iw);
    }
}
// IterateDistanceLimit (:330-376)
template <class R> void iterate_distance_limit(Constraint<R>& c, ConstraintBody<R>& A, ConstraintBody<R>& B,
                                               R inv_sub) {
    if (c.clamp == 0) return;
    R jv = dot(*A.vel, c.j0) + dot(*A.angvel, c.j1) + dot(*B.vel, c.j2) + dot(*B.angvel, c.j3);
    R soft = c.s[0] * c.p[7] * inv_sub;
    R lambda = -c.em_s * (jv + c.bias_s + soft);
    R old = c.s[0];
    R acc = old + lambda;
    if (c.clamp == 1) acc = std::min(acc, R(0.0));
    else if (c.clamp == 2) acc = std::max(acc, R(0.0));
    c.s[0] = acc;
    lambda = acc - old;
    if (A.dynamic) {
        *A.vel = *A.vel + cmul(lambda * c.j0, A.imv);
        *A.angvel = *A.angvel + sym_mul(lambda * c.j1, A.iw);
    }
    if (B.dynamic) {
        *B.vel = *B.vel + cmul(lambda * c.j2, B.imv);
        *B.angvel = *B.angvel + sym_mul(lambda * c.j3, B.iw);
    }
}

// TwistAngle (TwistAngle.cs): params B (0-2, body-2 frame), Q0 (3-6), Angle1 (7), Angle2 (8),
// FixedAngle (9), BiasFactor (10), Softness (11); state AccumulatedImpulse (0), Hemisphere (1).
// PrepareForIterationTwistAngle (:152-210)
template <class R> void prepare_twist_angle(Constraint<R>& c, ConstraintBody<R>& A, ConstraintBody<R>& B, R inv_dt,
                                            R inv_sub) {
    const R kPi = R(3.14159265358979323846);
    Q4<R> Q0{c.p[3], c.p[4], c.p[5], c.p[6]};
    int hemisphere;
    R error = calc_twist(qmul(Q0, qconj(A.q)), B.q, vecp(c.p), c.j0, hemisphere);
    R lower = c.p[7], upper = c.p[8], target = c.p[9];
    int old_h = int(c.s[1]);
    if (old_h != 0 && old_h != hemisphere) c.s[0] = R(0);
    c.s[1] = R(hemisphere);
    c.clamp = 0;
    bool full_range = lower <= -kPi && upper >= kPi;
    if (lower >= upper) {
        c.clamp = 3;
        error = shortest_angle(error - target, kPi);
    } else if (!full_range && error >= upper) {
        c.clamp = 1;
        error -= upper;
    } else if (!full_range && error <= lower) {
        c.clamp = 2;
        error -= lower;
    } else {
        c.s[0] = R(0.0);
        return;
    }
    if (c.clamp == 1) c.s[0] = std::min(c.s[0], R(0));
    else if (c.clamp == 2) c.s[0] = std::max(c.s[0], R(0));
    c.em_s = dot(sym_mul(c.j0, sym_add(A.iw, B.iw)), c.j0);
    c.em_s += c.p[11] * inv_sub;
    c.em_s = c.em_s > R(0) ? R(1.0) / c.em_s : R(0);
    if (c.em_s == R(0)) c.s[0] = R(0);
    c.bias_s = error * c.p[10] * inv_dt;
    R acc = c.s[0];
    if (A.dynamic) *A.angvel = *A.angvel + sym_mul(acc * c.j0, A.iw);
    if (B.dynamic) *B.angvel = *B.angvel - sym_mul(acc * c.j0, B.iw);
}
// IterateTwistAngle (:292-328)
template <class R> void iterate_twist_angle(Constraint<R>& c, ConstraintBody<R>& A, ConstraintBody<R>& B, R inv_sub) {
    if (c.clamp == 0) return;
    R jv = dot(*A.angvel - *B.angvel, c.j0);
    R soft = c.s[0] * (c.p[11] * inv_sub);
    R lambda = -c.em_s * (jv + c.bias_s + soft);
    R orig = c.s[0];
    R acc = orig + lambda;
    if (c.clamp == 1) acc = std::min(acc, R(0.0));
    else if (c.clamp == 2) acc = std::max(acc, R(0.0));
    c.s[0] = acc;
    lambda = acc - orig;
    if (A.dynamic) *A.angvel = *A.angvel + sym_mul(lambda * c.j0, A.iw);
    if (B.dynamic) *B.angvel = *B.angvel - sym_mul(lambda * c.j0, B.iw);
}

// ConeLimit (ConeLimit.cs): params LocalAxis1 (0-2), LocalAxis2 (3-5), BiasFactor (6), Softness (7),
// MinAngle (8), MaxAngle (9); state AccumulatedImpulse (0), the previous step's Jacobian J0 (1-3).
// PrepareForIterationConeLimit (:256-347)
template <class R> void prepare_cone_limit(Constraint<R>& c, ConstraintBody<R>& A, ConstraintBody<R>& B, R inv_dt,
                                           R inv_sub) {
    const R kPi = R(3.14159265358979323846);
    V3<R> a1 = rotate(vecp(c.p), A.q);
    V3<R> a2 = rotate(vecp(c.p + 3), B.q);
    V3<R> previous = vecp(c.s + 1);
    V3<R> j0 = cross(a2, a1);
    c.clamp = 0;
    R error = dot(a1, a2);
    R lower = -c.p[8], upper = -c.p[9];
    R sine = std::sqrt(len_sq(j0));
    if (sine > R(0)) {
        j0 = j0 * (R(1.0) / sine);
    } else {
        V3<R> axis = create_orthonormal(a1);
        V3<R> other = cross(a1, axis);
        Sym<R> ii = sym_add(A.iw, B.iw);
        j0 = dot(sym_mul(axis, ii), axis) >= dot(sym_mul(other, ii), other) ? axis : other;
    }
    V3<R> j1 = -j0;
    setv(c.s + 1, j0);
    if (dot(previous, j0) < R(0)) c.s[0] = R(0);
    error = -StableMath<R>::atan2(sine, error);
    if (lower == upper) {
        c.clamp = 3;
        error -= upper;
    } else if (error <= upper && c.p[9] < kPi) {
        c.clamp = 1;
        error -= upper;
    } else if (error >= lower && c.p[8] > R(0)) {
        c.clamp = 2;
        error -= lower;
    } else {
        c.s[0] = R(0.0);
        return;
    }
    c.j0 = j0;
    c.j1 = j1;
    if (c.clamp == 1) c.s[0] = std::min(c.s[0], R(0));
    else if (c.clamp == 2) c.s[0] = std::max(c.s[0], R(0));
    c.em_s = dot(sym_mul(j0, A.iw), j0) + dot(sym_mul(j1, B.iw), j1);
    c.em_s += c.p[7] * inv_sub;
    c.em_s = c.em_s > R(0) ? R(1.0) / c.em_s : R(0);
    if (c.em_s == R(0)) c.s[0] = R(0);
    c.bias_s = -error * c.p[6] * inv_dt;
    R acc = c.s[0];
    if (A.dynamic) *A.angvel = *A.angvel + sym_mul(acc * j0, A.iw);
    if (B.dynamic) *B.angvel = *B.angvel + sym_mul(acc * j1, B.iw);
}
// IterateConeLimit (:386-427)
template <class R> void iterate_cone_limit(Constraint<R>& c, ConstraintBody<R>& A, ConstraintBody<R>& B, R inv_sub) {
    if (c.clamp == 0) return;
    R jv = dot(*A.angvel, c.j0) + dot(*B.angvel, c.j1);
    R soft = c.s[0] * c.p[7] * inv_sub;
    R lambda = -c.em_s * (jv + c.bias_s + soft);
    R old = c.s[0];
    R acc = old + lambda;
    if (c.clamp == 1) acc = std::min(acc, R(0.0));
    else if (c.clamp == 2) acc = std::max(acc, R(0.0));
    c.s[0] = acc;
    lambda = acc - old;
    if (A.dynamic) *A.angvel = *A.angvel + sym_mul(lambda * c.j0, A.iw);
    if (B.dynamic) *B.angvel = *B.angvel + sym_mul(lambda * c.j1, B.iw);
}

// AngularMotor (AngularMotor.cs) / LinearMotor (LinearMotor.cs): params LocalAxis1 (0-2), LocalAxis2
// (3-5), Velocity (6), MaxForce (7); state AccumulatedImpulse (0).  MaxLambda is kept in bias_s.
// PrepareForIterationAngularMotor (AngularMotor.cs:150-175), PrepareForIterationLinearMotor
// (LinearMotor.cs:152-176)
template <class R> void prepare_motor(Constraint<R>& c, ConstraintBody<R>& A, ConstraintBody<R>& B, R sub_dt,
                                      bool linear) {
    V3<R> j1 = rotate(vecp(c.p), A.q);
    V3<R> j2 = rotate(vecp(c.p + 3), B.q);
    if (linear) c.em_s = dot(cmul(j1, A.imv), j1) + dot(cmul(j2, B.imv), j2);
    else c.em_s = dot(sym_mul(j1, A.iw), j1) + dot(sym_mul(j2, B.iw), j2);
    c.em_s = c.em_s > R(0) ? R(1.0) / c.em_s : R(0);
    if (c.em_s == R(0)) c.s[0] = R(0);
    c.bias_s = sub_dt * c.p[7];
    R acc = c.s[0];
    if (linear) {
        if (A.dynamic) *A.vel = *A.vel - cmul(j1 * acc, A.imv);
        if (B.dynamic) *B.vel = *B.vel + cmul(j2 * acc, B.imv);
    } else {
        if (A.dynamic) *A.angvel = *A.angvel - sym_mul(j1 * acc, A.iw);
        if (B.dynamic) *B.angvel = *B.angvel + sym_mul(j2 * acc, B.iw);
    }
}
// IterateAngularMotor (AngularMotor.cs:177-207), IterateLinearMotor (LinearMotor.cs:203-232)
template <class R> void iterate_motor(Constraint<R>& c, ConstraintBody<R>& A, ConstraintBody<R>& B, bool linear) {
    V3<R> j1 = rotate(vecp(c.p), A.q);
    V3<R> j2 = rotate(vecp(c.p + 3), B.q);
    R jv = linear ? dot(-j1, *A.vel) + dot(j2, *B.vel) : dot(-j1, *A.angvel) + dot(j2, *B.angvel);
    R lambda = -(jv - c.p[6]) * c.em_s;
    R old = c.s[0];
    R acc = clamp_cs(old + lambda, -c.bias_s, c.bias_s);
    c.s[0] = acc;
    lambda = acc - old;
    if (linear) {
        if (A.dynamic) *A.vel = *A.vel - cmul(j1 * lambda, A.imv);
        if (B.dynamic) *B.vel = *B.vel + cmul(j2 * lambda, B.imv);
    } else {
        if (A.dynamic) *A.angvel = *A.angvel - sym_mul(j1 * lambda, A.iw);
        if (B.dynamic) *B.angvel = *B.angvel + sym_mul(j2 * lambda, B.iw);
    }
}

// Linear rows of the slider constraints: for direction d, Jacobian (-d, -((r1 + u) x d), d, r2 x d).
template <class R> struct LinRow {
    V3<R> j0, j1, j2, j3;
};
template <class R> inline LinRow<R> lin_row(V3<R> r1, V3<R> r2, V3<R> u, V3<R> d) {
    return {-d, -cross(r1 + u, d), d, cross(r2, d)};
}
// JVector.Multiply(a, imv1) * b + JVector.Multiply(c, imv2) * e: the linear part of a mass entry
template <class R> inline R lin_mass(V3<R> a, V3<R> imv1, V3<R> b, V3<R> c, V3<R> imv2, V3<R> e) {
    return dot(cmul(a, imv1), b) + dot(cmul(c, imv2), e);
}
template <class R> inline R ang_mass(V3<R> a, Sym<R> i1, V3<R> b, V3<R> c, Sym<R> i2, V3<R> e) {
    return dot(sym_mul(a, i1), b) + dot(sym_mul(c, i2), e);
}

// PointOnPlane (PointOnPlane.cs): params LocalAxis (0-2, body-1 frame), LocalAnchor1 (3-5), LocalAnchor2
// (6-8), BiasFactor (9), Softness (10), Min (11), Max (12); state AccumulatedImpulse (0).
// PrepareForIterationPointOnPlane (:118-192)
template <class R> void prepare_point_on_plane(Constraint<R>& c, ConstraintBody<R>& A, ConstraintBody<R>& B, R inv_dt,
                                               R inv_sub) {
    V3<R> axis = rotate(vecp(c.p), A.q);
    V3<R> r1 = rotate(vecp(c.p + 3), A.q);
    V3<R> r2 = rotate(vecp(c.p + 6), B.q);
    V3<R> p1 = A.pos + r1;
    V3<R> p2 = B.pos + r2;
    c.clamp = 0;
    V3<R> u = p2 - p1;
    c.j0 = -axis;
    c.j1 = -cross(r1 + u, axis);
    c.j2 = axis;
    c.j3 = cross(r2, axis);
    R error = dot(u, axis);
    c.em_s = R(1.0);
    R mn = c.p[11], mx = c.p[12];
    if (mn >= mx) {
        error -= R(0.5) * mn + R(0.5) * mx;
        c.clamp = 3;
    } else if (error >= mx) {
        error -= mx;
        c.clamp = 1;
    } else if (error <= mn) {
        error -= mn;
        c.clamp = 2;
    } else {
        c.s[0] = R(0);
        return;
    }
    c.em_s = dot(cmul(c.j0, A.imv), c.j0) + dot(cmul(c.j2, B.imv), c.j2) + dot(sym_mul(c.j1, A.iw), c.j1) +
             dot(sym_mul(c.j3, B.iw), c.j3);
    c.em_s += (c.p[10] * inv_sub);
    c.em_s = c.em_s > R(0) ? R(1.0) / c.em_s : R(0);
    if (c.em_s == R(0)) c.s[0] = R(0);
    c.bias_s = error * c.p[9] * inv_dt;
    R acc = c.s[0];
    if (A.dynamic) {
        *A.vel = *A.vel + cmul(c.j0 * acc, A.imv);
        *A.angvel = *A.angvel + sym_mul(c.j1 * acc, A.iw);
    }
    if (B.dynamic) {
        *B.vel = *B.vel + cmul(c.j2 * acc, B.imv);
        *B.angvel = *B.angvel + sym_mul(c.j3 * acc, B.iw);
    }
}
// IteratePointOnPlane (:231-274)
template <class R> void iterate_point_on_plane(Constraint<R>& c, ConstraintBody<R>& A, ConstraintBody<R>& B,
                                               R inv_sub) {
    if (c.clamp == 0) return;
    R jv = dot(c.j0, *A.vel) + dot(c.j1, *A.angvel) + dot(c.j2, *B.vel) + dot(c.j3, *B.angvel);
    R soft = c.s[0] * c.p[10] * inv_sub;
    R lambda = -R(1.0) * (jv + c.bias_s + soft) * c.em_s;
    R orig = c.s[0];
    R acc = orig + lambda;
    if (c.clamp == 1) acc = std::min(acc, R(0.0));
    else if (c.clamp == 2) acc = std::max(acc, R(0.0));
    c.s[0] = acc;
    lambda = acc - orig;
    if (A.dynamic) {
        *A.vel = *A.vel + cmul(c.j0 * lambda, A.imv);
        *A.angvel = *A.angvel + sym_mul(c.j1 * lambda, A.iw);
    }
    if (B.dynamic) {
        *B.vel = *B.vel + cmul(c.j2 * lambda, B.imv);
        *B.angvel = *B.angvel + sym_mul(c.j3 * lambda, B.iw);
    }
}

// PointOnLine (PointOnLine.cs): params LocalAxis (0-2, body-1 frame), LocalAnchor1 (3-5), LocalAnchor2
// (6-8), BiasFactor (9), LimitBias (10), Softness (11), LimitSoftness (12), Min (13), Max (14); state
// AccumulatedImpulse (0-2).  The three rows (two across the line, one along it) are rebuilt from the
// bodies' poses in prepare and in every iteration, as the reference does.
template <class R> void point_on_line_rows(const Constraint<R>& c, const ConstraintBody<R>& A,
                                           const ConstraintBody<R>& B, LinRow<R>* rows, V3<R>& err) {
    V3<R> aw = rotate(vecp(c.p), A.q);
    V3<R> n1 = create_orthonormal(vecp(c.p));
    n1 = rotate(n1, A.q);
    V3<R> n2 = cross(aw, n1);
    V3<R> r1 = rotate(vecp(c.p + 3), A.q);
    V3<R> r2 = rotate(vecp(c.p + 6), B.q);
    V3<R> p1 = A.pos + r1;
    V3<R> p2 = B.pos + r2;
    V3<R> u = p2 - p1;
    rows[0] = lin_row(r1, r2, u, n1);
    rows[1] = lin_row(r1, r2, u, n2);
    rows[2] = lin_row(r1, r2, u, aw);
    err = V3<R>{dot(u, n1), dot(u, n2), dot(u, aw)};
}
// PrepareForIterationPointOnLine (:149-300)
template <class R> void prepare_point_on_line(Constraint<R>& c, ConstraintBody<R>& A, ConstraintBody<R>& B, R inv_dt,
                                              R inv_sub) {
    LinRow<R> J[3];
    V3<R> error;
    point_on_line_rows(c, A, B, J, error);
    c.clamp = 0;
    c.em = midentity<R>();
    c.em.m11 = lin_mass(J[0].j0, A.imv, J[0].j0, J[0].j2, B.imv, J[0].j2) + dot(sym_mul(J[0].j1, A.iw), J[0].j1) +
               dot(sym_mul(J[0].j3, B.iw), J[0].j3);
    c.em.m22 = lin_mass(J[1].j0, A.imv, J[1].j0, J[1].j2, B.imv, J[1].j2) + dot(sym_mul(J[1].j1, A.iw), J[1].j1) +
               dot(sym_mul(J[1].j3, B.iw), J[1].j3);
    c.em.m12 = ang_mass(J[0].j1, A.iw, J[1].j1, J[0].j3, B.iw, J[1].j3);
    c.em.m21 = ang_mass(J[1].j1, A.iw, J[0].j1, J[1].j3, B.iw, J[0].j3);
    R coupling12 = lin_mass(J[0].j0, A.imv, J[1].j0, J[0].j2, B.imv, J[1].j2);
    c.em.m12 += coupling12;
    c.em.m21 += coupling12;
    R mn = c.p[13], mx = c.p[14];
    if (mn >= mx) {
        error.z -= R(0.5) * mn + R(0.5) * mx;
        c.clamp = 3;
    } else if (error.z >= mx) {
        error.z -= mx;
        c.clamp = 1;
    } else if (error.z <= mn) {
        error.z -= mn;
        c.clamp = 2;
    } else {
        c.s[2] = R(0);
    }
    if (c.clamp != 0) {
        c.em.m33 = lin_mass(J[2].j0, A.imv, J[2].j0, J[2].j2, B.imv, J[2].j2) + dot(sym_mul(J[2].j1, A.iw), J[2].j1) +
                   dot(sym_mul(J[2].j3, B.iw), J[2].j3);
        c.em.m13 = ang_mass(J[0].j1, A.iw, J[2].j1, J[0].j3, B.iw, J[2].j3);
        c.em.m31 = ang_mass(J[2].j1, A.iw, J[0].j1, J[2].j3, B.iw, J[0].j3);
        c.em.m23 = ang_mass(J[1].j1, A.iw, J[2].j1, J[1].j3, B.iw, J[2].j3);
        c.em.m32 = ang_mass(J[2].j1, A.iw, J[1].j1, J[2].j3, B.iw, J[1].j3);
        R coupling13 = lin_mass(J[0].j0, A.imv, J[2].j0, J[0].j2, B.imv, J[2].j2);
        R coupling23 = lin_mass(J[1].j0, A.imv, J[2].j0, J[1].j2, B.imv, J[2].j2);
        c.em.m13 += coupling13;
        c.em.m31 += coupling13;
        c.em.m23 += coupling23;
        c.em.m32 += coupling23;
    }
    c.em.m11 += c.p[11] * inv_sub;
    c.em.m22 += c.p[11] * inv_sub;
    c.em.m33 += c.p[12] * inv_sub;
    if (c.clamp == 1 || c.clamp == 2) {
        c.bilateral = inverse_bilateral_block(c.em);
        c.cx = c.em.m13;
        c.cy = c.em.m23;
    }
    if (c.clamp == 0) c.em.m33 = R(0);
    c.em = inverse_symmetric(c.em);
    c.bias = error * inv_dt;
    c.bias.x *= c.p[9];
    c.bias.y *= c.p[9];
    c.bias.z *= c.p[10];
    if (c.clamp == 1) c.s[2] = std::min(R(0), c.s[2]);
    else if (c.clamp == 2) c.s[2] = std::max(R(0), c.s[2]);
    V3<R> acc = vecp(c.s);
    if (A.dynamic) {
        *A.vel = *A.vel + cmul(J[0].j0 * acc.x + J[1].j0 * acc.y + J[2].j0 * acc.z, A.imv);
        *A.angvel = *A.angvel + sym_mul(J[0].j1 * acc.x + J[1].j1 * acc.y + J[2].j1 * acc.z, A.iw);
    }
    if (B.dynamic) {
        *B.vel = *B.vel + cmul(J[0].j2 * acc.x + J[1].j2 * acc.y + J[2].j2 * acc.z, B.imv);
        *B.angvel = *B.angvel + sym_mul(J[0].j3 * acc.x + J[1].j3 * acc.y + J[2].j3 * acc.z, B.iw);
    }
}
// IteratePointOnLine (:372-465)
template <class R> void iterate_point_on_line(Constraint<R>& c, ConstraintBody<R>& A, ConstraintBody<R>& B,
                                              R inv_sub) {
    LinRow<R> J[3];
    V3<R> unused;
    point_on_line_rows(c, A, B, J, unused);
    V3<R> jv;
    jv.x = dot(J[0].j0, *A.vel) + dot(J[0].j1, *A.angvel) + dot(J[0].j2, *B.vel) + dot(J[0].j3, *B.angvel);
    jv.y = dot(J[1].j0, *A.vel) + dot(J[1].j1, *A.angvel) + dot(J[1].j2, *B.vel) + dot(J[1].j3, *B.angvel);
    jv.z = dot(J[2].j0, *A.vel) + dot(J[2].j1, *A.angvel) + dot(J[2].j2, *B.vel) + dot(J[2].j3, *B.angvel);
    V3<R> acc = vecp(c.s);
    V3<R> soft = acc * inv_sub;
    soft.x *= c.p[11];
    soft.y *= c.p[11];
    soft.z *= c.p[12];
    V3<R> residual = jv + c.bias + soft;
    V3<R> lambda = -R(1.0) * mat_mul(residual, c.em);
    V3<R> orig = acc;
    acc = acc + lambda;
    if ((c.clamp == 1 && acc.z > R(0)) || (c.clamp == 2 && acc.z < R(0))) {
        lambda.z = -orig.z;
        R x = residual.x + c.cx * lambda.z;
        R y = residual.y + c.cy * lambda.z;
        lambda.x = -(c.bilateral.x * x + c.bilateral.y * y);
        lambda.y = -(c.bilateral.y * x + c.bilateral.z * y);
        acc.x = orig.x + lambda.x;
        acc.y = orig.y + lambda.y;
        acc.z = R(0);
    } else if (c.clamp == 0) {
        acc.z = R(0.0);
        orig.z = R(0.0);
    }
    setv(c.s, acc);
    lambda = acc - orig;
    if (A.dynamic) {
        *A.vel = *A.vel + cmul(J[0].j0 * lambda.x + J[1].j0 * lambda.y + J[2].j0 * lambda.z, A.imv);
        *A.angvel = *A.angvel + sym_mul(J[0].j1 * lambda.x + J[1].j1 * lambda.y + J[2].j1 * lambda.z, A.iw);
    }
    if (B.dynamic) {
        *B.vel = *B.vel + cmul(J[0].j2 * lambda.x + J[1].j2 * lambda.y + J[2].j2 * lambda.z, B.imv);
        *B.angvel = *B.angvel + sym_mul(J[0].j3 * lambda.x + J[1].j3 * lambda.y + J[2].j3 * lambda.z, B.iw);
    }
}

// ---- Servos (Dynamics/Constraints/Internal/ServoMath.cs) ----------------------------------------
// MaximumImpulse (:30): the largest impulse a force or torque limit allows in one substep
template <class R> inline R servo_max_impulse(R maximum, R dt) { return dt > R(0) ? maximum * dt : R(0); }
// ClampMagnitude (:32-38)
template <class R> inline V3<R> servo_clamp_magnitude(V3<R> impulse, R maximum) {
    if (std::isinf(maximum) && maximum > R(0)) return impulse;
    R l2 = len_sq(impulse);
    if (l2 > maximum * maximum) impulse = impulse * (maximum / std::sqrt(l2));
    return impulse;
}
// ClampAngular (:40-48): limits the angular impulse J^T lambda a body receives
template <class R> inline V3<R> servo_clamp_angular(V3<R> impulse, const M3<R>& jac, R maximum) {
    if (std::isinf(maximum) && maximum > R(0)) return impulse;
    R l2 = len_sq(tmat_mul(impulse, jac));
    if (l2 > maximum * maximum) impulse = impulse * (maximum / std::sqrt(l2));
    return impulse;
}

// AngularServo (AngularServo.cs): params TargetRelativeOrientation (0-3, x y z w), MaximumTorque (4),
// Softness (5), BiasFactor (6); state AccumulatedImpulse (0-2), Hemisphere (3).  The substep's maximum
// impulse is kept in cx.
template <class R> inline void angular_servo_apply(Constraint<R>& c, ConstraintBody<R>& A, ConstraintBody<R>& B,
                                                   V3<R> impulse) {  // ApplyImpulse (:164-172)
    if (A.dynamic) *A.angvel = *A.angvel + sym_mul(impulse, A.iw);
    if (B.dynamic) *B.angvel = *B.angvel - sym_mul(impulse, B.iw);
}
// PrepareForIterationAngularServo (:125-149)
template <class R> void prepare_angular_servo(Constraint<R>& c, ConstraintBody<R>& A, ConstraintBody<R>& B, R inv_dt,
                                              R inv_sub, R sub_dt) {
    Q4<R> target{c.p[0], c.p[1], c.p[2], c.p[3]};
    V3<R> error;
    int hemisphere;
    c.jac = mscale(calc_jacobian(qmul(qconj(target), qconj(A.q)), B.q, error, hemisphere), -R(2));
    int old_h = int(c.s[3]);
    if (old_h != 0 && old_h != hemisphere) c.s[0] = c.s[1] = c.s[2] = R(0);
    c.s[3] = R(hemisphere);
    c.em = sym_to_matrix(sym_transform(sym_add(A.iw, B.iw), c.jac));
    R softness = c.p[5] * inv_sub;
    c.em.m11 += softness;
    c.em.m22 += softness;
    c.em.m33 += softness;
    c.em = inverse_symmetric(c.em);
    c.bias = -error * c.p[6] * inv_dt;
    c.cx = servo_max_impulse(c.p[4], sub_dt);
    V3<R> acc = servo_clamp_angular(vecp(c.s), c.jac, c.cx);
    setv(c.s, acc);
    angular_servo_apply(c, A, B, tmat_mul(acc, c.jac));
}
// IterateAngularServo (:151-162)
template <class R> void iterate_angular_servo(Constraint<R>& c, ConstraintBody<R>& A, ConstraintBody<R>& B,
                                              R inv_sub) {
    V3<R> jv = mat_mul(*A.angvel - *B.angvel, c.jac);
    V3<R> previous = vecp(c.s);
    V3<R> softness = previous * (c.p[5] * inv_sub);
    V3<R> lambda = -mat_mul(jv + c.bias + softness, c.em);
    V3<R> acc = servo_clamp_angular(previous + lambda, c.jac, c.cx);
    setv(c.s, acc);
    angular_servo_apply(c, A, B, tmat_mul(acc - previous, c.jac));
}

// LinearServo (LinearServo.cs): params LocalAnchor1 (0-2), LocalAnchor2 (3-5), TargetPosition (6-8, body-1
// frame), MaximumForce (9), Softness (10), BiasFactor (11); state AccumulatedImpulse (0-2).  The substep's
// maximum impulse is kept in cx.
template <class R> inline void linear_servo_apply(Constraint<R>& c, ConstraintBody<R>& A, ConstraintBody<R>& B,
                                                  V3<R> impulse) {  // ApplyImpulse (:189-203)
    if (A.dynamic) {
        *A.vel = *A.vel - cmul(impulse, A.imv);
        *A.angvel = *A.angvel - sym_mul(cross(c.r1, impulse), A.iw);
    }
    if (B.dynamic) {
        *B.vel = *B.vel + cmul(impulse, B.imv);
        *B.angvel = *B.angvel + sym_mul(cross(c.r2, impulse), B.iw);
    }
}
// PrepareForIterationLinearServo (:149-171)
template <class R> void prepare_linear_servo(Constraint<R>& c, ConstraintBody<R>& A, ConstraintBody<R>& B, R inv_dt,
                                             R inv_sub, R sub_dt) {
    c.r1 = rotate(vecp(c.p) + vecp(c.p + 6), A.q);
    c.r2 = rotate(vecp(c.p + 3), B.q);
    M3<R> cr1 = cross_matrix(c.r1), cr2 = cross_matrix(c.r2);
    Sym<R> s = sym_add(sym_add(sym_add(sym_scale_vec(A.imv), sym_transform(A.iw, cr1)), sym_scale_vec(B.imv)),
                       sym_transform(B.iw, cr2));
    c.em = sym_to_matrix(s);
    R softness = c.p[10] * inv_sub;
    c.em.m11 += softness;
    c.em.m22 += softness;
    c.em.m33 += softness;
    c.em = inverse_symmetric(c.em);
    c.bias = (B.pos + c.r2 - A.pos - c.r1) * (c.p[11] * inv_dt);
    c.cx = servo_max_impulse(c.p[9], sub_dt);
    V3<R> acc = servo_clamp_magnitude(vecp(c.s), c.cx);
    setv(c.s, acc);
    linear_servo_apply(c, A, B, acc);
}
// IterateLinearServo (:173-187)
template <class R> void iterate_linear_servo(Constraint<R>& c, ConstraintBody<R>& A, ConstraintBody<R>& B,
                                             R inv_sub) {
    V3<R> jv = -*A.vel - cross(*A.angvel, c.r1) + *B.vel + cross(*B.angvel, c.r2);
    V3<R> previous = vecp(c.s);
    V3<R> softness = previous * (c.p[10] * inv_sub);
    V3<R> lambda = -mat_mul(jv + c.bias + softness, c.em);
    V3<R> acc = servo_clamp_magnitude(previous + lambda, c.cx);
    setv(c.s, acc);
    linear_servo_apply(c, A, B, acc - previous);
}

// LinearAxisServo (LinearAxisServo.cs): params LocalAxis (0-2, body-1 frame), LocalAnchor1 (3-5),
// LocalAnchor2 (6-8), TargetPosition (9), MaximumForce (10), Softness (11), BiasFactor (12); state
// AccumulatedImpulse (0).  The substep's maximum impulse is kept in cx.
template <class R> inline void linear_axis_servo_apply(Constraint<R>& c, ConstraintBody<R>& A, ConstraintBody<R>& B,
                                                       R impulse) {  // ApplyImpulse (:211-226)
    if (A.dynamic) {
        *A.vel = *A.vel + cmul(c.j0 * impulse, A.imv);
        *A.angvel = *A.angvel + sym_mul(c.j1 * impulse, A.iw);
    }
    if (B.dynamic) {
        *B.vel = *B.vel + cmul(c.j2 * impulse, B.imv);
        *B.angvel = *B.angvel + sym_mul(c.j3 * impulse, B.iw);
    }
}
// PrepareForIterationLinearAxisServo (:169-194)
template <class R> void prepare_linear_axis_servo(Constraint<R>& c, ConstraintBody<R>& A, ConstraintBody<R>& B,
                                                  R inv_dt, R inv_sub, R sub_dt) {
    V3<R> axis = rotate(vecp(c.p), A.q);
    V3<R> r1 = rotate(vecp(c.p + 3), A.q);
    V3<R> r2 = rotate(vecp(c.p + 6), B.q);
    V3<R> separation = B.pos + r2 - A.pos - r1;
    c.j0 = -axis;
    c.j1 = -cross(r1 + separation, axis);
    c.j2 = axis;
    c.j3 = cross(r2, axis);
    R mass = dot(cmul(c.j0, A.imv), c.j0) + dot(cmul(c.j2, B.imv), c.j2) + dot(sym_mul(c.j1, A.iw), c.j1) +
             dot(sym_mul(c.j3, B.iw), c.j3) + c.p[11] * inv_sub;
    c.em_s = mass > R(0) ? R(1) / mass : R(0);
    if (c.em_s == R(0)) c.s[0] = R(0);
    c.bias_s = (dot(separation, axis) - c.p[9]) * c.p[12] * inv_dt;
    c.cx = servo_max_impulse(c.p[10], sub_dt);
    c.s[0] = clamp_cs(c.s[0], -c.cx, c.cx);
    linear_axis_servo_apply(c, A, B, c.s[0]);
}
// IterateLinearAxisServo (:196-209)
template <class R> void iterate_linear_axis_servo(Constraint<R>& c, ConstraintBody<R>& A, ConstraintBody<R>& B,
                                                  R inv_sub) {
    R jv = dot(c.j0, *A.vel) + dot(c.j1, *A.angvel) + dot(c.j2, *B.vel) + dot(c.j3, *B.angvel);
    R lambda = -(jv + c.bias_s + c.s[0] * c.p[11] * inv_sub) * c.em_s;
    R previous = c.s[0];
    c.s[0] = clamp_cs(previous + lambda, -c.cx, c.cx);
    linear_axis_servo_apply(c, A, B, c.s[0] - previous);
}

// TwistServo (TwistServo.cs): params LocalAxis2 (0-2, body-2 frame), ReferenceOrientation (3-6, x y z w),
// TargetAngle (7), MaximumTorque (8), Softness (9), BiasFactor (10); state AccumulatedImpulse (0),
// Hemisphere (1).  The substep's maximum impulse is kept in cx.
template <class R> inline void twist_servo_apply(Constraint<R>& c, ConstraintBody<R>& A, ConstraintBody<R>& B,
                                                 R impulse) {  // ApplyImpulse (:184-192)
    if (A.dynamic) *A.angvel = *A.angvel + sym_mul(c.j0 * impulse, A.iw);
    if (B.dynamic) *B.angvel = *B.angvel - sym_mul(c.j0 * impulse, B.iw);
}
// PrepareForIterationTwistServo (:150-170)
template <class R> void prepare_twist_servo(Constraint<R>& c, ConstraintBody<R>& A, ConstraintBody<R>& B, R inv_dt,
                                            R inv_sub, R sub_dt) {
    Q4<R> ref{c.p[3], c.p[4], c.p[5], c.p[6]};
    int hemisphere;
    R angle = calc_twist(qmul(ref, qconj(A.q)), B.q, vecp(c.p), c.j0, hemisphere);
    int old_h = int(c.s[1]);
    if (old_h != 0 && old_h != hemisphere) c.s[0] = R(0);
    c.s[1] = R(hemisphere);
    R mass = dot(sym_mul(c.j0, sym_add(A.iw, B.iw)), c.j0) + c.p[9] * inv_sub;
    c.em_s = mass > R(0) ? R(1) / mass : R(0);
    if (c.em_s == R(0)) c.s[0] = R(0);
    R error = angle - c.p[7];
    c.bias_s = error * c.p[10] * inv_dt;
    c.cx = servo_max_impulse(c.p[8], sub_dt);
    c.s[0] = clamp_cs(c.s[0], -c.cx, c.cx);
    twist_servo_apply(c, A, B, c.s[0]);
}
// IterateTwistServo (:172-182)
template <class R> void iterate_twist_servo(Constraint<R>& c, ConstraintBody<R>& A, ConstraintBody<R>& B, R inv_sub) {
    R jv = dot(*A.angvel - *B.angvel, c.j0);
    R lambda = -c.em_s * (jv + c.bias_s + c.s[0] * c.p[9] * inv_sub);
    R previous = c.s[0];
    c.s[0] = clamp_cs(previous + lambda, -c.cx, c.cx);
    twist_servo_apply(c, A, B, c.s[0] - previous);
}

// EllipticalConeLimit (EllipticalConeLimit.cs): params LocalAxis1 (0-2), LocalAxis2 (3-5), LocalReference1
// (6-8), MinAngle1 (9), MaxAngle1 (10), MinAngle2 (11), MaxAngle2 (12), Softness (13), BiasFactor (14);
// state AccumulatedImpulse (0), Jacobian (1-3), Clamp (4) — the last two persist because the next
// prepare compares against them.
template <class R> inline R min_nan(R a, R b) {  // MathR.Min
    if (a != a) return a;
    if (b != b) return b;
    return a < b ? a : b;
}
// EllipseRadius (:258-268): radius of the ellipse with semiaxes a, b at the azimuth (cosine, sine) and
// the logarithmic azimuth derivative of that radius
template <class R> inline R ellipse_radius(R a, R b, R cosine, R sine, R& derivative) {
    R scale = min_nan(a, b);
    R u = scale / a;
    R v = scale / b;
    R squared = cosine * cosine * u * u + sine * sine * v * v;
    derivative = cosine * sine * (v * v - u * u) / squared;
    return scale / std::sqrt(squared);
}
// PrepareForIterationEllipticalConeLimit (:270-389)
template <class R> void prepare_elliptical_cone(Constraint<R>& c, ConstraintBody<R>& A, ConstraintBody<R>& B,
                                                R inv_dt, R inv_sub) {
    const R* p = c.p;
    V3<R> axis = rotate(vecp(p), A.q);
    V3<R> reference = rotate(vecp(p + 6), A.q);
    V3<R> other = cross(axis, reference);
    V3<R> target = rotate(vecp(p + 3), B.q);
    R x = dot(target, reference);
    R y = dot(target, other);
    R z = dot(target, axis);
    R sine = std::sqrt(x * x + y * y);
    R angle = StableMath<R>::atan2(sine, z);
    R ca, sa;
    if (sine > R(0)) {
        ca = x / sine;
        sa = y / sine;
    } else {
        Sym<R> inertia = sym_add(A.iw, B.iw);
        bool use_reference = dot(sym_mul(other, inertia), other) >= dot(sym_mul(reference, inertia), reference);
        ca = use_reference ? R(1) : R(0);
        sa = use_reference ? R(0) : R(1);
    }
    R derivative;
    R outer = ellipse_radius(p[10], p[12], ca, sa, derivative);
    bool fixed_cone = p[9] == p[10] && p[11] == p[12];
    R boundary = outer;
    int clamp = 0;
    if (fixed_cone) {
        clamp = 3;
    } else if (angle >= outer) {
        clamp = 1;
    } else if (p[9] > R(0)) {
        R inner_derivative;
        R inner = ellipse_radius(p[9], p[11], ca, sa, inner_derivative);
        if (angle <= inner) {
            clamp = 2;
            boundary = inner;
            derivative = inner_derivative;
        }
    }
    int previous = int(c.s[4]);
    c.s[4] = R(clamp);
    c.clamp = clamp;
    if (clamp == 0) {
        c.s[0] = R(0);
        return;
    }
    V3<R> tilt = reference * ca + other * sa;
    V3<R> azimuth_dir = cross(axis, tilt);
    V3<R> angle_dir = tilt * z - axis * sine;
    R azimuth_factor = sine > R(0) ? angle / sine * derivative : R(0);
    R scale = max_nan(R(1), std::fabs(azimuth_factor));
    V3<R> jac = -azimuth_dir * (R(1) / scale) + angle_dir * (azimuth_factor / scale);
    R inv_len = R(1) / length(jac);
    jac = jac * inv_len;
    inv_len /= scale;
    if (previous != clamp || dot(vecp(c.s + 1), jac) < R(0)) c.s[0] = R(0);
    setv(c.s + 1, jac);
    if (clamp == 1) c.s[0] = min_nan(c.s[0], R(0));
    else if (clamp == 2) c.s[0] = max_nan(c.s[0], R(0));
    R mass = dot(sym_mul(jac, A.iw), jac) + dot(sym_mul(jac, B.iw), jac) + p[13] * inv_sub;
    c.em_s = mass > R(0) ? R(1) / mass : R(0);
    if (c.em_s == R(0)) c.s[0] = R(0);
    c.bias_s = (angle - boundary) * inv_len * p[14] * inv_dt;
    R acc = c.s[0];
    if (A.dynamic) *A.angvel = *A.angvel + sym_mul(jac * acc, A.iw);
    if (B.dynamic) *B.angvel = *B.angvel - sym_mul(jac * acc, B.iw);
}
// IterateEllipticalConeLimit (:391-425)
template <class R> void iterate_elliptical_cone(Constraint<R>& c, ConstraintBody<R>& A, ConstraintBody<R>& B,
                                                R inv_sub) {
    if (c.clamp == 0) return;
    V3<R> jac = vecp(c.s + 1);
    R jv = dot(*A.angvel - *B.angvel, jac);
    R lambda = -c.em_s * (jv + c.bias_s + c.s[0] * c.p[13] * inv_sub);
    R old = c.s[0];
    R acc = old + lambda;
    if (c.clamp == 1) acc = min_nan(acc, R(0));
    else if (c.clamp == 2) acc = max_nan(acc, R(0));
    c.s[0] = acc;
    lambda = acc - old;
    if (A.dynamic) *A.angvel = *A.angvel + sym_mul(jac * lambda, A.iw);
    if (B.dynamic) *B.angvel = *B.angvel - sym_mul(jac * lambda, B.iw);
}

template <class R> void constraint_prepare(Constraint<R>& c, ConstraintBody<R>& A, ConstraintBody<R>& B, R inv_dt,
                                           R inv_sub, R sub_dt) {
    switch (c.type) {
        case JH_CONSTRAINT_BALL_SOCKET: prepare_ball_socket(c, A, B, inv_dt, inv_sub); break;
        case JH_CONSTRAINT_HINGE_ANGLE: prepare_hinge_angle(c, A, B, inv_dt, inv_sub); break;
        case JH_CONSTRAINT_FIXED_ANGLE: prepare_fixed_angle(c, A, B, inv_dt, inv_sub); break;
        case JH_CONSTRAINT_DISTANCE_LIMIT: prepare_distance_limit(c, A, B, inv_dt, inv_sub); break;
        case JH_CONSTRAINT_TWIST_ANGLE: prepare_twist_angle(c, A, B, inv_dt, inv_sub); break;
        case JH_CONSTRAINT_CONE_LIMIT: prepare_cone_limit(c, A, B, inv_dt, inv_sub); break;
        case JH_CONSTRAINT_ANGULAR_MOTOR: prepare_motor(c, A, B, sub_dt, false); break;
        case JH_CONSTRAINT_LINEAR_MOTOR: prepare_motor(c, A, B, sub_dt, true); break;
        case JH_CONSTRAINT_POINT_ON_LINE: prepare_point_on_line(c, A, B, inv_dt, inv_sub); break;
        case JH_CONSTRAINT_POINT_ON_PLANE: prepare_point_on_plane(c, A, B, inv_dt, inv_sub); break;
        case JH_CONSTRAINT_ANGULAR_SERVO: prepare_angular_servo(c, A, B, inv_dt, inv_sub, sub_dt); break;
        case JH_CONSTRAINT_LINEAR_SERVO: prepare_linear_servo(c, A, B, inv_dt, inv_sub, sub_dt); break;
        case JH_CONSTRAINT_LINEAR_AXIS_SERVO: prepare_linear_axis_servo(c, A, B, inv_dt, inv_sub, sub_dt); break;
        case JH_CONSTRAINT_TWIST_SERVO: prepare_twist_servo(c, A, B, inv_dt, inv_sub, sub_dt); break;
        default: prepare_elliptical_cone(c, A, B, inv_dt, inv_sub); break;
    }
}
template <class R> void constraint_iterate(Constraint<R>& c, ConstraintBody<R>& A, ConstraintBody<R>& B, R inv_sub) {
    switch (c.type) {
        case JH_CONSTRAINT_BALL_SOCKET: iterate_ball_socket(c, A, B, inv_sub); break;
        case JH_CONSTRAINT_HINGE_ANGLE: iterate_hinge_angle(c, A, B, inv_sub); break;
        case JH_CONSTRAINT_FIXED_ANGLE: iterate_fixed_angle(c, A, B, inv_sub); break;
        case JH_CONSTRAINT_DISTANCE_LIMIT: iterate_distance_limit(c, A, B, inv_sub); break;
        case JH_CONSTRAINT_TWIST_ANGLE: iterate_twist_angle(c, A, B, inv_sub); break;
        case JH_CONSTRAINT_CONE_LIMIT: iterate_cone_limit(c, A, B, inv_sub); break;
        case JH_CONSTRAINT_ANGULAR_MOTOR: iterate_motor(c, A, B, false); break;
        case JH_CONSTRAINT_LINEAR_MOTOR: iterate_motor(c, A, B, true); break;
        case JH_CONSTRAINT_POINT_ON_LINE: iterate_point_on_line(c, A, B, inv_sub); break;
        case JH_CONSTRAINT_POINT_ON_PLANE: iterate_point_on_plane(c, A, B, inv_sub); break;
        case JH_CONSTRAINT_ANGULAR_SERVO: iterate_angular_servo(c, A, B, inv_sub); break;
        case JH_CONSTRAINT_LINEAR_SERVO: iterate_linear_servo(c, A, B, inv_sub); break;
        case JH_CONSTRAINT_LINEAR_AXIS_SERVO: iterate_linear_axis_servo(c, A, B, inv_sub); break;
        case JH_CONSTRAINT_TWIST_SERVO: iterate_twist_servo(c, A, B, inv_sub); break;
        default: iterate_elliptical_cone(c, A, B, inv_sub); break;
    }
}

}  // namespace orc
