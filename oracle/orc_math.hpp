// ORACLE — test infrastructure only.  CPU restatement of Jitter2's linear algebra and of its
// deterministic trigonometry, templated on the precision R (float ≙ Jitter2, double ≙
// Jitter2.Double, src/Jitter2/Precision.cs:14-28).  Only tests/, __graft_entry__.smoke() and
// bench.py's cpu_baseline leg may use this directory.  Every expression keeps the operation
// order of the cited C# source: the translation unit is compiled with -ffp-contract=off and
// without fast-math, so each line rounds exactly like RyuJIT's scalar SSE code.
#pragma once
#include <cmath>
#include <cstdint>
#include <cstring>
#include <limits>

namespace orc {

template <class R> struct V3 { R x, y, z; };
template <class R> struct Q4 { R x, y, z, w; };
// JSymmetricMatrix (LinearMath/JSymmetricMatrix.cs): M11 M12 M13 M22 M23 M33; M21=M12, M31=M13.
template <class R> struct Sym { R m11, m12, m13, m22, m23, m33; };
// JMatrix (LinearMath/JMatrix.cs), row-major.
template <class R> struct M3 { R m11, m12, m13, m21, m22, m23, m31, m32, m33; };

template <class R> inline V3<R> v3(R x, R y, R z) { return V3<R>{x, y, z}; }
template <class R> inline V3<R> operator+(V3<R> a, V3<R> b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
template <class R> inline V3<R> operator-(V3<R> a, V3<R> b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
template <class R> inline V3<R> operator-(V3<R> a) { return {-a.x, -a.y, -a.z}; }
// JVector * Real and Real * JVector (JVector.cs:696-720).
template <class R> inline V3<R> operator*(V3<R> a, R s) { return {a.x * s, a.y * s, a.z * s}; }
template <class R> inline V3<R> operator*(R s, V3<R> a) { return {s * a.x, s * a.y, s * a.z}; }
// JVector.Dot (JVector.cs:395-398): left-to-right sum.
template <class R> inline R dot(V3<R> a, V3<R> b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
// operator % (JVector.cs:674-681).
template <class R> inline V3<R> cross(V3<R> a, V3<R> b) {
    return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
// JVector.Multiply(a, b) component-wise (JVector.cs:651-654).
template <class R> inline V3<R> cmul(V3<R> a, V3<R> b) { return {a.x * b.x, a.y * b.y, a.z * b.z}; }
template <class R> inline R len_sq(V3<R> a) { return a.x * a.x + a.y * a.y + a.z * a.z; }
template <class R> inline R length(V3<R> a) { return std::sqrt(len_sq(a)); }
// JVector.Normalize (JVector.cs:600-607): 1/sqrt then multiply.
template <class R> inline V3<R> normalize(V3<R> v) {
    R n2 = v.x * v.x + v.y * v.y + v.z * v.z;
    R n = R(1.0) / std::sqrt(n2);
    return {v.x * n, v.y * n, v.z * n};
}
template <class R> inline bool veq(V3<R> a, V3<R> b) { return a.x == b.x && a.y == b.y && a.z == b.z; }

// JVector.Transform(v, q) (JVector.cs:326-339).
template <class R> inline V3<R> rotate(V3<R> v, Q4<R> q) {
    R n0 = R(2.0) * (q.y * v.z - q.z * v.y);
    R n1 = R(2.0) * (q.z * v.x - q.x * v.z);
    R n2 = R(2.0) * (q.x * v.y - q.y * v.x);
    R n00 = q.y * n2 - q.z * n1;
    R n11 = q.z * n0 - q.x * n2;
    R n22 = q.x * n1 - q.y * n0;
    return {v.x + q.w * n0 + n00, v.y + q.w * n1 + n11, v.z + q.w * n2 + n22};
}
// JVector.ConjugatedTransform(v, q) (JVector.cs:348-361).
template <class R> inline V3<R> rotate_inv(V3<R> v, Q4<R> q) {
    R n0 = R(2.0) * (q.z * v.y - q.y * v.z);
    R n1 = R(2.0) * (q.x * v.z - q.z * v.x);
    R n2 = R(2.0) * (q.y * v.x - q.x * v.y);
    R n00 = q.z * n1 - q.y * n2;
    R n11 = q.x * n2 - q.z * n0;
    R n22 = q.y * n0 - q.x * n1;
    return {v.x + q.w * n0 + n00, v.y + q.w * n1 + n11, v.z + q.w * n2 + n22};
}
// JQuaternion.Multiply (JQuaternion.cs:353-368).
template <class R> inline Q4<R> qmul(Q4<R> a, Q4<R> b) {
    R r1 = a.w, i1 = a.x, j1 = a.y, k1 = a.z;
    R r2 = b.w, i2 = b.x, j2 = b.y, k2 = b.z;
    Q4<R> o;
    o.w = r1 * r2 - (i1 * i2 + j1 * j2 + k1 * k2);
    o.x = r1 * i2 + r2 * i1 + j1 * k2 - k1 * j2;
    o.y = r1 * j2 + r2 * j1 + k1 * i2 - i1 * k2;
    o.z = r1 * k2 + r2 * k1 + i1 * j2 - j1 * i2;
    return o;
}
// JQuaternion.ConjugateMultiply (JQuaternion.cs:378-395): conj(a) * b.
template <class R> inline Q4<R> qconjmul(Q4<R> a, Q4<R> b) {
    Q4<R> c{-a.x, -a.y, -a.z, a.w};
    return qmul(c, b);
}
// JQuaternion.Normalize (JQuaternion.cs:521-529).
template <class R> inline Q4<R> qnormalize(Q4<R> q) {
    R n2 = q.x * q.x + q.y * q.y + q.z * q.z + q.w * q.w;
    R n = R(1.0) / std::sqrt(n2);
    return {q.x * n, q.y * n, q.z * n, q.w * n};
}
// JQuaternion.GetBasisY (JQuaternion.cs:233-240).
template <class R> inline V3<R> basis_y(Q4<R> q) {
    return {R(2.0) * (q.x * q.y - q.z * q.w), R(1.0) - R(2.0) * (q.x * q.x + q.z * q.z),
            R(2.0) * (q.y * q.z + q.x * q.w)};
}
// JMatrix.CreateFromQuaternion (JMatrix.cs:553-568).
template <class R> inline M3<R> mat_from_quat(Q4<R> q) {
    R r = q.w, i = q.x, j = q.y, k = q.z;
    M3<R> m;
    m.m11 = R(1.0) - R(2.0) * (j * j + k * k);
    m.m12 = R(2.0) * (i * j - k * r);
    m.m13 = R(2.0) * (i * k + j * r);
    m.m21 = R(2.0) * (i * j + k * r);
    m.m22 = R(1.0) - R(2.0) * (i * i + k * k);
    m.m23 = R(2.0) * (j * k - i * r);
    m.m31 = R(2.0) * (i * k - j * r);
    m.m32 = R(2.0) * (j * k + i * r);
    m.m33 = R(1.0) - R(2.0) * (i * i + j * j);
    return m;
}
// JMatrix.Absolute (JMatrix.cs:534-545).
template <class R> inline M3<R> mat_abs(M3<R> m) {
    return {std::fabs(m.m11), std::fabs(m.m12), std::fabs(m.m13), std::fabs(m.m21), std::fabs(m.m22),
            std::fabs(m.m23), std::fabs(m.m31), std::fabs(m.m32), std::fabs(m.m33)};
}
// JVector.Transform(v, JMatrix) (JVector.cs:290-298).
template <class R> inline V3<R> mat_mul(V3<R> v, M3<R> m) {
    return {v.x * m.m11 + v.y * m.m12 + v.z * m.m13, v.x * m.m21 + v.y * m.m22 + v.z * m.m23,
            v.x * m.m31 + v.y * m.m32 + v.z * m.m33};
}
// JVector.Transform(v, JSymmetricMatrix) (JVector.cs:233-241).
template <class R> inline V3<R> sym_mul(V3<R> v, Sym<R> m) {
    return {v.x * m.m11 + v.y * m.m12 + v.z * m.m13, v.x * m.m12 + v.y * m.m22 + v.z * m.m23,
            v.x * m.m13 + v.y * m.m23 + v.z * m.m33};
}
// Contact.TransformSymmetricInertia (Contact.cs:1027-1036): same sums written with M21/M31.
template <class R> inline V3<R> sym_mul_contact(V3<R> v, Sym<R> m) {
    V3<R> r;
    r.x = v.x * m.m11 + v.y * m.m12 + v.z * m.m13;
    r.y = v.x * m.m12 + v.y * m.m22 + v.z * m.m23;
    r.z = v.x * m.m13 + v.y * m.m23 + v.z * m.m33;
    return r;
}
// JSymmetricMatrix.Transform(M, T) = T M T^T (JSymmetricMatrix.cs:247-260).
template <class R> inline Sym<R> sym_transform(Sym<R> mi, M3<R> t) {
    V3<R> row1{t.m11, t.m12, t.m13}, row2{t.m21, t.m22, t.m23}, row3{t.m31, t.m32, t.m33};
    V3<R> p1 = sym_mul(row1, mi), p2 = sym_mul(row2, mi), p3 = sym_mul(row3, mi);
    return {dot(p1, row1), dot(p1, row2), dot(p1, row3), dot(p2, row2), dot(p2, row3), dot(p3, row3)};
}

// MathHelper.SignBit (MathHelper.cs:28-38): +1 for clear sign bit, -1 for set (incl. -0).
inline int signbit_i(float v) { int32_t b; std::memcpy(&b, &v, 4); return 1 | (b >> 31); }
inline int signbit_i(double v) { int64_t b; std::memcpy(&b, &v, 8); return 1 | int(b >> 63); }
// Math.Sign(float/double): -1, 0, +1.
template <class R> inline int sign_i(R v) { return v > R(0) ? 1 : (v < R(0) ? -1 : 0); }
// MathR.Max / MathR.Min for the non-NaN inputs of the hot path.
template <class R> inline R rmax(R a, R b) { return a > b ? a : b; }
template <class R> inline R rmin(R a, R b) { return a < b ? a : b; }
// Math.Clamp(v, lo, hi).
template <class R> inline R rclamp(R v, R lo, R hi) { return v < lo ? lo : (v > hi ? hi : v); }

// MathHelper.CreateOrthonormal (MathHelper.cs:204-255).
template <class R> inline V3<R> create_orthonormal(V3<R> v) {
    R ax = std::fabs(v.x), ay = std::fabs(v.y), az = std::fabs(v.z);
    V3<R> r;
    if (ax <= ay && ax <= az) {
        R y = v.z, z = -v.y;
        R inv = R(1.0) / std::sqrt(y * y + z * z);
        r.x = 0; r.y = y * inv; r.z = z * inv;
    } else if (ay <= az) {
        R x = -v.z, z = v.x;
        R inv = R(1.0) / std::sqrt(x * x + z * z);
        r.x = x * inv; r.y = 0; r.z = z * inv;
    } else {
        R x = v.y, y = -v.x;
        R inv = R(1.0) / std::sqrt(x * x + y * y);
        r.x = x * inv; r.y = y * inv; r.z = 0;
    }
    return r;
}

// ---- StableMath (LinearMath/StableMath.cs) -------------------------------------------------
template <class R> struct StableMath {
    static constexpr R Pi = R(3.141592653589793238462643383279502884);
    static constexpr R HalfPi = R(1.570796326794896619231321691639751442);
    static constexpr R QuarterPi = R(0.785398163397448309615660845819875721);
    static constexpr R TwoPi = R(6.283185307179586476925286766559005768);
    static constexpr R TanPiOver8 = R(0.414213562373095048801688724209698079);

    // FloorToInt (StableMath.cs:24-31).
    static int floor_to_int(R v) { int i = int(v); return v < R(i) ? i - 1 : i; }
    // ReduceAngle (StableMath.cs:33-47).
    static R reduce_angle(R a) {
        int periods = floor_to_int((a + Pi) / TwoPi);
        a -= R(periods) * TwoPi;
        if (a > Pi) a -= TwoPi;
        else if (a <= -Pi) a += TwoPi;
        return a;
    }
    // ReduceToQuadrant (StableMath.cs:49-59).
    static void reduce_to_quadrant(R a, int& quadrant, R& reduced) {
        a = reduce_angle(a);
        int q = floor_to_int((a + QuarterPi) / HalfPi);
        reduced = a - R(q) * HalfPi;
        quadrant = q & 3;
    }
    // SinPolynomial (StableMath.cs:61-78).
    static R sin_poly(R x) {
        R x2 = x * x;
        R p = -R(1.0 / 6227020800.0);
        p = p * x2 + R(1.0 / 39916800.0);
        p = p * x2 - R(1.0 / 362880.0);
        p = p * x2 + R(1.0 / 5040.0);
        p = p * x2 - R(1.0 / 120.0);
        p = p * x2 + R(1.0 / 6.0);
        p = p * x2 - R(1.0);
        return -x * p;
    }
    // CosPolynomial (StableMath.cs:80-95).
    static R cos_poly(R x) {
        R x2 = x * x;
        R p = -R(1.0 / 479001600.0);
        p = p * x2 + R(1.0 / 3628800.0);
        p = p * x2 - R(1.0 / 40320.0);
        p = p * x2 + R(1.0 / 720.0);
        p = p * x2 - R(1.0 / 24.0);
        p = p * x2 + R(1.0 / 2.0);
        p = p * x2 - R(1.0);
        return -p;
    }
    // SinCos (StableMath.cs:144-158) with ApplyQuadrant (:97-108).
    static void sincos(R a, R& s, R& c) {
        if (a >= -QuarterPi && a <= QuarterPi) { s = sin_poly(a); c = cos_poly(a); return; }
        int q; R r;
        reduce_to_quadrant(a, q, r);
        R sn = sin_poly(r), cs = cos_poly(r);
        switch (q) {
            case 0: s = sn; c = cs; break;
            case 1: s = cs; c = -sn; break;
            case 2: s = -sn; c = -cs; break;
            default: s = -cs; c = sn; break;
        }
    }
    // AtanTaylor (StableMath.cs:193-208).
    static R atan_taylor(R v) {
        R x2 = v * v;
        R p = R(1.0 / 17.0);
        p = p * x2 - R(1.0 / 15.0);
        p = p * x2 + R(1.0 / 13.0);
        p = p * x2 - R(1.0 / 11.0);
        p = p * x2 + R(1.0 / 9.0);
        p = p * x2 - R(1.0 / 7.0);
        p = p * x2 + R(1.0 / 5.0);
        p = p * x2 - R(1.0 / 3.0);
        p = p * x2 + R(1.0);
        return v * p;
    }
    // Atan (StableMath.cs:210-228).
    static R atan(R v) {
        if (v < R(0.0)) return -atan(-v);
        if (v > R(1.0)) return HalfPi - atan(R(1.0) / v);
        if (v > TanPiOver8) {
            R red = (v - R(1.0)) / (v + R(1.0));
            return QuarterPi + atan_taylor(red);
        }
        return atan_taylor(v);
    }
    // Atan2 (StableMath.cs:247-264).
    static R atan2(R y, R x) {
        if (x > R(0.0)) return atan(y / x);
        if (x < R(0.0)) return y >= R(0.0) ? atan(y / x) + Pi : atan(y / x) - Pi;
        if (y > R(0.0)) return HalfPi;
        if (y < R(0.0)) return -HalfPi;
        return R(0.0);
    }
};

// MathHelper.RotationQuaternion (MathHelper.cs:47-78).
template <class R> inline Q4<R> rotation_quaternion(V3<R> omega, R dt) {
    R angle = length(omega);
    R theta = angle * dt;
    if (theta < R(1e-3)) {
        R theta2 = theta * theta;
        R scale = R(0.5) * dt * (R(1.0) - theta2 / R(24.0));
        V3<R> ax = omega * scale;
        R c = R(1.0) - (R(1.0) / R(8.0)) * theta2;
        return {ax.x, ax.y, ax.z, c};
    }
    R half = R(0.5) * angle * dt;
    R s, c;
    StableMath<R>::sincos(half, s, c);
    R scale = s / angle;
    V3<R> ax = omega * scale;
    return {ax.x, ax.y, ax.z, c};
}

}  // namespace orc
