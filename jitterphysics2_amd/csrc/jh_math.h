// Device linear algebra for the gfx950 path.  Every expression keeps the operation order of the
// Jitter2 source it restates (file:line given per function) and the library is compiled with
// -ffp-contract=off and IEEE division/sqrt, so each FP32 result is bit-identical to the
// reference's scalar evaluation.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#define JH_HD __host__ __device__ __forceinline__

namespace jh {

struct vec3 {
    float x, y, z;
};
struct quat {
    float x, y, z, w;
};
// JSymmetricMatrix (LinearMath/JSymmetricMatrix.cs): M11 M12 M13 M22 M23 M33.
struct sym3 {
    float m11, m12, m13, m22, m23, m33;
};
struct mat3 {
    float m11, m12, m13, m21, m22, m23, m31, m32, m33;
};

JH_HD vec3 mk(float x, float y, float z) { return vec3{x, y, z}; }
JH_HD vec3 operator+(vec3 a, vec3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
JH_HD vec3 operator-(vec3 a, vec3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
JH_HD vec3 operator-(vec3 a) { return {-a.x, -a.y, -a.z}; }
JH_HD vec3 operator*(vec3 a, float s) { return {a.x * s, a.y * s, a.z * s}; }
JH_HD vec3 operator*(float s, vec3 a) { return {s * a.x, s * a.y, s * a.z}; }
// JVector.Dot (JVector.cs:395-398)
JH_HD float dot(vec3 a, vec3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
// JVector operator % (JVector.cs:674-681)
JH_HD vec3 cross(vec3 a, vec3 b) { return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
// JVector.Multiply(a, b) (JVector.cs:651-654)
JH_HD vec3 cmul(vec3 a, vec3 b) { return {a.x * b.x, a.y * b.y, a.z * b.z}; }
JH_HD float len_sq(vec3 a) { return a.x * a.x + a.y * a.y + a.z * a.z; }
JH_HD bool veq(vec3 a, vec3 b) { return a.x == b.x && a.y == b.y && a.z == b.z; }
// JVector.Normalize (JVector.cs:600-607)
JH_HD vec3 normalize(vec3 v) {
    float n2 = v.x * v.x + v.y * v.y + v.z * v.z;
    float n = 1.0f / sqrtf(n2);
    return {v.x * n, v.y * n, v.z * n};
}

// JVector.Transform(v, q) (JVector.cs:326-339)
JH_HD vec3 rot(vec3 v, quat q) {
    float n0 = 2.0f * (q.y * v.z - q.z * v.y);
    float n1 = 2.0f * (q.z * v.x - q.x * v.z);
    float n2 = 2.0f * (q.x * v.y - q.y * v.x);
    float n00 = q.y * n2 - q.z * n1;
    float n11 = q.z * n0 - q.x * n2;
    float n22 = q.x * n1 - q.y * n0;
    return {v.x + q.w * n0 + n00, v.y + q.w * n1 + n11, v.z + q.w * n2 + n22};
}
// JVector.ConjugatedTransform(v, q) (JVector.cs:348-361)
JH_HD vec3 rot_inv(vec3 v, quat q) {
    float n0 = 2.0f * (q.z * v.y - q.y * v.z);
    float n1 = 2.0f * (q.x * v.z - q.z * v.x);
    float n2 = 2.0f * (q.y * v.x - q.x * v.y);
    float n00 = q.z * n1 - q.y * n2;
    float n11 = q.x * n2 - q.z * n0;
    float n22 = q.y * n0 - q.x * n1;
    return {v.x + q.w * n0 + n00, v.y + q.w * n1 + n11, v.z + q.w * n2 + n22};
}
// JQuaternion.Multiply (JQuaternion.cs:353-368)
JH_HD quat qmul(quat a, quat b) {
    quat o;
    o.w = a.w * b.w - (a.x * b.x + a.y * b.y + a.z * b.z);
    o.x = a.w * b.x + b.w * a.x + a.y * b.z - a.z * b.y;
    o.y = a.w * b.y + b.w * a.y + a.z * b.x - a.x * b.z;
    o.z = a.w * b.z + b.w * a.z + a.x * b.y - a.y * b.x;
    return o;
}
// JQuaternion.ConjugateMultiply (JQuaternion.cs:378-395): conj(a) * b
JH_HD quat qconjmul(quat a, quat b) { return qmul(quat{-a.x, -a.y, -a.z, a.w}, b); }
// JQuaternion.Normalize (JQuaternion.cs:521-529)
JH_HD quat qnormalize(quat q) {
    float n2 = q.x * q.x + q.y * q.y + q.z * q.z + q.w * q.w;
    float n = 1.0f / sqrtf(n2);
    return {q.x * n, q.y * n, q.z * n, q.w * n};
}
// JQuaternion.GetBasisY (JQuaternion.cs:233-240)
JH_HD vec3 basis_y(quat q) {
    return {2.0f * (q.x * q.y - q.z * q.w), 1.0f - 2.0f * (q.x * q.x + q.z * q.z), 2.0f * (q.y * q.z + q.x * q.w)};
}
// JMatrix.CreateFromQuaternion (JMatrix.cs:553-568)
JH_HD mat3 mat_from_quat(quat q) {
    float r = q.w, i = q.x, j = q.y, k = q.z;
    mat3 m;
    m.m11 = 1.0f - 2.0f * (j * j + k * k);
    m.m12 = 2.0f * (i * j - k * r);
    m.m13 = 2.0f * (i * k + j * r);
    m.m21 = 2.0f * (i * j + k * r);
    m.m22 = 1.0f - 2.0f * (i * i + k * k);
    m.m23 = 2.0f * (j * k - i * r);
    m.m31 = 2.0f * (i * k - j * r);
    m.m32 = 2.0f * (j * k + i * r);
    m.m33 = 1.0f - 2.0f * (i * i + j * j);
    return m;
}
// JMatrix.Absolute then JVector.Transform(v, M) (JMatrix.cs:534-545, JVector.cs:290-298)
JH_HD vec3 abs_mat_mul(vec3 v, mat3 m) {
    return {v.x * fabsf(m.m11) + v.y * fabsf(m.m12) + v.z * fabsf(m.m13),
            v.x * fabsf(m.m21) + v.y * fabsf(m.m22) + v.z * fabsf(m.m23),
            v.x * fabsf(m.m31) + v.y * fabsf(m.m32) + v.z * fabsf(m.m33)};
}
// JVector.Transform(v, JSymmetricMatrix) (JVector.cs:233-241); Contact.TransformSymmetricInertia
// (Contact.cs:1027-1036) evaluates the same sums.
JH_HD vec3 sym_mul(vec3 v, sym3 m) {
    return {v.x * m.m11 + v.y * m.m12 + v.z * m.m13, v.x * m.m12 + v.y * m.m22 + v.z * m.m23,
            v.x * m.m13 + v.y * m.m23 + v.z * m.m33};
}
// JSymmetricMatrix.Transform(M, T) = T M T^T (JSymmetricMatrix.cs:247-260)
JH_HD sym3 sym_rotate(sym3 mi, mat3 t) {
    vec3 r1{t.m11, t.m12, t.m13}, r2{t.m21, t.m22, t.m23}, r3{t.m31, t.m32, t.m33};
    vec3 p1 = sym_mul(r1, mi), p2 = sym_mul(r2, mi), p3 = sym_mul(r3, mi);
    return {dot(p1, r1), dot(p1, r2), dot(p1, r3), dot(p2, r2), dot(p2, r3), dot(p3, r3)};
}

// MathHelper.SignBit (MathHelper.cs:28-32): +1 for a clear sign bit, -1 for a set one (incl. -0).
JH_HD float signbit_f(float v) { return (__float_as_uint(v) >> 31) ? -1.0f : 1.0f; }
// Math.Sign(float)
JH_HD float sign_f(float v) { return v > 0.0f ? 1.0f : (v < 0.0f ? -1.0f : 0.0f); }
JH_HD float fmaxr(float a, float b) { return a > b ? a : b; }
JH_HD float fminr(float a, float b) { return a < b ? a : b; }
JH_HD float clampr(float v, float lo, float hi) { return v < lo ? lo : (v > hi ? hi : v); }

// MathHelper.CreateOrthonormal (MathHelper.cs:204-255)
JH_HD vec3 create_orthonormal(vec3 v) {
    float ax = fabsf(v.x), ay = fabsf(v.y), az = fabsf(v.z);
    if (ax <= ay && ax <= az) {
        float y = v.z, z = -v.y;
        float inv = 1.0f / sqrtf(y * y + z * z);
        return {0.0f, y * inv, z * inv};
    } else if (ay <= az) {
        float x = -v.z, z = v.x;
        float inv = 1.0f / sqrtf(x * x + z * z);
        return {x * inv, 0.0f, z * inv};
    }
    float x = v.y, y = -v.x;
    float inv = 1.0f / sqrtf(x * x + y * y);
    return {x * inv, y * inv, 0.0f};
}

// ---- StableMath (LinearMath/StableMath.cs): the reference's own deterministic polynomials -------
namespace stable {
constexpr float Pi = float(3.141592653589793238462643383279502884);
constexpr float HalfPi = float(1.570796326794896619231321691639751442);
constexpr float QuarterPi = float(0.785398163397448309615660845819875721);
constexpr float TwoPi = float(6.283185307179586476925286766559005768);
constexpr float TanPiOver8 = float(0.414213562373095048801688724209698079);

JH_HD int floor_to_int(float v) {  // :24-31
    int i = int(v);
    return v < float(i) ? i - 1 : i;
}
JH_HD float sin_poly(float x) {  // :61-78
    float x2 = x * x;
    float p = -float(1.0 / 6227020800.0);
    p = p * x2 + float(1.0 / 39916800.0);
    p = p * x2 - float(1.0 / 362880.0);
    p = p * x2 + float(1.0 / 5040.0);
    p = p * x2 - float(1.0 / 120.0);
    p = p * x2 + float(1.0 / 6.0);
    p = p * x2 - 1.0f;
    return -x * p;
}
JH_HD float cos_poly(float x) {  // :80-95
    float x2 = x * x;
    float p = -float(1.0 / 479001600.0);
    p = p * x2 + float(1.0 / 3628800.0);
    p = p * x2 - float(1.0 / 40320.0);
    p = p * x2 + float(1.0 / 720.0);
    p = p * x2 - float(1.0 / 24.0);
    p = p * x2 + float(1.0 / 2.0);
    p = p * x2 - 1.0f;
    return -p;
}
// SinCos (:144-158) with ReduceAngle (:33-47), ReduceToQuadrant (:49-59), ApplyQuadrant (:97-108)
JH_HD void sincos(float a, float& s, float& c) {
    if (a >= -QuarterPi && a <= QuarterPi) {
        s = sin_poly(a);
        c = cos_poly(a);
        return;
    }
    int periods = floor_to_int((a + Pi) / TwoPi);
    a -= float(periods) * TwoPi;
    if (a > Pi) a -= TwoPi;
    else if (a <= -Pi) a += TwoPi;
    int nq = floor_to_int((a + QuarterPi) / HalfPi);
    float r = a - float(nq) * HalfPi;
    int q = nq & 3;
    float sn = sin_poly(r), cs = cos_poly(r);
    if (q == 0) { s = sn; c = cs; }
    else if (q == 1) { s = cs; c = -sn; }
    else if (q == 2) { s = -sn; c = -cs; }
    else { s = -cs; c = sn; }
}
JH_HD float atan_taylor(float v) {  // :193-208
    float x2 = v * v;
    float p = float(1.0 / 17.0);
    p = p * x2 - float(1.0 / 15.0);
    p = p * x2 + float(1.0 / 13.0);
    p = p * x2 - float(1.0 / 11.0);
    p = p * x2 + float(1.0 / 9.0);
    p = p * x2 - float(1.0 / 7.0);
    p = p * x2 + float(1.0 / 5.0);
    p = p * x2 - float(1.0 / 3.0);
    p = p * x2 + 1.0f;
    return v * p;
}
// Atan (:210-228), written without recursion: sign fold, reciprocal fold, pi/4 fold.
JH_HD float atan(float v) {
    bool neg = v < 0.0f;
    if (neg) v = -v;
    float r;
    if (v > 1.0f) {
        float w = 1.0f / v;
        float t = (w > TanPiOver8) ? QuarterPi + atan_taylor((w - 1.0f) / (w + 1.0f)) : atan_taylor(w);
        r = HalfPi - t;
    } else if (v > TanPiOver8) {
        r = QuarterPi + atan_taylor((v - 1.0f) / (v + 1.0f));
    } else {
        r = atan_taylor(v);
    }
    return neg ? -r : r;
}
JH_HD float atan2(float y, float x) {  // :247-264
    if (x > 0.0f) return atan(y / x);
    if (x < 0.0f) return y >= 0.0f ? atan(y / x) + Pi : atan(y / x) - Pi;
    if (y > 0.0f) return HalfPi;
    if (y < 0.0f) return -HalfPi;
    return 0.0f;
}
}  // namespace stable

// MathHelper.RotationQuaternion (MathHelper.cs:47-78)
JH_HD quat rotation_quaternion(vec3 w, float dt) {
    float angle = sqrtf(len_sq(w));
    float theta = angle * dt;
    if (theta < 1e-3f) {
        float theta2 = theta * theta;
        float scale = 0.5f * dt * (1.0f - theta2 / 24.0f);
        vec3 ax = w * scale;
        float c = 1.0f - (1.0f / 8.0f) * theta2;
        return {ax.x, ax.y, ax.z, c};
    }
    float half = 0.5f * angle * dt;
    float s, c;
    stable::sincos(half, s, c);
    float scale = s / angle;
    vec3 ax = w * scale;
    return {ax.x, ax.y, ax.z, c};
}

JH_HD vec3 ld3(float4 v) { return {v.x, v.y, v.z}; }
JH_HD float4 st3(vec3 v, float w = 0.0f) { return make_float4(v.x, v.y, v.z, w); }
JH_HD quat ldq(float4 v) { return {v.x, v.y, v.z, v.w}; }
JH_HD float4 stq(quat q) { return make_float4(q.x, q.y, q.z, q.w); }

}  // namespace jh
