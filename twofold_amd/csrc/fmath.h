// fmath.h - float32 vector, matrix and quaternion operations for the contact solver kernels.
//
// Every function performs exactly the float operations of the reference's math
// (include/twofold/math.h, src/internal.h:28-45), in the same order, so that with
// -ffp-contract=off (no fused multiply-add) and IEEE division and square root the kernels give
// the reference's bits. Nothing here may be rewritten into a cheaper but different sequence
// (reciprocal-multiply for division, rsqrt, fma); maxq/minq below are the min/max instructions
// where they give maxf/minf's bits.
#pragma once

#include <stdint.h>

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define TF_HD __host__ __device__ __forceinline__
#else
#include <math.h>
#define TF_HD inline
#endif

namespace tf {

struct V3 {
    float x, y, z;
};
struct Q4 {
    float x, y, z, w;
};
struct M3 {
    V3 r[3];
};

TF_HD V3 vec(float x, float y, float z) { return V3{ x, y, z }; }
TF_HD Q4 quat(float x, float y, float z, float w) { return Q4{ x, y, z, w }; }

// maxf / minf: src/internal.h:28-45, NaN and signed zeros included.
TF_HD float maxf(float a, float b) {
    if (a == b || a != a || b != b) {
        if (a != a) { return b; }
        if (b != b) { return a; }
        return a == 0.0f ? a + b : a;
    }
    return a > b ? a : b;
}
TF_HD float minf(float a, float b) {
    if (a == b || a != a || b != b) {
        if (a != a) { return b; }
        if (b != b) { return a; }
        return a == 0.0f ? -(-a + -b) : a;
    }
    return a < b ? a : b;
}

// maxf / minf as one instruction, for operands that are never signalling NaNs (results of float
// arithmetic, and values only ever written from such results). The gfx950 v_max_f32 / v_min_f32
// order -0 below +0 and return the other operand when one is NaN, as maxf / minf do; when both are
// NaN they return their first operand where maxf / minf return their second, hence the swapped
// operands. A signalling NaN operand would come back quieted, which maxf / minf do not do.
TF_HD float maxq(float a, float b) {
#if defined(__HIP_DEVICE_COMPILE__)
    float r;
    asm("v_max_f32 %0, %1, %2" : "=v"(r) : "v"(b), "v"(a));
    return r;
#else
    return maxf(a, b);
#endif
}
TF_HD float minq(float a, float b) {
#if defined(__HIP_DEVICE_COMPILE__)
    float r;
    asm("v_min_f32 %0, %1, %2" : "=v"(r) : "v"(b), "v"(a));
    return r;
#else
    return minf(a, b);
#endif
}

TF_HD float fsqrt(float x) { return sqrtf(x); }
TF_HD float fabs_(float x) { return fabsf(x); }

// math.h:40-62
TF_HD V3 operator+(V3 a, V3 b) { return vec(a.x + b.x, a.y + b.y, a.z + b.z); }
TF_HD V3 operator-(V3 a, V3 b) { return vec(a.x - b.x, a.y - b.y, a.z - b.z); }
TF_HD V3 operator*(V3 a, float s) { return vec(a.x * s, a.y * s, a.z * s); }
TF_HD float dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
TF_HD V3 cross(V3 a, V3 b) { return vec(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x); }
TF_HD float length(V3 v) { return fsqrt(dot(v, v)); }
TF_HD V3 normalize(V3 v) {
    const float l = length(v);
    return l > 0 ? v * (1.0f / l) : vec(0, 0, 0);
}

// math.h:76-138 (row-major)
TF_HD M3 mzero() { return M3{ { vec(0, 0, 0), vec(0, 0, 0), vec(0, 0, 0) } }; }
TF_HD M3 outer(V3 a, V3 b) { return M3{ { b * a.x, b * a.y, b * a.z } }; }
TF_HD V3 column(const M3& m, int i) {
    return i == 0 ? vec(m.r[0].x, m.r[1].x, m.r[2].x)
                  : i == 1 ? vec(m.r[0].y, m.r[1].y, m.r[2].y) : vec(m.r[0].z, m.r[1].z, m.r[2].z);
}
TF_HD V3 operator*(const M3& m, V3 v) { return vec(dot(m.r[0], v), dot(m.r[1], v), dot(m.r[2], v)); }
TF_HD M3 mul(const M3& a, const M3& b) {
    M3 m;
    for (int i = 0; i < 3; ++i) { m.r[i] = vec(dot(a.r[i], column(b, 0)), dot(a.r[i], column(b, 1)), dot(a.r[i], column(b, 2))); }
    return m;
}
TF_HD M3 operator+(const M3& a, const M3& b) { return M3{ { a.r[0] + b.r[0], a.r[1] + b.r[1], a.r[2] + b.r[2] } }; }
TF_HD M3 operator-(const M3& a, const M3& b) { return M3{ { a.r[0] - b.r[0], a.r[1] - b.r[1], a.r[2] - b.r[2] } }; }
TF_HD M3 operator*(const M3& a, float s) { return M3{ { a.r[0] * s, a.r[1] * s, a.r[2] * s } }; }
TF_HD M3 transpose(const M3& a) { return M3{ { column(a, 0), column(a, 1), column(a, 2) } }; }
TF_HD M3 inverse(const M3& m) {
    const V3 c0 = cross(m.r[1], m.r[2]);
    const V3 c1 = cross(m.r[2], m.r[0]);
    const V3 c2 = cross(m.r[0], m.r[1]);
    float det = dot(m.r[0], c0);
    if (det == 0.0f) { return mzero(); }
    det = 1.0f / det;
    return M3{ { vec(c0.x, c1.x, c2.x) * det, vec(c0.y, c1.y, c2.y) * det, vec(c0.z, c1.z, c2.z) * det } };
}

// math.h:162-194, 227-230
TF_HD Q4 operator*(Q4 a, Q4 b) {
    return quat(a.w * b.x + a.x * b.w + a.y * b.z - a.z * b.y,
                a.w * b.y - a.x * b.z + a.y * b.w + a.z * b.x,
                a.w * b.z + a.x * b.y - a.y * b.x + a.z * b.w,
                a.w * b.w - a.x * b.x - a.y * b.y - a.z * b.z);
}
TF_HD Q4 conjugate(Q4 q) { return quat(-q.x, -q.y, -q.z, q.w); }
TF_HD V3 rotate(Q4 q, V3 v) {
    const V3 u = vec(q.x, q.y, q.z);
    const V3 t = cross(u, v) * 2.0f;
    return v + t * q.w + cross(u, t);
}
TF_HD M3 toMat3(Q4 q) {
    const float xx = q.x * q.x, yy = q.y * q.y, zz = q.z * q.z;
    const float xy = q.x * q.y, xz = q.x * q.z, yz = q.y * q.z;
    const float wx = q.w * q.x, wy = q.w * q.y, wz = q.w * q.z;
    M3 m;
    m.r[0] = vec(1 - 2 * (yy + zz), 2 * (xy - wz), 2 * (xz + wy));
    m.r[1] = vec(2 * (xy + wz), 1 - 2 * (xx + zz), 2 * (yz - wx));
    m.r[2] = vec(2 * (xz - wy), 2 * (yz + wx), 1 - 2 * (xx + yy));
    return m;
}
TF_HD Q4 qnormalize(Q4 q) {
    const float l = fsqrt(q.x * q.x + q.y * q.y + q.z * q.z + q.w * q.w);
    return l > 0 ? quat(q.x / l, q.y / l, q.z / l, q.w / l) : quat(0, 0, 0, 1);
}
TF_HD Q4 integrate(Q4 q, V3 w, float h) {
    const Q4 d = quat(w.x, w.y, w.z, 0.0f) * q;
    return qnormalize(quat(q.x + 0.5f * h * d.x, q.y + 0.5f * h * d.y, q.z + 0.5f * h * d.z, q.w + 0.5f * h * d.w));
}

// Loads and stores of float arrays.
TF_HD V3 ld3(const float* p) { return vec(p[0], p[1], p[2]); }
TF_HD Q4 ld4(const float* p) { return quat(p[0], p[1], p[2], p[3]); }
TF_HD M3 ld9(const float* p) { return M3{ { ld3(p), ld3(p + 3), ld3(p + 6) } }; }
TF_HD void st3(float* p, V3 v) { p[0] = v.x; p[1] = v.y; p[2] = v.z; }
TF_HD void st4(float* p, Q4 q) { p[0] = q.x; p[1] = q.y; p[2] = q.z; p[3] = q.w; }
TF_HD void st9(float* p, const M3& m) { st3(p, m.r[0]); st3(p + 3, m.r[1]); st3(p + 6, m.r[2]); }

}  // namespace tf
