// Oracle — test infrastructure only.
//
// CPU restatement of the numeric primitives that Mufflon's hot path takes from
// the Epsilon library (deps/epsilon, an empty submodule in the reference, so
// their source is not available). Every definition here is this project's
// own, written from the published method it names, and is the authoritative
// semantics that the HIP kernels must reproduce bit for bit:
//   * component-wise min/max as a compare-select `(b < a) ? b : a`;
//   * Box surface 2(dx*dy + dy*dz + dz*dx) (ei::surface, used at
//     build_lbvh.cu:339,387-388);
//   * FastRay {invDir = 1/d, oDivDir = o*invDir} (ei::FastRay, used at
//     intersection.cu:47-77);
//   * Moeller-Trumbore 1997 ray/triangle with barycentric (w0, w1) where w0
//     weights v0 and w1 weights v1 (convention fixed by intersection.cu:568-569
//     and tangent_space.hpp:32-34);
//   * Mat3x4 point / direction transforms (ei::transform, transformDir).
// Compiled with -ffp-contract=off: no operation is fused unless written as fma.
// Only tests/, the smoke check and the benchmark's CPU baseline may use this.
#pragma once
#include <cstdint>
#include <cmath>

namespace orc {

using u32 = uint32_t;
using u64 = uint64_t;
using i32 = int32_t;

inline float mn(float a, float b) { return (b < a) ? b : a; }
inline float mx(float a, float b) { return (a < b) ? b : a; }

struct V3 {
    float x, y, z;
    V3() = default;
    constexpr V3(float a, float b, float c) : x(a), y(b), z(c) {}
    float operator[](int i) const { return i == 0 ? x : (i == 1 ? y : z); }
};
inline V3 operator+(V3 a, V3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
inline V3 operator-(V3 a, V3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline V3 operator*(V3 a, V3 b) { return {a.x * b.x, a.y * b.y, a.z * b.z}; }
inline V3 operator*(V3 a, float s) { return {a.x * s, a.y * s, a.z * s}; }
inline V3 operator*(float s, V3 a) { return {s * a.x, s * a.y, s * a.z}; }
inline V3 operator/(V3 a, float s) { return {a.x / s, a.y / s, a.z / s}; }
inline V3 operator-(V3 a) { return {-a.x, -a.y, -a.z}; }
inline float dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline V3 cross(V3 a, V3 b) {
    return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
inline V3 vmin(V3 a, V3 b) { return {mn(a.x, b.x), mn(a.y, b.y), mn(a.z, b.z)}; }
inline V3 vmax(V3 a, V3 b) { return {mx(a.x, b.x), mx(a.y, b.y), mx(a.z, b.z)}; }
inline float len(V3 a) { return std::sqrt(dot(a, a)); }

struct Box {
    V3 lo, hi;
};
// ei::surface(Box)
inline float surface(const Box& b) {
    V3 d = b.hi - b.lo;
    return 2.0f * (d.x * d.y + d.y * d.z + d.z * d.x);
}
inline Box tri_box(V3 a, V3 b, V3 c) { return {vmin(vmin(a, b), c), vmax(vmax(a, b), c)}; }

// Row-major 3x4 affine matrix (instance transforms, descriptors.hpp:143-222).
struct M34 {
    float m[12];
};
inline V3 xform_point(const M34& M, V3 p) {
    const float* m = M.m;
    return {m[0] * p.x + m[1] * p.y + m[2] * p.z + m[3],
            m[4] * p.x + m[5] * p.y + m[6] * p.z + m[7],
            m[8] * p.x + m[9] * p.y + m[10] * p.z + m[11]};
}
inline V3 xform_dir(const M34& M, V3 d) {
    const float* m = M.m;
    return {m[0] * d.x + m[1] * d.y + m[2] * d.z,
            m[4] * d.x + m[5] * d.y + m[6] * d.z,
            m[8] * d.x + m[9] * d.y + m[10] * d.z};
}
// transpose(Mat3x3(M)) * v : the normal transform of intersection.cu:612-614.
inline V3 xform_normal(const M34& M, V3 v) {
    const float* m = M.m;
    return {m[0] * v.x + m[4] * v.y + m[8] * v.z,
            m[1] * v.x + m[5] * v.y + m[9] * v.z,
            m[2] * v.x + m[6] * v.y + m[10] * v.z};
}
// Inverse of a worldToInstance transform: descriptors.hpp:108-129 (adjugate
// of the 3x3 part over its determinant; translation -(inv * t)).
inline M34 instance_to_world(const float* w) {
    const float r00 = w[0], r01 = w[1], r02 = w[2];
    const float r10 = w[4], r11 = w[5], r12 = w[6];
    const float r20 = w[8], r21 = w[9], r22 = w[10];
    const float m00 = r11 * r22 - r21 * r12;
    const float m01 = r21 * r02 - r01 * r22;
    const float m02 = r01 * r12 - r11 * r02;
    const float m10 = r20 * r12 - r10 * r22;
    const float m11 = r00 * r22 - r20 * r02;
    const float m12 = r10 * r02 - r00 * r12;
    const float m20 = r10 * r21 - r20 * r11;
    const float m21 = r20 * r01 - r00 * r21;
    const float m22 = r00 * r11 - r10 * r01;
    const float det = r00 * m00 + r01 * m10 + r02 * m20;
    const float inv = 1.0f / det;
    M34 o;
    const float a[9] = {m00, m01, m02, m10, m11, m12, m20, m21, m22};
    for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c) o.m[4 * r + c] = inv * a[3 * r + c];
    const float t[3] = {w[3], w[7], w[11]};
    for (int r = 0; r < 3; ++r)
        o.m[4 * r + 3] = -(o.m[4 * r] * t[0] + o.m[4 * r + 1] * t[1] + o.m[4 * r + 2] * t[2]);
    return o;
}

// Box transform over its 8 corners (ei::transform(Box, Mat3x4)).
inline Box xform_box(const M34& M, const Box& b) {
    Box r;
    bool first = true;
    for (int c = 0; c < 8; ++c) {
        V3 p{(c & 1) ? b.hi.x : b.lo.x, (c & 2) ? b.hi.y : b.lo.y, (c & 4) ? b.hi.z : b.lo.z};
        V3 q = xform_point(M, p);
        if (first) {
            r.lo = q;
            r.hi = q;
            first = false;
        } else {
            r.lo = vmin(r.lo, q);
            r.hi = vmax(r.hi, q);
        }
    }
    return r;
}

struct FastRay {
    V3 o, d, inv, odiv;
};
inline FastRay make_fast_ray(V3 o, V3 d) {
    FastRay r;
    r.o = o;
    r.d = d;
    r.inv = {1.0f / d.x, 1.0f / d.y, 1.0f / d.z};
    r.odiv = o * r.inv;
    return r;
}

// Moeller-Trumbore (1997). Returns t and (w0, w1); w2 = 1 - w0 - w1.
inline bool ray_triangle(const FastRay& r, V3 v0, V3 v1, V3 v2, float& t, float& w0, float& w1) {
    const V3 e1 = v1 - v0;
    const V3 e2 = v2 - v0;
    const V3 p = cross(r.d, e2);
    const float det = dot(e1, p);
    if (det == 0.0f) return false;
    const float invDet = 1.0f / det;
    const V3 s = r.o - v0;
    const float u = dot(s, p) * invDet;
    if (u < 0.0f || u > 1.0f) return false;
    const V3 q = cross(s, e1);
    const float v = dot(r.d, q) * invDet;
    if (v < 0.0f || u + v > 1.0f) return false;
    t = dot(e2, q) * invDet;
    w0 = 1.0f - u - v;
    w1 = u;
    return true;
}

}  // namespace orc
