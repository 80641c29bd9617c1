// Device-side numeric core shared by the HIP kernels of this library.
//
// The rounding rules are fixed so that every kernel computes exactly what the
// CPU path of Mufflon computes with the same inputs:
//   * all sources are compiled with -ffp-contract=off; a fused multiply-add
//     appears only where it is written as fmaf();
//   * division and sqrt are the correctly rounded IEEE operations (hipcc's
//     default -fhip-fp32-correctly-rounded-divide-sqrt);
//   * min/max are the compare-selects (b < a) ? b : a and (a < b) ? b : a,
//     whose result for +0/-0 depends on the operand order, so every caller
//     fixes the order.
// Mufflon takes these primitives from the Epsilon library, whose source is not
// part of the reference tree; the definitions below are this library's own,
// each after the published method it names.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace mfh {

using u32 = uint32_t;
using u64 = uint64_t;
using i32 = int32_t;

__device__ __forceinline__ float mn(float a, float b) { return (b < a) ? b : a; }
__device__ __forceinline__ float mx(float a, float b) { return (a < b) ? b : a; }

struct V3 {
    float x, y, z;
};
__device__ __forceinline__ V3 v3(float x, float y, float z) { return V3{x, y, z}; }
__device__ __forceinline__ V3 operator+(V3 a, V3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
__device__ __forceinline__ V3 operator-(V3 a, V3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
__device__ __forceinline__ V3 operator*(V3 a, float s) { return {a.x * s, a.y * s, a.z * s}; }
__device__ __forceinline__ V3 operator*(V3 a, V3 b) { return {a.x * b.x, a.y * b.y, a.z * b.z}; }
__device__ __forceinline__ V3 operator/(V3 a, float s) { return {a.x / s, a.y / s, a.z / s}; }
// ((a.x*b.x + a.y*b.y) + a.z*b.z), each product and sum rounded.
__device__ __forceinline__ float dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ __forceinline__ V3 cross(V3 a, V3 b) {
    return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
__device__ __forceinline__ V3 vmin(V3 a, V3 b) { return {mn(a.x, b.x), mn(a.y, b.y), mn(a.z, b.z)}; }
__device__ __forceinline__ V3 vmax(V3 a, V3 b) { return {mx(a.x, b.x), mx(a.y, b.y), mx(a.z, b.z)}; }
__device__ __forceinline__ float len(V3 a) { return __builtin_sqrtf(dot(a, a)); }

// Box surface area 2(dx*dy + dy*dz + dz*dx).
__device__ __forceinline__ float surface(V3 lo, V3 hi) {
    const V3 d = hi - lo;
    return 2.0f * (d.x * d.y + d.y * d.z + d.z * d.x);
}

// Row-major 3x4 affine transform, stored as 12 floats.
struct M34 {
    float m[12];
};
__device__ __forceinline__ V3 xform_point(const M34& M, V3 p) {
    const float* m = M.m;
    return {m[0] * p.x + m[1] * p.y + m[2] * p.z + m[3], m[4] * p.x + m[5] * p.y + m[6] * p.z + m[7],
            m[8] * p.x + m[9] * p.y + m[10] * p.z + m[11]};
}
// Inverse of a worldToInstance transform (descriptors.hpp:108-129: adjugate
// of the 3x3 part over its determinant, translation -(inv * t)).
__device__ __forceinline__ M34 inst_to_world(const float* w) {
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
    const float s = 1.0f / det;
    M34 o;
    o.m[0] = s * m00; o.m[1] = s * m01; o.m[2] = s * m02;
    o.m[4] = s * m10; o.m[5] = s * m11; o.m[6] = s * m12;
    o.m[8] = s * m20; o.m[9] = s * m21; o.m[10] = s * m22;
    const float tx = w[3], ty = w[7], tz = w[11];
    o.m[3] = -(o.m[0] * tx + o.m[1] * ty + o.m[2] * tz);
    o.m[7] = -(o.m[4] * tx + o.m[5] * ty + o.m[6] * tz);
    o.m[11] = -(o.m[8] * tx + o.m[9] * ty + o.m[10] * tz);
    return o;
}

__device__ __forceinline__ V3 xform_dir(const M34& M, V3 d) {
    const float* m = M.m;
    return {m[0] * d.x + m[1] * d.y + m[2] * d.z, m[4] * d.x + m[5] * d.y + m[6] * d.z,
            m[8] * d.x + m[9] * d.y + m[10] * d.z};
}
// transpose(Mat3x3(M)) * v
__device__ __forceinline__ V3 xform_normal(const M34& M, V3 v) {
    const float* m = M.m;
    return {m[0] * v.x + m[4] * v.y + m[8] * v.z, m[1] * v.x + m[5] * v.y + m[9] * v.z,
            m[2] * v.x + m[6] * v.y + m[10] * v.z};
}

// Ray with precomputed reciprocal direction and origin/direction product.
struct FastRay {
    V3 o, d, inv, odiv;
};
__device__ __forceinline__ FastRay make_fast_ray(V3 o, V3 d) {
    FastRay r;
    r.o = o;
    r.d = d;
    r.inv = {1.0f / d.x, 1.0f / d.y, 1.0f / d.z};
    r.odiv = o * r.inv;
    return r;
}

// Moeller & Trumbore, "Fast, minimum storage ray-triangle intersection", 1997.
// Returns t and the barycentric weights (w0, w1) of v0 and v1.
__device__ __forceinline__ bool ray_triangle(const FastRay& r, V3 v0, V3 v1, V3 v2, float& t, float& w0,
                                             float& w1) {
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

// 32-byte child record of a collapsed BVH node (two per 64-byte node):
// the layout of Mufflon's BvhNode (lbvh.hpp:13-18).
struct __attribute__((packed)) BvhChild {
    float lo[3];
    float hi[3];
    i32 index;
    i32 primCount;
};
static_assert(sizeof(BvhChild) == 32, "BvhChild must be 32 bytes");

}  // namespace mfh
