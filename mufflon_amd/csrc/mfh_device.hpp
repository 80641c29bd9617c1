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

// Component k (0, 1, 2) of v, as selects.
__device__ __forceinline__ float comp(V3 v, int k) { return k == 0 ? v.x : (k == 1 ? v.y : v.z); }

// Real roots x0 <= x1 of a x^2 + b x + c = 0 (the role of ei::solveSquarePoly):
// q = -(b + sgn(b) sqrt(b^2 - 4ac)) / 2, roots q / a and c / q, the
// cancellation-free form of Press et al., Numerical Recipes, 3rd ed., 5.6
// (sgn(0) = 1). a == 0: the root -c / b of the linear equation, twice.
__device__ __forceinline__ bool solve_square_poly(float a, float b, float c, float& x0, float& x1) {
    if (a == 0.0f) {
        if (b == 0.0f) return false;
        x0 = -c / b;
        x1 = x0;
        return true;
    }
    const float disc = b * b - 4.0f * a * c;
    if (disc < 0.0f) return false;
    const float r = __builtin_sqrtf(disc);
    const float q = b < 0.0f ? -0.5f * (b - r) : -0.5f * (b + r);
    if (q == 0.0f) {  // b == 0 and c == 0: the double root 0
        x0 = 0.0f;
        x1 = 0.0f;
        return true;
    }
    const float ra = q / a, rb = c / q;
    x0 = rb < ra ? rb : ra;
    x1 = rb < ra ? ra : rb;
    return true;
}

// u of the bilinear patch at a root v: of the two linear equations in u, the
// one with the larger coefficient (intersection.cu:86-95).
__device__ __forceinline__ float quad_u(float v, float A1, float A2, float B1, float B2, float C1, float C2, float D1,
                                        float D2) {
    const float a = v * A2 + B2;
    const float b = v * (A2 - A1) + B2 - B1;
    if (__builtin_fabsf(b) >= __builtin_fabsf(a)) return (v * (C1 - C2) + D1 - D2) / b;
    return (-v * C2 - D2) / a;
}

// Ramsey, Potter and Hansen, "Ray Bilinear Patch Intersections" (JGT 2004), in
// the steps of intersection.cu:99-179: the patch
//   P(u, v) = p00 + u (p10 - p00) + v (p01 - p00) + uv (p11 - p10 - p01 + p00),
// the two equations taken along the largest direction component, roots in v,
// u per root, t along the largest component. t: the nearer of the up to two
// hits with t > 0 (the first root's when the patch is planar), (u, v) its
// patch coordinates.
__device__ __forceinline__ bool ray_quad(const FastRay& r, V3 p00, V3 p10, V3 p11, V3 p01, float& t, float& u,
                                         float& v) {
    int d0 = 0, d1 = 1, d2 = 2;
    const float ax = __builtin_fabsf(r.d.x), ay = __builtin_fabsf(r.d.y), az = __builtin_fabsf(r.d.z);
    if (ax < ay || ax < az) {
        if (ay >= az) {
            d0 = 1;
            d1 = 0;
            d2 = 2;
        } else {
            d0 = 2;
            d1 = 1;
            d2 = 0;
        }
    }
    const V3 a = p11 - p10 - p01 + p00;
    const V3 b = p10 - p00;
    const V3 c = p01 - p00;
    const V3 d = p00 - r.o;
    const float r0 = comp(r.d, d0), r1 = comp(r.d, d1), r2 = comp(r.d, d2);
    const float A1 = comp(a, d1) * r0 - comp(a, d0) * r1;
    const float A2 = comp(a, d2) * r0 - comp(a, d0) * r2;
    const float B1 = comp(b, d1) * r0 - comp(b, d0) * r1;
    const float B2 = comp(b, d2) * r0 - comp(b, d0) * r2;
    const float C1 = comp(c, d1) * r0 - comp(c, d0) * r1;
    const float C2 = comp(c, d2) * r0 - comp(c, d0) * r2;
    const float D1 = comp(d, d1) * r0 - comp(d, d0) * r1;
    const float D2 = comp(d, d2) * r0 - comp(d, d0) * r2;
    float v0 = 0.0f, v1 = 0.0f;
    if (!solve_square_poly(A2 * C1 - A1 * C2, A2 * D1 - A1 * D2 + B2 * C1 - B1 * C2, B2 * D1 - B1 * D2, v0, v1))
        return false;
    const float ad = comp(a, d0), bd = comp(b, d0), cd = comp(c, d0), dd = comp(d, d0), inv = comp(r.inv, d0);
    float u0 = 0.0f, u1 = 0.0f, t0 = -1.0f, t1 = -1.0f;
    if (v0 >= 0.0f && v0 <= 1.0f) {
        u0 = quad_u(v0, A1, A2, B1, B2, C1, C2, D1, D2);
        if (u0 >= 0.0f && u0 <= 1.0f) t0 = (u0 * v0 * ad + u0 * bd + v0 * cd + dd) * inv;
    }
    if (v1 == v0) {  // one root (planar patch)
        t = t0;
        u = u0;
        v = v0;
    } else {
        if (v1 >= 0.0f && v1 <= 1.0f) {
            u1 = quad_u(v1, A1, A2, B1, B2, C1, C2, D1, D2);
            if (u1 >= 0.0f && u1 <= 1.0f) t1 = (u1 * v1 * ad + u1 * bd + v1 * cd + dd) * inv;
        }
        const bool second = !(t0 > 0.0f) || (t1 > 0.0f && t1 < t0);
        t = second ? t1 : t0;
        u = second ? u1 : u0;
        v = second ? v1 : v0;
    }
    return t > 0.0f;
}

// Ray / sphere after Haines, Guenther and Akenine-Moeller, "Precision
// Improvements for Ray/Sphere Intersection" (Ray Tracing Gems, 2019): the
// discriminant from the ray's closest approach l = f + b d (f = o - center,
// b = -f.d, unit d), roots c / q and q with q = b + sgn(b) sqrt(disc) (the
// role of ei::intersects(Ray, Sphere)). t: the nearer root when it is
// positive, else the farther one (origin inside); false when both are <= 0.
__device__ __forceinline__ bool ray_sphere(const FastRay& r, V3 center, float radius, float& t) {
    const V3 f = r.o - center;
    const float b = -dot(f, r.d);
    const V3 l = f + r.d * b;
    const float rr = radius * radius;
    const float disc = rr - dot(l, l);
    if (disc < 0.0f) return false;
    const float c = dot(f, f) - rr;
    const float s = __builtin_sqrtf(disc);
    const float q = b < 0.0f ? b - s : b + s;
    if (q == 0.0f) return false;  // then c == 0 too: the only root is t = 0
    const float ra = c / q;
    const float tn = ra < q ? ra : q, tf = ra < q ? q : ra;
    t = tn > 0.0f ? tn : tf;
    return tf > 0.0f;
}

// A vector perpendicular to v (the role of ei::perpendicular): the
// smallest-magnitude component set to zero, the other two swapped with one
// negated (Hughes and Moeller, "Building an Orthonormal Basis from a Unit
// Vector", JGT 1999).
__device__ __forceinline__ V3 perpendicular(V3 v) {
    const float ax = __builtin_fabsf(v.x), ay = __builtin_fabsf(v.y), az = __builtin_fabsf(v.z);
    if (ax <= ay && ax <= az) return {0.0f, -v.z, v.y};
    if (ay <= az) return {-v.z, 0.0f, v.x};
    return {-v.y, v.x, 0.0f};
}

// Bilinear interpolation over the corners x00 (s = t = 0), x01 (s = 1),
// x10 (t = 1), x11: lerp(lerp(x00, x01, s), lerp(x10, x11, s), t) with
// lerp(a, b, s) = (1 - s) a + s b, which returns the corners exactly (the
// role of ei::bilerp).
__device__ __forceinline__ float bilerp(float x00, float x01, float x10, float x11, float s, float t) {
    const float a = (1.0f - s) * x00 + s * x01;
    const float b = (1.0f - s) * x10 + s * x11;
    return (1.0f - t) * a + t * b;
}
__device__ __forceinline__ V3 bilerp(V3 x00, V3 x01, V3 x10, V3 x11, float s, float t) {
    return {bilerp(x00.x, x01.x, x10.x, x11.x, s, t), bilerp(x00.y, x01.y, x10.y, x11.y, s, t),
            bilerp(x00.z, x01.z, x10.z, x11.z, s, t)};
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
