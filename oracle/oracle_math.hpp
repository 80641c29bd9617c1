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

// Component k (0, 1, 2) of v.
inline float comp(V3 v, int k) { return k == 0 ? v.x : (k == 1 ? v.y : v.z); }

// Real roots x0 <= x1 of a x^2 + b x + c = 0 (the role of ei::solveSquarePoly):
// q = -(b + sgn(b) sqrt(b^2 - 4ac)) / 2, roots q / a and c / q, the
// cancellation-free form of Press et al., Numerical Recipes, 3rd ed., 5.6
// (sgn(0) = 1). a == 0: the root -c / b of the linear equation, twice.
inline bool solve_square_poly(float a, float b, float c, float& x0, float& x1) {
    if (a == 0.0f) {
        if (b == 0.0f) return false;
        x0 = -c / b;
        x1 = x0;
        return true;
    }
    const float disc = b * b - 4.0f * a * c;
    if (disc < 0.0f) return false;
    const float r = std::sqrt(disc);
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

// intersection.cu:86-95 (computeU)
inline float quad_u(float v, float A1, float A2, float B1, float B2, float C1, float C2, float D1, float D2) {
    const float a = v * A2 + B2;
    const float b = v * (A2 - A1) + B2 - B1;
    if (std::fabs(b) >= std::fabs(a)) return (v * (C1 - C2) + D1 - D2) / b;
    return (-v * C2 - D2) / a;
}

// Ramsey, Potter and Hansen, "Ray Bilinear Patch Intersections" (JGT 2004), in
// the steps of intersection.cu:99-179 (intersectQuad): the patch
//   P(u, v) = p00 + u (p10 - p00) + v (p01 - p00) + uv (p11 - p10 - p01 + p00),
// equations along the largest direction component, roots in v, u per root, t
// along the largest component. Returns the nearer hit with t > 0 (the first
// root's when the patch is planar) and its (u, v).
inline bool ray_quad(const FastRay& r, V3 p00, V3 p10, V3 p11, V3 p01, float& t, float& u, float& v) {
    int d0 = 0, d1 = 1, d2 = 2;
    if (std::fabs(r.d.x) < std::fabs(r.d.y) || std::fabs(r.d.x) < std::fabs(r.d.z)) {
        if (std::fabs(r.d.y) >= std::fabs(r.d.z)) {
            d0 = 1; d1 = 0; d2 = 2;
        } else {
            d0 = 2; d1 = 1; d2 = 0;
        }
    }
    const V3 a = p11 - p10 - p01 + p00;
    const V3 b = p10 - p00;
    const V3 c = p01 - p00;
    const V3 d = p00 - r.o;
    const V3& dir = r.d;
    const float A1 = a[d1] * dir[d0] - a[d0] * dir[d1];
    const float A2 = a[d2] * dir[d0] - a[d0] * dir[d2];
    const float B1 = b[d1] * dir[d0] - b[d0] * dir[d1];
    const float B2 = b[d2] * dir[d0] - b[d0] * dir[d2];
    const float C1 = c[d1] * dir[d0] - c[d0] * dir[d1];
    const float C2 = c[d2] * dir[d0] - c[d0] * dir[d2];
    const float D1 = d[d1] * dir[d0] - d[d0] * dir[d1];
    const float D2 = d[d2] * dir[d0] - d[d0] * dir[d2];
    float v0 = 0.0f, v1 = 0.0f;
    if (!solve_square_poly(A2 * C1 - A1 * C2, A2 * D1 - A1 * D2 + B2 * C1 - B1 * C2, B2 * D1 - B1 * D2, v0, v1))
        return false;
    float u0 = 0.0f, u1 = 0.0f, t0 = -1.0f, t1 = -1.0f;
    if (v0 >= 0.0f && v0 <= 1.0f) {
        u0 = quad_u(v0, A1, A2, B1, B2, C1, C2, D1, D2);
        if (u0 >= 0.0f && u0 <= 1.0f) t0 = (u0 * v0 * a[d0] + u0 * b[d0] + v0 * c[d0] + d[d0]) * r.inv[d0];
    }
    if (v1 == v0) {
        t = t0; u = u0; v = v0;
    } else {
        if (v1 >= 0.0f && v1 <= 1.0f) {
            u1 = quad_u(v1, A1, A2, B1, B2, C1, C2, D1, D2);
            if (u1 >= 0.0f && u1 <= 1.0f) t1 = (u1 * v1 * a[d0] + u1 * b[d0] + v1 * c[d0] + d[d0]) * r.inv[d0];
        }
        if (t0 > 0.0f) {
            if (t1 > 0.0f && t1 < t0) {
                t = t1; u = u1; v = v1;
            } else {
                t = t0; u = u0; v = v0;
            }
        } else {
            t = t1; u = u1; v = v1;
        }
    }
    return t > 0.0f;
}

// Ray / sphere after Haines, Guenther and Akenine-Moeller, "Precision
// Improvements for Ray/Sphere Intersection" (Ray Tracing Gems, 2019), the role
// of ei::intersects(Ray, Sphere): discriminant from the closest approach
// l = f + b d (f = o - center, b = -f.d, unit d), roots c / q and q with
// q = b + sgn(b) sqrt(disc). t: the nearer root when positive, else the
// farther one (origin inside); false when both are <= 0.
inline bool ray_sphere(const FastRay& r, V3 center, float radius, float& t) {
    const V3 f = r.o - center;
    const float b = -dot(f, r.d);
    const V3 l = f + r.d * b;
    const float rr = radius * radius;
    const float disc = rr - dot(l, l);
    if (disc < 0.0f) return false;
    const float c = dot(f, f) - rr;
    const float s = std::sqrt(disc);
    const float q = b < 0.0f ? b - s : b + s;
    if (q == 0.0f) return false;
    const float ra = c / q;
    const float tn = ra < q ? ra : q, tf = ra < q ? q : ra;
    t = tn > 0.0f ? tn : tf;
    return tf > 0.0f;
}

// A vector perpendicular to v (the role of ei::perpendicular): the
// smallest-magnitude component set to zero, the other two swapped with one
// negated (Hughes and Moeller, "Building an Orthonormal Basis from a Unit
// Vector", JGT 1999).
inline V3 perpendicular(V3 v) {
    const float ax = std::fabs(v.x), ay = std::fabs(v.y), az = std::fabs(v.z);
    if (ax <= ay && ax <= az) return {0.0f, -v.z, v.y};
    if (ay <= az) return {-v.z, 0.0f, v.x};
    return {-v.y, v.x, 0.0f};
}

// lerp(lerp(x00, x01, s), lerp(x10, x11, s), t), lerp(a, b, s) = (1 - s) a + s b
// (the role of ei::bilerp; exact at the corners).
inline float bilerp(float x00, float x01, float x10, float x11, float s, float t) {
    const float a = (1.0f - s) * x00 + s * x01;
    const float b = (1.0f - s) * x10 + s * x11;
    return (1.0f - t) * a + t * b;
}
inline V3 bilerp(V3 x00, V3 x01, V3 x10, V3 x11, float s, float t) {
    return {bilerp(x00.x, x01.x, x10.x, x11.x, s, t), bilerp(x00.y, x01.y, x10.y, x11.y, s, t),
            bilerp(x00.z, x01.z, x10.z, x11.z, s, t)};
}

// Box of four points, of a sphere (center -+ radius: the role of ei::Box(Sphere)).
inline Box quad_box(V3 a, V3 b, V3 c, V3 d) { return {vmin(vmin(vmin(a, b), c), d), vmax(vmax(vmax(a, b), c), d)}; }
inline Box sphere_box(V3 c, float r) { return {{c.x - r, c.y - r, c.z - r}, {c.x + r, c.y + r, c.z + r}}; }

}  // namespace orc
