// Device shading core of the path tracer: RNG, texture sampling, material
// fetch / sample / evaluate (EMISSIVE, LAMBERT, TORRANCE, WALTER, MICROFACET), media,
// background and light-tree next-event estimation. Used by the megakernel and
// the wavefront stages alike, so both produce the same bits.
//
// Numerics: every float expression keeps Mufflon's operand order; the library
// is compiled with -ffp-contract=off and correctly rounded division / sqrt.
// Functions of the absent Epsilon library and libm calls have these
// definitions: normalize(v) = v / len(v); sgn(x) = x < 0 ? -1 : 1; min / max
// are the compare-selects mn / mx; sdiv(x, d) = x / (d + 1e-20 sgn d)
// (util/types.hpp:121-127); sin, cos, acos, atan2, exp, log and pow of a float
// are evaluated in double and rounded once; float -> u64 conversion maps NaN
// and negatives to 0 and saturates at 2^64 - 1; mod is the non-negative
// remainder. Media are carried by value: each material holds its outer medium
// and (WALTER, MICROFACET) its inner medium; the camera and point lights sit in vacuum.
#pragma once
#include "pt_device.hpp"

namespace mfh {
namespace shade {

__device__ __forceinline__ float sgnf(float x) { return x < 0.0f ? -1.0f : 1.0f; }
__device__ __forceinline__ float clampf(float v, float lo, float hi) { return mn(mx(v, lo), hi); }
__device__ __forceinline__ float sq(float x) { return x * x; }
__device__ __forceinline__ V3 normalize(V3 v) { return v / len(v); }
__device__ __forceinline__ V3 neg(V3 v) { return {-v.x, -v.y, -v.z}; }
__device__ __forceinline__ V3 smul(float s, V3 a) { return {s * a.x, s * a.y, s * a.z}; }
__device__ __forceinline__ float sdiv(float x, float d) { return x / (d + 1e-20f * sgnf(d)); }
__device__ __forceinline__ bool nonzero(V3 v) { return v.x != 0.0f || v.y != 0.0f || v.z != 0.0f; }
__device__ __forceinline__ float fsin(float x) { return (float)sin((double)x); }
__device__ __forceinline__ float fcos(float x) { return (float)cos((double)x); }
__device__ __forceinline__ float facos(float x) { return (float)acos((double)x); }
__device__ __forceinline__ float fatan2(float y, float x) { return (float)atan2((double)y, (double)x); }
__device__ __forceinline__ float fexp(float x) { return (float)exp((double)x); }
__device__ __forceinline__ float flog(float x) { return (float)log((double)x); }
__device__ __forceinline__ float fpow(float x, float y) { return (float)pow((double)x, (double)y); }
__device__ __forceinline__ float fsqrt(float x) { return __builtin_sqrtf(x); }
__device__ __forceinline__ int ifloor(float x) { return (int)floorf(x); }
__device__ __forceinline__ int imod(int a, int m) {
    const int r = a % m;
    return r < 0 ? r + m : r;
}
__device__ __forceinline__ u64 f2u64(float v) {
    if (!(v > 0.0f)) return 0;
    if (v >= 18446744073709551616.0f) return ~0ull;
    return (u64)v;
}
__device__ __forceinline__ u64 d2u64(double v) {
    if (!(v > 0.0)) return 0;
    if (v >= 18446744073709551616.0) return ~0ull;
    return (u64)v;
}

// ------------------------------------------------------------------ RNG
// PCG64 with 64-bit state and RXS-M-XS output (rng.hpp:70-93).
__device__ __forceinline__ u64 pcg_next(u64& s) {
    u64 x = s;
    const u64 count = s >> 59ull;
    s = s * 6364136223846793005ull + 1442695040888963407ull;
    x ^= x >> (5 + count);
    x *= 12605985483714917081ull;
    return x ^ (x >> 43);
}
__device__ __forceinline__ u64 scramble_seed(u32 x) {
    u64 z = (u64)x + 0x9e3779b97f4a7c15ull;
    z = (z ^ (z >> 30u)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27u)) * 0x94d049bb133111ebull;
    return z ^ (z >> 31u);
}
__device__ __forceinline__ float sample_uniform(u32 u) { return (float)u / 4294967810.0f; }
// sampling.hpp:39-46
__device__ __forceinline__ u64 percentage_of(u64 num, float p) {
    const u64 pfix = f2u64(p * 2147483648.0f);
    return ((pfix * (num & 0x7fffffff)) >> 31) + (pfix * (num >> 31));
}
// sampling.hpp:72-85 (an empty interval maps to its left end)
__device__ __forceinline__ u64 rescale_sample(u64 x, u64 pLeft, u64 pRight) {
    const u64 interval = pRight - pLeft;
    const u64 xrel = x - pLeft;
    if (interval == 0) return 0;
    return xrel * (~0ull / interval) + d2u64((double)xrel * (double)(~0ull % interval) / (double)interval);
}

// ------------------------------------------------------------------ textures
__device__ __forceinline__ float4 texel(const DevTexture& t, int x, int y) {
    return t.texels[(size_t)imod(y, t.height) * t.width + imod(x, t.width)];
}
// CPU texture sampling (cputexture.cpp:372-411): nearest, linear without
// half-texel offset, 1x1 textures return their texel.
__device__ __forceinline__ float4 tex_sample(const DevTexture& t, float u, float v) {
    if (t.width == 1 && t.height == 1) return t.texels[0];
    if (t.sampling == 0) return texel(t, ifloor(u * (float)t.width), ifloor(v * (float)t.height));
    float fx = u * (float)t.width, fy = v * (float)t.height;
    const int ix = ifloor(fx), iy = ifloor(fy);
    fx -= (float)ix;
    fy -= (float)iy;
    const float4 a = texel(t, ix, iy), b = texel(t, ix + 1, iy), c = texel(t, ix, iy + 1), d = texel(t, ix + 1, iy + 1);
    const float wa0 = 1.0f - fx, wa1 = 1.0f - fy;
    float4 s;
    s.x = a.x * wa0 * wa1;
    s.y = a.y * wa0 * wa1;
    s.z = a.z * wa0 * wa1;
    s.w = a.w * wa0 * wa1;
    s.x += b.x * fx * wa1;
    s.y += b.y * fx * wa1;
    s.z += b.z * fx * wa1;
    s.w += b.w * fx * wa1;
    s.x += c.x * wa0 * fy;
    s.y += c.y * wa0 * fy;
    s.z += c.z * wa0 * fy;
    s.w += c.w * wa0 * fy;
    s.x += d.x * fx * fy;
    s.y += d.y * fx * fy;
    s.z += d.z * fx * fy;
    s.w += d.w * fx * fy;
    return s;
}

// ------------------------------------------------------------------ materials
struct Medium {
    float ior;
    V3 absorption;
};
// Medium::get_transmission (medium.hpp:40-42)
__device__ __forceinline__ V3 transmission(const Medium& m, float distance) {
    const V3 e = m.absorption * -distance;
    return {fexp(e.x), fexp(e.y), fexp(e.z)};
}

// Fetched ParameterPack of one material at one uv (material_sampling.hpp:33-53).
struct MatParams {
    int type;
    V3 albedo;    // LAMBERT, TORRANCE
    V3 radiance;  // EMISSIVE (scaled)
    float rx, ry;
    int ndf, shadowing;
    Medium inner, outer;
    __device__ __forceinline__ const Medium& medium(float xDotN) const { return xDotN < 0.0f ? inner : outer; }
    __device__ __forceinline__ bool reflective() const {
        return type == kMatLambert || type == kMatTorrance || type == kMatWalter || type == kMatMicrofacet;
    }
    __device__ __forceinline__ bool refractive() const { return mat_has_inner_medium(type); }
};

__device__ __forceinline__ MatParams fetch(const PtScene& S, int idx, float u, float v) {
    const PtMaterialRecord& m = S.materials[idx];
    MatParams p;
    p.type = m.type;
    p.outer.ior = m.outerIor;
    p.outer.absorption = v3(m.outerAbsorption[0], m.outerAbsorption[1], m.outerAbsorption[2]);
    if (mat_has_inner_medium(p.type)) {
        p.inner.ior = m.ior;
        p.inner.absorption = v3(m.absorption[0], m.absorption[1], m.absorption[2]);
    } else {
        p.inner = p.outer;
    }
    p.albedo = v3(0, 0, 0);
    p.radiance = v3(0, 0, 0);
    p.rx = p.ry = 0.0f;
    p.ndf = m.ndf;
    p.shadowing = m.shadowing;
    if (p.type == kMatLambert || p.type == kMatTorrance) {
        if (m.albedoTexture >= 0) {
            const float4 t = tex_sample(S.textures[m.albedoTexture], u, v);
            p.albedo = v3(t.x, t.y, t.z);
        } else {
            p.albedo = v3(m.albedo[0], m.albedo[1], m.albedo[2]);
        }
    }
    if (mat_has_roughness(p.type)) {
        if (m.roughnessTexture >= 0) {
            const DevTexture& tx = S.textures[m.roughnessTexture];
            const float4 t = tex_sample(tx, u, v);
            p.rx = t.x;
            p.ry = tx.channels > 1 ? t.y : t.x;
        } else {
            p.rx = m.roughness[0];
            p.ry = m.roughness[1];
        }
    }
    if (p.type == kMatEmissive)
        p.radiance = v3(m.radiance[0] * m.scale[0], m.radiance[1] * m.scale[1], m.radiance[2] * m.scale[2]);
    return p;
}

// emissive.hpp:39-44
__device__ __forceinline__ V3 emission(const MatParams& m, V3 geoN, V3 excident) {
    if (m.type != kMatEmissive) return v3(0, 0, 0);
    const float cosOut = dot(geoN, excident);
    if (cosOut <= 0.0f) return v3(0, 0, 0);
    return m.radiance;
}

// microfacet_base.hpp:15-27 via Boundary::get_halfTS (:45-52)
__device__ __forceinline__ V3 boundary_half(const Medium& incident, const Medium& other, V3 in, V3 ex) {
    const float nEx = in.z * ex.z > 0.0f ? incident.ior : other.ior;
    const V3 h = smul(incident.ior, in) + smul(nEx, ex);
    const float l = len(h) * sgnf(h.z);
    return v3(sdiv(h.x, l), sdiv(h.y, l), sdiv(h.z, l));
}

__device__ __forceinline__ float g_vcavity(float wDotH, float wDotN, float hDotN) {  // :73-80
    if (wDotH * wDotN <= 0.0f) return 0.0f;
    return mn(1.0f, 2.0f * fabsf(sdiv(hDotN * wDotN, wDotH)));
}
__device__ __forceinline__ float g_smith(float wDotH, V3 w, float rx, float ry) {  // :83-102, GGX
    if (wDotH * w.z <= 0.0f) return 0.0f;
    const float oneOverASquared = sdiv(sq(w.x * rx) + sq(w.y * ry), w.z * w.z);
    return 2.0f / (1.0f + fsqrt(1.f + oneOverASquared));
}
__device__ __forceinline__ float modify_g(float g) {  // :113-120
    return g < 0.11f ? fpow(g + 0.74299714456847f, 8.0f) - 0.092874643071059f
                     : ((1.139400367544197f * g - 3.317658394008854f) * g + 3.317115685385115f) * g -
                           0.13885765892046f;
}
__device__ __forceinline__ float g_reflection(int shadowing, float gi, float ge) {
    return shadowing == kShadowingSmith ? gi * ge : modify_g(mn(gi, ge));
}
__device__ __forceinline__ float g_transmission(int shadowing, float gi, float ge) {
    return shadowing == kShadowingSmith ? gi * ge : mx(0.0f, modify_g(gi) + modify_g(ge) - 1.0f);
}
__device__ __forceinline__ float g_term(const MatParams& p, float wDotH, V3 w, float hDotN) {
    return p.shadowing == kShadowingSmith ? g_smith(wDotH, w, p.rx, p.ry) : g_vcavity(wDotH, w.z, hDotN);
}

// eval_ndf (microfacet_base.hpp:281-310)
__device__ __forceinline__ float eval_ndf(const MatParams& p, V3 h) {
    const float rx = sdiv(h.x, p.rx);
    const float ry = sdiv(h.y, p.ry);
    const float cz2 = h.z * h.z;
    if (p.ndf == kNdfGgx) {
        const float norm = mx(kPi * p.rx * p.ry, 1e-30f);
        const float tmp = rx * rx + ry * ry + cz2;
        return sdiv(1.0f, norm * tmp * tmp);
    }
    const float norm = kPi * p.rx * p.ry;
    const float e = fexp(-(rx * rx + ry * ry) / cz2);
    return sdiv(e, norm * cz2 * cz2);
}

// Half vector and its pdf for TORRANCE / WALTER / MICROFACET sampling:
// Smith visible normals (GGX, microfacet_base.hpp:213-278) or V-cavity
// (sample_ndf :312-366 + sample_visible_normal_vcavity :137-156).
__device__ V3 sample_half(const MatParams& p, V3 in, float u0, float u1, u64& rnd, float& cavityPdf, float& iDotH) {
    if (p.shadowing == kShadowingSmith) {
        V3 st = normalize(in * v3(p.rx, p.ry, 1.0f));
        if (in.z < 0.0f) st = neg(st);
        const float sinTheta = fsqrt((1.0f - st.z) * (1.0f + st.z));
        const float tanTheta = sinTheta / st.z;
        const float cosPhi = sinTheta > 1e-5f ? st.x / sinTheta : 0.0f;
        const float sinPhi = sinTheta > 1e-5f ? st.y / sinTheta : 0.0f;
        float sx, sy;
        if (tanTheta < 0.0001f) {
            const float r = fsqrt(sdiv(u0, (1.f - u0)));
            const float phi = 6.28318530718f * u1;
            sx = r * fcos(phi);
            sy = r * fsin(phi);
        } else {
            const float a1 = 1.0f / tanTheta;
            const float g1 = 2.f / (1.f + fsqrt(1.0f + 1.0f / (a1 * a1)));
            const float a = 2.0f * u0 / g1 - 1.0f;
            const float tmp = sdiv(1.0f, a * a - 1.0f);
            const float d = tmp * fsqrt(tanTheta * tanTheta - (a * a - tanTheta * tanTheta) / tmp);
            const float x1 = tanTheta * tmp - d;
            const float x2 = tanTheta * tmp + d;
            sx = (a < 0.f || x2 > 1.0f / tanTheta) ? x1 : x2;
            const float s = (rnd & (1ull << 63)) ? 1.0f : -1.0f;
            rnd *= 2;
            const float z = sdiv((u1 * (u1 * (u1 * 0.27385f - 0.73369f) + 0.46341f)),
                                 (u1 * (u1 * (u1 * 0.093073f + 0.309420f) - 1.000000f) + 0.597999f));
            sy = s * z * fsqrt(1.0f + sx * sx);
        }
        const float rot = cosPhi * sx - sinPhi * sy;
        sy = sinPhi * sx + cosPhi * sy;
        sx = rot;
        sx *= p.rx;
        sy *= p.ry;
        const V3 h = normalize(v3(-sx, -sy, 1.0f));
        cavityPdf = eval_ndf(p, h);
        cavityPdf *= h.z;
        iDotH = dot(in, h);
        return h;
    }
    V3 cav;
    if (p.ndf == kNdfGgx) {
        const float phi = 2.0f * kPi * u0;
        const float e = fsqrt(u1 / (1.0f - u1));
        float slopeX = e * fcos(phi);
        float slopeY = e * fsin(phi);
        const float norm = kPi * p.rx * p.ry;
        const float tmp = 1.0f + slopeX * slopeX + slopeY * slopeY;
        slopeX *= p.rx;
        slopeY *= p.ry;
        cav = normalize(v3(-slopeX, -slopeY, 1.0f));
        cavityPdf = 1.0f / mx(norm * tmp * tmp * cav.z * cav.z * cav.z, 1e-20f);
    } else {
        const float phi = 2 * kPi * u0;
        const float xi = u1 + 1e-20f;
        const float e = fsqrt(-flog(xi));
        const float slopeX = e * fcos(phi);
        const float slopeY = e * fsin(phi);
        cav = normalize(v3(-p.rx * slopeX, -p.ry * slopeY, 1.0f));
        cavityPdf = xi / mx(kPi * p.rx * p.ry * cav.z * cav.z * cav.z, 1e-20f);
    }
    const V3 adj = v3(-cav.x, -cav.y, cav.z);
    const float iDotHOut = dot(in, cav);
    const float iDotHadj = dot(in, adj);
    const float t = in.z > 0.0f ? mx(0.0f, iDotHadj) / (mx(0.0f, iDotHOut) + mx(0.0f, iDotHadj))
                                : mx(0.0f, -iDotHadj) / (mx(0.0f, -iDotHOut) + mx(0.0f, -iDotHadj));
    const u64 probAdj = percentage_of(~0ull, t);
    if (rnd < probAdj) {
        rnd = rescale_sample(rnd, 0, probAdj - 1);
        iDotH = iDotHadj;
        return adj;
    }
    rnd = rescale_sample(rnd, probAdj, ~0ull);
    iDotH = iDotHOut;
    return cav;
}

// blend_fresnel.hpp:16-23
__device__ __forceinline__ float fresnel_dielectric(float n_i, float n_t, float cosIAbs, float& cosTAbs) {
    const float eta = sq(n_i / n_t);
    cosTAbs = fsqrt(mx(0.0f, 1.0f - eta * (1.0f - cosIAbs * cosIAbs)));
    const float rParl = sdiv(n_t * cosIAbs - n_i * cosTAbs, n_t * cosIAbs + n_i * cosTAbs);
    const float rPerp = sdiv(n_t * cosTAbs - n_i * cosIAbs, n_t * cosTAbs + n_i * cosIAbs);
    return 0.5f * (rParl * rParl + rPerp * rPerp);
}

// Tangent-space BSDF sample of lambert.hpp:18-35, microfacet_specular.hpp:26-83,
// microfacet_full.hpp:35-126. Returns false for an invalid sample.
__device__ bool sample_ts(const MatParams& p, V3 in, float u0, float u1, u64 rnd, V3& throughput, V3& ex,
                          float& pdfF, bool adjoint = false) {
    if (p.type == kMatLambert) {
        const float cosT = fsqrt(u0);
        const float sinT = fsqrt(1.0f - u0);
        const float phi = u1 * 2 * kPi;
        const float sg = sgnf(in.z);
        ex = v3(sinT * fsin(phi) * sg, sinT * fcos(phi) * sg, cosT * sg);
        throughput = p.albedo;
        pdfF = cosT / kPi;
        return true;
    }
    if (!mat_has_roughness(p.type)) return false;
    float cavityPdf, iDotH;
    const V3 h = sample_half(p, in, u0, u1, rnd, cavityPdf, iDotH);
    if (p.type == kMatTorrance) {
        ex = smul(2.0f * iDotH, h) - in;
        const float ge = g_term(p, iDotH, ex, h.z);
        const float gi = g_term(p, iDotH, in, h.z);
        const float g = g_reflection(p.shadowing, gi, ge);
        if (ge == 0.0f || gi == 0.0f) return false;
        throughput = p.albedo * sdiv(g, gi);
        pdfF = cavityPdf * sdiv(gi, fabsf(4.0f * in.z * h.z));
        return true;
    }
    const Medium& incident = p.medium(in.z);
    const Medium& other = p.medium(-in.z);
    const float n_i = incident.ior;
    const float n_t = other.ior;
    const float eta = n_i / n_t;
    const float iDotHabs = fabsf(iDotH);
    if (p.type == kMatWalter) {  // microfacet_refractive.hpp:26-121: refraction, reflection only when total
        const float t = 1.0f - eta * eta * (1.0f - iDotHabs * iDotHabs);
        float eDotH, eDotHabs;
        if (t <= 0.0f) {
            ex = smul(2.0f * iDotH, h) - in;
            eDotH = iDotH;
            eDotHabs = iDotHabs;
        } else {
            eDotHabs = fsqrt(t);
            eDotH = eDotHabs * -sgnf(iDotH);
            ex = smul(sgnf(iDotH) * (eta * iDotHabs - eDotHabs), h) - smul(eta, in);
        }
        const float ge = g_term(p, eDotH, ex, h.z);
        const float gi = g_term(p, iDotH, in, h.z);
        if (ge == 0.0f || gi == 0.0f) return false;
        float tp;
        if (t <= 0.0f) {
            tp = sdiv(g_reflection(p.shadowing, gi, ge), gi);
            pdfF = cavityPdf * sdiv(gi, fabsf(4.0f * in.z * h.z));
        } else {
            tp = sdiv(g_transmission(p.shadowing, gi, ge), gi);
            if (adjoint) tp *= eta * eta;  // light sub-paths (microfacet_refractive.hpp:109-110)
            const float common = cavityPdf * sdiv(iDotHabs * eDotHabs, sq(n_i * iDotH + n_t * eDotH) * h.z);
            pdfF = common * sdiv(gi * n_t * n_t, fabsf(in.z));
        }
        throughput = v3(tp, tp, tp);
        return true;
    }
    float cosTAbs;
    const float f = fresnel_dielectric(n_i, n_t, iDotHabs, cosTAbs);
    const bool reflect = ((float)rnd / (float)(~0ull - 1)) < f;
    float eDotH;
    if (reflect) {
        ex = smul(2.0f * iDotH, h) - in;
        eDotH = iDotH;
    } else {
        eDotH = cosTAbs * -sgnf(iDotH);
        ex = smul(sgnf(iDotH) * (eta * iDotHabs - cosTAbs), h) - smul(eta, in);
    }
    const float ge = g_term(p, eDotH, ex, h.z);
    const float gi = g_term(p, iDotH, in, h.z);
    if (ge == 0.0f || gi == 0.0f) return false;
    float tp;
    if (reflect) {
        tp = sdiv(g_reflection(p.shadowing, gi, ge), gi);
        const float common = cavityPdf * sdiv(f, 4.0f * h.z);
        pdfF = common * sdiv(gi, fabsf(in.z));
    } else {
        tp = sdiv(g_transmission(p.shadowing, gi, ge), gi);
        if (adjoint) tp *= eta * eta;  // light sub-paths (microfacet_full.hpp:114-115)
        const float common = cavityPdf * (1.0f - f) * sdiv(iDotHabs * cosTAbs, sq(n_i * iDotH + n_t * eDotH) * h.z);
        pdfF = common * sdiv(gi * n_t * n_t, fabsf(in.z));
    }
    throughput = v3(tp, tp, tp);
    return true;
}

// Tangent-space BSDF value and forward pdf (lambert.hpp:37-49,
// microfacet_specular.hpp:86-122, microfacet_full.hpp:129-187).
__device__ V3 evaluate_ts(const MatParams& p, V3 in, V3 ex, float& pdfF) {
    pdfF = 0.0f;
    if (p.type == kMatLambert) {
        if (in.z * ex.z < 0.0f) return v3(0, 0, 0);
        pdfF = fabsf(ex.z) / kPi;
        return p.albedo / kPi;
    }
    if (!mat_has_roughness(p.type)) return v3(0, 0, 0);
    const bool isReflection = in.z * ex.z > 0.0f;
    if (p.type == kMatTorrance && in.z * ex.z < 0.0f) return v3(0, 0, 0);
    const Medium& incident = p.medium(in.z);
    const Medium& other = p.medium(-in.z);
    const V3 h = boundary_half(incident, other, in, ex);
    const float iDotH = dot(in, h);
    if (p.type == kMatTorrance) {
        const float ge = g_term(p, iDotH, ex, h.z);
        const float gi = g_term(p, iDotH, in, h.z);
        const float g = g_reflection(p.shadowing, gi, ge);
        if (ge == 0.0f || gi == 0.0f) return v3(0, 0, 0);
        const float d = eval_ndf(p, h);
        pdfF = sdiv(gi * d, fabsf(4.0f * in.z));
        return p.albedo * sdiv(g * d, fabsf(4.0f * in.z * ex.z));
    }
    const float eDotH = dot(ex, h);
    const float n_i = incident.ior;
    const float n_t = other.ior;
    // Walter (microfacet_refractive.hpp:124-187): a reflection only under total internal reflection
    if (p.type == kMatWalter && isReflection && sq(n_i / n_t) * (1.0f - iDotH * iDotH) < 1.0f) return v3(0, 0, 0);
    const float ge = g_term(p, eDotH, ex, h.z);
    const float gi = g_term(p, iDotH, in, h.z);
    if (ge == 0.0f || gi == 0.0f) return v3(0, 0, 0);
    const float d = eval_ndf(p, h);
    if (p.type == kMatWalter) {
        float val;
        if (isReflection) {
            val = sdiv(g_reflection(p.shadowing, gi, ge) * d, 4.0f * in.z * ex.z);
            pdfF = sdiv(gi * d, 4.0f * fabsf(in.z));
        } else {
            const float common = sdiv(fabsf(d * iDotH * eDotH), sq(n_i * iDotH + n_t * eDotH));
            val = g_transmission(p.shadowing, gi, ge) * common * sdiv(n_t * n_t, fabsf(in.z * ex.z));
            pdfF = gi * common * sdiv(n_t * n_t, fabsf(in.z));
        }
        return v3(val, val, val);
    }
    float cosTAbs;
    const float f = fresnel_dielectric(n_i, n_t, fabsf(iDotH), cosTAbs);
    float val;
    if (isReflection) {
        const float common = d * f / 4.0f;
        val = sdiv(g_reflection(p.shadowing, gi, ge) * common, in.z * ex.z);
        pdfF = sdiv(gi * common, fabsf(in.z));
    } else {
        const float common = sdiv((1 - f) * fabsf(d * iDotH * eDotH), sq(n_i * iDotH + n_t * eDotH));
        val = g_transmission(p.shadowing, gi, ge) * common * sdiv(n_t * n_t, fabsf(in.z * ex.z));
        pdfF = gi * common * sdiv(n_t * n_t, fabsf(in.z));
    }
    return v3(val, val, val);
}

struct TS {
    V3 shadingN, geoN, shadingTX, shadingTY;
};

constexpr float kShadingNormalEps = 1e-3f;  // SHADING_NORMAL_EPS (material_sampling.hpp:56)

// materials::sample (material_sampling.hpp:68-146): view paths, or with
// `adjoint` light sub-paths (refraction scaled by eta^2, and the shading
// normal correction). On success writes the world-space excident, throughput
// and forward pdf.
__device__ __forceinline__ bool mat_sample(const TS& ts, const MatParams& m, V3 incident, float u0, float u1, u64 i0,
                                           V3& throughput, V3& excident, float& pdfF, bool adjoint = false) {
    const float iDotG = -dot(incident, ts.geoN);
    const float iDotN = -dot(incident, ts.shadingN);
    if (iDotG * iDotN <= 0.0f) return false;
    const V3 inTS = v3(-dot(incident, ts.shadingTX), -dot(incident, ts.shadingTY), iDotN);
    V3 ex;
    if (!sample_ts(m, inTS, u0, u1, i0, throughput, ex, pdfF, adjoint)) return false;
    if (!nonzero(throughput)) return false;
    const V3 g = normalize(smul(ex.x, ts.shadingTX) + smul(ex.y, ts.shadingTY) + smul(ex.z, ts.shadingN));
    const float eDotG = dot(g, ts.geoN);
    if (eDotG * ex.z <= 0.0f) return false;
    excident = g;
    pdfF = mn(pdfF, 1e9f);
    if (adjoint)
        throughput = throughput * ((kShadingNormalEps + fabsf(iDotN * eDotG)) / (kShadingNormalEps + fabsf(iDotG * ex.z)));
    return true;
}

// materials::evaluate (material_sampling.hpp:160-237), connections: value,
// cosOut = |eDotN| and forward pdf; value 0 for rejected pairs.
__device__ __forceinline__ V3 mat_evaluate(const TS& ts, const MatParams& m, V3 incident, V3 excident, float& cosOut,
                                           float& pdfF) {
    cosOut = 0.0f;
    pdfF = 0.0f;
    const float iDotN = -dot(incident, ts.shadingN);
    const float eDotN = dot(excident, ts.shadingN);
    if (!m.reflective() && iDotN * eDotN > 0.0f) return v3(0, 0, 0);
    if (!m.refractive() && iDotN * eDotN < 0.0f) return v3(0, 0, 0);
    const float iDotG = -dot(incident, ts.geoN);
    const float eDotG = dot(excident, ts.geoN);
    if (eDotG * eDotN <= 0.0f || iDotG * iDotN <= 0.0f) return v3(0, 0, 0);
    const V3 inTS = v3(-dot(incident, ts.shadingTX), -dot(incident, ts.shadingTY), iDotN);
    const V3 exTS = v3(dot(excident, ts.shadingTX), dot(excident, ts.shadingTY), eDotN);
    float pdf;
    const V3 value = evaluate_ts(m, inTS, exTS, pdf);
    if (!nonzero(value)) return v3(0, 0, 0);
    cosOut = fabsf(eDotN);
    pdfF = mn(pdf, 1e9f);
    return value;
}

// materials::evaluate with adjoint = true (material_sampling.hpp:160-237), the
// connections of light sub-paths: the BSDF of the swapped pair times the
// shading normal correction; cosOut = |eDotN|; 0 for rejected pairs.
__device__ __forceinline__ V3 mat_evaluate_adjoint(const TS& ts, const MatParams& m, V3 incident, V3 excident,
                                                   float& cosOut) {
    cosOut = 0.0f;
    const float iDotN = -dot(incident, ts.shadingN);
    const float eDotN = dot(excident, ts.shadingN);
    if (!m.reflective() && iDotN * eDotN > 0.0f) return v3(0, 0, 0);
    if (!m.refractive() && iDotN * eDotN < 0.0f) return v3(0, 0, 0);
    const float iDotG = -dot(incident, ts.geoN);
    const float eDotG = dot(excident, ts.geoN);
    if (eDotG * eDotN <= 0.0f || iDotG * iDotN <= 0.0f) return v3(0, 0, 0);
    const V3 inTS = v3(-dot(incident, ts.shadingTX), -dot(incident, ts.shadingTY), iDotN);
    const V3 exTS = v3(dot(excident, ts.shadingTX), dot(excident, ts.shadingTY), eDotN);
    float pdf;
    V3 value = evaluate_ts(m, exTS, inTS, pdf);
    if (!nonzero(value)) return v3(0, 0, 0);
    value = value * ((kShadingNormalEps + fabsf(iDotN * eDotG)) / (kShadingNormalEps + fabsf(iDotG * eDotN)));
    cosOut = fabsf(eDotN);
    return value;
}

// ------------------------------------------------------------------ roughness derivatives
// Logarithmic derivatives with respect to the roughness (rx, ry) for the
// adjoint. The BSDF value of TORRANCE / WALTER / MICROFACET is G * D * (terms free of
// roughness) and its pdf is G1(in) * D * (terms free of roughness)
// (evaluate_ts), so d ln value = d ln G + d ln D and d ln pdf = d ln G1(in)
// + d ln D. The V-cavity G does not depend on the roughness; Smith's is a
// product of G1 terms.

// d ln D / d(rx, ry) of eval_ndf at h.
__device__ __forceinline__ void dln_ndf(const MatParams& p, V3 h, float& dx, float& dy) {
    const float x = h.x / p.rx, y = h.y / p.ry;
    const float cz2 = h.z * h.z;
    if (p.ndf == kNdfGgx) {
        // D = 1 / (pi rx ry t^2), t = x^2 + y^2 + cz2
        const float t = x * x + y * y + cz2;
        dx = -1.0f / p.rx + 4.0f * x * x / (p.rx * t);
        dy = -1.0f / p.ry + 4.0f * y * y / (p.ry * t);
    } else {
        // D = exp(-(x^2 + y^2) / cz2) / (pi rx ry cz2^2)
        dx = 2.0f * x * x / (p.rx * cz2) - 1.0f / p.rx;
        dy = 2.0f * y * y / (p.ry * cz2) - 1.0f / p.ry;
    }
}

// d ln G1 / d(rx, ry) of g_smith for direction w (zero for the V-cavity).
__device__ __forceinline__ void dln_g1(const MatParams& p, V3 w, float& dx, float& dy) {
    dx = dy = 0.0f;
    if (p.shadowing != kShadowingSmith) return;
    // G1 = 2 / (1 + s), s = sqrt(1 + A), A = ((wx rx)^2 + (wy ry)^2) / wz^2
    const float A = (sq(w.x * p.rx) + sq(w.y * p.ry)) / (w.z * w.z);
    const float s = fsqrt(1.0f + A);
    const float k = -1.0f / (2.0f * s * (1.0f + s));
    dx = k * 2.0f * w.x * w.x * p.rx / (w.z * w.z);
    dy = k * 2.0f * w.y * w.y * p.ry / (w.z * w.z);
}

// d ln value and d ln pdf of mat_evaluate(ts, m, incident, excident) with
// respect to (rx, ry); false where the value is zero, the material has no
// roughness, or the pdf is clamped (then the pdf derivative is zero).
__device__ __forceinline__ bool dln_mat_evaluate(const TS& ts, const MatParams& m, V3 incident, V3 excident,
                                                 float dv[2], float dp[2]) {
    dv[0] = dv[1] = dp[0] = dp[1] = 0.0f;
    if (!mat_has_roughness(m.type)) return false;
    if (!(m.rx > 0.0f && m.ry > 0.0f)) return false;
    float cosOut, pdfF;
    const V3 value = mat_evaluate(ts, m, incident, excident, cosOut, pdfF);
    if (!nonzero(value)) return false;
    const V3 in = v3(-dot(incident, ts.shadingTX), -dot(incident, ts.shadingTY), -dot(incident, ts.shadingN));
    const V3 ex = v3(dot(excident, ts.shadingTX), dot(excident, ts.shadingTY), dot(excident, ts.shadingN));
    const V3 h = boundary_half(m.medium(in.z), m.medium(-in.z), in, ex);
    float ddx, ddy, gix, giy, gex, gey;
    dln_ndf(m, h, ddx, ddy);
    dln_g1(m, in, gix, giy);
    dln_g1(m, ex, gex, gey);
    dv[0] = ddx + gix + gex;
    dv[1] = ddy + giy + gey;
    if (pdfF < 1e9f) {
        dp[0] = ddx + gix;
        dp[1] = ddy + giy;
    }
    return true;
}

// ------------------------------------------------------------------ background
__device__ __forceinline__ float box_diag_len(const float* box) {
    return len(v3(box[3] - box[0], box[4] - box[1], box[5] - box[2]));
}

// lower_bound over a SAT row or column (texture_sampling.hpp:21-37)
__device__ __forceinline__ int sat_lower_bound(const float* sat, int width, float val, int max, int constCoord,
                                               bool row) {
    int left = 0;
    int count = max + 1;
    while (count > 0) {
        const int step = count / 2;
        const int curr = left + step;
        const float x = row ? sat[(size_t)curr * width + constCoord] : sat[(size_t)constCoord * width + curr];
        if (x < val) {
            left = curr + 1;
            count -= step + 1;
        } else {
            count = step;
        }
    }
    return left;
}

// sample_light_dir (light_sampling.hpp:225-290) for COLORED and polar ENVMAP
// backgrounds, with importance_sample_texture (texture_sampling.hpp:119-155).
__device__ void background_sample_dir(const PtScene& S, float u0, float u1, V3& dir, float& pdf, V3& radiance) {
    const DevBackground& bg = S.background;
    if (bg.type == 0) {
        const float cosTheta = u0 * 2.0f - 1.0f;
        const float sinTheta = fsqrt((1.0f - cosTheta) * (1.0f + cosTheta));
        const float phi = u1 * 2.0f * kPi;
        dir = v3(sinTheta * fsin(phi), sinTheta * fcos(phi), cosTheta);
        pdf = 1.f / (4.f * kPi);
        radiance = v3(bg.color[0], bg.color[1], bg.color[2]) * v3(bg.scale[0], bg.scale[1], bg.scale[2]);
        return;
    }
    const DevTexture& env = S.textures[bg.envmapTexture];
    const float* sat = bg.sat;
    const int W = env.width, H = env.height, bx = W - 1, by = H - 1;
    const float highestRowwise = sat[(size_t)by * W + bx];
    const float x = highestRowwise * u0;
    const int row = sat_lower_bound(sat, W, x, by, bx, true);
    const float vr0 = row == 0 ? 0.0f : sat[(size_t)(row - 1) * W + bx];
    const float vr1 = sat[(size_t)row * W + bx];
    const float rowVal = ((float)row + (x - vr0) / (vr1 - vr0)) / (float)H;
    const float y = (vr1 - vr0) * u1;
    const int column = sat_lower_bound(sat, W, y, bx, row, false);
    const float vc0 = column == 0 ? 0.f : sat[(size_t)row * W + column - 1];
    const float vc1 = column == bx ? highestRowwise : sat[(size_t)row * W + column];
    const float columnVal = ((float)column + (y - vc0) / (vc1 - vc0)) / (float)W;
    const float texPdf = (vc1 - vc0) * (float)W * (float)H / highestRowwise;
    const float phi = columnVal * 2.f * kPi;
    const float theta = (1.f - rowVal) * kPi;
    const float sinTheta = fsin(theta);
    const float cosTheta = fcos(theta);
    dir = v3(sinTheta * fcos(phi), cosTheta, sinTheta * fsin(phi));
    pdf = texPdf / (2.f * kPi * kPi * sinTheta);
    const float4 t = texel(env, column, row);
    radiance = v3(t.x, t.y, t.z) * v3(bg.scale[0], bg.scale[1], bg.scale[2]);
}

// evaluate_background (light_sampling.hpp:467-515): radiance and pdf.back.
__device__ V3 evaluate_background(const PtScene& S, V3 direction, float& pdfBack) {
    const DevBackground& bg = S.background;
    if (bg.type == 0) {
        pdfBack = 1.0f / (4.0f * kPi);
        return v3(bg.color[0], bg.color[1], bg.color[2]) * v3(bg.scale[0], bg.scale[1], bg.scale[2]);
    }
    const DevTexture& env = S.textures[bg.envmapTexture];
    const float v0 = facos(direction.y) / kPi;
    const float u = fatan2(direction.z, direction.x) / (kPi * 2.0f);
    const float halfTexel = 0.5f / (float)env.height;
    const float v = clampf(1.0f - v0, halfTexel, 1.0f - halfTexel);
    const float4 t = tex_sample(env, u, v);
    const V3 radiance = v3(t.x, t.y, t.z);
    const int W = env.width, H = env.height;
    const float cdf = bg.sat[(size_t)(H - 1) * W + (W - 1)] / (float)(H * W);
    float pdfScale = dot(v3(0.212671f, 0.715160f, 0.072169f), radiance);
    const int pixelY = ifloor(v * (float)H);
    const float sinPixel = fsin(kPi * (float)((float)pixelY + 0.5f) / (float)H);
    const float sinJac = fsqrt(1.0f - direction.y * direction.y);
    pdfScale *= sinPixel / (2.0f * kPi * kPi * sinJac);
    pdfBack = pdfScale / cdf;
    return radiance * v3(bg.scale[0], bg.scale[1], bg.scale[2]);
}

// ------------------------------------------------------------------ lights
struct Nee {
    V3 position;
    float creationPdf;
    V3 dir;
    float cosOut;
    V3 diffIrradiance;
    V3 geoNormal;
    float distSq;
    int lightMaterial;  // material of an area light, -1 for point lights and the background
};

__device__ __forceinline__ float4 ld4(const uint8_t* p) { return *reinterpret_cast<const float4*>(p); }

// AreaLightQuad record (lights.hpp:108-124, 96 bytes): posV[4] = v0, v3 - v0,
// v1 - v0, v2 - v3 - (v1 - v0) at 0, uvV[4] at 48, material at 80.
__device__ __forceinline__ void quad_record(const uint8_t* light, V3 pv[4]) {
    const float4 q0 = ld4(light), q1 = ld4(light + 16), q2 = ld4(light + 32);
    pv[0] = v3(q0.x, q0.y, q0.z);
    pv[1] = v3(q0.w, q1.x, q1.y);
    pv[2] = v3(q1.z, q1.w, q2.x);
    pv[3] = v3(q2.y, q2.z, q2.w);
}

// sample_light_pos + connect_light of a quad (light_sampling.hpp:142-171,
// 376-399): (u0, u1) are the patch coordinates of the sample.
__device__ Nee connect_quad(const PtScene& S, const uint8_t* light, V3 pos, float u0, float u1) {
    V3 pv[4];
    quad_record(light, pv);
    const float4 w0 = ld4(light + 48), w1 = ld4(light + 64);
    const int matIdx = (int)(*reinterpret_cast<const uint16_t*>(light + 80));
    const V3 tangentX = pv[1] + smul(u0, pv[3]);
    const V3 tangentY = pv[2] + smul(u1, pv[3]);
    V3 normal = cross(tangentY, tangentX);
    const float area = len(normal);
    normal = normal / area;
    const V3 position = pv[0] + tangentX * u1 + pv[2] * u0;
    // uvV[0..3]: (w0.x, w0.y), (w0.z, w0.w), (w1.x, w1.y), (w1.z, w1.w)
    const float lu = w0.x + w0.z * u1 + w1.x * u0 + w1.z * (u0 * u1);
    const float lv = w0.y + w0.w * u1 + w1.y * u0 + w1.w * (u0 * u1);
    const MatParams m = fetch(S, matIdx, lu, lv);
    const V3 intensity = emission(m, v3(0, 0, 1), v3(0, 0, 1)) * area;
    V3 d = position - pos;
    const float distSq = dot(d, d) + kDistSqEps;
    const float dist = fsqrt(distSq);
    d = d / dist;
    const float cosOut = dot(normal, neg(d));
    V3 value = v3(0, 0, 0);
    float vc = 0.0f;
    if (cosOut > 0.0f) {
        value = intensity;
        vc = cosOut;
    }
    Nee n;
    n.position = position;
    n.creationPdf = 1.0f / area;
    n.dir = d;
    n.cosOut = vc;
    n.diffIrradiance = v3(value.x / distSq, value.y / distSq, value.z / distSq) * transmission(m.outer, dist);
    n.geoNormal = normal;
    n.distSq = distSq;
    n.lightMaterial = matIdx;
    return n;
}

// connect_light of a sphere (light_sampling.hpp:400-432): a point of the cap
// visible from pos, uniform in area (a uniform cone of directions from the
// center, sample_cone_uniform, sampling.hpp:112-119). Record
// (AreaLightSphere, lights.hpp:125-137): center, radius, material at 16.
__device__ Nee connect_sphere(const PtScene& S, const uint8_t* light, V3 pos, float u0, float u1) {
    const float4 q0 = ld4(light);
    const int matIdx = (int)(*reinterpret_cast<const uint16_t*>(light + 16));
    const V3 c = v3(q0.x, q0.y, q0.z);
    const float radius = q0.w;
    Nee n;
    n.lightMaterial = matIdx;
    V3 centerDir = pos - c;
    const float cd = len(centerDir);
    if (cd <= radius) {  // inside: no estimate
        n.position = v3(0, 0, 0);
        n.creationPdf = 0.0f;
        n.dir = v3(0, 0, 0);
        n.cosOut = 0.0f;
        n.diffIrradiance = v3(0, 0, 0);
        n.geoNormal = v3(0, 0, 0);
        n.distSq = 0.0f;
        return n;
    }
    centerDir = centerDir / cd;
    const float cosSphere = radius / cd;
    const float cosTheta = (1.f - u0) + u0 * cosSphere;
    const float sinTheta = fsqrt(1.f - cosTheta * cosTheta);
    const float phi = 2.f * u1 * kPi;
    const float lx = fcos(phi) * sinTheta, ly = fsin(phi) * sinTheta;
    const float conePdf = 1.f / (2.f * kPi * (1.f - cosSphere));
    const V3 tangentX = normalize(perpendicular(centerDir));
    const V3 tangentY = cross(centerDir, tangentX);
    const V3 globalDir = smul(lx, tangentX) + smul(ly, tangentY) + smul(cosTheta, centerDir);
    const V3 surfPos = c + smul(radius, globalDir);
    V3 cdir = surfPos - pos;
    const float cDistSq = dot(cdir, cdir) + kDistSqEps;
    const float cDist = fsqrt(cDistSq);
    cdir = cdir / cDist;
    // direction_to_uv (textures/interface.hpp:169-174)
    const float tv = facos(globalDir.y) / kPi;
    const float tu = fatan2(globalDir.z, globalDir.x) / (kPi * 2.0f);
    const MatParams m = fetch(S, matIdx, tu, 1.0f - tv);
    V3 radiance = emission(m, v3(0, 0, 1), v3(0, 0, 1));
    const float sampleArea = radius * radius / conePdf;
    const float cosOut = mx(0.0f, -dot(globalDir, cdir));
    radiance = radiance * (sampleArea / cDistSq);
    radiance = radiance * transmission(m.outer, cDist);
    n.position = surfPos;
    n.creationPdf = 1.0f / sampleArea;
    n.dir = cdir;
    n.cosOut = cosOut;
    n.diffIrradiance = radiance;
    n.geoNormal = globalDir;
    n.distSq = cDistSq;
    return n;
}

// connect_light for point lights (light_sampling.hpp:308-327, vacuum),
// area triangles (:118-140, :350-375), quads and spheres (above).
__device__ Nee connect_light(const PtScene& S, uint16_t type, u32 offset, V3 pos, float u0, float u1) {
    const uint8_t* light = S.lightMem + offset;
    Nee n;
    if (type == kLightPoint) {
        const float4 a = ld4(light), b = ld4(light + 16);
        const V3 lp = v3(a.x, a.y, a.z);
        V3 d = lp - pos;
        const float distSq = dot(d, d) + kDistSqEps;
        const float dist = fsqrt(distSq);
        d = d / dist;
        n.position = lp;
        n.creationPdf = kInfPdf;
        n.dir = d;
        n.cosOut = 0.0f;
        n.diffIrradiance = v3(b.x / distSq, b.y / distSq, b.z / distSq);
        n.geoNormal = v3(0, 0, 0);
        n.distSq = distSq;
        n.lightMaterial = -1;
        return n;
    }
    if (type == kLightAreaQuad) return connect_quad(S, light, pos, u0, u1);
    if (type == kLightAreaSphere) return connect_sphere(S, light, pos, u0, u1);
    const float4 q0 = ld4(light), q1 = ld4(light + 16), q2 = ld4(light + 32), q3 = ld4(light + 48);
    const V3 p0 = v3(q0.x, q0.y, q0.z), e1 = v3(q0.w, q1.x, q1.y), e2 = v3(q1.z, q1.w, q2.x);
    const int matIdx = (int)(__float_as_uint(q2.y) & 0xFFFFu);
    V3 normal = cross(e1, e2);
    const float area2 = len(normal);
    normal = normal / area2;
    const float area = area2 * 0.5f;
    const float s = 1.f - fsqrt(1.f - u0);
    const float t = (1.f - s) * u1;
    const V3 position = p0 + e1 * s + e2 * t;
    // uvV[0..2] at byte 40: (q2.z, q2.w), (q3.x, q3.y), (q3.z, q3.w)
    const float lu = q2.z + q3.x * s + q3.z * t;
    const float lv = q2.w + q3.y * s + q3.w * t;
    const MatParams m = fetch(S, matIdx, lu, lv);
    const V3 rad = emission(m, v3(0, 0, 1), v3(0, 0, 1));
    const float pdf = 1.0f / area;
    const V3 intensity = rad * area;
    V3 d = position - pos;
    const float distSq = dot(d, d) + kDistSqEps;
    const float dist = fsqrt(distSq);
    d = d / dist;
    const float cosOut = dot(normal, neg(d));
    V3 value = v3(0, 0, 0);
    float vc = 0.0f;
    if (cosOut > 0.0f) {
        value = intensity;
        vc = cosOut;
    }
    n.position = position;
    n.creationPdf = pdf;
    n.dir = d;
    n.cosOut = vc;
    n.diffIrradiance = v3(value.x / distSq, value.y / distSq, value.z / distSq) * transmission(m.outer, dist);
    n.geoNormal = normal;
    n.distSq = distSq;
    n.lightMaterial = matIdx;
    return n;
}

// get_center (light_tree.hpp:284-296, lights.hpp:210-228) of a child of a
// light tree node: an internal node's stored center, a point light's
// position, a triangle's p0 + (e1 + e2) / 3.
__device__ __forceinline__ V3 light_center(const PtScene& S, u32 offset, uint16_t type) {
    const uint8_t* p = S.lightMem + offset;
    if (type == kLightInternalNode) {
        const float4 b = ld4(p + 16);
        return v3(b.y, b.z, b.w);
    }
    const float4 q0 = ld4(p);
    if (type == kLightPoint || type == kLightAreaSphere) return v3(q0.x, q0.y, q0.z);
    if (type == kLightAreaQuad) {  // posV[0] + (posV[1] + posV[2]) / 2 + posV[3] / 4
        V3 pv[4];
        quad_record(p, pv);
        return pv[0] + (pv[1] + pv[2]) * 0.5f + pv[3] * 0.25f;
    }
    const float4 q1 = ld4(p + 16), q2 = ld4(p + 32);
    const V3 p0 = v3(q0.x, q0.y, q0.z), e1 = v3(q0.w, q1.x, q1.y), e2 = v3(q1.z, q1.w, q2.x);
    return p0 + (e1 + e2) / 3.0f;
}

// Probability of the left child of a node {a, b} (the node's two float4s):
// guide_flux, or guide_flux_pos with the "NEE position guide" parameter
// (light_tree_sampling.hpp:58-71, :218-226, :356-364).
__device__ __forceinline__ float guide_left(const PtScene& S, float4 a, float4 b, V3 ref) {
    float lFlux = a.x, rFlux = b.x;
    if (S.posGuide) {
        const u32 types = __float_as_uint(a.z);
        const V3 dl = light_center(S, __float_as_uint(a.y), (uint16_t)(types & 0xFFFFu)) - ref;
        const V3 dr = light_center(S, __float_as_uint(a.w), (uint16_t)(types >> 16)) - ref;
        lFlux /= dot(dl, dl);
        rFlux /= dot(dr, dr);
    }
    return lFlux / (lFlux + rFlux);
}

// lights::connect (light_tree_sampling.hpp:195-329): background first, then
// the (empty) directional subtree, then the positional subtree.
__device__ Nee connect(const PtScene& S, u64 index, u64 numIndices, u64 seed, V3 pos, float u0, float u1) {
    u64 rndChoice = numIndices > 0 ? seed + index * (~0ull / numIndices) : seed;
    if (rndChoice == ~0ull) --rndChoice;
    const float fluxSum = S.bgFlux + S.dirRootFlux + S.posRootFlux;
    const float envProb = S.bgFlux / fluxSum;
    const float dirProb = S.dirRootFlux / fluxSum;
    const float posProb = S.posRootFlux / fluxSum;
    const u64 rightEnv = percentage_of(~0ull, envProb);
    Nee n;
    float lightProb;
    if (rndChoice < rightEnv) {
        V3 dir, radiance;
        float pdf;
        background_sample_dir(S, u0, u1, dir, pdf, radiance);
        n.position = pos + dir * box_diag_len(S.sceneBox);
        n.creationPdf = pdf * fabsf(1.0f) / (kMaxSceneSize * kMaxSceneSize);
        n.dir = dir;
        n.cosOut = 1.0f;
        n.diffIrradiance = radiance / pdf;
        n.geoNormal = v3(0, 0, 0);
        n.distSq = kMaxSceneSize * kMaxSceneSize;
        n.lightMaterial = -1;
        lightProb = envProb;
    } else {
        u64 left = percentage_of(~0ull, envProb + dirProb);
        u64 right = ~0ull;
        lightProb = posProb;
        uint16_t type = (uint16_t)S.posRootType;
        u32 offset = 0;
        for (int guard = 0; type == kLightInternalNode && guard < 64; ++guard) {
            const float4 a = ld4(S.lightMem + offset), b = ld4(S.lightMem + offset + 16);
            const u32 types = __float_as_uint(a.z);
            const float probLeft = guide_left(S, a, b, pos);
            const u64 boundary = left + percentage_of(right - left, probLeft);
            if (rndChoice < boundary) {
                type = (uint16_t)(types & 0xFFFFu);
                offset = __float_as_uint(a.y);
                right = boundary;
                lightProb *= probLeft;
            } else {
                type = (uint16_t)(types >> 16);
                offset = __float_as_uint(a.w);
                left = boundary;
                lightProb *= (1.0f - probLeft);
            }
        }
        n = connect_light(S, type, offset, pos, u0, u1);
    }
    n.creationPdf *= lightProb;  // adjustPdf (light_tree_sampling.hpp:16-19)
    n.diffIrradiance = n.diffIrradiance / lightProb;
    return n;
}

// light_pdf (light_tree_sampling.hpp:340-413): connect pdf of a hit emitter
// seen from refPos (the previous vertex of the path); (sp0, sp1): the hit's
// surface parameters (a quad's (u, v)).
__device__ float light_pdf(const PtScene& S, int inst, int prim, V3 refPos, float sp0, float sp1) {
    const float treeFlux = S.dirRootFlux + S.posRootFlux + S.bgFlux;
    float p = S.posRootFlux / treeFlux;
    const u64 key = ((u64)(u32)inst << 32) | (u64)(u32)prim;
    int lo = 0, hi = S.mapCount - 1, found = -1;
    while (lo <= hi) {
        const int mid = (lo + hi) >> 1;
        const u64 k = S.mapKeys[mid];
        if (k == key) {
            found = mid;
            break;
        }
        if (k < key)
            lo = mid + 1;
        else
            hi = mid - 1;
    }
    if (found < 0) return 0.0f;
    u32 code = S.mapCodes[found];
    u32 offset = 0;
    uint16_t type = (uint16_t)S.posRootType;
    for (int guard = 0; type == kLightInternalNode && guard < 64; ++guard) {
        const float4 a = ld4(S.lightMem + offset), b = ld4(S.lightMem + offset + 16);
        const float pLeft = guide_left(S, a, b, refPos);
        const u32 types = __float_as_uint(a.z);
        if (code & 0x80000000u) {
            p *= (1.0f - pLeft);
            type = (uint16_t)(types >> 16);
            offset = __float_as_uint(a.w);
        } else {
            p *= pLeft;
            type = (uint16_t)(types & 0xFFFFu);
            offset = __float_as_uint(a.y);
        }
        code <<= 1;
    }
    if (type == kLightAreaQuad) {  // the local area density at (sp0, sp1)
        V3 pv[4];
        quad_record(S.lightMem + offset, pv);
        const V3 tangentX = pv[1] + smul(sp0, pv[3]);
        const V3 tangentY = pv[2] + smul(sp1, pv[3]);
        return p / len(cross(tangentY, tangentX));
    }
    if (type == kLightAreaSphere) {  // uniform over the cap visible from refPos
        const float4 q = ld4(S.lightMem + offset);
        const float cosSphere = q.w / len(refPos - v3(q.x, q.y, q.z));
        const float solidAngle = 2 * kPi * (1.0f - cosSphere);
        return p / (solidAngle * (q.w * q.w));
    }
    const float4 q0 = ld4(S.lightMem + offset), q1 = ld4(S.lightMem + offset + 16);
    const V3 e1 = v3(q0.w, q1.x, q1.y), e2 = v3(q1.z, q1.w, *reinterpret_cast<const float*>(S.lightMem + offset + 32));
    const float area = len(cross(e1, e2)) / 2.0f;
    return p / area;
}

// background_pdf (light_tree_sampling.hpp:419-423)
__device__ __forceinline__ float background_pdf(const PtScene& S, float pdfBack) {
    return pdfBack * (S.bgFlux / (S.dirRootFlux + S.posRootFlux + S.bgFlux));
}

}  // namespace shade
}  // namespace mfh
