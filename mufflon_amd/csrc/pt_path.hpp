// Path-level building blocks of the path tracer shared by the megakernel and
// the wavefront stages: the surface / void vertex made from a hit, the
// per-iteration target accumulator, emission and target contributions, next-
// event candidates and the BSDF walk. Both drivers call exactly these
// functions in the order of pt_sample (pt_common.hpp:75-156), so they produce
// the same bits.
#pragma once
#include "pt_shading.hpp"
#include "traverse.hpp"

namespace mfh {
namespace path {

using namespace shade;

__device__ __forceinline__ M34 load_m34(const float* w) {
    const float4* q = reinterpret_cast<const float4*>(w);
    const float4 a = q[0], b = q[1], c = q[2];
    return M34{{a.x, a.y, a.z, a.w, b.x, b.y, b.z, b.w, c.x, c.y, c.z, c.w}};
}

// intersection.cu:16-24
__device__ __forceinline__ V3 add_epsilon(V3 o, V3 dir, V3 geoN) {
    const V3 off = geoN * kSceneEps;
    return dot(geoN, dir) >= 0.0f ? o + off : o - off;
}

__device__ __forceinline__ V3 lod_vert(const DevPtLod& L, u32 i) {
    return v3(L.verts[3 * (size_t)i], L.verts[3 * (size_t)i + 1], L.verts[3 * (size_t)i + 2]);
}
__device__ __forceinline__ V3 lod_normal(const DevPtLod& L, u32 i) {
    return v3(L.normals[3 * (size_t)i], L.normals[3 * (size_t)i + 1], L.normals[3 * (size_t)i + 2]);
}

// Hit attributes of a sphere (intersection.cu:524-541; the sphere of
// primitive p is sphere p - (triangles + quads) of the LoD) and its frame
// (tangent_space.hpp:14-21: the geometric one). position: the world hit.
__device__ TS sphere_frame(const DevPtLod& L, const M34& M, int prim, V3 position, float& hu, float& hv) {
    const float4 sp = L.spheres[prim - L.numTris - L.numQuads];
    const V3 n = normalize(xform_point(M, position) - v3(sp.x, sp.y, sp.z));
    V3 tangentX = v3(1.0f, 0.0f, 0.0f);
    if (!(n.x == 0.0f && n.y == 0.0f)) {
        const float l2 = fsqrt(n.y * n.y + n.x * n.x);
        tangentX = v3(n.y / l2, -n.x / l2, 0.0f);
    }
    hu = fatan2(n.z, n.x) / (2.0f * kPi) + 0.5f;
    hv = facos(-n.y) / kPi;
    TS ts;
    ts.geoN = normalize(xform_normal(M, n));
    ts.shadingN = ts.geoN;
    ts.shadingTX = normalize(xform_normal(M, tangentX));
    ts.shadingTY = cross(ts.geoN, ts.shadingTX);
    return ts;
}

// Hit attributes of a bilinear quad at patch coordinates (su, sv)
// (intersection.cu:572-610) and its frame (tangent_space.hpp:36-47, 50-60).
__device__ TS quad_frame(const DevPtLod& L, const M34& M, int prim, float su, float sv, float& hu, float& hv) {
    const uint32_t* q = L.quads + 4 * (size_t)(prim - L.numTris);
    const u32 i0 = q[0], i1 = q[1], i2 = q[2], i3 = q[3];
    const V3 v0 = lod_vert(L, i0), v1 = lod_vert(L, i1), v2 = lod_vert(L, i2), v3v = lod_vert(L, i3);
    const float* uv = L.uvs;
    const float u0x = uv[2 * (size_t)i0], u0y = uv[2 * (size_t)i0 + 1];
    const float u1x = uv[2 * (size_t)i1], u1y = uv[2 * (size_t)i1 + 1];
    const float u2x = uv[2 * (size_t)i2], u2y = uv[2 * (size_t)i2 + 1];
    const float u3x = uv[2 * (size_t)i3], u3y = uv[2 * (size_t)i3 + 1];
    const V3 dxds = smul(1.f - su, v3v - v0) + smul(su, v2 - v1);
    const V3 dxdt = smul(1.f - sv, v1 - v0) + smul(sv, v2 - v3v);
    const float dudsx = (1.f - su) * (u3x - u0x) + su * (u2x - u1x);
    const float dudsy = (1.f - su) * (u3y - u0y) + su * (u2y - u1y);
    const float dudtx = (1.f - sv) * (u1x - u0x) + sv * (u2x - u3x);
    const float dudty = (1.f - sv) * (u1y - u0y) + sv * (u2y - u3y);
    const float det = dudsx * dudty - dudtx * dudsy;
    V3 tangentX;
    if (det >= 1e-9f || det <= -1e-9f) {
        const float a = dudty / det, b = -dudsy / det;
        tangentX = v3(dxds.x * a + dxdt.x * b, dxds.y * a + dxdt.y * b, dxds.z * a + dxdt.z * b);
    } else {
        tangentX = dxds;
    }
    const V3 geoN0 = cross(dxdt, dxds);
    hu = bilerp(u0x, u1x, u3x, u2x, su, sv);
    hv = bilerp(u0y, u1y, u3y, u2y, su, sv);
    TS ts;
    ts.geoN = normalize(xform_normal(M, geoN0));
    tangentX = normalize(xform_normal(M, tangentX));
    V3 sn = bilerp(lod_normal(L, i0), lod_normal(L, i1), lod_normal(L, i3), lod_normal(L, i2), su, sv);
    sn = normalize(xform_normal(M, sn));
    ts.shadingN = sn;
    ts.shadingTX = normalize(tangentX - sn * dot(tangentX, sn));
    ts.shadingTY = cross(sn, ts.shadingTX);
    return ts;
}

// Hit attributes of first_intersection (intersection.cu:524-621): shading
// frame of tangent_space_geom_to_shader (tangent_space.hpp:10-60) and the
// interpolated uv. position: the world-space hit point (spheres use it).
__device__ __forceinline__ TS hit_frame(const PtScene& S, int inst, int prim, float w0, float w1, V3 position,
                                        float& hu, float& hv) {
    const int lod = S.trace.instLod[inst];
    const DevPtLod L = S.ptLods[lod];
    if (prim >= L.numTris) {  // only LoDs with quads or spheres have primitives past their triangles
        const M34 M = load_m34(S.trace.worldToInst + 12 * (size_t)inst);
        if (prim >= L.numTris + L.numQuads) return sphere_frame(L, M, prim, position, hu, hv);
        return quad_frame(L, M, prim, w0, w1, hu, hv);
    }
    const u32 i0 = L.tris[3 * (size_t)prim], i1 = L.tris[3 * (size_t)prim + 1], i2 = L.tris[3 * (size_t)prim + 2];
    const V3 v0 = v3(L.verts[3 * (size_t)i0], L.verts[3 * (size_t)i0 + 1], L.verts[3 * (size_t)i0 + 2]);
    const V3 v1 = v3(L.verts[3 * (size_t)i1], L.verts[3 * (size_t)i1 + 1], L.verts[3 * (size_t)i1 + 2]);
    const V3 v2 = v3(L.verts[3 * (size_t)i2], L.verts[3 * (size_t)i2 + 1], L.verts[3 * (size_t)i2 + 2]);
    const float* uv = L.uvs;
    const float u0x = uv[2 * (size_t)i0], u0y = uv[2 * (size_t)i0 + 1];
    const float u1x = uv[2 * (size_t)i1], u1y = uv[2 * (size_t)i1 + 1];
    const float u2x = uv[2 * (size_t)i2], u2y = uv[2 * (size_t)i2 + 1];
    const V3 dx0 = v1 - v0;
    const V3 dx1 = v2 - v0;
    const float du0x = u1x - u0x, du0y = u1y - u0y;
    const float du1x = u2x - u0x, du1y = u2y - u0y;
    const float det = du0x * du1y - du0y * du1x;
    V3 tangentX;
    if (det >= 1e-5f || det <= -1e5f)
        tangentX = (dx0 * du1y - dx1 * du0y) / det;
    else
        tangentX = dx0;
    const V3 geoN0 = cross(dx0, dx1);
    const float w2 = 1.f - w0 - w1;
    hu = u0x * w0 + u1x * w1 + u2x * w2;
    hv = u0y * w0 + u1y * w1 + u2y * w2;
    const M34 M = load_m34(S.trace.worldToInst + 12 * (size_t)inst);
    TS ts;
    ts.geoN = normalize(xform_normal(M, geoN0));
    tangentX = normalize(xform_normal(M, tangentX));
    const float* nr = L.normals;
    const V3 n0 = v3(nr[3 * (size_t)i0], nr[3 * (size_t)i0 + 1], nr[3 * (size_t)i0 + 2]);
    const V3 n1 = v3(nr[3 * (size_t)i1], nr[3 * (size_t)i1 + 1], nr[3 * (size_t)i1 + 2]);
    const V3 n2 = v3(nr[3 * (size_t)i2], nr[3 * (size_t)i2 + 1], nr[3 * (size_t)i2 + 2]);
    V3 sn = smul(w0, n0) + smul(w1, n1) + smul(1.f - (w0 + w1), n2);
    sn = normalize(xform_normal(M, sn));
    ts.shadingN = sn;
    ts.shadingTX = normalize(tangentX - sn * dot(tangentX, sn));
    ts.shadingTY = cross(sn, ts.shadingTX);
    return ts;
}

__device__ __forceinline__ int material_index(const PtScene& S, int inst, int prim) {
    const int lod = S.trace.instLod[inst];
    return (int)S.ptLods[lod].matIdx[prim];
}

// Per-iteration target values of one path: Radiance 0-2, Hit ID 3-4,
// Position 5-7, Depth 8, Normal 9-11, UV 12-13, Albedo 14-16, Lightness 17.
// contribute + check_nan of RenderBuffer (render_target.hpp:62-76,189-195).
struct Acc {
    float v[18];
    u32 mask;
    u32 nans;
    __device__ __forceinline__ void init(u32 m) {
#pragma unroll
        for (int k = 0; k < 18; ++k) v[k] = 0.0f;
        mask = m;
        nans = 0;
    }
    __device__ __forceinline__ void add(int target, int slot, float x) {
        if (!((mask >> target) & 1u)) return;
        if (x != x) {
            ++nans;
            x = 0.0f;
        }
        v[slot] += x;
    }
};

// The vertex a walk ends in: a surface hit or the void.
struct Vertex {
    bool isVoid;
    V3 position;
    V3 incEnc;  // (x, y, dist * sgn(z)) for surfaces, the raw direction for the void
    TS ts;
    int inst, prim;
    MatParams mat;
    float incidentPdf;
    float hu, hv;
    float sp0, sp1;  // surface parameters of the hit (a quad's (u, v))
    __device__ __forceinline__ V3 incident() const {
        const float iz = sgnf(incEnc.z) * fsqrt(clampf(1.0f - incEnc.x * incEnc.x - incEnc.y * incEnc.y, 0.0f, 1.0f));
        return v3(incEnc.x, incEnc.y, iz);
    }
};

// create_surface / create_void + PtVertexExt::update (path_util.hpp:629-665,
// pt_common.hpp:45-61) for the segment rayOrigin + t * dir with forward pdf
// pdfF; inst == -1 is a miss.
__device__ __forceinline__ void make_vertex(const PtScene& S, V3 rayOrigin, V3 dir, float pdfF, float t, int inst,
                                            int prim, float w0, float w1, Vertex& vx) {
    vx.isVoid = inst == -1;
    vx.inst = inst;
    vx.prim = prim;
    vx.hu = 0.0f;
    vx.hv = 0.0f;
    vx.sp0 = 0.0f;
    vx.sp1 = 0.0f;
    if (vx.isVoid) {
        vx.incEnc = dir;
        vx.incidentPdf = pdfF * 1.0f;
        vx.mat.type = -1;
        return;
    }
    vx.position = rayOrigin + dir * t;
    vx.sp0 = w0;
    vx.sp1 = w1;
    vx.ts = hit_frame(S, inst, prim, w0, w1, vx.position, vx.hu, vx.hv);
    vx.incEnc = v3(dir.x, dir.y, t * sgnf(dir.z));
    vx.mat = fetch(S, material_index(S, inst, prim), vx.hu, vx.hv);
    const float inCosAbs = fabsf(dot(dir, vx.ts.shadingN));
    vx.incidentPdf = pdfF * fabsf(inCosAbs) / (t * t);
}

// Radiance a new vertex contributes (pt_common.hpp:134-143): emission (or the
// background) with the MIS weight of hitting the emitter, times the
// throughput; em receives the weighted emission.
__device__ __forceinline__ V3 emission_contribution(const PtScene& S, const Vertex& vx, int pathLen, V3 throughput,
                                                    V3& em, V3 prevPos, float& mis) {
    float connectPdf = 0.0f;
    mis = 1.0f;
    if (vx.isVoid) {  // get_emission, VOID (path_util.hpp:415-424)
        float pdfBack;
        em = evaluate_background(S, vx.incEnc, pdfBack);
        if (nonzero(em)) connectPdf = background_pdf(S, pdfBack);
    } else {
        em = emission(vx.mat, vx.ts.geoN, neg(vx.incident()));
        if (nonzero(em)) connectPdf = light_pdf(S, vx.inst, vx.prim, prevPos, vx.sp0, vx.sp1);
    }
    if (nonzero(em) && pathLen > 1) {
        mis = 1.0f / (1.0f + (float)S.neeCount * (connectPdf / vx.incidentPdf));
        em = em * mis;
    }
    return throughput * em;
}
__device__ __forceinline__ V3 emission_contribution(const PtScene& S, const Vertex& vx, int pathLen, V3 throughput,
                                                    V3& em, V3 prevPos) {
    float mis;
    return emission_contribution(S, vx, pathLen, throughput, em, prevPos, mis);
}

// Contributions of a new vertex at path length pathLen (pt_common.hpp:128-151):
// Hit ID on the first segment, then emission with MIS and the guide-weighted
// auxiliary targets.
__device__ __forceinline__ void vertex_targets(const PtScene& S, const Vertex& vx, int pathLen, V3 throughput,
                                               float guideWeight, Acc& acc, V3 prevPos) {
    if (pathLen == 1) {
        acc.add(1, 3, (float)vx.inst);
        acc.add(1, 4, (float)vx.prim);
    }
    if (pathLen < S.minPathLength) return;
    V3 em;
    const V3 cc = emission_contribution(S, vx, pathLen, throughput, em, prevPos);
    acc.add(0, 0, cc.x);
    acc.add(0, 1, cc.y);
    acc.add(0, 2, cc.z);
    if (acc.mask & 0xFCu) {
        const V3 pos = vx.isVoid ? v3(0, 0, 0) : vx.position;
        acc.add(2, 5, guideWeight * pos.x);
        acc.add(2, 6, guideWeight * pos.y);
        acc.add(2, 7, guideWeight * pos.z);
        acc.add(3, 8, guideWeight * (vx.isVoid ? 0.0f : fabsf(vx.incEnc.z)));
        const V3 nrm = vx.isVoid ? v3(0, 0, 0) : vx.ts.shadingN;
        acc.add(4, 9, guideWeight * nrm.x);
        acc.add(4, 10, guideWeight * nrm.y);
        acc.add(4, 11, guideWeight * nrm.z);
        acc.add(5, 12, guideWeight * vx.hu);
        acc.add(5, 13, guideWeight * vx.hv);
        V3 alb = v3(0, 0, 0);  // materials::albedo
        if (!vx.isVoid) {
            if (vx.mat.type == kMatLambert || vx.mat.type == kMatTorrance)
                alb = vx.mat.albedo;
            else if (mat_has_inner_medium(vx.mat.type))
                alb = v3(1.0f / (1.0f + vx.mat.inner.absorption.x), 1.0f / (1.0f + vx.mat.inner.absorption.y),
                         1.0f / (1.0f + vx.mat.inner.absorption.z));
        }
        acc.add(6, 14, guideWeight * alb.x);
        acc.add(6, 15, guideWeight * alb.y);
        acc.add(6, 16, guideWeight * alb.z);
        acc.add(7, 17, guideWeight * ((em.x + em.y + em.z) / 3.0f));
    }
}

// One next-event estimate of pt_common.hpp:95-124 up to the visibility test:
// on success the shadow segment (origin a, unit direction, length) and the
// Radiance / Lightness values to add when it is unoccluded.
struct NeeCandidate {
    V3 a, dir;
    float tmax;
    V3 radiance;
    float lightness;
    int lightMaterial;
    V3 lightDir;       // direction to the light sample (the BSDF's excident)
    float misFactor;   // -d ln(mis) / d ln(BSDF pdf) = mis * hitPdf / creationPdf
};
__device__ __forceinline__ bool nee_candidate(const PtScene& S, const Vertex& vx, V3 throughput, float guideWeight,
                                              int i, u64 neeSeed, u64 x, NeeCandidate& c) {
    const Nee nee = connect(S, (u64)i, (u64)S.neeCount, neeSeed, vx.position, sample_uniform((u32)x),
                            sample_uniform((u32)(x >> 32)));
    float vCos, vPdf;
    const V3 value = mat_evaluate(vx.ts, vx.mat, vx.incident(), nee.dir, vCos, vPdf);
    if (nee.cosOut != 0) vCos *= nee.cosOut;
    const V3 radiance = value * nee.diffIrradiance;
    if (!((radiance.x > 0.0f || radiance.y > 0.0f || radiance.z > 0.0f) && vCos > 0.0f)) return false;
    // any_intersection (intersection.cu:624-637)
    const V3 a = add_epsilon(vx.position, nee.dir, vx.ts.geoN);
    const V3 b = add_epsilon(nee.position, neg(nee.dir), nee.geoNormal);
    const V3 d = b - a;
    c.a = a;
    c.tmax = len(d);
    c.dir = d / c.tmax;
    const float hitPdf = vPdf * fabsf(nee.cosOut) / nee.distSq;
    const float mis = 1.0f / ((float)S.neeCount + hitPdf / nee.creationPdf);
    c.radiance = throughput * vCos * radiance * mis;
    c.lightness = guideWeight * vCos;
    c.lightMaterial = nee.lightMaterial;
    c.lightDir = nee.dir;
    c.misFactor = mis * (hitPdf / nee.creationPdf);
    return true;
}

__device__ __forceinline__ void add_nee(const NeeCandidate& c, Acc& acc) {
    acc.add(0, 0, c.radiance.x);
    acc.add(0, 1, c.radiance.y);
    acc.add(0, 2, c.radiance.z);
    acc.add(7, 17, c.lightness);
}

// walk (random_walk.hpp:38-112) from a surface vertex up to the next ray:
// BSDF sample, throughput, Russian roulette. Returns false when the path ends.
// The walk also reports the throughput before the roulette (tPrime) and the
// continuation probability, which the adjoint replay differentiates.
__device__ __forceinline__ bool walk_surface(const Vertex& vx, u64 x0, u64 x1, float rndRoulette, V3& throughput,
                                             V3& rayOrigin, V3& excident, float& pdfF, V3& absorption, V3& tPrime,
                                             float& cont) {
    const float u0 = sample_uniform((u32)x0), u1 = sample_uniform((u32)(x0 >> 32));
    V3 sThroughput;
    cont = 1.0f;
    if (!mat_sample(vx.ts, vx.mat, vx.incident(), u0, u1, x1, sThroughput, excident, pdfF)) return false;
    absorption = vx.mat.medium(dot(excident, vx.ts.geoN)).absorption;  // path_util.hpp:346-351
    throughput = throughput * sThroughput;
    tPrime = throughput;
    cont = mn(mx(mx(throughput.x, throughput.y), throughput.z) + 0.1f, 1.0f);
    if (rndRoulette >= cont) return false;
    throughput = throughput / cont;
    rayOrigin = add_epsilon(vx.position, excident, vx.ts.geoN);
    return true;
}
__device__ __forceinline__ bool walk_surface(const Vertex& vx, u64 x0, u64 x1, float rndRoulette, V3& throughput,
                                             V3& rayOrigin, V3& excident, float& pdfF, V3& absorption) {
    V3 tPrime;
    float cont;
    return walk_surface(vx, x0, x1, rndRoulette, throughput, rayOrigin, excident, pdfF, absorption, tPrime, cont);
}

// The camera's walk (pinholecam_sample_ray, pinhole.hpp:56-84): the first
// ray of pixel (px, py); Russian roulette always continues at throughput 1.
__device__ __forceinline__ void camera_ray(const PtScene& S, int px, int py, u64 x0, V3& dir, float& pdfF) {
    const float u0 = sample_uniform((u32)x0), u1 = sample_uniform((u32)(x0 >> 32));
    const V3 viewDir = v3(S.camDir[0], S.camDir[1], S.camDir[2]);
    const V3 up = v3(S.camUp[0], S.camUp[1], S.camUp[2]);
    const float w = (float)S.width, h = (float)S.height;
    float cx = ((float)px + u0) / w * 2.0f - 1.0f;
    float cy = ((float)py + u1) / h * 2.0f - 1.0f;
    cx *= S.tanVFov;
    cy *= S.tanVFov;
    const float aspect = w / h;
    const V3 xAxis = cross(viewDir, up);
    dir = normalize(xAxis * (cx * aspect) + up * cy + viewDir * 1.0f);
    const float pixelArea = (2 * S.tanVFov) * (2 * S.tanVFov) * aspect;
    const float cosOut = dot(viewDir, dir);
    pdfF = 1.0f / (pixelArea * cosOut * cosOut * cosOut);
}

// Running means of the enabled targets (output_handler_variance.hpp:7-31).
__device__ __forceinline__ void accumulate(const PtScene& S, size_t pixel, int iteration, const float* v, u32 mask) {
    const float denom = mx(1.0f, (float)iteration + 1);
    const float varDenom = mx(1.0f, (float)iteration);
    const u32 varMask = S.varianceMask & mask;
    int slot = 0;
#pragma unroll
    for (int t = 0; t < kNumTargets; ++t) {
        const int nc = kTargetChannels[t];
        if ((mask >> t) & 1u) {
            float* c = S.cum[t] + (size_t)nc * pixel;
            if ((varMask >> t) & 1u) {
                // UpdateIter (output_handler_variance.hpp:7-31): mean and
                // variance with the iteration index as the divisor base
                float* var = S.var[t] + (size_t)nc * pixel;
#pragma unroll
                for (int k = 0; k < nc; ++k) {
                    const float iter = v[slot + k];
                    const float diff = iter - c[k];
                    const float cum = c[k] + diff / denom;
                    c[k] = cum;
                    var[k] += (diff * (iter - cum) - var[k]) / varDenom;
                }
            } else {
#pragma unroll
                for (int k = 0; k < nc; ++k) c[k] += (v[slot + k] - c[k]) / denom;
            }
        }
        slot += nc;
    }
}

}  // namespace path
}  // namespace mfh
