// Path tracer megakernel for gfx950: the GPU device of Mufflon's pure path
// tracer (GpuPathTracer, gpu_pt.cu:54-152), one iteration per launch.
//
// One thread per pixel, 16x16 pixel tiles per 256-thread block (a wave covers
// a 16x4 pixel patch, which keeps primary rays coherent). Per pixel it runs
// pt_sample (pt_common.hpp:75-156): pinhole camera sample, BSDF sampling with
// Russian roulette (random_walk.hpp:38-112), next-event estimation through the
// light tree with MIS (light_tree_sampling.hpp:195-329, light_sampling.hpp),
// emission hits weighted with light_pdf (:340-412), and accumulates the
// iteration's Radiance and Hit ID into the running means of
// output_handler_variance.hpp:7-31 in its epilogue (no separate update pass,
// no atomics: each pixel belongs to one thread). The RNG is Mufflon's PCG64,
// one 64-bit state per pixel, seeded as in gpu_pt.cu:140-152.
//
// Float expressions keep the reference's operand order; transcendental
// functions of the samplers are evaluated in double and rounded once
// (correctly rounded sin/cos), the same definition as the CPU oracle's.
#include "pt_device.hpp"
#include "traverse.hpp"

namespace mfh {

namespace {

constexpr int kPtTile = 16;
constexpr int kPtThreads = kPtTile * kPtTile;

__device__ __forceinline__ float sgnf(float x) { return x < 0.0f ? -1.0f : 1.0f; }
__device__ __forceinline__ float clampf(float v, float lo, float hi) { return mn(mx(v, lo), hi); }
__device__ __forceinline__ V3 normalize(V3 v) { return v / len(v); }
__device__ __forceinline__ V3 neg(V3 v) { return {-v.x, -v.y, -v.z}; }
__device__ __forceinline__ V3 smul(float s, V3 a) { return {s * a.x, s * a.y, s * a.z}; }

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
__device__ __forceinline__ u64 percentage_of(u64 num, float p) {
    const u64 pfix = (u64)(p * 2147483648.0f);
    return ((pfix * (num & 0x7fffffff)) >> 31) + (pfix * (num >> 31));
}

struct TS {
    V3 shadingN, geoN, shadingTX, shadingTY;
};

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

struct PtMat {
    int type;
    V3 albedo;
    V3 radiance;  // already scaled
};
__device__ __forceinline__ PtMat load_material(const PtScene& S, int idx) {
    const PtMaterialRecord& m = S.materials[idx];
    PtMat r;
    r.type = m.type;
    r.albedo = v3(m.albedo[0], m.albedo[1], m.albedo[2]);
    r.radiance = v3(m.radiance[0] * m.scale, m.radiance[1] * m.scale, m.radiance[2] * m.scale);
    return r;
}

// emissive.hpp:39-44
__device__ __forceinline__ V3 mat_emission(const PtMat& m, V3 geoN, V3 excident) {
    if (m.type != kMatEmissive) return v3(0, 0, 0);
    const float cosOut = dot(geoN, excident);
    if (cosOut <= 0.0f) return v3(0, 0, 0);
    return m.radiance;
}

struct Nee {
    V3 position;
    float creationPdf;
    V3 dir;
    float cosOut;
    V3 diffIrradiance;
    V3 geoNormal;
    float distSq;
};

__device__ Nee connect_light(const PtScene& S, uint16_t type, u32 offset, V3 pos, float u0, float u1) {
    const uint8_t* light = S.lightMem + offset;
    Nee n;
    if (type == kLightPoint) {  // light_sampling.hpp:308-327
        const float4 a = *reinterpret_cast<const float4*>(light);
        const float4 b = *reinterpret_cast<const float4*>(light + 16);
        const V3 lp = v3(a.x, a.y, a.z);
        V3 d = lp - pos;
        const float distSq = dot(d, d) + kDistSqEps;
        const float dist = __builtin_sqrtf(distSq);
        d = d / dist;
        n.position = lp;
        n.creationPdf = kInfPdf;
        n.dir = d;
        n.cosOut = 0.0f;
        n.diffIrradiance = v3(b.x / distSq, b.y / distSq, b.z / distSq);
        n.geoNormal = v3(0, 0, 0);
        n.distSq = distSq;
        return n;
    }
    // area triangle (light_sampling.hpp:118-140, 350-375)
    const float4 q0 = *reinterpret_cast<const float4*>(light);
    const float4 q1 = *reinterpret_cast<const float4*>(light + 16);
    const float4 q2 = *reinterpret_cast<const float4*>(light + 32);
    const V3 p0 = v3(q0.x, q0.y, q0.z), e1 = v3(q0.w, q1.x, q1.y), e2 = v3(q1.z, q1.w, q2.x);
    const int matIdx = (int)(__float_as_uint(q2.y) & 0xFFFFu);
    V3 normal = cross(e1, e2);
    const float area2 = len(normal);
    normal = normal / area2;
    const float area = area2 * 0.5f;
    const float s = 1.f - __builtin_sqrtf(1.f - u0);
    const float t = (1.f - s) * u1;
    const V3 position = p0 + e1 * s + e2 * t;
    const PtMat m = load_material(S, matIdx);
    const V3 rad = mat_emission(m, v3(0, 0, 1), v3(0, 0, 1));
    const float pdf = 1.0f / area;
    const V3 intensity = rad * area;
    V3 d = position - pos;
    const float distSq = dot(d, d) + kDistSqEps;
    const float dist = __builtin_sqrtf(distSq);
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
    n.diffIrradiance = v3(value.x / distSq, value.y / distSq, value.z / distSq);
    n.geoNormal = normal;
    n.distSq = distSq;
    return n;
}

// lights::connect (light_tree_sampling.hpp:195-329), positional subtree.
__device__ Nee connect(const PtScene& S, u64 index, u64 numIndices, u64 seed, V3 pos, float u0, float u1) {
    u64 rndChoice = numIndices > 0 ? seed + index * (~0ull / numIndices) : seed;
    if (rndChoice == ~0ull) --rndChoice;
    const float fluxSum = S.bgFlux + S.dirRootFlux + S.posRootFlux;
    const float envProb = S.bgFlux / fluxSum;
    const float dirProb = S.dirRootFlux / fluxSum;
    const float posProb = S.posRootFlux / fluxSum;
    u64 left = percentage_of(~0ull, envProb + dirProb);
    u64 right = ~0ull;
    float lightProb = posProb;
    uint16_t type = (uint16_t)S.posRootType;
    u32 offset = 0;
    for (int guard = 0; type == kLightInternalNode && guard < 64; ++guard) {
        const float4 a = *reinterpret_cast<const float4*>(S.lightMem + offset);
        const float4 b = *reinterpret_cast<const float4*>(S.lightMem + offset + 16);
        const float lFlux = a.x, rFlux = b.x;
        const u32 lOff = __float_as_uint(a.y), rOff = __float_as_uint(a.w);
        const u32 types = __float_as_uint(a.z);
        const float probLeft = lFlux / (lFlux + rFlux);
        const u64 boundary = left + percentage_of(right - left, probLeft);
        if (rndChoice < boundary) {
            type = (uint16_t)(types & 0xFFFFu);
            offset = lOff;
            right = boundary;
            lightProb *= probLeft;
        } else {
            type = (uint16_t)(types >> 16);
            offset = rOff;
            left = boundary;
            lightProb *= (1.0f - probLeft);
        }
    }
    Nee n = connect_light(S, type, offset, pos, u0, u1);
    n.creationPdf *= lightProb;
    n.diffIrradiance = n.diffIrradiance / lightProb;
    return n;
}

// light_pdf (light_tree_sampling.hpp:340-390): connect pdf of a hit emitter.
__device__ float light_pdf(const PtScene& S, int inst, int prim) {
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
        const float4 a = *reinterpret_cast<const float4*>(S.lightMem + offset);
        const float4 b = *reinterpret_cast<const float4*>(S.lightMem + offset + 16);
        const float lFlux = a.x, rFlux = b.x;
        const float pLeft = lFlux / (lFlux + rFlux);
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
    const float4 q0 = *reinterpret_cast<const float4*>(S.lightMem + offset);
    const float4 q1 = *reinterpret_cast<const float4*>(S.lightMem + offset + 16);
    const V3 e1 = v3(q0.w, q1.x, q1.y), e2 = v3(q1.z, q1.w, *reinterpret_cast<const float*>(S.lightMem + offset + 32));
    const float area = len(cross(e1, e2)) / 2.0f;
    return p / area;
}

struct SurfaceHit {
    V3 position;
    TS ts;
    int inst, prim;
};

// Hit attributes of first_intersection (intersection.cu:524-621) and the
// shading frame of tangent_space_geom_to_shader (tangent_space.hpp:10-60).
__device__ __forceinline__ TS hit_frame(const PtScene& S, const TraceHit& h) {
    const int lod = S.trace.instLod[h.inst];
    const DevPtLod L = S.ptLods[lod];
    const u32 i0 = L.tris[3 * (size_t)h.prim], i1 = L.tris[3 * (size_t)h.prim + 1], i2 = L.tris[3 * (size_t)h.prim + 2];
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
    const M34 M = load_m34(S.trace.worldToInst + 12 * (size_t)h.inst);
    TS ts;
    ts.geoN = normalize(xform_normal(M, geoN0));
    tangentX = normalize(xform_normal(M, tangentX));
    const float* nr = L.normals;
    const V3 n0 = v3(nr[3 * (size_t)i0], nr[3 * (size_t)i0 + 1], nr[3 * (size_t)i0 + 2]);
    const V3 n1 = v3(nr[3 * (size_t)i1], nr[3 * (size_t)i1 + 1], nr[3 * (size_t)i1 + 2]);
    const V3 n2 = v3(nr[3 * (size_t)i2], nr[3 * (size_t)i2 + 1], nr[3 * (size_t)i2 + 2]);
    V3 sn = smul(h.w0, n0) + smul(h.w1, n1) + smul(1.f - (h.w0 + h.w1), n2);
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

template <bool kSingle>
__global__ __launch_bounds__(kPtThreads) void k_pt_iterate(PtScene S) {
    __shared__ int sStack[kLdsDepth * kPtThreads];
    const int px = blockIdx.x * kPtTile + (threadIdx.x % kPtTile);
    const int py = blockIdx.y * kPtTile + (threadIdx.x / kPtTile);
    if (px >= S.width || py >= S.height) return;
    const size_t pixel = (size_t)px + (size_t)py * S.width;
    int* const lstack = sStack + threadIdx.x;
    u64 rng = S.rngs[pixel];

    float accR = 0.0f, accG = 0.0f, accB = 0.0f, accH0 = 0.0f, accH1 = 0.0f;
    u32 nans = 0;
    auto add = [&](float& dst, float v) {
        if (v != v) {
            ++nans;
            v = 0.0f;
        }
        dst += v;
    };

    const V3 viewDir = v3(S.camDir[0], S.camDir[1], S.camDir[2]);
    const V3 up = v3(S.camUp[0], S.camUp[1], S.camUp[2]);
    V3 throughput = v3(1.0f, 1.0f, 1.0f);
    bool isCamera = true;
    V3 position = v3(S.camPos[0], S.camPos[1], S.camPos[2]);
    V3 incEnc = v3(0, 0, 0);
    TS ts;
    int hInst = -1, hPrim = -1;
    PtMat mat;
    mat.type = -1;
    pcg_next(rng);  // create_camera's RndSet2, unused by the pinhole
    int pathLen = 0;
    do {
        if (pathLen > 0 && pathLen + 1 >= S.minPathLength && pathLen + 1 <= S.maxPathLength) {
            const u64 neeSeed = pcg_next(rng);
            const float iz = sgnf(incEnc.z) *
                             __builtin_sqrtf(clampf(1.0f - incEnc.x * incEnc.x - incEnc.y * incEnc.y, 0.0f, 1.0f));
            const V3 incident = v3(incEnc.x, incEnc.y, iz);
            for (int i = 0; i < S.neeCount; ++i) {
                const u64 x = pcg_next(rng);
                const Nee nee = connect(S, (u64)i, (u64)S.neeCount, neeSeed, position, sample_uniform((u32)x),
                                        sample_uniform((u32)(x >> 32)));
                // materials::evaluate (material_sampling.hpp:160-237), Lambert
                const float iDotN = -dot(incident, ts.shadingN);
                const float eDotN = dot(nee.dir, ts.shadingN);
                bool ok = mat.type == kMatLambert ? !(iDotN * eDotN < 0.0f) : !(iDotN * eDotN != 0.0f);
                V3 value = v3(0, 0, 0);
                float vCos = 0.0f, vPdf = 0.0f;
                if (ok) {
                    const float iDotG = -dot(incident, ts.geoN);
                    const float eDotG = dot(nee.dir, ts.geoN);
                    ok = !(eDotG * eDotN <= 0.0f || iDotG * iDotN <= 0.0f);
                    if (ok && mat.type == kMatLambert) {
                        const float inZ = iDotN, exZ = eDotN;
                        if (!(inZ * exZ < 0.0f)) {
                            value = v3(mat.albedo.x / kPi, mat.albedo.y / kPi, mat.albedo.z / kPi);
                            vPdf = fabsf(exZ) / kPi;
                            if (value.x == 0.0f && value.y == 0.0f && value.z == 0.0f) {
                                vPdf = 0.0f;
                            } else {
                                vPdf = mn(vPdf, 1e9f);
                                vCos = fabsf(eDotN);
                            }
                        }
                    }
                }
                if (nee.cosOut != 0) vCos *= nee.cosOut;
                const V3 radiance = value * nee.diffIrradiance;
                if ((radiance.x > 0.0f || radiance.y > 0.0f || radiance.z > 0.0f) && vCos > 0.0f) {
                    // any_intersection (intersection.cu:624-637)
                    const V3 a = add_epsilon(position, nee.dir, ts.geoN);
                    const V3 b = add_epsilon(nee.position, neg(nee.dir), nee.geoNormal);
                    const V3 c = b - a;
                    const float tmax = len(c);
                    const TraceHit sh = trace_ray<true, false, kSingle, kPtThreads>(S.trace, a, c / tmax, tmax, lstack);
                    if (sh.inst == -1) {
                        const float hitPdf = vPdf * fabsf(nee.cosOut) / nee.distSq;
                        const float mis = 1.0f / ((float)S.neeCount + hitPdf / nee.creationPdf);
                        const V3 cc = throughput * vCos * radiance * mis;
                        add(accR, cc.x);
                        add(accG, cc.y);
                        add(accB, cc.z);
                    }
                }
            }
        }
        // walk (random_walk.hpp:38-112)
        const u64 x0 = pcg_next(rng);
        pcg_next(rng);  // the i0 word of RndSet2_1 (unused by these samplers)
        const float u0 = sample_uniform((u32)x0), u1 = sample_uniform((u32)(x0 >> 32));
        const float rndRoulette = sample_uniform((u32)pcg_next(rng));
        V3 excident, sThroughput;
        float pdfF;
        if (isCamera) {  // pinholecam_sample_ray (pinhole.hpp:56-84)
            const float w = (float)S.width, h = (float)S.height;
            float cx = ((float)px + u0) / w * 2.0f - 1.0f;
            float cy = ((float)py + u1) / h * 2.0f - 1.0f;
            cx *= S.tanVFov;
            cy *= S.tanVFov;
            const float aspect = w / h;
            const V3 xAxis = cross(viewDir, up);
            excident = normalize(xAxis * (cx * aspect) + up * cy + viewDir * 1.0f);
            const float pixelArea = (2 * S.tanVFov) * (2 * S.tanVFov) * aspect;
            const float cosOut = dot(viewDir, excident);
            pdfF = 1.0f / (pixelArea * cosOut * cosOut * cosOut);
            sThroughput = v3(1.0f, 1.0f, 1.0f);
        } else {
            // materials::sample (material_sampling.hpp:77-146), Lambert (lambert.hpp:18-35)
            const float iz = sgnf(incEnc.z) *
                             __builtin_sqrtf(clampf(1.0f - incEnc.x * incEnc.x - incEnc.y * incEnc.y, 0.0f, 1.0f));
            const V3 incident = v3(incEnc.x, incEnc.y, iz);
            const float iDotG = -dot(incident, ts.geoN);
            const float iDotN = -dot(incident, ts.shadingN);
            if (iDotG * iDotN <= 0.0f) break;
            if (mat.type != kMatLambert) break;
            const float cosT = __builtin_sqrtf(u0);
            const float sinT = __builtin_sqrtf(1.0f - u0);
            const float phi = u1 * 2 * kPi;
            const float sphi = (float)sin((double)phi), cphi = (float)cos((double)phi);
            const float sg = sgnf(iDotN);
            const V3 exTS = v3(sinT * sphi * sg, sinT * cphi * sg, cosT * sg);
            sThroughput = mat.albedo;
            pdfF = cosT / kPi;
            if (sThroughput.x == 0.0f && sThroughput.y == 0.0f && sThroughput.z == 0.0f) break;
            V3 g = smul(exTS.x, ts.shadingTX) + smul(exTS.y, ts.shadingTY) + smul(exTS.z, ts.shadingN);
            g = normalize(g);
            const float eDotG = dot(g, ts.geoN);
            const float eDotN = exTS.z;
            if (eDotG * eDotN <= 0.0f) break;
            excident = g;
            pdfF = mn(pdfF, 1e9f);
        }
        throughput = throughput * sThroughput;
        const float cont = mn(mx(mx(throughput.x, throughput.y), throughput.z) + 0.1f, 1.0f);
        if (rndRoulette >= cont) break;
        throughput = throughput / cont;
        const V3 rayOrigin = add_epsilon(position, excident, isCamera ? v3(0, 0, 0) : ts.geoN);
        const TraceHit hit =
            trace_ray<false, false, kSingle, kPtThreads>(S.trace, rayOrigin, excident, kMaxSceneSize, lstack);
        const bool isVoid = hit.inst == -1;
        hInst = hit.inst;
        hPrim = hit.prim;
        float incidentPdf;
        if (isVoid) {
            incEnc = excident;
            incidentPdf = pdfF * 1.0f;
        } else {
            position = rayOrigin + excident * hit.t;
            ts = hit_frame(S, hit);
            incEnc = v3(excident.x, excident.y, hit.t * sgnf(excident.z));
            mat = load_material(S, material_index(S, hit.inst, hit.prim));
            const float inCosAbs = fabsf(dot(excident, ts.shadingN));
            incidentPdf = pdfF * fabsf(inCosAbs) / (hit.t * hit.t);
        }
        isCamera = false;
        if (pathLen == 0) {
            add(accH0, (float)hInst);
            add(accH1, (float)hPrim);
        }
        ++pathLen;
        if (pathLen >= S.minPathLength && !isVoid) {
            const float iz = sgnf(incEnc.z) *
                             __builtin_sqrtf(clampf(1.0f - incEnc.x * incEnc.x - incEnc.y * incEnc.y, 0.0f, 1.0f));
            V3 emission = mat_emission(mat, ts.geoN, neg(v3(incEnc.x, incEnc.y, iz)));
            if (emission.x != 0.0f || emission.y != 0.0f || emission.z != 0.0f) {
                const float connectPdf = light_pdf(S, hInst, hPrim);
                if (pathLen > 1) {
                    const float mis = 1.0f / (1.0f + (float)S.neeCount * (connectPdf / incidentPdf));
                    emission = emission * mis;
                }
                const V3 cc = throughput * emission;
                add(accR, cc.x);
                add(accG, cc.y);
                add(accB, cc.z);
            }
        }
        if (isVoid) break;
    } while (pathLen < S.maxPathLength);

    S.rngs[pixel] = rng;
    // running mean (output_handler_variance.hpp:7-31)
    const float denom = mx(1.0f, (float)S.iteration + 1);
    float* cr = S.cumRadiance + 3 * pixel;
    cr[0] += (accR - cr[0]) / denom;
    cr[1] += (accG - cr[1]) / denom;
    cr[2] += (accB - cr[2]) / denom;
    float* ch = S.cumHitId + 2 * pixel;
    ch[0] += (accH0 - ch[0]) / denom;
    ch[1] += (accH1 - ch[1]) / denom;
    if (nans) atomicAdd(S.nanCounter, nans);
}

// rngs[i] = PCG64(u32(i + seed)) (gpu_pt.cu:140-144, rng.hpp:76-81)
__global__ void k_pt_init_rng(u64* rngs, int num, int seed) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= num) return;
    const u32 s = (u32)(i + seed);
    u64 st = scramble_seed(s);
    pcg_next(st);
    st += s;
    pcg_next(st);
    rngs[i] = st;
}

}  // namespace

void launch_pt_init_rng(u64* rngs, int num, int seed, hipStream_t stream) {
    hipLaunchKernelGGL(k_pt_init_rng, dim3((num + 255) / 256), dim3(256), 0, stream, rngs, num, seed);
    MFH_HIP(hipGetLastError());
}

void launch_pt_iterate(const PtScene& S, hipStream_t stream) {
    const dim3 grid((S.width + kPtTile - 1) / kPtTile, (S.height + kPtTile - 1) / kPtTile);
    if (S.trace.numInstances == 1)
        hipLaunchKernelGGL(HIP_KERNEL_NAME(k_pt_iterate<true>), grid, dim3(kPtThreads), 0, stream, S);
    else
        hipLaunchKernelGGL(HIP_KERNEL_NAME(k_pt_iterate<false>), grid, dim3(kPtThreads), 0, stream, S);
    MFH_HIP(hipGetLastError());
}

}  // namespace mfh
