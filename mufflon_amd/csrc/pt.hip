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
#include "pt_shading.hpp"
#include "traverse.hpp"

namespace mfh {

namespace {

using namespace shade;

constexpr int kPtTile = 16;
constexpr int kPtThreads = kPtTile * kPtTile;

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

// Hit attributes of first_intersection (intersection.cu:524-621) and the
// shading frame of tangent_space_geom_to_shader (tangent_space.hpp:10-60);
// also the interpolated uv.
__device__ __forceinline__ TS hit_frame(const PtScene& S, const TraceHit& h, float& hu, float& hv) {
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
    const float w2 = 1.f - h.w0 - h.w1;
    hu = u0x * h.w0 + u1x * h.w1 + u2x * w2;
    hv = u0y * h.w0 + u1y * h.w1 + u2y * w2;
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

    // per-iteration target values: Radiance 0-2, Hit ID 3-4, Position 5-7,
    // Depth 8, Normal 9-11, UV 12-13, Albedo 14-16, Lightness 17
    float acc[18];
#pragma unroll
    for (int k = 0; k < 18; ++k) acc[k] = 0.0f;
    const u32 mask = S.targetMask;
    u32 nans = 0;
    // RenderBuffer::contribute + check_nan (render_target.hpp:62-76,189-195)
    auto add = [&](int target, int slot, float v) {
        if (!((mask >> target) & 1u)) return;
        if (v != v) {
            ++nans;
            v = 0.0f;
        }
        acc[slot] += v;
    };
    float guideWeight = 1.0f;
    float footA = 0.0f;  // FootprintV0::m_a (footprint.hpp:137-170)

    const V3 viewDir = v3(S.camDir[0], S.camDir[1], S.camDir[2]);
    const V3 up = v3(S.camUp[0], S.camUp[1], S.camUp[2]);
    V3 throughput = v3(1.0f, 1.0f, 1.0f);
    bool isCamera = true;
    V3 position = v3(S.camPos[0], S.camPos[1], S.camPos[2]);
    V3 incEnc = v3(0, 0, 0);
    TS ts;
    int hInst = -1, hPrim = -1;
    MatParams mat;
    mat.type = -1;
    pcg_next(rng);  // create_camera's RndSet2, unused by the pinhole
    int pathLen = 0;
    do {
        if (pathLen > 0 && pathLen + 1 >= S.minPathLength && pathLen + 1 <= S.maxPathLength) {
            const u64 neeSeed = pcg_next(rng);
            const float iz = sgnf(incEnc.z) *
                             fsqrt(clampf(1.0f - incEnc.x * incEnc.x - incEnc.y * incEnc.y, 0.0f, 1.0f));
            const V3 incident = v3(incEnc.x, incEnc.y, iz);
            for (int i = 0; i < S.neeCount; ++i) {
                const u64 x = pcg_next(rng);
                const Nee nee = connect(S, (u64)i, (u64)S.neeCount, neeSeed, position, sample_uniform((u32)x),
                                        sample_uniform((u32)(x >> 32)));
                float vCos, vPdf;
                const V3 value = mat_evaluate(ts, mat, incident, nee.dir, vCos, vPdf);
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
                        add(0, 0, cc.x);
                        add(0, 1, cc.y);
                        add(0, 2, cc.z);
                        add(7, 17, guideWeight * vCos);
                    }
                }
            }
        }
        // walk (random_walk.hpp:38-112)
        const u64 x0 = pcg_next(rng);
        const u64 x1 = pcg_next(rng);
        const float u0 = sample_uniform((u32)x0), u1 = sample_uniform((u32)(x0 >> 32));
        const float rndRoulette = sample_uniform((u32)pcg_next(rng));
        V3 excident, sThroughput;
        float pdfF;
        Medium medium{1.0f, v3(0.0f, 0.0f, 0.0f)};
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
            const float iz = sgnf(incEnc.z) *
                             fsqrt(clampf(1.0f - incEnc.x * incEnc.x - incEnc.y * incEnc.y, 0.0f, 1.0f));
            if (!mat_sample(ts, mat, v3(incEnc.x, incEnc.y, iz), u0, u1, x1, sThroughput, excident, pdfF)) break;
            medium = mat.medium(dot(excident, ts.geoN));  // path_util.hpp:346-351
        }
        throughput = throughput * sThroughput;
        const float cont = mn(mx(mx(throughput.x, throughput.y), throughput.z) + 0.1f, 1.0f);
        if (rndRoulette >= cont) break;
        throughput = throughput / cont;
        const V3 rayOrigin = add_epsilon(position, excident, isCamera ? v3(0, 0, 0) : ts.geoN);
        const TraceHit hit =
            trace_ray<false, false, kSingle, kPtThreads>(S.trace, rayOrigin, excident, kMaxSceneSize, lstack);
        const bool isVoid = hit.inst == -1;
        throughput = throughput * transmission(medium, isVoid ? kMaxSceneSize : hit.t);
        // PtVertexExt::update (pt_common.hpp:45-61): footprint and guide weight
        footA += 1.0f / fsqrt(pdfF);
        guideWeight = 1.f / (1.f + footA * footA / 1.0f);
        float hu = 0.0f, hv = 0.0f;
        hInst = hit.inst;
        hPrim = hit.prim;
        float incidentPdf;
        if (isVoid) {
            incEnc = excident;
            incidentPdf = pdfF * 1.0f;
        } else {
            position = rayOrigin + excident * hit.t;
            ts = hit_frame(S, hit, hu, hv);
            incEnc = v3(excident.x, excident.y, hit.t * sgnf(excident.z));
            mat = fetch(S, material_index(S, hit.inst, hit.prim), hu, hv);
            const float inCosAbs = fabsf(dot(excident, ts.shadingN));
            incidentPdf = pdfF * fabsf(inCosAbs) / (hit.t * hit.t);
        }
        isCamera = false;
        if (pathLen == 0) {
            add(1, 3, (float)hInst);
            add(1, 4, (float)hPrim);
        }
        ++pathLen;
        if (pathLen >= S.minPathLength) {
            V3 em;
            float connectPdf = 0.0f;
            if (isVoid) {  // get_emission, VOID (path_util.hpp:415-424)
                float pdfBack;
                em = evaluate_background(S, incEnc, pdfBack);
                if (nonzero(em)) connectPdf = background_pdf(S, pdfBack);
            } else {
                const float iz = sgnf(incEnc.z) *
                                 fsqrt(clampf(1.0f - incEnc.x * incEnc.x - incEnc.y * incEnc.y, 0.0f, 1.0f));
                em = emission(mat, ts.geoN, neg(v3(incEnc.x, incEnc.y, iz)));
                if (nonzero(em)) connectPdf = light_pdf(S, hInst, hPrim);
            }
            if (nonzero(em) && pathLen > 1) {
                const float mis = 1.0f / (1.0f + (float)S.neeCount * (connectPdf / incidentPdf));
                em = em * mis;
            }
            const V3 cc = throughput * em;  // contributed even when zero
            add(0, 0, cc.x);
            add(0, 1, cc.y);
            add(0, 2, cc.z);
            if (mask & 0xFCu) {  // Position .. Lightness
                const V3 pos = isVoid ? v3(0, 0, 0) : position;
                add(2, 5, guideWeight * pos.x);
                add(2, 6, guideWeight * pos.y);
                add(2, 7, guideWeight * pos.z);
                add(3, 8, guideWeight * (isVoid ? 0.0f : fabsf(incEnc.z)));
                const V3 nrm = isVoid ? v3(0, 0, 0) : ts.shadingN;
                add(4, 9, guideWeight * nrm.x);
                add(4, 10, guideWeight * nrm.y);
                add(4, 11, guideWeight * nrm.z);
                add(5, 12, guideWeight * hu);
                add(5, 13, guideWeight * hv);
                V3 alb = v3(0, 0, 0);
                if (!isVoid) {
                    if (mat.type == kMatLambert || mat.type == kMatTorrance)
                        alb = mat.albedo;
                    else if (mat.type == kMatMicrofacet)
                        alb = v3(1.0f / (1.0f + mat.inner.absorption.x), 1.0f / (1.0f + mat.inner.absorption.y),
                                 1.0f / (1.0f + mat.inner.absorption.z));
                }
                add(6, 14, guideWeight * alb.x);
                add(6, 15, guideWeight * alb.y);
                add(6, 16, guideWeight * alb.z);
                add(7, 17, guideWeight * ((em.x + em.y + em.z) / 3.0f));
            }
        }
        if (isVoid) break;
    } while (pathLen < S.maxPathLength);

    S.rngs[pixel] = rng;
    // running means (output_handler_variance.hpp:7-31)
    const float denom = mx(1.0f, (float)S.iteration + 1);
    int slot = 0;
#pragma unroll
    for (int t = 0; t < kNumTargets; ++t) {
        const int nc = kTargetChannels[t];
        if ((mask >> t) & 1u) {
            float* c = S.cum[t] + (size_t)nc * pixel;
#pragma unroll
            for (int k = 0; k < nc; ++k) c[k] += (acc[slot + k] - c[k]) / denom;
        }
        slot += nc;
    }
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
