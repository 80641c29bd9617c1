// Light tracer for gfx950: the GPU device of Mufflon's "Lighttracer"
// (GpuLightTracer, gpu_lt.cpp:25-37, gpu_lt.cu:19-101, lt_sample of
// lt_common.hpp:17-78) over the scene, camera and positional lights of a
// path tracer.
//
// One thread per light path; path i of an iteration draws from generator i
// (seeded like the path tracer's) and emits a photon from the light tree
// with index i of numPixels (emit, light_tree_sampling.hpp:81-187: the index
// spreads the photons over the tree's flux intervals), then walks the
// scene on adjoint BSDF samples (walk, random_walk.hpp:38-112). Every
// vertex, the light's own included, is connected to the pinhole camera:
// the direction is projected to its pixel (pinholecam_project,
// pinhole.hpp:86-116), the adjoint BSDF (or the emitter's distribution)
// weights it, a shadow ray checks it, and the contribution is added to the
// pixel's Radiance, Lightness and (when recorded) Density with float
// atomics. The per-iteration splat image then goes into the running means
// (and variances) of the three targets like every target of the output
// handler. The order of the atomic additions is not fixed, so sums agree
// with a sequential evaluation to rounding, not bit for bit.
#include "pt_path.hpp"

namespace mfh {

namespace {

using namespace shade;
using namespace path;

constexpr int kLtThreads = 256;

// Sampler result of a positional light (Emitter, light_sampling.hpp:21-52).
struct LtEmitter {
    V3 position, intensity, normal;
    V3 absorption;  // of the medium the photon leaves into (vacuum around point lights)
    bool area;
};

// sample_light_pos (light_sampling.hpp:80-199) of a light of the tree:
// point, area triangle, quad and sphere.
__device__ LtEmitter light_position(const PtScene& S, uint16_t type, u32 offset, float u0, float u1) {
    const uint8_t* light = S.lightMem + offset;
    LtEmitter e;
    if (type == kLightPoint) {
        const float4 a = ld4(light), b = ld4(light + 16);
        e.position = v3(a.x, a.y, a.z);
        e.intensity = v3(b.x, b.y, b.z);
        e.normal = v3(0, 0, 0);
        e.absorption = v3(0, 0, 0);
        e.area = false;
        return e;
    }
    e.area = true;
    if (type == kLightAreaQuad) {
        V3 pv[4];
        quad_record(light, pv);
        const float4 w0 = ld4(light + 48), w1 = ld4(light + 64);
        const int matIdx = (int)(*reinterpret_cast<const uint16_t*>(light + 80));
        const V3 tangentX = pv[1] + smul(u0, pv[3]);
        const V3 tangentY = pv[2] + smul(u1, pv[3]);
        V3 normal = cross(tangentY, tangentX);
        const float area = len(normal);
        e.normal = normal / area;
        e.position = pv[0] + tangentX * u1 + pv[2] * u0;
        const float lu = w0.x + w0.z * u1 + w1.x * u0 + w1.z * (u0 * u1);
        const float lv = w0.y + w0.w * u1 + w1.y * u0 + w1.w * (u0 * u1);
        const MatParams m = fetch(S, matIdx, lu, lv);
        e.intensity = emission(m, v3(0, 0, 1), v3(0, 0, 1)) * area;
        e.absorption = m.outer.absorption;
        return e;
    }
    if (type == kLightAreaSphere) {
        const float4 q0 = ld4(light);
        const int matIdx = (int)(*reinterpret_cast<const uint16_t*>(light + 16));
        // sample_dir_sphere_uniform (sampling.hpp:101-109)
        const float cosTheta = u0 * 2.0f - 1.0f;
        const float sinTheta = fsqrt((1.0f - cosTheta) * (1.0f + cosTheta));
        const float phi = u1 * 2.0f * kPi;
        const V3 n = v3(sinTheta * fsin(phi), sinTheta * fcos(phi), cosTheta);
        // the emitter's uv is (u1, acos(n.z)), the polar angle unnormalized
        // (light_sampling.hpp:181)
        const MatParams m = fetch(S, matIdx, u1, facos(n.z));
        const V3 radiance = emission(m, v3(0, 0, 1), v3(0, 0, 1));
        e.absorption = m.outer.absorption;
        const float area = 4 * kPi * q0.w * q0.w;
        e.position = v3(q0.x, q0.y, q0.z) + n * q0.w;
        e.normal = n;
        e.intensity = radiance * area;
        return e;
    }
    const float4 q0 = ld4(light), q1 = ld4(light + 16), q2 = ld4(light + 32), q3 = ld4(light + 48);
    const V3 p0 = v3(q0.x, q0.y, q0.z), e1 = v3(q0.w, q1.x, q1.y), e2 = v3(q1.z, q1.w, q2.x);
    const int matIdx = (int)(__float_as_uint(q2.y) & 0xFFFFu);
    V3 normal = cross(e1, e2);
    const float area2 = len(normal);
    e.normal = normal / area2;
    const float area = area2 * 0.5f;
    const float s = 1.f - fsqrt(1.f - u0);
    const float t = (1.f - s) * u1;
    e.position = p0 + e1 * s + e2 * t;
    const float lu = q2.z + q3.x * s + q3.z * t;
    const float lv = q2.w + q3.y * s + q3.w * t;
    const MatParams m = fetch(S, matIdx, lu, lv);
    e.intensity = emission(m, v3(0, 0, 1), v3(0, 0, 1)) * area;
    e.absorption = m.outer.absorption;
    return e;
}

// emit (light_tree_sampling.hpp:81-187) for trees of positional lights (the
// host refuses backgrounds that emit): the flux intervals of the tree's
// nodes partition the 64-bit range, which path `index` of numIndices enters
// at seed + index * (2^64-1) / numIndices; the intensity is divided by the
// probability of the chosen light (adjustPdf).
__device__ LtEmitter emit(const PtScene& S, u64 index, u64 numIndices, u64 seed, float u0, float u1) {
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
        const float4 a = ld4(S.lightMem + offset), b = ld4(S.lightMem + offset + 16);
        const u32 types = __float_as_uint(a.z);
        const float probLeft = a.x / (a.x + b.x);
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
    LtEmitter e = light_position(S, type, offset, u0, u1);
    e.intensity = e.intensity / lightProb;
    return e;
}

// pinholecam_project (pinhole.hpp:86-116): the pixel a direction from the
// camera falls on and the camera's importance w (the pdf of sampling it).
__device__ __forceinline__ bool project(const PtScene& S, V3 dir, int& px, int& py, float& w) {
    const V3 viewDir = v3(S.camDir[0], S.camDir[1], S.camDir[2]);
    const V3 up = v3(S.camUp[0], S.camUp[1], S.camUp[2]);
    const float cosOut = dot(viewDir, dir);
    if (cosOut < 0.0f) return false;
    const V3 xAxis = cross(viewDir, up);
    const float den = cosOut * S.tanVFov;
    float ux = dot(xAxis, dir) / den;
    const float uy = dot(up, dir) / den;
    const float aspect = (float)S.width / (float)S.height;
    ux /= aspect;
    if (!(ux > -1 && ux <= 1 && uy > -1 && uy <= 1)) return false;
    px = (int)floorf((ux * 0.5f + 0.5f) * (float)S.width);
    py = (int)floorf((uy * 0.5f + 0.5f) * (float)S.height);
    if (px >= S.width) px = S.width - 1;
    if (py >= S.height) py = S.height - 1;
    const float pixelArea = (2.0f * S.tanVFov) * (2.0f * S.tanVFov) * aspect;
    w = 1.0f / (pixelArea * cosOut * cosOut * cosOut);
    return true;
}

// Splat image of one iteration: Radiance (3), Lightness, Density per pixel.
constexpr int kLtChannels = 5;

template <int kMode>
__global__ __launch_bounds__(kLtThreads) void k_lt_iterate(PtScene S, float* __restrict__ splat, int minPathLength,
                                                          int maxPathLength, int density) {
    __shared__ int sStack[kLdsDepth * kLtThreads];
    const int num = S.width * S.height;
    const int idx = blockIdx.x * kLtThreads + (int)threadIdx.x;
    if (idx >= num) return;
    int* const lstack = sStack + threadIdx.x;
    u64 rng = S.rngs[idx];
    const V3 camPos = v3(S.camPos[0], S.camPos[1], S.camPos[2]);
    V3 throughput = v3(1.0f, 1.0f, 1.0f);
    const u64 xStart = pcg_next(rng);
    const u64 treeSeed = pcg_next(rng);
    const LtEmitter light =
        emit(S, (u64)idx, (u64)num, treeSeed, sample_uniform((u32)xStart), sample_uniform((u32)(xStart >> 32)));
    pcg_next(rng);  // create_camera's RndSet2, unused by the pinhole
    Vertex vx;
    int pathLen = 0;
    do {
        if (pathLen + 1 >= minPathLength && pathLen + 1 <= maxPathLength) {
            // get_connection(camera, vertex) (path_util.hpp:384-399)
            const V3 vpos = pathLen == 0 ? light.position : vx.position;
            const V3 conn = vpos - camPos;
            const float distSq = dot(conn, conn);
            const float dist = fsqrt(distSq);
            const float dden = dist + 1e-20f * (dist < 0.0f ? -1.0f : 1.0f);
            const V3 dir = conn / dden;
            int px, py;
            float w;
            if (project(S, dir, px, py, w)) {
                V3 lval;
                float lcos;
                V3 geoN;
                float geomFactor;
                if (pathLen == 0) {
                    if (light.area) {  // evaluate_area (lights.hpp:286-295)
                        const float c = dot(light.normal, neg(dir));
                        lval = c <= 0.0f ? v3(0, 0, 0) : light.intensity;
                        lcos = c <= 0.0f ? 0.0f : c;
                        geoN = light.normal;
                        geomFactor = dot(dir, light.normal);
                    } else {  // evaluate_point (lights.hpp:262-265)
                        lval = light.intensity;
                        lcos = 1.0f;
                        geoN = v3(0, 0, 0);
                        geomFactor = 0.0f;
                    }
                } else {
                    lval = mat_evaluate_adjoint(vx.ts, vx.mat, vx.incident(), neg(dir), lcos);
                    geoN = vx.ts.geoN;
                    geomFactor = dot(dir, vx.ts.shadingN);
                }
                V3 bxdfProd = v3(w * lval.x, w * lval.y, w * lval.z);
                const float cosProd = 1.0f * lcos;
                if (density || ((bxdfProd.x > 0.0f || bxdfProd.y > 0.0f || bxdfProd.z > 0.0f) && cosProd > 0.0f)) {
                    // any_intersection (intersection.cu:624-637); the camera has no normal
                    const V3 a = add_epsilon(camPos, dir, v3(0, 0, 0));
                    const V3 b = add_epsilon(vpos, neg(dir), geoN);
                    const V3 d = b - a;
                    const float tmax = len(d);
                    const TraceHit sh =
                        trace_mode<kMode, true, kLtThreads, kLdsDepth, 1>(S.trace, a, d / tmax, tmax, lstack);
                    if (sh.inst == -1) {
                        bxdfProd = bxdfProd / distSq;
                        const V3 rad = throughput * cosProd * bxdfProd;
                        float* o = splat + (size_t)kLtChannels * ((size_t)px + (size_t)py * S.width);
                        atomicAdd(o, rad.x);
                        atomicAdd(o + 1, rad.y);
                        atomicAdd(o + 2, rad.z);
                        atomicAdd(o + 3, (throughput.x + throughput.y + throughput.z) / 3.0f * cosProd);
                        if (density) atomicAdd(o + 4, w * 1.0f / distSq * fabsf(geomFactor));
                    }
                }
            }
        }
        // walk (random_walk.hpp:38-112), adjoint
        const u64 x0 = pcg_next(rng);
        const u64 x1 = pcg_next(rng);
        const float rndRoulette = sample_uniform((u32)pcg_next(rng));
        V3 rayOrigin, dirOut, sThroughput, absorption, geoOut;
        if (pathLen == 0) {
            // the light's direction (sample_light_dir_point / _area,
            // light_sampling.hpp:85-89, 201-211)
            const float u0 = sample_uniform((u32)x0), u1 = sample_uniform((u32)(x0 >> 32));
            if (light.area) {
                const float cosT = fsqrt(u0);
                const float sinT = fsqrt(1.0f - u0);
                const float phi = u1 * 2 * kPi;
                const float lx = sinT * fsin(phi), ly = sinT * fcos(phi);
                const V3 tx = normalize(perpendicular(light.normal));
                const V3 ty = cross(light.normal, tx);
                dirOut = smul(lx, tx) + smul(ly, ty) + smul(cosT, light.normal);
                sThroughput = light.intensity * kPi;
                geoOut = light.normal;
            } else {
                const float cosT = u0 * 2.0f - 1.0f;
                const float sinT = fsqrt((1.0f - cosT) * (1.0f + cosT));
                const float phi = u1 * 2.0f * kPi;
                dirOut = v3(sinT * fsin(phi), sinT * fcos(phi), cosT);
                sThroughput = light.intensity * 4 * kPi;
                geoOut = v3(0, 0, 0);
            }
            rayOrigin = light.position;
            absorption = light.absorption;
        } else {
            float pdfF;
            if (!mat_sample(vx.ts, vx.mat, vx.incident(), sample_uniform((u32)x0), sample_uniform((u32)(x0 >> 32)),
                            x1, sThroughput, dirOut, pdfF, true))
                break;
            absorption = vx.mat.medium(dot(dirOut, vx.ts.geoN)).absorption;
            rayOrigin = vx.position;
            geoOut = vx.ts.geoN;
        }
        throughput = throughput * sThroughput;
        const float cont = mn(mx(mx(throughput.x, throughput.y), throughput.z) + 0.1f, 1.0f);
        if (rndRoulette >= cont) break;
        throughput = throughput / cont;
        rayOrigin = add_epsilon(rayOrigin, dirOut, geoOut);
        const TraceHit hit = trace_mode<kMode, false, kLtThreads, kLdsDepth, 1>(S.trace, rayOrigin, dirOut,
                                                                                 kMaxSceneSize, lstack);
        throughput = throughput * transmission(Medium{1.0f, absorption}, hit.inst == -1 ? kMaxSceneSize : hit.t);
        if (hit.inst == -1) break;
        make_vertex(S, rayOrigin, dirOut, 0.0f, hit.t, hit.inst, hit.prim, hit.w0, hit.w1, vx);
        ++pathLen;
    } while (pathLen < maxPathLength);
    S.rngs[idx] = rng;
}

__global__ void k_lt_clear(float* p, size_t n) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0.0f;
}

// Running mean (and variance) of the splat image (UpdateIter,
// output_handler_variance.hpp:7-31).
__global__ void k_lt_accumulate(const float* __restrict__ splat, float* __restrict__ mean, float* __restrict__ var,
                                size_t n, int iteration) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float denom = mx(1.0f, (float)iteration + 1);
    const float iter = splat[i];
    const float diff = iter - mean[i];
    const float cum = mean[i] + diff / denom;
    mean[i] = cum;
    if (var) var[i] += (diff * (iter - cum) - var[i]) / mx(1.0f, (float)iteration);
}

}  // namespace

void launch_lighttracer(const PtScene& S, float* splat, float* mean, float* var, int minPathLength,
                        int maxPathLength, bool density, hipStream_t stream) {
    const size_t n = (size_t)kLtChannels * S.width * S.height;
    const int clearBlocks = (int)((n + 255) / 256);
    hipLaunchKernelGGL(k_lt_clear, dim3(clearBlocks), dim3(256), 0, stream, splat, n);
    const int blocks = (S.width * S.height + kLtThreads - 1) / kLtThreads;
    const int mode = trace_mode_of(S.trace);
#define MFH_LT(M)                                                                                              \
    hipLaunchKernelGGL(HIP_KERNEL_NAME(k_lt_iterate<M>), dim3(blocks), dim3(kLtThreads), 0, stream, S, splat, \
                       minPathLength, maxPathLength, density ? 1 : 0)
    if (mode == kTraceSingle)
        MFH_LT(kTraceSingle);
    else if (mode == kTraceMixed)
        MFH_LT(kTraceMixed);
    else
        MFH_LT(kTraceMulti);
#undef MFH_LT
    hipLaunchKernelGGL(k_lt_accumulate, dim3(clearBlocks), dim3(256), 0, stream, splat, mean, var, n, S.iteration);
    MFH_HIP(hipGetLastError());
}

}  // namespace mfh
