// Wireframe renderer for gfx950: the GPU device of Mufflon's "Wireframe"
// renderer (GpuWireframe, gpu_wireframe.cu:19-51, sample_wireframe of
// wireframe_common.hpp:94-181) over the scene, camera and background of a
// path tracer.
//
// One thread per pixel, 16x16 tiles. The pixel's PCG64 is seeded with this
// iteration's seed (drawn on the host from the renderer's own generator),
// the pinhole camera ray is jittered as the path tracer's, and the ray walks
// through the scene hit after hit (up to maxTraceDepth traversals): at each
// hit, the closest point on the hit triangle's or quad's edges (in world
// space) is measured against the camera ray (for a sphere, 0.05 r times the
// cosine between the ray and the direction to the center); within lineWidth
// pixel footprints the
// pixel gets the border colour 1, otherwise the walk continues from the hit
// (at least 1e-3 further). A walk that leaves the scene adds the background
// radiance if it is positive. Values go into the running mean (and
// variance) of the Border target like every target of the output handler.
#include "pt_path.hpp"

namespace mfh {

namespace {

using namespace shade;
using namespace path;

constexpr int kWfTile = 16;
constexpr int kWfThreads = kWfTile * kWfTile;

// Closest point of the triangle's three edge lines to `hit`
// (computeClosestLinePoint, wireframe_common.hpp:16-40).
__device__ __forceinline__ V3 closest_line_point(V3 A, V3 B, V3 C, V3 hit) {
    const V3 AB = B - A, AC = C - A, BC = C - B;
    const V3 AP = hit - A, BP = hit - B;
    const V3 onAB = A + AB * (dot(AP, AB) / dot(AB, AB));
    const V3 onAC = A + AC * (dot(AP, AC) / dot(AC, AC));
    const V3 onBC = B + BC * (dot(BP, BC) / dot(BC, BC));
    const float dAB = dot(onAB - hit, onAB - hit);
    const float dAC = dot(onAC - hit, onAC - hit);
    const float dBC = dot(onBC - hit, onBC - hit);
    if (dAB <= dAC && dAB <= dBC) return onAB;
    if (dAC <= dAB && dAC <= dBC) return onAC;
    return onBC;
}

// Closest point of a quad's four edge lines (A = v0, B = v1, C = v3, D = v2:
// edges AB, AC, BD, CD; computeClosestLinePoint, wireframe_common.hpp:42-77).
__device__ __forceinline__ V3 closest_quad_line_point(V3 A, V3 B, V3 C, V3 D, V3 hit) {
    const V3 AB = B - A, AC = C - A, BD = D - B, CD = D - C;
    const V3 AP = hit - A, BP = hit - B, CP = hit - C;
    const V3 onAB = A + AB * (dot(AP, AB) / dot(AB, AB));
    const V3 onAC = A + AC * (dot(AP, AC) / dot(AC, AC));
    const V3 onBD = B + BD * (dot(BP, BD) / dot(BD, BD));
    const V3 onCD = C + CD * (dot(CP, CD) / dot(CD, CD));
    const float dAB = dot(onAB - hit, onAB - hit);
    const float dAC = dot(onAC - hit, onAC - hit);
    const float dBD = dot(onBD - hit, onBD - hit);
    const float dCD = dot(onCD - hit, onCD - hit);
    if (dAB <= dAC && dAB <= dBD && dAB <= dCD) return onAB;
    if (dAC <= dAB && dAC <= dBD && dAC <= dCD) return onAC;
    if (dBD <= dAB && dBD <= dAC && dBD <= dCD) return onBD;
    return onCD;
}

template <int kMode>
__global__ __launch_bounds__(kWfThreads) void k_wireframe(PtScene S, const u32* __restrict__ seeds, float lineWidth,
                                                          int maxTraceDepth) {
    __shared__ int sStack[kLdsDepth * kWfThreads];
    const int px = blockIdx.x * kWfTile + (threadIdx.x % kWfTile);
    const int py = blockIdx.y * kWfTile + (threadIdx.x / kWfTile);
    if (px >= S.width || py >= S.height) return;
    const size_t pixel = (size_t)px + (size_t)py * S.width;
    int* const lstack = sStack + threadIdx.x;
    // PCG64(u32 seed) (rng.hpp:76-81)
    const u32 seed = seeds[pixel];
    u64 rng = scramble_seed(seed);
    pcg_next(rng);
    rng += seed;
    pcg_next(rng);
    pcg_next(rng);  // create_camera's RndSet2, unused by the pinhole
    const u64 x0 = pcg_next(rng);
    V3 dir;
    float pdfF;
    camera_ray(S, px, py, x0, dir, pdfF);
    const V3 camPos = v3(S.camPos[0], S.camPos[1], S.camPos[2]);
    V3 origin = camPos;  // the camera vertex has no normal: no epsilon offset
    Acc acc;
    acc.init(1u << kTargetRadiance);  // the Border target rides in the first slot
    float totalDistance = 0.0f;
    for (int i = 0; i < maxTraceDepth; ++i) {
        const TraceHit h = trace_mode<kMode, false, kWfThreads, kLdsDepth, kPtLeafBatch, false>(
            S.trace, origin, dir, kMaxSceneSize, lstack);
        if (h.inst == -1) {
            float pdfBack;
            const V3 bg = evaluate_background(S, dir, pdfBack);
            if (bg.x > 0.0f || bg.y > 0.0f || bg.z > 0.0f) {
                acc.add(kTargetRadiance, 0, bg.x);
                acc.add(kTargetRadiance, 1, bg.y);
                acc.add(kTargetRadiance, 2, bg.z);
            }
            break;
        }
        const float hitDist = h.t < 0.001f ? 0.001f : h.t;
        const V3 hit = origin + dir * hitDist;
        totalDistance += hitDist;
        const float pixelSize = totalDistance * S.tanVFov / (float)S.height;
        // the hit triangle in world space (compute_instance_to_world_transformation)
        const M34 M = inst_to_world(S.trace.worldToInst + 12 * (size_t)h.inst);
        const DevPtLod L = S.ptLods[S.trace.instLod[h.inst]];
        float projDistToRim;
        if (h.prim >= L.numTris + L.numQuads) {
            // sphere (computeDistToRim, wireframe_common.hpp:79-89, scaled by 0.05 at :168)
            const float4 sp = L.spheres[h.prim - L.numTris - L.numQuads];
            const V3 center = xform_point(M, v3(sp.x, sp.y, sp.z));
            projDistToRim = 0.05f * (sp.w * dot(normalize(center - hit), dir));
        } else {
            V3 closest;
            if (h.prim >= L.numTris) {
                const uint32_t* q = L.quads + 4 * (size_t)(h.prim - L.numTris);
                const V3 A = xform_point(M, lod_vert(L, q[0])), B = xform_point(M, lod_vert(L, q[1]));
                const V3 C = xform_point(M, lod_vert(L, q[3])), D = xform_point(M, lod_vert(L, q[2]));
                closest = closest_quad_line_point(A, B, C, D, hit);
            } else {
                const u32 i0 = L.tris[3 * (size_t)h.prim], i1 = L.tris[3 * (size_t)h.prim + 1],
                          i2 = L.tris[3 * (size_t)h.prim + 2];
                closest = closest_line_point(xform_point(M, lod_vert(L, i0)), xform_point(M, lod_vert(L, i1)),
                                             xform_point(M, lod_vert(L, i2)), hit);
            }
            const V3 lineSegment = closest - camPos;
            const V3 projected = lineSegment - dir * dot(lineSegment, dir);
            projDistToRim = len(projected);
        }
        if (projDistToRim > lineWidth * pixelSize) {
            origin = origin + dir * hitDist;
        } else {
            acc.add(kTargetRadiance, 0, 1.0f);
            acc.add(kTargetRadiance, 1, 1.0f);
            acc.add(kTargetRadiance, 2, 1.0f);
            break;
        }
    }
    accumulate(S, pixel, S.iteration, acc.v, acc.mask);
    if (acc.nans) atomicAdd(S.nanCounter, acc.nans);
}

}  // namespace

void launch_wireframe(const PtScene& S, const uint32_t* seeds, float lineWidth, int maxTraceDepth,
                      hipStream_t stream) {
    const dim3 grid((S.width + kWfTile - 1) / kWfTile, (S.height + kWfTile - 1) / kWfTile);
    const int mode = trace_mode_of(S.trace);
    if (mode == kTraceSingle)
        hipLaunchKernelGGL(HIP_KERNEL_NAME(k_wireframe<kTraceSingle>), grid, dim3(kWfThreads), 0, stream, S, seeds,
                           lineWidth, maxTraceDepth);
    else if (mode == kTraceMulti)
        hipLaunchKernelGGL(HIP_KERNEL_NAME(k_wireframe<kTraceMulti>), grid, dim3(kWfThreads), 0, stream, S, seeds,
                           lineWidth, maxTraceDepth);
    else
        hipLaunchKernelGGL(HIP_KERNEL_NAME(k_wireframe<kTraceMixed>), grid, dim3(kWfThreads), 0, stream, S, seeds,
                           lineWidth, maxTraceDepth);
    MFH_HIP(hipGetLastError());
}

}  // namespace mfh
