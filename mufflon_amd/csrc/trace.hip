// Closest-hit and any-hit trace kernels over ray buffers (one ray per lane),
// built on the traversal of traverse.hpp.
#include "traverse.hpp"

namespace mfh {

namespace {

constexpr int kThreads = kTraceThreads;

// kAny: stop at the first accepted hit. kStats: count per ray the internal
// nodes visited (an object-box test at instance entry counts as one) and the
// triangles tested. kSingle: the scene has one instance (no instance level).
// kMixed: the scene's LoDs hold quads or spheres.
template <bool kAny, bool kStats, bool kSingle, bool kMixed>
__global__ __launch_bounds__(kThreads) void k_trace(TraceParams p, const float4* __restrict__ rays, int n,
                                                   float4* __restrict__ hits, float2* __restrict__ bary,
                                                   int* __restrict__ occluded, uint2* __restrict__ stats) {
    __shared__ int sStack[kLdsDepth * kThreads];
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float4 r0 = rays[2 * (size_t)i], r1 = rays[2 * (size_t)i + 1];
    TraceHit h;
    if constexpr (kSingle && !kAny && !kStats)
        h = trace_closest_held<kThreads, kLdsDepth, 1, kMixed>(p, v3(r0.x, r0.y, r0.z), v3(r1.x, r1.y, r1.z), r0.w,
                                                               sStack + threadIdx.x);
    else
        h = trace_ray<kAny, kStats, kSingle, kThreads, kLdsDepth, 1, kMixed>(p, v3(r0.x, r0.y, r0.z),
                                                                           v3(r1.x, r1.y, r1.z), r0.w,
                                                                           sStack + threadIdx.x);
    if (kAny) {
        occluded[i] = h.overflow ? -1 : (h.inst != -1 ? 1 : 0);
    } else {
        hits[i] = make_float4(h.t, __int_as_float(h.inst), __int_as_float(h.prim), h.overflow ? 1.0f : 0.0f);
        if (bary) bary[i] = make_float2(h.w0, h.w1);
    }
    if (kStats) stats[i] = make_uint2(h.nodeVisits, h.primTests);
}

template <bool kAny, bool kStats>
void launch_trace(const TraceParams& p, const float* rays, int n, float4* hits, float2* bary, int* occ, uint2* stats,
                  hipStream_t stream) {
    const dim3 grid((n + kThreads - 1) / kThreads), block(kThreads);
    const float4* r = reinterpret_cast<const float4*>(rays);
    if (p.numInstances == 1 && p.mixed)
        hipLaunchKernelGGL(HIP_KERNEL_NAME(k_trace<kAny, kStats, true, true>), grid, block, 0, stream, p, r, n, hits,
                           bary, occ, stats);
    else if (p.numInstances == 1)
        hipLaunchKernelGGL(HIP_KERNEL_NAME(k_trace<kAny, kStats, true, false>), grid, block, 0, stream, p, r, n, hits,
                           bary, occ, stats);
    else if (p.mixed)
        hipLaunchKernelGGL(HIP_KERNEL_NAME(k_trace<kAny, kStats, false, true>), grid, block, 0, stream, p, r, n, hits,
                           bary, occ, stats);
    else
        hipLaunchKernelGGL(HIP_KERNEL_NAME(k_trace<kAny, kStats, false, false>), grid, block, 0, stream, p, r, n,
                           hits, bary, occ, stats);
    MFH_HIP(hipGetLastError());
}

}  // namespace

TraceParams trace_params(const Scene& sc, bool reference) {
    TraceParams p;
    p.numInstances = sc.numInstances;
    p.worldToInst = sc.dWorldToInst.as<float>();
    p.instLod = sc.dInstLod.as<int32_t>();
    p.lods = (reference ? sc.dLodsRef : sc.dLods).as<DevLod>();
    p.tlasNodes = sc.tlasNodes.ptr;
    p.tlasPrimIds = sc.tlasPrimIds.as<int32_t>();
    p.tlasNumInternal = sc.tlasNumInternal;
    p.tlasCoordMax = 0.0f;
    for (int k = 0; k < 6; ++k) p.tlasCoordMax = std::max(p.tlasCoordMax, std::fabs(sc.box[k]));
    p.mixed = sc.mixed ? 1 : 0;
    return p;
}

void launch_trace_closest(const TraceParams& p, const float* rays, int n, float4* hits, float2* bary, uint2* stats,
                          hipStream_t stream) {
    if (n <= 0) return;
    if (stats)
        launch_trace<false, true>(p, rays, n, hits, bary, nullptr, stats, stream);
    else
        launch_trace<false, false>(p, rays, n, hits, bary, nullptr, nullptr, stream);
}

void launch_trace_any(const TraceParams& p, const float* rays, int n, int32_t* occluded, hipStream_t stream) {
    if (n <= 0) return;
    launch_trace<true, false>(p, rays, n, nullptr, nullptr, occluded, nullptr, stream);
}

}  // namespace mfh
