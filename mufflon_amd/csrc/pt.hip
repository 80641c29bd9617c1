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
#include "pt_path.hpp"

namespace mfh {

namespace {

using namespace shade;
using namespace path;

constexpr int kPtTile = 16;
constexpr int kPtThreads = kPtTile * kPtTile;

// kLeaf: triangles of a BVH leaf loaded ahead of their tests (1 or 8).
// Three waves per SIMD: left to itself the compiler gives the whole path loop
// 256 VGPRs + ~80 AGPRs (one wave per SIMD), too few waves to hide the
// traversal's memory latency; at three it spills some state to scratch and
// the 10M-triangle terrain iteration drops from 27.4 to 14.7 ms (leaf batch
// 8; 34.8 -> 18.3 ms with 1), two waves 17.6 ms, four 17.6 ms (more spills).
// kMode: traversal specialization (trace_mode_of: one instance, instance
// level, or a scene with quads or spheres).
template <int kMode, int kLeaf>
__global__ __launch_bounds__(kPtThreads) __attribute__((amdgpu_waves_per_eu(3))) void k_pt_iterate(PtScene S) {
    __shared__ int sStack[kLdsDepth * kPtThreads];
    const int px = blockIdx.x * kPtTile + (threadIdx.x % kPtTile);
    const int py = blockIdx.y * kPtTile + (threadIdx.x / kPtTile);
    if (px >= S.width || py >= S.height) return;
    const size_t pixel = (size_t)px + (size_t)py * S.width;
    int* const lstack = sStack + threadIdx.x;
    u64 rng = S.rngs[pixel];
    Acc acc;
    acc.init(S.targetMask);
    float guideWeight = 1.0f;
    float footA = 0.0f;  // FootprintV0::m_a (footprint.hpp:137-170)
    V3 throughput = v3(1.0f, 1.0f, 1.0f);
    Vertex vx;
    pcg_next(rng);  // create_camera's RndSet2, unused by the pinhole
    int pathLen = 0;
    do {
        if (pathLen > 0 && pathLen + 1 >= S.minPathLength && pathLen + 1 <= S.maxPathLength) {
            const u64 neeSeed = pcg_next(rng);
            for (int i = 0; i < S.neeCount; ++i) {
                const u64 x = pcg_next(rng);
                NeeCandidate c;
                if (nee_candidate(S, vx, throughput, guideWeight, i, neeSeed, x, c)) {
                    const TraceHit sh =
                        trace_mode<kMode, true, kPtThreads, kLdsDepth, kLeaf>(S.trace, c.a, c.dir, c.tmax, lstack);
                    if (sh.inst == -1) add_nee(c, acc);
                }
            }
        }
        const u64 x0 = pcg_next(rng);
        const u64 x1 = pcg_next(rng);
        const float rndRoulette = sample_uniform((u32)pcg_next(rng));
        V3 rayOrigin, dir, absorption;
        float pdfF;
        if (pathLen == 0) {
            camera_ray(S, px, py, x0, dir, pdfF);
            rayOrigin = v3(S.camPos[0], S.camPos[1], S.camPos[2]);
            absorption = v3(0.0f, 0.0f, 0.0f);
        } else if (!walk_surface(vx, x0, x1, rndRoulette, throughput, rayOrigin, dir, pdfF, absorption)) {
            break;
        }
        const TraceHit hit =
            trace_mode<kMode, false, kPtThreads, kLdsDepth, kLeaf>(S.trace, rayOrigin, dir, kMaxSceneSize, lstack);
        throughput = throughput * transmission(Medium{1.0f, absorption}, hit.inst == -1 ? kMaxSceneSize : hit.t);
        footA += 1.0f / fsqrt(pdfF);
        guideWeight = 1.f / (1.f + footA * footA / 1.0f);
        const V3 prevPos = pathLen == 0 ? rayOrigin : vx.position;
        make_vertex(S, rayOrigin, dir, pdfF, hit.t, hit.inst, hit.prim, hit.w0, hit.w1, vx);
        ++pathLen;
        vertex_targets(S, vx, pathLen, throughput, guideWeight, acc, prevPos);
        if (vx.isVoid) break;
    } while (pathLen < S.maxPathLength);

    S.rngs[pixel] = rng;
    accumulate(S, pixel, S.iteration, acc.v, acc.mask);
    if (acc.nans) atomicAdd(S.nanCounter, acc.nans);
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

void launch_pt_iterate(const PtScene& S, int leafBatch, hipStream_t stream) {
    const dim3 grid((S.width + kPtTile - 1) / kPtTile, (S.height + kPtTile - 1) / kPtTile);
    const int mode = trace_mode_of(S.trace);
#define MFH_ITERATE(M, L) hipLaunchKernelGGL(HIP_KERNEL_NAME(k_pt_iterate<M, L>), grid, dim3(kPtThreads), 0, stream, S)
    if (mode == kTraceSingle) {
        if (leafBatch == 8) MFH_ITERATE(kTraceSingle, 8); else MFH_ITERATE(kTraceSingle, 1);
    } else if (mode == kTraceMulti) {
        if (leafBatch == 8) MFH_ITERATE(kTraceMulti, 8); else MFH_ITERATE(kTraceMulti, 1);
    } else {
        if (leafBatch == 8) MFH_ITERATE(kTraceMixed, 8); else MFH_ITERATE(kTraceMixed, 1);
    }
#undef MFH_ITERATE
    MFH_HIP(hipGetLastError());
}

}  // namespace mfh
