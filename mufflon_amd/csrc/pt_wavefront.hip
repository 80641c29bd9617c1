// Wavefront form of the path tracer for gfx950: one iteration of pt_sample
// (pt_common.hpp:75-156) split into stages over ray queues.
//
//   begin     zero the queue / work counters, advance the iteration counter
//   primary   camera ray + closest hit for every pixel (queue 0 = all paths)
//   per bounce k:
//     shade   hit -> vertex, Hit ID / emission / auxiliary targets, next-event
//             candidates into the shadow queue, BSDF walk into queue k+1
//     shadow  any-hit of the candidates; unoccluded ones are added
//     extend  closest hit of queue k+1
//   finish    running means of the enabled targets
//
// The trace, shade and shadow kernels are persistent: a grid sized to the
// chip's occupancy drains its queue in chunks grabbed by one atomic per wave,
// so the launch shape does not depend on the queue length (which lives in
// device memory) and a whole iteration can be captured in a HIP graph.
// Queue appends are compacted per wave64 with ballot + mbcnt and one atomic.
// Every stage calls the same path:: functions as the megakernel, and each
// path's contributions reach its accumulator in the megakernel's order
// (emission of vertex k, then its next-event estimates, then vertex k+1), so
// both drivers produce the same bits.
#include <array>
#include <cstring>

#include "pt_path.hpp"
#include "pt_wavefront.hpp"

namespace mfh {

namespace {

using namespace shade;
using namespace path;

__device__ __forceinline__ int lane_id() { return (int)__lane_id(); }

// Wave-wide append of the lanes with `take` set to a queue with length at
// *count: returns this lane's slot (valid where take is set).
__device__ __forceinline__ int wave_append(int* count, bool take) {
    const unsigned long long m = __ballot(take);
    if (m == 0ull) return -1;
    const int leader = __ffsll((long long)m) - 1;
    int base = 0;
    if (lane_id() == leader) base = atomicAdd(count, (int)__popcll(m));
    base = __shfl(base, leader);
    const unsigned lo = (unsigned)m, hi = (unsigned)(m >> 32);
    const int rank = (int)__builtin_amdgcn_mbcnt_hi(hi, __builtin_amdgcn_mbcnt_lo(lo, 0u));
    return base + rank;
}

// Grab the next chunk of a queue for this wave; every lane gets the same base.
__device__ __forceinline__ int wave_grab(int* work, int chunk) {
    int base = 0;
    if (lane_id() == 0) base = atomicAdd(work, chunk);
    return __shfl(base, 0);
}

// Queue walk of a persistent grid: a wave's first chunk is its wave index
// (no atomic, so waves beyond a short queue exit at once); later chunks come
// from the work counter, offset past the statically assigned ones.
struct QueueWalk {
    int chunk, waves, base;
    __device__ __forceinline__ QueueWalk(int n, int maxChunk);
    __device__ __forceinline__ void next(int* work) { base = waves * chunk + wave_grab(work, chunk); }
};

// Chunk for a queue of n entries: about kChunksPerWave grabs per wave of the
// grid (a multiple of 64, at least 64, at most the tuned maximum), so short
// queues still spread over every wave and long ones need few atomics.
constexpr int kChunksPerWave = 4;
__device__ __forceinline__ int queue_chunk(int n, int maxChunk) {
    const int waves = (int)(gridDim.x * blockDim.x) / 64;
    int c = (n + waves * kChunksPerWave - 1) / (waves * kChunksPerWave);
    c = (c + 63) & ~63;
    return c < 64 ? 64 : (c > maxChunk ? maxChunk : c);
}

__device__ __forceinline__ QueueWalk::QueueWalk(int n, int maxChunk) {
    chunk = queue_chunk(n, maxChunk);
    waves = (int)(gridDim.x * blockDim.x) / 64;
    base = (int)((blockIdx.x * blockDim.x + threadIdx.x) / 64) * chunk;
}

// Per-path accumulator in HBM: slot-major (slot * N + path).
__device__ __forceinline__ void load_acc(const WfState& W, u32 mask, int p, Acc& acc) {
    acc.init(mask);
    int slot = 0;
#pragma unroll
    for (int t = 0; t < kNumTargets; ++t) {
        const int nc = kTargetChannels[t];
        if ((mask >> t) & 1u) {
#pragma unroll
            for (int k = 0; k < nc; ++k) acc.v[slot + k] = W.acc[(size_t)(slot + k) * W.numPaths + p];
        }
        slot += nc;
    }
}
__device__ __forceinline__ void store_acc(const WfState& W, u32 mask, int p, const Acc& acc) {
    int slot = 0;
#pragma unroll
    for (int t = 0; t < kNumTargets; ++t) {
        const int nc = kTargetChannels[t];
        if ((mask >> t) & 1u) {
#pragma unroll
            for (int k = 0; k < nc; ++k) W.acc[(size_t)(slot + k) * W.numPaths + p] = acc.v[slot + k];
        }
        slot += nc;
    }
}

// Scene of a path: the kernel argument for one scene, or the batch entry
// whose path range holds p (binary search over the scene offsets).
__device__ __forceinline__ int scene_index(const WfState& W, int p, int& local) {
    int lo = 0, hi = W.numScenes - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (W.sceneOffset[mid] <= p)
            lo = mid;
        else
            hi = mid - 1;
    }
    local = p - W.sceneOffset[lo];
    return lo;
}
template <bool kBatch>
__device__ __forceinline__ const PtScene& scene_ref(const PtScene& S, const WfState& W, int s) {
    if constexpr (kBatch)
        return W.scenes[s];
    else
        return S;
}
template <bool kBatch>
__device__ __forceinline__ int locate(const WfState& W, int p, int& local) {
    if constexpr (kBatch) return scene_index(W, p, local);
    local = p;
    return 0;
}

__global__ void k_wf_begin(WfState W) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < W.numCounters) W.counters[i] = 0;
    if (i == 0) {
        W.counters[W.qBase] = W.numPaths;  // queue 0 holds every path
        ++*W.iteration;
    }
}

// Camera rays and their closest hits, one thread per pixel: 16x16 tiles for
// one scene, one path per thread in path order for a batch (kMode 2).
// kLeaf: triangles of a leaf loaded ahead of their tests (trace_ray's kBatch).
template <int kMode, int kLeaf>
__global__ __launch_bounds__(256) void k_wf_primary(PtScene S0, WfState W) {
    __shared__ int sStack[kLdsDepth * 256];
    constexpr bool kBatch = kMode == 2;
    int px, py, p, local, s = 0;
    if constexpr (kBatch) {
        p = blockIdx.x * 256 + threadIdx.x;
        if (p >= W.numPaths) return;
        s = scene_index(W, p, local);
        const int w = W.scenes[s].width;
        px = local % w;
        py = local / w;
    } else {
        px = blockIdx.x * 16 + (threadIdx.x % 16);
        py = blockIdx.y * 16 + (threadIdx.x / 16);
        if (px >= S0.width || py >= S0.height) return;
        p = local = px + py * S0.width;
    }
    const PtScene& S = scene_ref<kBatch>(S0, W, s);
    u64 rng = S.rngs[local];
    pcg_next(rng);  // create_camera's RndSet2
    const u64 x0 = pcg_next(rng);
    pcg_next(rng);  // the i0 word of the camera's RndSet2_1
    pcg_next(rng);  // roulette (always continues at throughput 1)
    V3 dir;
    float pdfF;
    camera_ray(S, px, py, x0, dir, pdfF);
    const V3 o = v3(S.camPos[0], S.camPos[1], S.camPos[2]);
    const TraceHit h =
        trace_mode<kMode, false, 256, kLdsDepth, kLeaf>(S.trace, o, dir, kMaxSceneSize, sStack + threadIdx.x);
    S.rngs[local] = rng;
    W.rayO[0][p] = float4{o.x, o.y, o.z, pdfF};
    W.rayD[0][p] = float4{dir.x, dir.y, dir.z, __int_as_float(p)};
    W.rayM[0][p] = float4{0.0f, 0.0f, 0.0f, h.w1};
    if (W.rayP[0]) W.rayP[0][p] = float4{o.x, o.y, o.z, 0.0f};
    W.hit[0][p] = float4{h.t, __int_as_float(h.inst), __int_as_float(h.prim), h.w0};
    W.tp[p] = float4{1.0f, 1.0f, 1.0f, 0.0f};
#pragma unroll
    for (int k = 0; k < 18; ++k)
        if ((W.slotMask >> k) & 1u) W.acc[(size_t)k * W.numPaths + p] = 0.0f;
}

// Closest hit of queue `buf` (length at counters[count]).
template <int kMode, int kBlock, int kLds, int kLeaf>
__global__ __launch_bounds__(kBlock) void k_wf_extend(PtScene S0, WfState W, int buf, int count, int work,
                                                      const int* perm) {
    __shared__ int sStack[kLds * kBlock];
    constexpr bool kBatch = kMode == 2;
    const int n = W.counters[count];
    int* const lstack = sStack + threadIdx.x;
    for (QueueWalk wk(n, W.chunk); wk.base < n; wk.next(W.counters + work)) {
        const int end = min(wk.base + wk.chunk, n);
        for (int qs = wk.base + lane_id(); qs < end; qs += 64) {
            const int q = perm ? perm[qs] : qs;
            const float4 o = W.rayO[buf][q];
            const float4 d = W.rayD[buf][q];
            int local;
            const int s = locate<kBatch>(W, __float_as_int(d.w), local);
            const PtScene& S = scene_ref<kBatch>(S0, W, s);
            const TraceHit h = trace_mode<kMode, false, kBlock, kLds, kLeaf>(S.trace, v3(o.x, o.y, o.z),
                                                                             v3(d.x, d.y, d.z), kMaxSceneSize, lstack);
            W.hit[buf][q] = float4{h.t, __int_as_float(h.inst), __int_as_float(h.prim), h.w0};
            W.rayM[buf][q].w = h.w1;
        }
    }
}

// Any hit of the shadow candidates of one bounce; visible ones are added to
// their path's accumulator in candidate order.
template <int kMode, int kBlock, int kLds, int kLeaf>
__global__ __launch_bounds__(kBlock) void k_wf_shadow(PtScene S0, WfState W, int count, int work, const int* perm) {
    __shared__ int sStack[kLds * kBlock];
    constexpr bool kBatch = kMode == 2;
    const int n = W.counters[count];
    int* const lstack = sStack + threadIdx.x;
    for (QueueWalk wk(n, W.chunk); wk.base < n; wk.next(W.counters + work)) {
        const int end = min(wk.base + wk.chunk, n);
        for (int ss = wk.base + lane_id(); ss < end; ss += 64) {
            const int sl = perm ? perm[ss] : ss;
            const int2 pm = W.shPath[sl];
            const u32 cand = (u32)pm.y;
            int local;
            const int si = locate<kBatch>(W, pm.x, local);
            const PtScene& S = scene_ref<kBatch>(S0, W, si);
            u32 visible = 0;
            for (int i = 0; i < S.neeCount; ++i) {
                if (!((cand >> i) & 1u)) continue;
                const size_t e = (size_t)sl * W.neeStride + i;
                const float4 a = W.shO[e];
                const float4 d = W.shD[e];
                const TraceHit h = trace_mode<kMode, true, kBlock, kLds, kLeaf>(S.trace, v3(a.x, a.y, a.z),
                                                                                v3(d.x, d.y, d.z), a.w, lstack);
                if (h.inst == -1) visible |= 1u << i;
            }
            if (visible) {  // add the unoccluded estimates in candidate order
                Acc acc;
                load_acc(W, S.targetMask, pm.x, acc);
                for (int i = 0; i < S.neeCount; ++i) {
                    if (!((visible >> i) & 1u)) continue;
                    const float4 c = W.shC[(size_t)sl * W.neeStride + i];
                    NeeCandidate nc;
                    nc.radiance = v3(c.x, c.y, c.z);
                    nc.lightness = c.w;
                    add_nee(nc, acc);
                }
                store_acc(W, S.targetMask, pm.x, acc);
                if (acc.nans) atomicAdd(S.nanCounter, acc.nans);
            }
        }
    }
}

// ---- trace stages with lane refill
// The persistent trace kernels above give each lane whole rays from the
// wave's chunk; a lane whose ray ends early idles until the wave's longest
// traversal in the chunk is done. The refill forms below traverse in rounds
// of `budget` steps (Traversal::run) and, between rounds, hand new rays to
// idle lanes as soon as a quarter of the wave is idle: one atomic on the
// queue's work counter per refill, lanes ranked with mbcnt. The queue is
// consumed in order of the grabs, and a ray's result does not depend on the
// rounds it took.
constexpr int kRefillMinIdle = 16;

__device__ __forceinline__ int lane_rank(unsigned long long mask) {
    return (int)__builtin_amdgcn_mbcnt_hi((unsigned)(mask >> 32), __builtin_amdgcn_mbcnt_lo((unsigned)mask, 0u));
}

// Refill of the idle lanes (mask `idle`, wave-uniform): first queue index of
// this lane or -1; drained is set once the queue has been handed out.
__device__ __forceinline__ int refill_slot(int* work, int n, unsigned long long idle, bool& drained) {
    const int want = __popcll(idle);
    const int leader = __ffsll((long long)idle) - 1;
    int base = 0;
    if (lane_id() == leader) base = atomicAdd(work, want);
    base = __shfl(base, leader);
    if (base + want >= n) drained = true;
    if (!((idle >> lane_id()) & 1ull)) return -1;
    const int slot = base + lane_rank(idle);
    return slot < n ? slot : -1;
}

template <int kMode, int kBlock, int kLds>
__global__ __launch_bounds__(kBlock) void k_wf_extend_refill(PtScene S0, WfState W, int buf, int count, int work,
                                                             const int* perm, int budget) {
    __shared__ int sStack[kLds * kBlock];
    constexpr bool kBatch = kMode == 2;
    using T = Traversal<false, false, traversal_mode<kMode>(), kBlock, kLds, kMode == kTraceMixed>;
    const int n = W.counters[count];
    int* const lstack = sStack + threadIdx.x;
    int spill[T::kSpill];
    T tr;
    int q = -1, si = 0;
    bool drained = false;
    while (true) {
        const unsigned long long idle = __ballot(q < 0);
        if (!drained && __popcll(idle) >= kRefillMinIdle) {
            const int slot = refill_slot(W.counters + work, n, idle, drained);
            if (slot >= 0) {
                q = perm ? perm[slot] : slot;
                const float4 o = W.rayO[buf][q];
                const float4 d = W.rayD[buf][q];
                int local;
                si = locate<kBatch>(W, __float_as_int(d.w), local);
                tr.init(scene_ref<kBatch>(S0, W, si).trace, v3(o.x, o.y, o.z), v3(d.x, d.y, d.z), kMaxSceneSize);
            }
        }
        if (__ballot(q >= 0) == 0) break;  // nothing in flight: the queue is drained
        if (q >= 0 && tr.template run<true>(scene_ref<kBatch>(S0, W, si).trace, lstack, spill, budget)) {
            const TraceHit h = tr.result();
            W.hit[buf][q] = float4{h.t, __int_as_float(h.inst), __int_as_float(h.prim), h.w0};
            W.rayM[buf][q].w = h.w1;
            q = -1;
        }
    }
}

template <int kMode, int kBlock, int kLds>
__global__ __launch_bounds__(kBlock) void k_wf_shadow_refill(PtScene S0, WfState W, int count, int work,
                                                             const int* perm, int budget) {
    __shared__ int sStack[kLds * kBlock];
    constexpr bool kBatch = kMode == 2;
    using T = Traversal<true, false, traversal_mode<kMode>(), kBlock, kLds, kMode == kTraceMixed>;
    const int n = W.counters[count];
    int* const lstack = sStack + threadIdx.x;
    int spill[T::kSpill];
    T tr;
    int sl = -1, si = 0, path = 0, cur = 0;
    u32 remaining = 0, visible = 0;
    bool drained = false;
    // next candidate of the slot, or its result once every candidate is traced
    auto advance = [&]() {
        while (sl >= 0) {
            if (remaining) {
                cur = __ffs(remaining) - 1;
                remaining &= remaining - 1;
                const size_t e = (size_t)sl * W.neeStride + cur;
                const float4 a = W.shO[e];
                const float4 d = W.shD[e];
                tr.init(scene_ref<kBatch>(S0, W, si).trace, v3(a.x, a.y, a.z), v3(d.x, d.y, d.z), a.w);
                return;
            }
            if (visible) {  // add the unoccluded estimates in candidate order
                const PtScene& S = scene_ref<kBatch>(S0, W, si);
                Acc acc;
                load_acc(W, S.targetMask, path, acc);
                for (int i = 0; i < S.neeCount; ++i) {
                    if (!((visible >> i) & 1u)) continue;
                    const float4 c = W.shC[(size_t)sl * W.neeStride + i];
                    NeeCandidate nc;
                    nc.radiance = v3(c.x, c.y, c.z);
                    nc.lightness = c.w;
                    add_nee(nc, acc);
                }
                store_acc(W, S.targetMask, path, acc);
                if (acc.nans) atomicAdd(S.nanCounter, acc.nans);
            }
            sl = -1;
        }
    };
    while (true) {
        const unsigned long long idle = __ballot(sl < 0);
        if (!drained && __popcll(idle) >= kRefillMinIdle) {
            const int slot = refill_slot(W.counters + work, n, idle, drained);
            if (slot >= 0) {
                sl = perm ? perm[slot] : slot;
                const int2 pm = W.shPath[sl];
                path = pm.x;
                remaining = (u32)pm.y;
                visible = 0;
                int local;
                si = locate<kBatch>(W, path, local);
                advance();
            }
        }
        if (__ballot(sl >= 0) == 0) break;
        if (sl >= 0 && tr.template run<true>(scene_ref<kBatch>(S0, W, si).trace, lstack, spill, budget)) {
            if (tr.result().inst == -1) visible |= 1u << cur;
            advance();
        }
    }
}

// Shade the hits of queue `buf` at path length k + 1.
template <bool kBatch>
__global__ __launch_bounds__(256) void k_wf_shade(PtScene S0, WfState W, int k, int buf, int count, int nextCount,
                                                  int shCount, int work) {
    const int n = W.counters[count];
    const int pathLen = k + 1;
    const int nbuf = buf ^ 1;
    for (QueueWalk wk(n, W.chunk); wk.base < n; wk.next(W.counters + work)) {
        const int end = wk.base + wk.chunk;
        for (int q0 = wk.base; q0 < end; q0 += 64) {
            const int q = q0 + lane_id();
            const bool active = q < n;
            bool doNee = false, cont = false;
            int p = 0, local = 0, si = 0;
            u64 rng = 0;
            V3 throughput = v3(0, 0, 0), rayOrigin = v3(0, 0, 0), dir = v3(0, 0, 0), absorption = v3(0, 0, 0);
            float footA = 0.0f, guideWeight = 0.0f, pdfF = 0.0f;
            Vertex vx;
            vx.isVoid = true;
            Acc acc;
            if (active) {
                const float4 o = W.rayO[buf][q];
                const float4 d = W.rayD[buf][q];
                const float4 m = W.rayM[buf][q];
                const float4 h = W.hit[buf][q];
                p = __float_as_int(d.w);
                si = locate<kBatch>(W, p, local);
                const PtScene& S = scene_ref<kBatch>(S0, W, si);
                rng = S.rngs[local];
                const float4 t4 = W.tp[p];
                throughput = v3(t4.x, t4.y, t4.z);
                footA = t4.w;
                const int inst = __float_as_int(h.y);
                throughput = throughput * transmission(Medium{1.0f, v3(m.x, m.y, m.z)}, inst == -1 ? kMaxSceneSize : h.x);
                footA += 1.0f / fsqrt(o.w);
                guideWeight = 1.f / (1.f + footA * footA / 1.0f);
                make_vertex(S, v3(o.x, o.y, o.z), v3(d.x, d.y, d.z), o.w, h.x, inst, __float_as_int(h.z), h.w, m.w, vx);
                load_acc(W, S.targetMask, p, acc);
                V3 prevPos = v3(0.0f, 0.0f, 0.0f);
                if (W.rayP[buf]) {
                    const float4 pp = W.rayP[buf][q];
                    prevPos = v3(pp.x, pp.y, pp.z);
                }
                vertex_targets(S, vx, pathLen, throughput, guideWeight, acc, prevPos);
                const bool more = !vx.isVoid && pathLen < S.maxPathLength;
                doNee = more && pathLen + 1 >= S.minPathLength && pathLen + 1 <= S.maxPathLength;
                cont = more;
            }
            // next-event candidates: one shadow slot per path that estimates
            const int sh = wave_append(W.counters + shCount, doNee);
            if (doNee) {
                const PtScene& S = scene_ref<kBatch>(S0, W, si);
                const u64 neeSeed = pcg_next(rng);
                u32 cand = 0;
                for (int i = 0; i < S.neeCount; ++i) {
                    const u64 x = pcg_next(rng);
                    NeeCandidate c;
                    if (nee_candidate(S, vx, throughput, guideWeight, i, neeSeed, x, c)) {
                        const size_t e = (size_t)sh * W.neeStride + i;
                        W.shO[e] = float4{c.a.x, c.a.y, c.a.z, c.tmax};
                        W.shD[e] = float4{c.dir.x, c.dir.y, c.dir.z, 0.0f};
                        W.shC[e] = float4{c.radiance.x, c.radiance.y, c.radiance.z, c.lightness};
                        cand |= 1u << i;
                    }
                }
                W.shPath[sh] = int2{p, (int)cand};
            }
            if (cont) {
                const u64 x0 = pcg_next(rng);
                const u64 x1 = pcg_next(rng);
                const float rndRoulette = sample_uniform((u32)pcg_next(rng));
                cont = walk_surface(vx, x0, x1, rndRoulette, throughput, rayOrigin, dir, pdfF, absorption);
            }
            const int nq = wave_append(W.counters + nextCount, cont);
            if (cont) {
                W.rayO[nbuf][nq] = float4{rayOrigin.x, rayOrigin.y, rayOrigin.z, pdfF};
                W.rayD[nbuf][nq] = float4{dir.x, dir.y, dir.z, __int_as_float(p)};
                W.rayM[nbuf][nq] = float4{absorption.x, absorption.y, absorption.z, 0.0f};
                if (W.rayP[nbuf]) W.rayP[nbuf][nq] = float4{vx.position.x, vx.position.y, vx.position.z, 0.0f};
                W.tp[p] = float4{throughput.x, throughput.y, throughput.z, footA};
            }
            if (active) {
                const PtScene& S = scene_ref<kBatch>(S0, W, si);
                S.rngs[local] = rng;
                store_acc(W, S.targetMask, p, acc);
                if (acc.nans) atomicAdd(S.nanCounter, acc.nans);
            }
        }
    }
}

// The rest of every path of queue `buf` (rays of bounce k traced to their
// hits) in one launch: a lane takes a path and runs the megakernel's loop
// from the vertex of that hit to the end of the path, tracing its next-event
// rays and extensions itself. Few but long paths (glass, deep bounces) then
// cost one launch instead of three per bounce, each waiting for its slowest
// ray. Random numbers are drawn, and contributions reach the accumulator, in
// the order of the staged kernels and of the megakernel (emission of a vertex,
// its next-event estimates, the next vertex), so the bits are the same.
template <int kMode, int kLeaf>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(3))) void k_wf_tail(PtScene S0, WfState W, int k,
                                                                                       int buf, int count, int work) {
    __shared__ int sStack[kLdsDepth * 256];
    constexpr bool kBatch = kMode == 2;
    const int n = W.counters[count];
    int* const lstack = sStack + threadIdx.x;
    for (QueueWalk wk(n, W.chunk); wk.base < n; wk.next(W.counters + work)) {
        const int end = min(wk.base + wk.chunk, n);
        for (int q = wk.base + lane_id(); q < end; q += 64) {
            const float4 o = W.rayO[buf][q];
            const float4 d = W.rayD[buf][q];
            const float4 m = W.rayM[buf][q];
            const float4 h = W.hit[buf][q];
            const int p = __float_as_int(d.w);
            int local;
            const int si = locate<kBatch>(W, p, local);
            const PtScene& S = scene_ref<kBatch>(S0, W, si);
            u64 rng = S.rngs[local];
            const float4 t4 = W.tp[p];
            V3 throughput = v3(t4.x, t4.y, t4.z);
            float footA = t4.w;
            Acc acc;
            load_acc(W, S.targetMask, p, acc);
            V3 prevPos = v3(0.0f, 0.0f, 0.0f);
            if (W.rayP[buf]) {
                const float4 pp = W.rayP[buf][q];
                prevPos = v3(pp.x, pp.y, pp.z);
            }
            V3 rayOrigin = v3(o.x, o.y, o.z), dir = v3(d.x, d.y, d.z), absorption = v3(m.x, m.y, m.z);
            float pdfF = o.w;
            TraceHit hit;
            hit.t = h.x;
            hit.inst = __float_as_int(h.y);
            hit.prim = __float_as_int(h.z);
            hit.w0 = h.w;
            hit.w1 = m.w;
            int pathLen = k;
            Vertex vx;
            while (true) {
                throughput = throughput * transmission(Medium{1.0f, absorption}, hit.inst == -1 ? kMaxSceneSize : hit.t);
                footA += 1.0f / fsqrt(pdfF);
                const float guideWeight = 1.f / (1.f + footA * footA / 1.0f);
                make_vertex(S, rayOrigin, dir, pdfF, hit.t, hit.inst, hit.prim, hit.w0, hit.w1, vx);
                ++pathLen;
                vertex_targets(S, vx, pathLen, throughput, guideWeight, acc, prevPos);
                if (vx.isVoid || pathLen >= S.maxPathLength) break;
                if (pathLen + 1 >= S.minPathLength && pathLen + 1 <= S.maxPathLength) {
                    const u64 neeSeed = pcg_next(rng);
                    for (int i = 0; i < S.neeCount; ++i) {
                        const u64 x = pcg_next(rng);
                        NeeCandidate c;
                        if (nee_candidate(S, vx, throughput, guideWeight, i, neeSeed, x, c)) {
                            const TraceHit sh =
                                trace_mode<kMode, true, 256, kLdsDepth, kLeaf>(S.trace, c.a, c.dir, c.tmax, lstack);
                            if (sh.inst == -1) add_nee(c, acc);
                        }
                    }
                }
                const u64 x0 = pcg_next(rng);
                const u64 x1 = pcg_next(rng);
                const float rndRoulette = sample_uniform((u32)pcg_next(rng));
                if (!walk_surface(vx, x0, x1, rndRoulette, throughput, rayOrigin, dir, pdfF, absorption)) break;
                prevPos = vx.position;
                hit = trace_mode<kMode, false, 256, kLdsDepth, kLeaf>(S.trace, rayOrigin, dir, kMaxSceneSize, lstack);
            }
            S.rngs[local] = rng;
            store_acc(W, S.targetMask, p, acc);
            if (acc.nans) atomicAdd(S.nanCounter, acc.nans);
        }
    }
}

// 9-bit-per-axis Morton code of the origin in the scene box, then the
// direction octant: rays from nearby origins heading the same way become
// neighbours. Entries past the queue length sort last.
__device__ __forceinline__ u32 spread9(u32 x) {
    x &= 0x1FFu;
    x = (x | (x << 16)) & 0x030000FFu;
    x = (x | (x << 8)) & 0x0300F00Fu;
    x = (x | (x << 4)) & 0x030C30C3u;
    x = (x | (x << 2)) & 0x09249249u;
    return x;
}
__device__ __forceinline__ u64 ray_key(const PtScene& S, float4 o, float4 d) {
    u32 q[3];
    const float* b = S.sceneBox;
    const float oc[3] = {o.x, o.y, o.z};
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        const float span = b[3 + a] - b[a];
        const float f = span > 0.0f ? (oc[a] - b[a]) / span : 0.0f;
        const float c = f * 512.0f;
        q[a] = c <= 0.0f ? 0u : (c >= 511.0f ? 511u : (u32)c);  // NaN -> 0
    }
    const u32 m = spread9(q[0]) | (spread9(q[1]) << 1) | (spread9(q[2]) << 2);
    const u32 oct = (d.x < 0.0f ? 1u : 0u) | (d.y < 0.0f ? 2u : 0u) | (d.z < 0.0f ? 4u : 0u);
    return (u64)((m << 3) | oct);
}
// shadow = false: extension queue `buf`; true: the shadow slots (first candidate's segment)
__global__ void k_wf_sort_keys(PtScene S, WfState W, int buf, int count, bool shadow) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= W.numPaths) return;
    const int n = W.counters[count];
    // every entry of the queue sorts before every entry past it, so the first n
    // sorted ids are a permutation of [0, n)
    u64 key = 0xFFFFFFFFull;
    if (i < n) {
        key = 0xFFFFFFFEull;
        if (!shadow) {
            key = ray_key(S, W.rayO[buf][i], W.rayD[buf][i]);
        } else {
            const u32 cand = (u32)W.shPath[i].y;
            if (cand) {
                const int first = __ffs((int)cand) - 1;
                const size_t e = (size_t)i * W.neeStride + first;
                key = ray_key(S, W.shO[e], W.shD[e]);
            }
        }
    }
    W.sortKeys[0][i] = key;
    W.sortIds[0][i] = i;
}

template <bool kBatch>
__global__ void k_wf_finish(PtScene S0, WfState W) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= W.numPaths) return;
    int local;
    const int s = locate<kBatch>(W, p, local);
    const PtScene& S = scene_ref<kBatch>(S0, W, s);
    float v[18];
#pragma unroll
    for (int k = 0; k < 18; ++k) v[k] = ((W.slotMask >> k) & 1u) ? W.acc[(size_t)k * W.numPaths + p] : 0.0f;
    accumulate(S, (size_t)local, *W.iteration, v, S.targetMask);
}

template <int kMode, int kBlock, int kLds, int kLeaf>
struct TraceKernels {
    static void extend(const PtScene& S, const WfState& W, int grid, int buf, int count, int work, const int* perm,
                       int refill, hipStream_t st) {
        if (refill > 0)
            hipLaunchKernelGGL(HIP_KERNEL_NAME(k_wf_extend_refill<kMode, kBlock, kLds>), dim3(grid), dim3(kBlock), 0,
                               st, S, W, buf, count, work, perm, refill);
        else
            hipLaunchKernelGGL(HIP_KERNEL_NAME(k_wf_extend<kMode, kBlock, kLds, kLeaf>), dim3(grid), dim3(kBlock), 0,
                               st, S, W, buf, count, work, perm);
    }
    static void shadow(const PtScene& S, const WfState& W, int grid, int count, int work, const int* perm, int refill,
                       hipStream_t st) {
        if (refill > 0)
            hipLaunchKernelGGL(HIP_KERNEL_NAME(k_wf_shadow_refill<kMode, kBlock, kLds>), dim3(grid), dim3(kBlock), 0,
                               st, S, W, count, work, perm, refill);
        else
            hipLaunchKernelGGL(HIP_KERNEL_NAME(k_wf_shadow<kMode, kBlock, kLds, kLeaf>), dim3(grid), dim3(kBlock), 0,
                               st, S, W, count, work, perm);
    }
    static int occupancy(bool refill) {
        int a = 0, b = 0;
        if (refill) {
            MFH_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&a, k_wf_extend_refill<kMode, kBlock, kLds>, kBlock,
                                                                 0));
            MFH_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&b, k_wf_shadow_refill<kMode, kBlock, kLds>, kBlock,
                                                                 0));
        } else {
            MFH_HIP(
                hipOccupancyMaxActiveBlocksPerMultiprocessor(&a, k_wf_extend<kMode, kBlock, kLds, kLeaf>, kBlock, 0));
            MFH_HIP(
                hipOccupancyMaxActiveBlocksPerMultiprocessor(&b, k_wf_shadow<kMode, kBlock, kLds, kLeaf>, kBlock, 0));
        }
        return std::max(1, std::min(a, b));
    }
};

using ExtendFn = void (*)(const PtScene&, const WfState&, int, int, int, int, const int*, int, hipStream_t);
using ShadowFn = void (*)(const PtScene&, const WfState&, int, int, int, const int*, int, hipStream_t);
using OccFn = int (*)(bool);

struct TraceVariant {
    ExtendFn extend;
    ShadowFn shadow;
    OccFn occupancy;
};

template <int kMode, int kBlock, int kLds, int kLeaf>
constexpr TraceVariant variant() {
    return {&TraceKernels<kMode, kBlock, kLds, kLeaf>::extend, &TraceKernels<kMode, kBlock, kLds, kLeaf>::shadow,
            &TraceKernels<kMode, kBlock, kLds, kLeaf>::occupancy};
}

// Tuning space of the trace stages: block size x LDS stack depth x leaf
// batch (1 or 8 triangles loaded ahead; the lane-refill kernels test one at a
// time), for multi-instance scenes (0), single-instance scenes (1), batches
// (2) and scenes with quads or spheres (3).
template <int kMode, int kLeaf>
constexpr std::array<TraceVariant, 7> shapes() {
    return {variant<kMode, 64, 8, kLeaf>(),   variant<kMode, 64, 16, kLeaf>(),  variant<kMode, 128, 8, kLeaf>(),
            variant<kMode, 128, 16, kLeaf>(), variant<kMode, 256, 8, kLeaf>(),  variant<kMode, 256, 16, kLeaf>(),
            variant<kMode, 256, 24, kLeaf>()};
}
template <int kMode>
std::array<TraceVariant, kWfNumTraceVariants> variants() {
    const std::array<TraceVariant, 7> a = shapes<kMode, 1>(), b = shapes<kMode, 8>();
    std::array<TraceVariant, kWfNumTraceVariants> v{};
    for (int i = 0; i < 7; ++i) {
        v[i] = a[i];
        v[7 + i] = b[i];
    }
    return v;
}
const std::array<TraceVariant, kWfNumTraceVariants> kVariants[4] = {variants<0>(), variants<1>(), variants<2>(),
                                                                     variants<3>()};

}  // namespace

const WfTraceShape kWfTraceShapes[kWfNumTraceVariants] = {
    {64, 8, 1}, {64, 16, 1}, {128, 8, 1}, {128, 16, 1}, {256, 8, 1}, {256, 16, 1}, {256, 24, 1},
    {64, 8, 8}, {64, 16, 8}, {128, 8, 8}, {128, 16, 8}, {256, 8, 8}, {256, 16, 8}, {256, 24, 8}};

int wf_grid(const PtScene& S, const WfTuning& tune, int numCUs) {
    const TraceVariant& v = kVariants[trace_mode_of(S.trace)][tune.traceVariant];
    return v.occupancy(tune.refill > 0) * numCUs;
}

int wf_grid_batch(const WfTuning& tune, int numCUs) {
    return kVariants[2][tune.traceVariant].occupancy(tune.refill > 0) * numCUs;
}

int wf_shade_grid(int numCUs) {
    int a = 0, b = 0;
    MFH_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&a, k_wf_shade<false>, 256, 0));
    MFH_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&b, k_wf_shade<true>, 256, 0));
    return std::max(1, std::min(a, b)) * numCUs;
}

int wf_tail_grid(int numCUs) {
    int a = 0, b = 0, c = 0, d = 0;
    MFH_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&a, k_wf_tail<0, 1>, 256, 0));
    MFH_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&b, k_wf_tail<1, 8>, 256, 0));
    MFH_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&c, k_wf_tail<2, 1>, 256, 0));
    MFH_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&d, k_wf_tail<3, 8>, 256, 0));
    return std::max(1, std::min(std::min(a, d), std::min(b, c))) * numCUs;
}

namespace {
void launch_tail(int mode, bool leaf8, const PtScene& S, const WfState& W, int grid, int k, int buf, int count,
                 int work, hipStream_t st) {
#define MFH_TAIL(M, L) \
    hipLaunchKernelGGL(HIP_KERNEL_NAME(k_wf_tail<M, L>), dim3(grid), dim3(256), 0, st, S, W, k, buf, count, work)
    if (mode == 0) {
        if (leaf8) MFH_TAIL(0, 8); else MFH_TAIL(0, 1);
    } else if (mode == 1) {
        if (leaf8) MFH_TAIL(1, 8); else MFH_TAIL(1, 1);
    } else if (mode == 2) {
        if (leaf8) MFH_TAIL(2, 8); else MFH_TAIL(2, 1);
    } else {
        if (leaf8) MFH_TAIL(3, 8); else MFH_TAIL(3, 1);
    }
#undef MFH_TAIL
    MFH_HIP(hipGetLastError());
}
}  // namespace

// Sorted order of a queue: keys, then a 4-pass LSD radix sort over every path
// slot (the queue length stays on the device); the permutation ends in sortIds[0].
void sort_queue(const PtScene& S, const WfState& W, int buf, int count, bool shadow, BuildScratch& scratch,
                hipStream_t stream) {
    hipLaunchKernelGGL(k_wf_sort_keys, dim3((W.numPaths + 255) / 256), dim3(256), 0, stream, S, W, buf, count, shadow);
    radix_sort_pairs(W.sortKeys[0], W.sortKeys[1], W.sortIds[0], W.sortIds[1], W.numPaths, 0, 32, scratch, stream);
}

void launch_pt_wavefront(const PtScene& S, const WfState& W, const WfTuning& tune, BuildScratch& scratch,
                         hipStream_t stream) {
    const int mode = trace_mode_of(S.trace);
    const TraceVariant& v = kVariants[mode][tune.traceVariant];
    hipLaunchKernelGGL(k_wf_begin, dim3((W.numCounters + 255) / 256), dim3(256), 0, stream, W);
    const dim3 pgrid((S.width + 15) / 16, (S.height + 15) / 16);
    const bool leaf8 = kWfTraceShapes[tune.traceVariant].leafBatch == 8;
#define MFH_PRIMARY(M, L) hipLaunchKernelGGL(HIP_KERNEL_NAME(k_wf_primary<M, L>), pgrid, dim3(256), 0, stream, S, W)
    if (mode == kTraceSingle) {
        if (leaf8) MFH_PRIMARY(1, 8); else MFH_PRIMARY(1, 1);
    } else if (mode == kTraceMulti) {
        if (leaf8) MFH_PRIMARY(0, 8); else MFH_PRIMARY(0, 1);
    } else {
        if (leaf8) MFH_PRIMARY(3, 8); else MFH_PRIMARY(3, 1);
    }
#undef MFH_PRIMARY
    // pt_sample's loop is a do-while: the first hit is shaded even at maximum path length 0
    const int L = S.maxPathLength > 1 ? S.maxPathLength : 1;
    int work = W.workBase;
    for (int k = 0; k < L; ++k) {
        const int buf = k & 1;
        hipLaunchKernelGGL(HIP_KERNEL_NAME(k_wf_shade<false>), dim3(tune.shadeGrid), dim3(256), 0, stream, S, W, k,
                           buf, W.qBase + k, W.qBase + k + 1, W.shBase + k, work++);
        const bool sortNow = k < tune.sortBounces;
        if (sortNow) sort_queue(S, W, 0, W.shBase + k, true, scratch, stream);
        v.shadow(S, W, tune.traceGrid, W.shBase + k, work++, sortNow ? W.sortIds[0] : nullptr, tune.refill, stream);
        if (k + 1 < L) {
            if (sortNow) sort_queue(S, W, buf ^ 1, W.qBase + k + 1, false, scratch, stream);
            v.extend(S, W, tune.traceGrid, buf ^ 1, W.qBase + k + 1, work++, sortNow ? W.sortIds[0] : nullptr,
                     tune.refill, stream);
            if (tune.tailAfter > 0 && k + 1 == tune.tailAfter) {  // the rest of every path in one launch
                launch_tail(mode, leaf8, S, W, tune.tailGrid, k + 1, buf ^ 1, W.qBase + k + 1, work++, stream);
                break;
            }
        }
    }
    hipLaunchKernelGGL(HIP_KERNEL_NAME(k_wf_finish<false>), dim3((W.numPaths + 255) / 256), dim3(256), 0, stream, S,
                       W);
    MFH_HIP(hipGetLastError());
}

void launch_pt_wavefront_batch(const WfState& W, const WfTuning& tune, int maxLen, hipStream_t stream) {
    PtScene none;
    std::memset(&none, 0, sizeof(none));
    const TraceVariant& v = kVariants[2][tune.traceVariant];
    hipLaunchKernelGGL(k_wf_begin, dim3((W.numCounters + 255) / 256), dim3(256), 0, stream, W);
    if (kWfTraceShapes[tune.traceVariant].leafBatch == 8)
        hipLaunchKernelGGL(HIP_KERNEL_NAME(k_wf_primary<2, 8>), dim3((W.numPaths + 255) / 256), dim3(256), 0, stream,
                           none, W);
    else
        hipLaunchKernelGGL(HIP_KERNEL_NAME(k_wf_primary<2, 1>), dim3((W.numPaths + 255) / 256), dim3(256), 0, stream,
                           none, W);
    const int L = maxLen > 1 ? maxLen : 1;
    int work = W.workBase;
    for (int k = 0; k < L; ++k) {
        const int buf = k & 1;
        hipLaunchKernelGGL(HIP_KERNEL_NAME(k_wf_shade<true>), dim3(tune.shadeGrid), dim3(256), 0, stream, none, W, k,
                           buf, W.qBase + k, W.qBase + k + 1, W.shBase + k, work++);
        v.shadow(none, W, tune.traceGrid, W.shBase + k, work++, nullptr, tune.refill, stream);
        if (k + 1 < L) {
            v.extend(none, W, tune.traceGrid, buf ^ 1, W.qBase + k + 1, work++, nullptr, tune.refill, stream);
            if (tune.tailAfter > 0 && k + 1 == tune.tailAfter) {
                launch_tail(2, kWfTraceShapes[tune.traceVariant].leafBatch == 8, none, W, tune.tailGrid, k + 1,
                            buf ^ 1, W.qBase + k + 1, work++, stream);
                break;
            }
        }
    }
    hipLaunchKernelGGL(HIP_KERNEL_NAME(k_wf_finish<true>), dim3((W.numPaths + 255) / 256), dim3(256), 0, stream, none,
                       W);
    MFH_HIP(hipGetLastError());
}

}  // namespace mfh
