// Closest-hit and any-hit BVH traversal for gfx950: the GPU device of
// Mufflon's first_intersection / any_intersection (intersection.cu:366-753).
//
// One ray per lane. The traversal stack keeps its lowest kLdsDepth entries in
// LDS (one column per lane, so the 64 lanes of a wave touch 64 consecutive
// dwords) and spills deeper entries to a per-lane private array. Nodes are
// Mufflon's 64-byte collapsed LBVH nodes, read as four 16-byte loads; the
// triangles of a leaf are contiguous 48-byte records in BVH order
// ({v0, primId}, {v1}, {v2}), so a leaf test needs no index indirection.
//
// Visiting order and box test are those of the reference CPU path: the slab
// test of intersection.cu:57-75 (computed without its early outs, which give
// the same accept/reject and entry distance), nearer child first with ties
// to the left, the farther child pushed, leaves pushed as (count, index), and
// the two-level instance transition of intersection.cu:321-361. Closest hit is
// the lexicographic minimum of (t, instance, primitive).
#include "mfh_device.hpp"
#include "mfh_internal.hpp"

namespace mfh {

namespace {

constexpr int kSentinel = (int)0xFFFFFFFF;
constexpr int kStackCap = 128;
constexpr int kThreads = 256;
constexpr int kLdsDepth = 16;

// Slab test, reference CPU semantics (intersection.cu:57-75).
__device__ __forceinline__ bool slab(const float* lo, const float* hi, const FastRay& r, float tmax, float& cmin) {
    const float x0 = lo[0] * r.inv.x - r.odiv.x, x1 = hi[0] * r.inv.x - r.odiv.x;
    const float y0 = lo[1] * r.inv.y - r.odiv.y, y1 = hi[1] * r.inv.y - r.odiv.y;
    const float z0 = lo[2] * r.inv.z - r.odiv.z, z1 = hi[2] * r.inv.z - r.odiv.z;
    const float mnx = mn(x0, x1), mxx = mx(x0, x1);
    const float mny = mn(y0, y1), mxy = mx(y0, y1);
    const float mnz = mn(z0, z1), mxz = mx(z0, z1);
    const float cmx = mn(mn(mxx, mxy), mxz);
    cmin = mx(mx(mnx, mny), mnz);
    return (cmx >= 0.0f) && (cmin <= tmax) && (cmin <= cmx);
}

// Accept a hit closer than the current one; at equal distance the lower
// (instance, primitive) pair wins. A hit exactly at the initial tmax (no hit
// yet) is rejected, as the reference's strict t < tmax test does.
__device__ __forceinline__ bool better(float t, int inst, int prim, float hitT, int hInst, int hPrim) {
    if (t < hitT) return true;
    if (t > hitT || hInst == -1) return false;
    if (inst != hInst) return inst < hInst;
    return prim < hPrim;
}

struct Stack {
    int* lds;  // column of this lane: lds[k * kThreads]
    int spill[kStackCap - kLdsDepth];
    int sp;
    bool overflow;
    __device__ __forceinline__ void push(int v) {
        if (sp < kLdsDepth)
            lds[sp * kThreads] = v;
        else if (sp < kStackCap)
            spill[sp - kLdsDepth] = v;
        else {
            overflow = true;
            return;
        }
        ++sp;
    }
    __device__ __forceinline__ int pop() {
        --sp;
        return (sp < kLdsDepth) ? lds[sp * kThreads] : spill[sp - kLdsDepth];
    }
};

struct Level {
    const BvhChild* bvh;
    const float4* tris;
    const int* primIds;
    int numInternal;
};

// kStats: also count per ray the internal nodes visited (object-box tests at
// instance entry included) and the triangles tested.
template <bool kAny, bool kStats>
__global__ __launch_bounds__(kThreads) void k_trace(TraceParams p, const float4* __restrict__ rays, int n,
                                                   float4* __restrict__ hits, float2* __restrict__ bary,
                                                   int* __restrict__ occluded, uint2* __restrict__ stats) {
    __shared__ int sStack[kLdsDepth * kThreads];
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float4 r0 = rays[2 * (size_t)i], r1 = rays[2 * (size_t)i + 1];
    const FastRay world = make_fast_ray(v3(r0.x, r0.y, r0.z), v3(r1.x, r1.y, r1.z));
    float hitT = r0.w;
    int hInst = -1, hPrim = -1;
    float hw0 = 0.0f, hw1 = 0.0f;

    Stack st;
    st.lds = sStack + threadIdx.x;
    st.sp = 0;
    st.overflow = false;

    const Level top{(const BvhChild*)p.tlasNodes, nullptr, p.tlasPrimIds, p.tlasNumInternal};
    Level lv = top;
    FastRay cur = world;
    float scale = 1.0f;
    bool inObj = false;
    int curInst = -1;
    int objPrims = 0;
    const float4* objTris = nullptr;
    int nodeAddr = 0, primCount = 2, primOffset = 0;
    bool done = false;
    unsigned nodeVisits = 0, primTests = 0;

    // world_to_object_space: transform the ray, test the object box.
    auto to_object = [&](int inst) -> bool {
        M34 M;
        const float4* w = reinterpret_cast<const float4*>(p.worldToInst + 12 * (size_t)inst);
        const float4 a = w[0], b = w[1], c = w[2];
        M.m[0] = a.x; M.m[1] = a.y; M.m[2] = a.z; M.m[3] = a.w;
        M.m[4] = b.x; M.m[5] = b.y; M.m[6] = b.z; M.m[7] = b.w;
        M.m[8] = c.x; M.m[9] = c.y; M.m[10] = c.z; M.m[11] = c.w;
        const V3 d = xform_dir(M, world.d);
        const float sc = len(d);
        const FastRay fr = make_fast_ray(xform_point(M, world.o), d / sc);
        const float objTMax = hitT * sc;
        const DevLod& L = p.lods[p.instLod[inst]];
        float ot;
        if (kStats) ++nodeVisits;
        if (!slab(L.box, L.box + 3, fr, objTMax, ot)) return false;
        cur = fr;
        scale = sc;
        hitT = objTMax;
        objPrims = L.numPrims;
        objTris = L.triVerts;
        lv.bvh = (const BvhChild*)L.nodes;
        lv.tris = L.triVerts;
        lv.numInternal = L.numInternal;
        return true;
    };
    auto to_world = [&]() {
        cur = world;
        hitT /= scale;
        scale = 1.0f;
        lv = top;
        inObj = false;
    };
    auto test_tri = [&](const float4* tris, int pos, int inst) -> bool {
        const float4 a = tris[3 * (size_t)pos], b = tris[3 * (size_t)pos + 1], c = tris[3 * (size_t)pos + 2];
        float t, w0, w1;
        const int prim = __float_as_int(a.w);
        if (kStats) ++primTests;
        if (ray_triangle(cur, v3(a.x, a.y, a.z), v3(b.x, b.y, b.z), v3(c.x, c.y, c.z), t, w0, w1) && t > 0.0f &&
            better(t, inst, prim, hitT, hInst, hPrim)) {
            hitT = t;
            hInst = inst;
            hPrim = prim;
            hw0 = w0;
            hw1 = w1;
            return true;
        }
        return false;
    };

    if (p.numInstances == 1) {
        if (!to_object(0))
            primCount = 0;
        else {
            inObj = true;
            curInst = 0;
            if (objPrims == 1) primCount = 1;
        }
    }

    while (!done && (st.sp > 0 || primCount > 0)) {
        if (primCount == 0) {
            nodeAddr = st.pop();
            if (nodeAddr == kSentinel) {
                if (st.sp == 0) break;
                to_world();
                nodeAddr = st.pop();
            }
            if (nodeAddr >= lv.numInternal)
                primCount = st.pop();
            else
                primCount = 2;
            primOffset = 0;
        }
        while (nodeAddr < lv.numInternal && primCount > 0) {
            const float4* nd = reinterpret_cast<const float4*>(lv.bvh + 2 * (size_t)nodeAddr);
            if (kStats) ++nodeVisits;
            const float4 q0 = nd[0], q1 = nd[1], q2 = nd[2], q3 = nd[3];
            const float lLo[3] = {q0.x, q0.y, q0.z}, lHi[3] = {q0.w, q1.x, q1.y};
            const float rLo[3] = {q2.x, q2.y, q2.z}, rHi[3] = {q2.w, q3.x, q3.y};
            const int lIdx = __float_as_int(q1.z), lCnt = __float_as_int(q1.w);
            const int rIdx = __float_as_int(q3.z), rCnt = __float_as_int(q3.w);
            float c0, c1;
            const bool t0 = slab(lLo, lHi, cur, hitT, c0);
            const bool t1 = slab(rLo, rHi, cur, hitT, c1);
            if (!t0) c0 = 1e38f;
            if (!t1) c1 = 1e38f;
            if (t0 && t1) {
                const bool farLeft = c1 < c0;
                const int push = farLeft ? lIdx : rIdx;
                if (push >= lv.numInternal) st.push(farLeft ? lCnt : rCnt);
                st.push(push);
            }
            const bool nearLeft = c0 <= c1;
            nodeAddr = nearLeft ? lIdx : rIdx;
            if (nodeAddr >= lv.numInternal) {
                primCount = nearLeft ? lCnt : rCnt;
                primOffset = 0;
            } else
                primCount = 2;
            if (!t0 && !t1) primCount = 0;
        }
        if (st.overflow) break;
        if (nodeAddr >= lv.numInternal && primCount > 0) {
            const int leafId = nodeAddr - lv.numInternal + primOffset;
            ++primOffset;
            --primCount;
            if (!inObj) {
                const int inst = p.tlasPrimIds[leafId];
                if (to_object(inst)) {
                    if (objPrims == 1) {
                        if (test_tri(objTris, 0, inst) && kAny) done = true;
                        // back to the instance level at once
                        cur = world;
                        hitT /= scale;
                        scale = 1.0f;
                        lv = top;
                    } else {
                        if (primCount > 0) {
                            st.push(primCount);
                            st.push(nodeAddr + primOffset);
                        }
                        st.push(kSentinel);
                        inObj = true;
                        nodeAddr = 0;
                        primCount = 2;
                        primOffset = 0;
                        curInst = inst;
                    }
                }
            } else if (test_tri(lv.tris, leafId, curInst) && kAny) {
                done = true;
            }
        }
    }
    if (kAny) {
        occluded[i] = st.overflow ? -1 : (hInst != -1 ? 1 : 0);
    } else {
        if (inObj) to_world();
        hits[i] = make_float4(hitT, __int_as_float(hInst), __int_as_float(hPrim), st.overflow ? 1.0f : 0.0f);
        if (bary) bary[i] = make_float2(hw0, hw1);
    }
    if (kStats) stats[i] = make_uint2(nodeVisits, primTests);
}

}  // namespace

TraceParams trace_params(const Scene& sc) {
    TraceParams p;
    p.numInstances = sc.numInstances;
    p.worldToInst = sc.dWorldToInst.as<float>();
    p.instLod = sc.dInstLod.as<int32_t>();
    p.lods = sc.dLods.as<DevLod>();
    p.tlasNodes = sc.tlasNodes.ptr;
    p.tlasPrimIds = sc.tlasPrimIds.as<int32_t>();
    p.tlasNumInternal = sc.tlasNumInternal;
    return p;
}

void launch_trace_closest(const TraceParams& p, const float* rays, int n, float4* hits, float2* bary, uint2* stats,
                          hipStream_t stream) {
    if (n <= 0) return;
    const dim3 grid((n + kThreads - 1) / kThreads);
    if (stats)
        hipLaunchKernelGGL(HIP_KERNEL_NAME(k_trace<false, true>), grid, dim3(kThreads), 0, stream, p,
                           reinterpret_cast<const float4*>(rays), n, hits, bary, nullptr, stats);
    else
        hipLaunchKernelGGL(HIP_KERNEL_NAME(k_trace<false, false>), grid, dim3(kThreads), 0, stream, p,
                           reinterpret_cast<const float4*>(rays), n, hits, bary, nullptr, nullptr);
    MFH_HIP(hipGetLastError());
}

void launch_trace_any(const TraceParams& p, const float* rays, int n, int32_t* occluded, hipStream_t stream) {
    if (n <= 0) return;
    hipLaunchKernelGGL(HIP_KERNEL_NAME(k_trace<true, false>), dim3((n + kThreads - 1) / kThreads), dim3(kThreads), 0,
                       stream, p, reinterpret_cast<const float4*>(rays), n, nullptr, nullptr, occluded, nullptr);
    MFH_HIP(hipGetLastError());
}

}  // namespace mfh
