// BVH traversal (closest / any hit) as a device function shared by the trace
// kernels and the path tracer: the GPU device of Mufflon's first_intersection /
// any_intersection (intersection.cu:366-753).
//
// One ray per lane. The stack keeps its lowest kLdsDepth entries in LDS (one
// column per lane, so the 64 lanes of a wave touch 64 consecutive dwords) and
// spills deeper entries to a per-lane private array. Nodes are Mufflon's
// 64-byte collapsed LBVH nodes, read as four 16-byte loads; the primitives of
// a leaf are contiguous 48-byte records in BVH order ({v0, primId}, {v1, kind},
// {v2} for triangles and quads, {center, primId}, {radius, .., kind} for
// spheres).
//
// Visiting order and box test are those of the reference CPU path: the slab
// test of intersection.cu:57-75 (computed without its early outs, which give
// the same accept/reject and entry distance), nearer child first with ties to
// the left, the farther child pushed, leaves pushed as (count, index), and the
// two-level instance transition of intersection.cu:321-361. Closest hit is the
// lexicographic minimum of (t, instance, primitive).
#pragma once
#include "mfh_device.hpp"
#include "mfh_internal.hpp"

namespace mfh {

constexpr int kSentinel = (int)0xFFFFFFFF;
constexpr int kStackCap = 128;
constexpr int kTraceThreads = 256;
constexpr int kLdsDepth = 16;
// Leaf batch (trace_ray's kBatch) of the path-tracing kernels: on scenes with
// long leaves (a 10M-triangle heightfield leaves up to ~1400 triangle tests on
// one ray) the slowest rays of a launch are bound by one memory latency per
// triangle; the plain trace kernels keep 1, which is faster on small leaves.
constexpr int kPtLeafBatch = 8;

// A load of BVH-node or triangle data as a global-memory load: the pointers
// arrive in structs as generic pointers, whose loads would compile to flat
// loads that also wait on the LDS counter.
typedef float GlobalF4 __attribute__((ext_vector_type(4)));
__device__ __forceinline__ float4 gload(const float4* p) {
    const GlobalF4 v = *(const __attribute__((address_space(1))) GlobalF4*)p;
    return make_float4(v.x, v.y, v.z, v.w);
}

// An entry of a lane's LDS stack column, addressed as LDS so that a pop,
// which reads either the LDS column or the private spill array, keeps a
// ds_read and a scratch load instead of one flat load through a merged
// pointer.
__device__ __forceinline__ __attribute__((address_space(3))) int& lds_at(int* column, int i) {
    return ((__attribute__((address_space(3))) int*)column)[i];
}
__device__ __forceinline__ int lds_at(const int* column, int i) {
    return ((const __attribute__((address_space(3))) int*)column)[i];
}

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

// The same slab test with the hardware min / max (v_min_f32, v_max_f32,
// v_min3_f32, v_max3_f32: about half the instructions of the compare-selects
// above, and no VCC wait states). Where no x0 .. z1 is NaN its accept / reject
// and the order of its entry distances equal slab()'s: the two min / max
// forms then differ at most in the sign of a zero result, and every use of
// cmx, cmin and the entry distances is a value comparison. slab_finite()
// tells a lane where that holds.
__device__ __forceinline__ bool slab_fast(const float* lo, const float* hi, const FastRay& r, float tmax,
                                          float& cmin) {
    const float x0 = lo[0] * r.inv.x - r.odiv.x, x1 = hi[0] * r.inv.x - r.odiv.x;
    const float y0 = lo[1] * r.inv.y - r.odiv.y, y1 = hi[1] * r.inv.y - r.odiv.y;
    const float z0 = lo[2] * r.inv.z - r.odiv.z, z1 = hi[2] * r.inv.z - r.odiv.z;
    const float cmx = __builtin_fminf(__builtin_fminf(__builtin_fmaxf(x0, x1), __builtin_fmaxf(y0, y1)),
                                      __builtin_fmaxf(z0, z1));
    cmin = __builtin_fmaxf(__builtin_fmaxf(__builtin_fminf(x0, x1), __builtin_fminf(y0, y1)),
                           __builtin_fminf(z0, z1));
    return (cmx >= 0.0f) && (cmin <= tmax) && (cmin <= cmx);
}

// True when the slab tests of ray r against boxes whose coordinates are at
// most coordMax in magnitude can produce no NaN: with |1/d| <= 2^64 on every
// axis and |origin|, coordMax <= 2^62 (node boxes are enlarged by 1.0001
// about their centre, so they stay within 2^63) every product is below
// 2^127 and every difference below 2^128, hence finite. A direction with a
// zero component (1/d infinite) or NaN input fails it.
__device__ __forceinline__ bool slab_finite(const FastRay& r, float coordMax) {
    constexpr float kInvMax = 18446744073709551616.0f;   // 2^64
    constexpr float kCoordMax = 4611686018427387904.0f;  // 2^62
    const float inv = __builtin_fmaxf(__builtin_fmaxf(fabsf(r.inv.x), fabsf(r.inv.y)), fabsf(r.inv.z));
    const float org = __builtin_fmaxf(__builtin_fmaxf(fabsf(r.o.x), fabsf(r.o.y)), fabsf(r.o.z));
    // written so that a NaN anywhere makes it false
    return inv <= kInvMax && org <= kCoordMax && coordMax <= kCoordMax && r.inv.x == r.inv.x &&
           r.inv.y == r.inv.y && r.inv.z == r.inv.z && r.o.x == r.o.x && r.o.y == r.o.y && r.o.z == r.o.z;
}

// Both child boxes of a node; fast (wave-uniform) selects slab_fast.
__device__ __forceinline__ void slab_pair(bool fast, const float* lLo, const float* lHi, const float* rLo,
                                          const float* rHi, const FastRay& r, float tmax, bool& t0, bool& t1,
                                          float& c0, float& c1) {
    if (fast) {
        t0 = slab_fast(lLo, lHi, r, tmax, c0);
        t1 = slab_fast(rLo, rHi, r, tmax, c1);
    } else {
        t0 = slab(lLo, lHi, r, tmax, c0);
        t1 = slab(rLo, rHi, r, tmax, c1);
    }
}

// Quad or sphere of a mixed LoD (intersects_primitve, intersection.cu:220-300,
// without alpha testing): the record (a, b, c) at sorted position pos, the
// quad's fourth vertex from quadV3.
struct PrimHit {
    float t, w0, w1;
    int hit;
};
__device__ __forceinline__ PrimHit prim_hit_other(const FastRay& r, float4 a, float4 b, float4 c,
                                                  const float4* quadV3, int pos) {
    PrimHit h;
    h.t = 0.0f;
    h.w0 = 0.0f;
    h.w1 = 0.0f;
    if (__float_as_int(b.w) == kPrimQuad) {
        const float4 d = quadV3[pos];
        h.hit = ray_quad(r, v3(a.x, a.y, a.z), v3(b.x, b.y, b.z), v3(c.x, c.y, c.z), v3(d.x, d.y, d.z), h.t, h.w0,
                         h.w1)
                    ? 1
                    : 0;
    } else {
        h.hit = ray_sphere(r, v3(a.x, a.y, a.z), b.x, h.t) ? 1 : 0;
    }
    return h;
}

// Test of the primitive whose 48-byte record (a, b, c) sits at sorted
// position pos of a LoD: true with t > 0 and the surface parameters (the
// barycentric weights of v0 and v1, or a quad's (u, v)). quadV3() yields the
// LoD's fourth quad vertices; it is only evaluated for a quad. kMixed: the
// scene has LoDs with quads or spheres; without, every record is a triangle
// and the kernels carry no quad or sphere code (the quad test alone would
// raise the single-instance trace kernel from 53 to 70 VGPRs, 8 to 7 waves).
template <bool kMixed, class QuadV3>
__device__ __forceinline__ bool prim_hit(const FastRay& r, float4 a, float4 b, float4 c, int pos, QuadV3 quadV3,
                                         float& t, float& w0, float& w1) {
    if (!kMixed || __float_as_int(b.w) == kPrimTriangle)
        return ray_triangle(r, v3(a.x, a.y, a.z), v3(b.x, b.y, b.z), v3(c.x, c.y, c.z), t, w0, w1) && t > 0.0f;
    const PrimHit h = prim_hit_other(r, a, b, c, __float_as_int(b.w) == kPrimQuad ? quadV3() : nullptr, pos);
    t = h.t;
    w0 = h.w0;
    w1 = h.w1;
    return h.hit != 0;
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

// The instance transition of world_to_object_space (intersection.cu:321-349):
// transform the world ray into the instance, scale tmax, test the object box.
struct ObjectEntry {
    FastRay ray;
    float scale, tmax;
    bool hit;
};
__device__ __forceinline__ ObjectEntry enter_object(const TraceParams& p, int inst, V3 wo, V3 wd, float hitT,
                                                    DevLod& lod) {
    ObjectEntry e;
    const float4* w = reinterpret_cast<const float4*>(p.worldToInst + 12 * (size_t)inst);
    const float4 a = w[0], b = w[1], c = w[2];
    const M34 M{{a.x, a.y, a.z, a.w, b.x, b.y, b.z, b.w, c.x, c.y, c.z, c.w}};
    const V3 d = xform_dir(M, wd);
    e.scale = len(d);
    e.ray = make_fast_ray(xform_point(M, wo), d / e.scale);
    e.tmax = hitT * e.scale;
    lod = p.lods[p.instLod[inst]];
    float ot;
    e.hit = slab(lod.box, lod.box + 3, e.ray, e.tmax, ot);
    return e;
}


struct TraceHit {
    float t;
    int inst, prim;
    float w0, w1;
    bool overflow;
    unsigned nodeVisits, primTests;
};

// Resumable traversal of one ray: trace_ray's loops (below) with their state
// kept in a struct, so that run() can stop after a budget of steps (node
// visits and primitive tests) and continue later. The persistent wavefront
// kernels use it to refill lanes whose ray has finished while the rest of
// the wave is still traversing. It takes the same steps as trace_ray, so the
// result of a ray does not depend on how its traversal is split; trace_ray
// stays a separate straight-line form because the resumable one needs more
// registers.
// kSingleMode: 1 = one instance (no instance level), 0 = instance level,
// 2 = chosen per ray at init (batches mix both kinds of scenes).
// The stack's deeper entries live in a private array of the caller
// (kSpill ints), passed to run() so that the struct itself holds only
// scalars and stays in registers.
template <bool kAny, bool kStats, int kSingleMode, int kStride, int kLds = kLdsDepth, bool kMixed = false>
struct Traversal {
    static constexpr int kSpill = kStackCap - kLds;
    FastRay cur;
    V3 wo, wd;
    float hitT, hw0, hw1, scale;
    int hInst, hPrim;
    const BvhChild* bvh;
    const float4* tris;
    int numInternal, objPrims, curInst;
    int nodeAddr, primCount, primOffset, sp;
    bool inObj, overflow, done, single;
    unsigned nodeVisits, primTests;

    __device__ __forceinline__ bool is_single() const {
        if constexpr (kSingleMode == 2)
            return single;
        else
            return kSingleMode == 1;
    }
    __device__ __forceinline__ void push(int* lstack, int* spill, int v) {
        if (sp < kLds)
            lds_at(lstack, sp * kStride) = v;
        else if (sp < kStackCap)
            spill[sp - kLds] = v;
        if (sp < kStackCap)
            ++sp;
        else
            overflow = true;
    }
    __device__ __forceinline__ int pop(const int* lstack, const int* spill) {
        --sp;
        return sp < kLds ? lds_at(lstack, sp * kStride) : spill[sp - kLds];
    }

    __device__ __forceinline__ void init(const TraceParams& p, V3 o, V3 d, float tmax) {
        wo = o;
        wd = d;
        sp = 0;
        overflow = done = false;
        cur = make_fast_ray(wo, wd);
        hitT = tmax;
        hInst = hPrim = -1;
        hw0 = hw1 = 0.0f;
        scale = 1.0f;
        nodeVisits = primTests = 0;
        bvh = (const BvhChild*)p.tlasNodes;
        tris = nullptr;
        numInternal = p.tlasNumInternal;
        objPrims = 0;
        inObj = false;
        curInst = -1;
        nodeAddr = 0;
        primCount = 2;
        primOffset = 0;
        if constexpr (kSingleMode == 2) single = p.numInstances == 1;
        if (is_single()) {
            DevLod lod;
            const ObjectEntry e = enter_object(p, 0, wo, wd, hitT, lod);
            if (kStats) ++nodeVisits;
            if (!e.hit) {
                primCount = 0;
            } else {
                cur = e.ray;
                scale = e.scale;
                hitT = e.tmax;
                bvh = (const BvhChild*)lod.nodes;
                tris = lod.triVerts;
                numInternal = lod.numInternal;
                inObj = true;
                curInst = 0;
                if (lod.numPrims == 1) primCount = 1;
            }
        }
    }

    // Advance by at most `budget` steps (kBudget) or to the end; true when
    // the ray is finished.
    template <bool kBudget>
    __device__ __forceinline__ bool run(const TraceParams& p, int* lstack, int* spill, int budget) {
        while (true) {
            // descend through internal nodes, nearer child first, farther pushed
            while (nodeAddr < numInternal && primCount > 0) {
                if (kBudget && budget-- <= 0) return false;
                const float4* nd = reinterpret_cast<const float4*>(bvh + 2 * (size_t)nodeAddr);
                const float4 q0 = gload(nd), q1 = gload(nd + 1), q2 = gload(nd + 2), q3 = gload(nd + 3);
                if (kStats) ++nodeVisits;
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
                    const int pushed = farLeft ? lIdx : rIdx;
                    if (pushed >= numInternal) push(lstack, spill, farLeft ? lCnt : rCnt);
                    push(lstack, spill, pushed);
                }
                const bool nearLeft = c0 <= c1;
                nodeAddr = nearLeft ? lIdx : rIdx;
                primCount = (nodeAddr >= numInternal) ? (nearLeft ? lCnt : rCnt) : 2;
                primOffset = 0;
                if (!t0 && !t1) primCount = 0;
            }
            if (overflow) return true;
            // the primitives of a leaf, one after the other
            bool descend = false;
            while (nodeAddr >= numInternal && primCount > 0) {
                if (kBudget && budget-- <= 0) return false;
                const int leafId = nodeAddr - numInternal + primOffset;
                ++primOffset;
                --primCount;
                if (!is_single() && !inObj) {
                    const int inst = p.tlasPrimIds[leafId];
                    DevLod lod;
                    const ObjectEntry e = enter_object(p, inst, wo, wd, hitT, lod);
                    if (kStats) ++nodeVisits;
                    if (!e.hit) continue;
                    if (lod.numPrims == 1) {
                        // single-primitive object: test it and return to the instance level
                        const float4 a = lod.triVerts[0], b = lod.triVerts[1], c = lod.triVerts[2];
                        float t, w0, w1;
                        if (kStats) ++primTests;
                        const float objT = e.tmax;
                        float ht = objT;
                        if (prim_hit<kMixed>(e.ray, a, b, c, 0, [&] { return lod.quadV3; }, t, w0, w1) &&
                            better(t, inst, __float_as_int(a.w), ht, hInst, hPrim)) {
                            ht = t;
                            hInst = inst;
                            hPrim = __float_as_int(a.w);
                            hw0 = w0;
                            hw1 = w1;
                            if (kAny) done = true;
                        }
                        hitT = ht / e.scale;
                        if (done) return true;
                        continue;
                    }
                    if (primCount > 0) {
                        push(lstack, spill, primCount);
                        push(lstack, spill, nodeAddr + primOffset);
                    }
                    push(lstack, spill, kSentinel);
                    cur = e.ray;
                    scale = e.scale;
                    hitT = e.tmax;
                    bvh = (const BvhChild*)lod.nodes;
                    tris = lod.triVerts;
                    numInternal = lod.numInternal;
                    objPrims = lod.numPrims;
                    inObj = true;
                    curInst = inst;
                    nodeAddr = 0;
                    primCount = objPrims == 1 ? 1 : 2;
                    primOffset = 0;
                    descend = true;
                    break;
                }
                const float4 a = gload(tris + 3 * (size_t)leafId), b = gload(tris + 3 * (size_t)leafId + 1),
                             c = gload(tris + 3 * (size_t)leafId + 2);
                float t, w0, w1;
                const int prim = __float_as_int(a.w);
                if (kStats) ++primTests;
                if (prim_hit<kMixed>(cur, a, b, c, leafId, [&] { return p.lods[p.instLod[curInst]].quadV3; }, t, w0,
                                     w1) &&
                    better(t, curInst, prim, hitT, hInst, hPrim)) {
                    hitT = t;
                    hInst = curInst;
                    hPrim = prim;
                    hw0 = w0;
                    hw1 = w1;
                    if (kAny) {
                        done = true;
                        return true;
                    }
                }
            }
            if (overflow) return true;
            if (descend) continue;
            // pop the next subtree (or return from an object to the instance level)
            if (sp == 0) return true;
            nodeAddr = pop(lstack, spill);
            if (!is_single() && nodeAddr == kSentinel) {
                if (sp == 0) return true;
                cur = make_fast_ray(wo, wd);
                hitT /= scale;
                scale = 1.0f;
                bvh = (const BvhChild*)p.tlasNodes;
                numInternal = p.tlasNumInternal;
                inObj = false;
                nodeAddr = pop(lstack, spill);
            }
            primCount = (nodeAddr >= numInternal) ? pop(lstack, spill) : 2;
            primOffset = 0;
        }
    }

    // On return t is the world-space distance (or the round-tripped tmax on
    // a miss) and inst/prim are -1 on a miss.
    __device__ __forceinline__ TraceHit result() const {
        TraceHit h;
        h.t = (!kAny && inObj) ? hitT / scale : hitT;  // object_to_world_space
        h.inst = hInst;
        h.prim = hPrim;
        h.w0 = hw0;
        h.w1 = hw1;
        h.overflow = overflow;
        h.nodeVisits = nodeVisits;
        h.primTests = primTests;
        return h;
    }
};

// Traverse from (wo, wd) up to tmax. lstack: this lane's LDS stack column
// (stride kStride ints). On return t is the world-space distance (or the
// round-tripped tmax on a miss) and inst/prim are -1 on a miss.
template <bool kAny, bool kStats, bool kSingle, int kStride, int kLds = kLdsDepth, int kBatch = 1,
          bool kMixed = false>
__device__ __forceinline__ TraceHit trace_ray(const TraceParams& p, V3 wo, V3 wd, float tmax, int* const lstack) {
    int spill[kStackCap - kLds];
    int sp = 0;
    bool overflow = false;
#define MFH_PUSH(v)                                     \
    do {                                                \
        if (sp < kLds)                                  \
            lds_at(lstack, sp * kStride) = (v);                 \
        else if (sp < kStackCap)                        \
            spill[sp - kLds] = (v);                     \
        if (sp < kStackCap) ++sp; else overflow = true; \
    } while (0)
#define MFH_POP() (--sp, (sp < kLds) ? lds_at(lstack, sp * kStride) : spill[sp - kLds])

    FastRay cur = make_fast_ray(wo, wd);
    float hitT = tmax;
    int hInst = -1, hPrim = -1;
    float hw0 = 0.0f, hw1 = 0.0f;
    float scale = 1.0f;
    unsigned nodeVisits = 0, primTests = 0;

    // current level: the instance BVH or the BVH of the entered object
    const BvhChild* bvh = (const BvhChild*)p.tlasNodes;
    const float4* tris = nullptr;
    int numInternal = p.tlasNumInternal;
    int objPrims = 0;
    bool inObj = false;
    int curInst = -1;
    int nodeAddr = 0, primCount = 2, primOffset = 0;
    bool done = false;
    // no NaN can reach this lane's slab tests at the current level
    bool finite = slab_finite(cur, p.tlasCoordMax);

    if (kSingle) {
        DevLod lod;
        const ObjectEntry e = enter_object(p, 0, wo, wd, hitT, lod);
        if (kStats) ++nodeVisits;
        if (!e.hit) {
            primCount = 0;
        } else {
            cur = e.ray;
            finite = slab_finite(cur, lod.coordMax);
            scale = e.scale;
            hitT = e.tmax;
            bvh = (const BvhChild*)lod.nodes;
            tris = lod.triVerts;
            numInternal = lod.numInternal;
            inObj = true;
            curInst = 0;
            if (lod.numPrims == 1) primCount = 1;
        }
    }

    while (true) {
        // descend through internal nodes, nearer child first, farther pushed
        while (nodeAddr < numInternal && primCount > 0) {
            const float4* nd = reinterpret_cast<const float4*>(bvh + 2 * (size_t)nodeAddr);
            const float4 q0 = gload(nd), q1 = gload(nd + 1), q2 = gload(nd + 2), q3 = gload(nd + 3);
            if (kStats) ++nodeVisits;
            const float lLo[3] = {q0.x, q0.y, q0.z}, lHi[3] = {q0.w, q1.x, q1.y};
            const float rLo[3] = {q2.x, q2.y, q2.z}, rHi[3] = {q2.w, q3.x, q3.y};
            const int lIdx = __float_as_int(q1.z), lCnt = __float_as_int(q1.w);
            const int rIdx = __float_as_int(q3.z), rCnt = __float_as_int(q3.w);
            float c0, c1;
            bool t0, t1;
            slab_pair(__all(finite), lLo, lHi, rLo, rHi, cur, hitT, t0, t1, c0, c1);
            if (!t0) c0 = 1e38f;
            if (!t1) c1 = 1e38f;
            if (t0 && t1) {
                const bool farLeft = c1 < c0;
                const int push = farLeft ? lIdx : rIdx;
                if (push >= numInternal) MFH_PUSH(farLeft ? lCnt : rCnt);
                MFH_PUSH(push);
            }
            const bool nearLeft = c0 <= c1;
            nodeAddr = nearLeft ? lIdx : rIdx;
            primCount = (nodeAddr >= numInternal) ? (nearLeft ? lCnt : rCnt) : 2;
            primOffset = 0;
            if (!t0 && !t1) primCount = 0;
        }
        if (overflow) break;
        // the primitives of a leaf, one after the other
        bool descend = false;
        while (nodeAddr >= numInternal && primCount > 0) {
            if (kBatch > 1 && (kSingle || inObj)) {
                // Triangles in batches of kBatch: every load of the batch is
                // issued before the first test, so a long leaf costs one memory
                // latency per batch instead of one per triangle. The tests and
                // hit updates still run in leaf order (the result and the counts
                // are those of testing one at a time).
                const int nb = primCount < kBatch ? primCount : kBatch;
                const int first = nodeAddr - numInternal + primOffset;
                float4 ta[kBatch], tb[kBatch], tc[kBatch];
#pragma unroll
                for (int j = 0; j < kBatch; ++j) {
                    if (j < nb) {
                        const size_t id = (size_t)(first + j);
                        ta[j] = gload(tris + 3 * id);
                        tb[j] = gload(tris + 3 * id + 1);
                        tc[j] = gload(tris + 3 * id + 2);
                    }
                }
#pragma unroll
                for (int j = 0; j < kBatch; ++j) {
                    if (j >= nb || done) continue;
                    float t, w0, w1;
                    const int prim = __float_as_int(ta[j].w);
                    if (kStats) ++primTests;
                    if (prim_hit<kMixed>(cur, ta[j], tb[j], tc[j], first + j,
                                 [&] { return p.lods[p.instLod[kSingle ? 0 : curInst]].quadV3; }, t, w0, w1) &&
                        better(t, curInst, prim, hitT, hInst, hPrim)) {
                        hitT = t;
                        hInst = curInst;
                        hPrim = prim;
                        hw0 = w0;
                        hw1 = w1;
                        if (kAny) done = true;
                    }
                }
                primOffset += nb;
                primCount -= nb;
                if (done) break;
                continue;
            }
            const int leafId = nodeAddr - numInternal + primOffset;
            ++primOffset;
            --primCount;
            if (!kSingle && !inObj) {
                // an instance of the top level: enter its object
                const int inst = p.tlasPrimIds[leafId];
                DevLod lod;
                const ObjectEntry e = enter_object(p, inst, wo, wd, hitT, lod);
                if (kStats) ++nodeVisits;
                if (!e.hit) continue;
                if (lod.numPrims == 1) {
                    // single-primitive object: test it and return to the instance level
                    const float4 a = lod.triVerts[0], b = lod.triVerts[1], c = lod.triVerts[2];
                    float t, w0, w1;
                    if (kStats) ++primTests;
                    const float objT = e.tmax;
                    float ht = objT;
                    if (prim_hit<kMixed>(e.ray, a, b, c, 0, [&] { return lod.quadV3; }, t, w0, w1) &&
                        better(t, inst, __float_as_int(a.w), ht, hInst, hPrim)) {
                        ht = t;
                        hInst = inst;
                        hPrim = __float_as_int(a.w);
                        hw0 = w0;
                        hw1 = w1;
                        if (kAny) done = true;
                    }
                    hitT = ht / e.scale;
                    if (done) break;
                    continue;
                }
                if (primCount > 0) {
                    MFH_PUSH(primCount);
                    MFH_PUSH(nodeAddr + primOffset);
                }
                MFH_PUSH(kSentinel);
                cur = e.ray;
                finite = slab_finite(cur, lod.coordMax);
                scale = e.scale;
                hitT = e.tmax;
                bvh = (const BvhChild*)lod.nodes;
                tris = lod.triVerts;
                numInternal = lod.numInternal;
                objPrims = lod.numPrims;
                inObj = true;
                curInst = inst;
                nodeAddr = 0;
                primCount = objPrims == 1 ? 1 : 2;
                primOffset = 0;
                descend = true;
                break;
            }
            // one triangle at a time (kBatch == 1)
            const float4 a = gload(tris + 3 * (size_t)leafId), b = gload(tris + 3 * (size_t)leafId + 1),
                         c = gload(tris + 3 * (size_t)leafId + 2);
            float t, w0, w1;
            const int prim = __float_as_int(a.w);
            if (kStats) ++primTests;
            if (prim_hit<kMixed>(cur, a, b, c, leafId, [&] { return p.lods[p.instLod[kSingle ? 0 : curInst]].quadV3; },
                                 t, w0, w1) &&
                better(t, curInst, prim, hitT, hInst, hPrim)) {
                hitT = t;
                hInst = curInst;
                hPrim = prim;
                hw0 = w0;
                hw1 = w1;
                if (kAny) {
                    done = true;
                    break;
                }
            }
        }
        if (done || overflow) break;
        if (descend) continue;
        // pop the next subtree (or return from an object to the instance level)
        if (sp == 0) break;
        nodeAddr = MFH_POP();
        if (!kSingle && nodeAddr == kSentinel) {
            if (sp == 0) break;
            cur = make_fast_ray(wo, wd);
            finite = slab_finite(cur, p.tlasCoordMax);
            hitT /= scale;
            scale = 1.0f;
            bvh = (const BvhChild*)p.tlasNodes;
            numInternal = p.tlasNumInternal;
            inObj = false;
            nodeAddr = MFH_POP();
        }
        primCount = (nodeAddr >= numInternal) ? MFH_POP() : 2;
        primOffset = 0;
    }
#undef MFH_PUSH
#undef MFH_POP
    if (!kAny && inObj) hitT /= scale;  // object_to_world_space
    TraceHit h;
    h.t = hitT;
    h.inst = hInst;
    h.prim = hPrim;
    h.w0 = hw0;
    h.w1 = hw1;
    h.overflow = overflow;
    h.nodeVisits = nodeVisits;
    h.primTests = primTests;
    return h;
}

// Closest hit of a single-instance scene with postponed leaves (Aila and
// Laine 2009, "speculative while-while"): a lane that reaches a leaf keeps it
// and goes on descending through inner nodes until it reaches a second leaf
// or half of the active lanes of its wave hold one; then the held leaves are
// tested (kBatch triangles loaded at a time, as in trace_ray). Node steps and
// leaf tests of a wave thus run with more lanes active. (1M-triangle
// benchmark, test when: every lane holds 976, 7/8 994, 3/4 1041, 5/8 1086,
// 1/2 1122, 3/8 1106, 1/4 1073 Mrays/s. On the two-level terrain, holding
// inside objects was slower, so two-level scenes keep trace_ray.)
// The visiting order differs from trace_ray's, which does not change the
// result: inside one object the closest hit is the lexicographic minimum of
// (t, primitive) over every triangle whose box path passes the slab tests,
// and a box is culled only when its entry distance exceeds the current best
// (the collapse enlarges boxes, so a triangle never lies closer than its
// box's entry). No counts (the stats kernels use trace_ray).
template <int kStride, int kLds = kLdsDepth, int kBatch = 1, bool kMixed = false>
__device__ __forceinline__ TraceHit trace_closest_held(const TraceParams& p, V3 wo, V3 wd, float tmax,
                                                       int* const lstack) {
    int spill[kStackCap - kLds];
    int sp = 0;
    bool overflow = false;
    auto push = [&](int v) {
        if (sp < kLds)
            lds_at(lstack, sp * kStride) = v;
        else if (sp < kStackCap)
            spill[sp - kLds] = v;
        if (sp < kStackCap)
            ++sp;
        else
            overflow = true;
    };
    auto pop = [&]() {
        --sp;
        return sp < kLds ? lds_at(lstack, sp * kStride) : spill[sp - kLds];
    };
    TraceHit h;
    h.inst = h.prim = -1;
    h.w0 = h.w1 = 0.0f;
    h.nodeVisits = h.primTests = 0;
    DevLod lod;
    const ObjectEntry e = enter_object(p, 0, wo, wd, tmax, lod);
    if (!e.hit) {
        h.t = tmax;
        h.overflow = false;
        return h;
    }
    const FastRay cur = e.ray;
    // every lane of the wave that traverses is active here: the choice holds for the whole traversal
    const bool fast = __all(slab_finite(cur, lod.coordMax));
    float hitT = e.tmax;
    const BvhChild* const bvh = (const BvhChild*)lod.nodes;
    const float4* const tris = lod.triVerts;
    const int numInternal = lod.numInternal;
    // current node (count 2 = inner node, n = leaf of n triangles, 0 = none)
    int nodeAddr = 0, primCount = lod.numPrims == 1 ? 1 : 2;
    int heldAddr = 0, heldCount = 0;  // the postponed leaf
    while (true) {
        while (true) {
            if (primCount > 0 && nodeAddr < numInternal) {
                const float4* nd = reinterpret_cast<const float4*>(bvh + 2 * (size_t)nodeAddr);
                const float4 q0 = gload(nd), q1 = gload(nd + 1), q2 = gload(nd + 2), q3 = gload(nd + 3);
                const float lLo[3] = {q0.x, q0.y, q0.z}, lHi[3] = {q0.w, q1.x, q1.y};
                const float rLo[3] = {q2.x, q2.y, q2.z}, rHi[3] = {q2.w, q3.x, q3.y};
                const int lIdx = __float_as_int(q1.z), lCnt = __float_as_int(q1.w);
                const int rIdx = __float_as_int(q3.z), rCnt = __float_as_int(q3.w);
                float c0, c1;
                bool t0, t1;
                slab_pair(fast, lLo, lHi, rLo, rHi, cur, hitT, t0, t1, c0, c1);
                if (!t0) c0 = 1e38f;
                if (!t1) c1 = 1e38f;
                if (t0 && t1) {
                    const bool farLeft = c1 < c0;
                    const int far = farLeft ? lIdx : rIdx;
                    if (far >= numInternal) push(farLeft ? lCnt : rCnt);
                    push(far);
                }
                const bool nearLeft = c0 <= c1;
                nodeAddr = nearLeft ? lIdx : rIdx;
                primCount = (!t0 && !t1) ? 0 : (nodeAddr >= numInternal ? (nearLeft ? lCnt : rCnt) : 2);
                if (overflow) break;
            }
            if (primCount > 0 && nodeAddr >= numInternal) {
                if (heldCount > 0) break;  // a second leaf: test the held one first
                heldAddr = nodeAddr;
                heldCount = primCount;
                primCount = 0;
            }
            if (primCount == 0) {
                if (sp == 0) break;
                nodeAddr = pop();
                primCount = nodeAddr >= numInternal ? pop() : 2;
            }
            // half of the active lanes hold a leaf: test them together
            if (2 * __popcll(__ballot(heldCount > 0)) >= __popcll(__builtin_amdgcn_read_exec())) break;
        }
        if (overflow) break;
        for (int k = 0; k < heldCount; k += kBatch) {
            const int nb = heldCount - k < kBatch ? heldCount - k : kBatch;
            const int first = heldAddr - numInternal + k;
            float4 ta[kBatch], tb[kBatch], tc[kBatch];
#pragma unroll
            for (int j = 0; j < kBatch; ++j) {
                if (j < nb) {
                    const size_t id = (size_t)(first + j);
                    ta[j] = gload(tris + 3 * id);
                    tb[j] = gload(tris + 3 * id + 1);
                    tc[j] = gload(tris + 3 * id + 2);
                }
            }
#pragma unroll
            for (int j = 0; j < kBatch; ++j) {
                if (j >= nb) continue;
                float t, w0, w1;
                const int prim = __float_as_int(ta[j].w);
                if (prim_hit<kMixed>(cur, ta[j], tb[j], tc[j], first + j, [&] { return lod.quadV3; }, t, w0, w1) &&
                    better(t, 0, prim, hitT, h.inst, h.prim)) {
                    hitT = t;
                    h.inst = 0;
                    h.prim = prim;
                    h.w0 = w0;
                    h.w1 = w1;
                }
            }
        }
        heldCount = 0;
        if (primCount == 0 && sp == 0) break;
    }
    h.t = hitT / e.scale;  // object_to_world_space
    h.overflow = overflow;
    return h;
}

// Traversal specializations of the kernels that trace (their kMode): fixed
// for one scene, or chosen per ray (a batch of scenes mixes single- and
// multi-instance scenes; a scene with quads or spheres takes the general
// form).
constexpr int kTraceMulti = 0;   // instance level, triangle LoDs
constexpr int kTraceSingle = 1;  // one instance, triangle LoDs
constexpr int kTraceBatch = 2;   // per ray, triangle LoDs
constexpr int kTraceMixed = 3;   // per ray, LoDs with quads and spheres
// The mode of a kernel tracing one scene.
inline int trace_mode_of(const TraceParams& p) {
    return p.mixed ? kTraceMixed : (p.numInstances == 1 ? kTraceSingle : kTraceMulti);
}
// Traversal kind of the resumable Traversal for a mode.
template <int kMode>
constexpr int traversal_mode() {
    return kMode == kTraceSingle ? 1 : (kMode == kTraceMulti ? 0 : 2);
}

// Closest (kAny = false) or any hit in mode kMode; closest hits of
// single-instance scenes hold leaves (trace_closest_held) unless kHeld is
// false.
template <int kMode, bool kAny, int kStride, int kLds, int kLeaf, bool kHeld = true>
__device__ __forceinline__ TraceHit trace_mode(const TraceParams& tp, V3 o, V3 d, float tmax, int* lstack) {
    constexpr bool kMixed = kMode == kTraceMixed;
    if constexpr (kMode == kTraceSingle && !kAny && kHeld)
        return trace_closest_held<kStride, kLds, kLeaf>(tp, o, d, tmax, lstack);
    if constexpr (kMode == kTraceSingle) return trace_ray<kAny, false, true, kStride, kLds, kLeaf>(tp, o, d, tmax, lstack);
    if constexpr (kMode == kTraceMulti) return trace_ray<kAny, false, false, kStride, kLds, kLeaf>(tp, o, d, tmax, lstack);
    if (tp.numInstances == 1) {
        if constexpr (!kAny && kHeld) return trace_closest_held<kStride, kLds, kLeaf, kMixed>(tp, o, d, tmax, lstack);
        return trace_ray<kAny, false, true, kStride, kLds, kLeaf, kMixed>(tp, o, d, tmax, lstack);
    }
    return trace_ray<kAny, false, false, kStride, kLds, kLeaf, kMixed>(tp, o, d, tmax, lstack);
}

}  // namespace mfh
