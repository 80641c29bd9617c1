// LBVH builder for gfx950: the GPU device of Mufflon's LBVHBuilder
// (build_lbvh.cu:586-811) for triangle LoDs and for the instance level.
//
// Stages (one kernel each unless noted):
//   k_morton    63-bit Morton code of each primitive centroid (build_lbvh.cu:70-125)
//   radix sort  stable LSD sort of (code, id)           (sort_scan.hip)
//   k_karras    Karras 2012 split hierarchy; writes parents[] in Mufflon's
//               encoding plus explicit child lists and each node's refit
//               level                                     (build_lbvh.cu:164-245)
//   k_refit_tiles, k_refit_level
//               bottom-up AABB + SAH cost climb with one arrival counter per
//               node; the second arriving thread merges   (build_lbvh.cu:326-411);
//               per tile of leaves, then per span of each higher level, with
//               the climbing state in LDS
//   k_mark      collapse decision per internal node      (build_lbvh.cu:417-463)
//   scan        inclusive sum of the keep marks          (build_lbvh.cu:748-758)
//   k_collapse  64-byte-node BVH with enlarged boxes     (build_lbvh.cu:503-566)
//
// Determinism. The refit merges a node in a fixed operand order whatever
// thread arrives second: box = min/max(right, left) and
//   cost = Ct + SA(R)/SA * c(R) + SA(L)/SA * c(L),
// which is the order of the reference's serial CPU path (the right subtree's
// thread always arrives second there). Whether a child is a leaf comes from
// the explicit child list, not from the value left in the arrival counter,
// so the reference's ambiguous encoding (thread N-1 == numInternalNodes,
// build_lbvh.cu:371-374) cannot arise. The collapse marks use the equivalent
// per-node form: an internal node X (not the root) is removed iff X has fewer
// than 1023 primitives of every type and some node on the path from X up to
// (excluding) the first node with 1023 or more, or the root, is cheaper as a
// leaf than split. This is the set of nodes the per-leaf walks of the
// reference mark, because the counts only grow towards the root.
#include "mfh_device.hpp"
#include "mfh_internal.hpp"
#include "mufflon_hip.h"

#include <algorithm>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <string>
#include <type_traits>
#include <vector>

namespace mfh {

namespace {

constexpr i32 kTreeHead = 0x10000000;
constexpr int kThreads = 256;
// Upper bound on any root path: 64 key bits + 32 index bits of a Karras tree.
constexpr int kMaxDepth = 128;

struct Box6 {
    float v[6];
};

__device__ __forceinline__ float clampf(float v, float lo, float hi) { return mn(mx(v, lo), hi); }

__device__ __forceinline__ u64 part_by_two21(u32 x) {
    u64 r = x;
    r = (r | (r << 32)) & 0x001f00000000ffffull;
    r = (r | (r << 16)) & 0x001f0000ff0000ffull;
    r = (r | (r << 8)) & 0x100f00f00f00f00full;
    r = (r | (r << 4)) & 0x10c30c30c30c30c3ull;
    r = (r | (r << 2)) & 0x1249249249249249ull;
    return r;
}

// Mufflon's sdiv (util/types.hpp): x / (d + 1e-20 * sgn(d)), sgn(0) = 1.
__device__ __forceinline__ float sdiv(float x, float d) { return x / (d + 1e-20f * (d < 0.0f ? -1.0f : 1.0f)); }

__device__ __forceinline__ u64 morton_of_point(V3 c, const Box6& bb) {
    const V3 lo = v3(bb.v[0], bb.v[1], bb.v[2]);
    const V3 span = v3(bb.v[3], bb.v[4], bb.v[5]) - lo;
    const V3 num = c - lo;
    const float nx = sdiv(num.x, span.x);
    const float ny = sdiv(num.y, span.y);
    const float nz = sdiv(num.z, span.z);
    const u32 qx = (u32)clampf(nx * 2097152.0f, 0.0f, 2097151.0f);
    const u32 qy = (u32)clampf(ny * 2097152.0f, 0.0f, 2097151.0f);
    const u32 qz = (u32)clampf(nz * 2097152.0f, 0.0f, 2097151.0f);
    return part_by_two21(qx) * 4 + part_by_two21(qy) * 2 + part_by_two21(qz);
}

// ------------------------------------------------------------- primitive sets
struct TriPrims {
    const float* verts;
    const u32* tris;
    __device__ __forceinline__ V3 vert(u32 i) const { return v3(verts[3 * i], verts[3 * i + 1], verts[3 * i + 2]); }
    __device__ __forceinline__ void tri(i32 p, V3& a, V3& b, V3& c) const {
        const u32 i0 = tris[3 * p], i1 = tris[3 * p + 1], i2 = tris[3 * p + 2];
        a = vert(i0);
        b = vert(i1);
        c = vert(i2);
    }
    __device__ __forceinline__ V3 centroid(i32 p) const {
        V3 a, b, c;
        tri(p, a, b, c);
        return (a + b + c) / 3.0f;
    }
    __device__ __forceinline__ void box(i32 p, V3& lo, V3& hi) const {
        V3 a, b, c;
        tri(p, a, b, c);
        lo = vmin(vmin(a, b), c);
        hi = vmax(vmax(a, b), c);
    }
    __device__ __forceinline__ bool masked(i32) const { return false; }
    __device__ __forceinline__ int kind(i32) const { return kPrimTriangle; }
    static constexpr bool kRecords = true;
    // Staging: the Morton pass, which reads every triangle in order anyway,
    // writes its record in primitive order; the refit then gathers one
    // 48-byte record per leaf instead of an index row and three vertices.
    static constexpr bool kStaged = true;
    __device__ __forceinline__ V3 centroid_staged(i32 p, float4* stage) const {
        V3 a, b, c;
        tri(p, a, b, c);
        stage[3 * (size_t)p] = make_float4(a.x, a.y, a.z, __int_as_float(p));
        stage[3 * (size_t)p + 1] = make_float4(b.x, b.y, b.z, 0.0f);
        stage[3 * (size_t)p + 2] = make_float4(c.x, c.y, c.z, 0.0f);
        return (a + b + c) / 3.0f;
    }
    // The record of `prim` from the staged copy, moved to sorted position idx; its box.
    __device__ __forceinline__ void record_staged(i32 prim, int idx, const float4* stage, float4* recs, V3& lo,
                                                  V3& hi) const {
        const float4 r0 = stage[3 * (size_t)prim], r1 = stage[3 * (size_t)prim + 1], r2 = stage[3 * (size_t)prim + 2];
        const V3 a = v3(r0.x, r0.y, r0.z), b = v3(r1.x, r1.y, r1.z), c = v3(r2.x, r2.y, r2.z);
        lo = vmin(vmin(a, b), c);
        hi = vmax(vmax(a, b), c);
        recs[3 * (size_t)idx] = r0;
        recs[3 * (size_t)idx + 1] = r1;
        recs[3 * (size_t)idx + 2] = r2;
    }
    // The traversal record of triangle `prim` at sorted position idx; its box.
    __device__ __forceinline__ void record(i32 prim, int idx, float4* recs, V3& lo, V3& hi) const {
        V3 a, b, c;
        tri(prim, a, b, c);
        lo = vmin(vmin(a, b), c);
        hi = vmax(vmax(a, b), c);
        recs[3 * (size_t)idx] = make_float4(a.x, a.y, a.z, __int_as_float(prim));
        recs[3 * (size_t)idx + 1] = make_float4(b.x, b.y, b.z, 0.0f);
        recs[3 * (size_t)idx + 2] = make_float4(c.x, c.y, c.z, 0.0f);
    }
};

// A LoD with quads and / or spheres besides triangles, primitive ids in the
// order triangles, quads, spheres (descriptors.hpp:33-68). Centroids and
// boxes of accel_structs_commons.hpp:56-100: a quad's centroid is the mean of
// its four vertices, a sphere's its center, its box center -+ radius.
struct MixedPrims {
    const float* verts;
    const u32* tris;
    const u32* quads;
    const float4* spheres;
    float4* quadV3;  // out: fourth quad vertex per sorted position
    i32 nt, nq;
    __device__ __forceinline__ V3 vert(u32 i) const { return v3(verts[3 * i], verts[3 * i + 1], verts[3 * i + 2]); }
    __device__ __forceinline__ int kind(i32 p) const {
        return p < nt ? kPrimTriangle : (p < nt + nq ? kPrimQuad : kPrimSphere);
    }
    __device__ __forceinline__ void quad(i32 q, V3& a, V3& b, V3& c, V3& d) const {
        const u32* ix = quads + 4 * (size_t)q;
        a = vert(ix[0]);
        b = vert(ix[1]);
        c = vert(ix[2]);
        d = vert(ix[3]);
    }
    __device__ __forceinline__ V3 centroid(i32 p) const {
        const int k = kind(p);
        if (k == kPrimTriangle) {
            const u32* ix = tris + 3 * (size_t)p;
            return (vert(ix[0]) + vert(ix[1]) + vert(ix[2])) / 3.0f;
        }
        if (k == kPrimQuad) {
            V3 a, b, c, d;
            quad(p - nt, a, b, c, d);
            return (a + b + c + d) / 4.0f;
        }
        const float4 s = spheres[p - nt - nq];
        return v3(s.x, s.y, s.z);
    }
    __device__ __forceinline__ bool masked(i32) const { return false; }
    static constexpr bool kRecords = true;
    static constexpr bool kStaged = false;
    __device__ __forceinline__ void record(i32 prim, int idx, float4* recs, V3& lo, V3& hi) const {
        const int k = kind(prim);
        float4* r = recs + 3 * (size_t)idx;
        if (k == kPrimTriangle) {
            const u32* ix = tris + 3 * (size_t)prim;
            const V3 a = vert(ix[0]), b = vert(ix[1]), c = vert(ix[2]);
            lo = vmin(vmin(a, b), c);
            hi = vmax(vmax(a, b), c);
            r[0] = make_float4(a.x, a.y, a.z, __int_as_float(prim));
            r[1] = make_float4(b.x, b.y, b.z, __int_as_float(kPrimTriangle));
            r[2] = make_float4(c.x, c.y, c.z, 0.0f);
        } else if (k == kPrimQuad) {
            V3 a, b, c, d;
            quad(prim - nt, a, b, c, d);
            lo = vmin(vmin(vmin(a, b), c), d);
            hi = vmax(vmax(vmax(a, b), c), d);
            r[0] = make_float4(a.x, a.y, a.z, __int_as_float(prim));
            r[1] = make_float4(b.x, b.y, b.z, __int_as_float(kPrimQuad));
            r[2] = make_float4(c.x, c.y, c.z, 0.0f);
            quadV3[idx] = make_float4(d.x, d.y, d.z, 0.0f);
        } else {
            const float4 s = spheres[prim - nt - nq];
            lo = v3(s.x - s.w, s.y - s.w, s.z - s.w);
            hi = v3(s.x + s.w, s.y + s.w, s.z + s.w);
            r[0] = make_float4(s.x, s.y, s.z, __int_as_float(prim));
            r[1] = make_float4(s.w, 0.0f, 0.0f, __int_as_float(kPrimSphere));
            r[2] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        }
    }
};

struct InstPrims {
    const float* worldToInst;  // 12 per instance
    const i32* instLod;
    const DevLod* lods;
    __device__ __forceinline__ void obox(i32 p, V3& lo, V3& hi) const {
        const float* b = lods[instLod[p]].box;
        lo = v3(b[0], b[1], b[2]);
        hi = v3(b[3], b[4], b[5]);
    }
    __device__ __forceinline__ V3 centroid(i32 p) const {
        V3 lo, hi;
        obox(p, lo, hi);
        const M34 M = inst_to_world(worldToInst + 12 * (size_t)p);
        return xform_point(M, (lo + hi) * 0.5f);
    }
    // Transform of the 8 corners, folded in corner order x-fastest.
    __device__ __forceinline__ void box(i32 p, V3& rlo, V3& rhi) const {
        V3 lo, hi;
        obox(p, lo, hi);
        const M34 M = inst_to_world(worldToInst + 12 * (size_t)p);
#pragma unroll
        for (int c = 0; c < 8; ++c) {
            const V3 q = xform_point(M, v3((c & 1) ? hi.x : lo.x, (c & 2) ? hi.y : lo.y, (c & 4) ? hi.z : lo.z));
            if (c == 0) {
                rlo = q;
                rhi = q;
            } else {
                rlo = vmin(rlo, q);
                rhi = vmax(rhi, q);
            }
        }
    }
    __device__ __forceinline__ bool masked(i32) const { return false; }
    __device__ __forceinline__ int kind(i32) const { return 0; }
    static constexpr bool kRecords = false;
    static constexpr bool kStaged = false;
    __device__ __forceinline__ void record(i32, int, float4*, V3&, V3&) const {}
};

// ------------------------------------------------------------- cost models
// Triangle LoD (build_lbvh.cu:34-40): Ct = 2, Cp = (1.2 tri, 4 quad, 1 sphere);
// counts packed 3 x 10 bits, each saturating at 1023.
struct LodCost {
    static constexpr float Ct = 2.0f;
    // count code of one primitive of kind k (get_count, build_lbvh.cu:285-295)
    __device__ static i32 leaf(int k) { return k == kPrimTriangle ? 1 << 20 : (k == kPrimQuad ? 1 << 10 : 1); }
    __device__ static float cost(i32 e) {
        const float x = (float)((e & 0x3FF00000) >> 20), y = (float)((e & 0x000FFC00) >> 10),
                    z = (float)(e & 0x3FF);
        return Ct + (x * 1.2f + y * 4.0f + z * 1.0f);
    }
    __device__ static i32 add(i32 a, i32 b) {
        const i32 x = min(((a & 0x3FF00000) >> 20) + ((b & 0x3FF00000) >> 20), 0x3FF);
        const i32 y = min(((a & 0x000FFC00) >> 10) + ((b & 0x000FFC00) >> 10), 0x3FF);
        const i32 z = min((a & 0x3FF) + (b & 0x3FF), 0x3FF);
        return (x << 20) | (y << 10) | z;
    }
    __device__ static bool capped(i32 e) {
        return ((e & 0x3FF00000) >> 20) >= 1023 || ((e & 0x000FFC00) >> 10) >= 1023 || (e & 0x3FF) >= 1023;
    }
    __device__ static i32 sum(i32 e) { return ((e & 0x3FF00000) >> 20) + ((e & 0x000FFC00) >> 10) + (e & 0x3FF); }
};
// Instance level (build_lbvh.cu:41-48): Ct = 1, Cp = 8.5, plain count.
struct InstCost {
    static constexpr float Ct = 1.0f;
    __device__ static i32 leaf(int) { return 1; }
    __device__ static float cost(i32 e) { return Ct + (float)e * 8.5f; }
    __device__ static i32 add(i32 a, i32 b) { return a + b; }
    __device__ static bool capped(i32 e) { return e >= 1023; }
    __device__ static i32 sum(i32 e) { return e; }
};

// ------------------------------------------------------------- kernels
template <class P>
__global__ __launch_bounds__(kThreads) void k_morton(P set, int n, Box6 bb, u64* __restrict__ keys,
                                                    i32* __restrict__ ids, float4* __restrict__ stage) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    V3 c;
    if constexpr (P::kStaged)
        c = stage ? set.centroid_staged(i, stage) : set.centroid(i);
    else
        c = set.centroid(i);
    keys[i] = set.masked(i) ? ~0ull : morton_of_point(c, bb);
    ids[i] = i;
}

__device__ __forceinline__ int lcp(const u64* keys, int n, int i1, int i2, u64 k1) {
    if (i2 < 0 || i2 >= n) return -1;
    const u64 k2 = keys[i2];
    if (k1 == k2) return 64 + __clz((int)(u32)(i1 ^ i2));
    return __clzll((long long)(k1 ^ k2));
}

// Refit levels. Level 0 is a tile of kRefitTile consecutive leaves; the
// spans of level L are kRefitTile * kSpanGrowth^L leaves. An internal node
// belongs to the lowest level L whose span grid holds its leaf range inside
// one span. A Karras node's index is one end of its range, so a node of level
// L lies in the span (index / span) of that level. Nodes of level 1 and above
// go to the straddle list as (level << 29) | node.
constexpr int kRefitThreads = 1024;
constexpr int kRefitIpt = 2;
constexpr int kRefitTile = kRefitThreads * kRefitIpt;
constexpr int kSpanGrowth = 64;
constexpr int kMaxPrims = 1 << 29;  // node indices fit the 29-bit field of a straddle entry

// Ordering of LDS accesses between the threads of a workgroup (no wait for
// global memory operations in flight).
__device__ __forceinline__ void lds_release() { __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local"); }
__device__ __forceinline__ void lds_acquire() { __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local"); }

constexpr int kTileShift = 11;  // log2 kRefitTile
constexpr int kGrowthShift = 6;  // log2 kSpanGrowth
static_assert((1 << kTileShift) == kRefitTile && (1 << kGrowthShift) == kSpanGrowth, "refit spans");

// Level of a node with leaf range [lo, hi] (indices below 2^29: at most 3).
__device__ __forceinline__ int refit_level(int lo, int hi) {
    int level = 0;
    for (int shift = kTileShift; (lo >> shift) != (hi >> shift); shift += kGrowthShift) ++level;
    return level;
}

__global__ __launch_bounds__(kThreads) void k_karras(const u64* __restrict__ keys, int n, i32* __restrict__ parents,
                                                    i32* __restrict__ children, uint8_t* __restrict__ levels) {
    const int idx = blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= n - 1) return;
    const u64 k1 = keys[idx];
    const int l1 = lcp(keys, n, idx, idx + 1, k1);
    const int l2 = lcp(keys, n, idx, idx - 1, k1);
    const int d = (l1 > l2) ? 1 : -1;  // never equal for distinct (key, index) pairs
    const int minLcp = lcp(keys, n, idx, idx - d, k1);
    int lMax = 2;
    while (lcp(keys, n, idx, idx + lMax * d, k1) > minLcp) lMax *= 2;
    int l = 0;
    for (int t = lMax; t > 1;) {
        t >>= 1;
        if (lcp(keys, n, idx, idx + (l + t) * d, k1) > minLcp) l += t;
    }
    const int j = idx + l * d;
    const int nodeLcp = lcp(keys, n, idx, j, k1);
    int s = 0;
    const int maxDiv = 1 << (32 - __clz(l));
    for (int div = 2; div <= maxDiv; div <<= 1) {
        const int t = (l + div - 1) / div;
        if (lcp(keys, n, idx, idx + (s + t) * d, k1) > nodeLcp) s += t;
    }
    const int split = idx + s * d + min(d, 0);
    const int left = (min(idx, j) == split) ? split + n - 1 : split;
    const int right = (max(idx, j) == split + 1) ? split + n : split + 1;
    parents[left] = ~idx;
    parents[right] = idx;
    children[2 * idx] = left;
    children[2 * idx + 1] = right;
    levels[idx] = (uint8_t)refit_level(min(idx, j), max(idx, j));
    if (idx == 0) parents[0] = kTreeHead;
}

__device__ __forceinline__ int parent_of(const i32* parents, int node) {
    const int p = parents[node];
    return p < 0 ? ~p : p;
}

// Merge of an internal node from its two finished children's boxes, in the
// fixed operand order (right, left): box, count, cost and the collapse flags.
template <class C>
__device__ __forceinline__ void merge_boxes(float4 lLo, float4 lHi, float4 rLo, float4 rHi, float4& oLo, float4& oHi,
                                            uint8_t& flag) {
    const V3 Llo = v3(lLo.x, lLo.y, lLo.z), Lhi = v3(lHi.x, lHi.y, lHi.z);
    const V3 Rlo = v3(rLo.x, rLo.y, rLo.z), Rhi = v3(rHi.x, rHi.y, rHi.z);
    const V3 nlo = vmin(Rlo, Llo), nhi = vmax(Rhi, Lhi);
    const float sa = surface(nlo, nhi);
    const float saR = surface(Rlo, Rhi), saL = surface(Llo, Lhi);
    const i32 cR = __float_as_int(rLo.w), cL = __float_as_int(lLo.w);
    const float cost = C::Ct + saR / sa * C::cost(cR) + saL / sa * C::cost(cL);
    const i32 c = C::add(cR, cL);
    oLo = make_float4(nlo.x, nlo.y, nlo.z, __int_as_float(c));
    oHi = make_float4(nhi.x, nhi.y, nhi.z, cost);
    flag = (uint8_t)((C::capped(c) ? 1 : 0) | (C::cost(c) < cost ? 2 : 0));
}

// Merge of internal node `node` (children L, R) with the boxes in global memory.
template <class C>
__device__ __forceinline__ void merge_node(int node, int L, int R, float4* boxes, uint8_t* flags) {
    float4 oLo, oHi;
    uint8_t f;
    merge_boxes<C>(boxes[2 * (size_t)L], boxes[2 * (size_t)L + 1], boxes[2 * (size_t)R], boxes[2 * (size_t)R + 1],
                   oLo, oHi, f);
    boxes[2 * (size_t)node] = oLo;
    boxes[2 * (size_t)node + 1] = oHi;
    flags[node] = f;
}

// Refit, level 0. Block b owns the leaves [b*kRefitTile, (b+1)*kRefitTile).
// Every level-0 node of the tile has its index inside the tile and both its
// subtrees inside it, so the whole climb runs in LDS: the tile's topology
// (tile-local 16-bit parent and child references), the leaf and internal
// boxes and byte-sized arrival counters. The climb goes in rounds: a node
// whose second child finishes is queued (LDS) for the next round, and each
// round merges its queue one node per thread between two barriers, so the
// waves work on packed queues instead of on divergent per-leaf climbs.
// Boxes and flags go to global memory with coalesced stores at the end.
// Nodes of level 1 and above are merged by the level kernels; the block
// lists the tile's such nodes in the straddle list (one global atomic).
constexpr u32 kLeafRef = 0x8000u;
constexpr short kNoParent = -1;
constexpr uint8_t kHigher = 0x80;  // tile byte: node of level 1 or above (low bits: collapse flags)

template <class P, class C>
__global__ __launch_bounds__(kRefitThreads) void k_refit_tiles(P set, int n, const i32* __restrict__ ids,
                                                              const i32* __restrict__ parents,
                                                              const i32* __restrict__ children,
                                                              const uint8_t* __restrict__ levels, float4* boxes,
                                                              uint8_t* flags, float4* triVerts,
                                                              const float4* __restrict__ stage,
                                                              i32* __restrict__ primIds, u32* straddle,
                                                              u32* straddleCount) {
    __shared__ float4 sLeafBox[kRefitTile][2];
    __shared__ float4 sIntBox[kRefitTile][2];
    __shared__ u32 sChild[kRefitTile];  // two 16-bit refs: local internal index, or kLeafRef | local leaf
    __shared__ short sParentInt[kRefitTile];
    __shared__ short sParentLeaf[kRefitTile];
    __shared__ uint8_t sNodeByte[kRefitTile];  // kHigher, or the collapse flags of a merged node
    __shared__ u32 sArrived[kRefitTile / 4];
    __shared__ short sQueue[2][kRefitTile / 2];
    __shared__ u32 sQCount[2];
    __shared__ u32 sListCount, sListBase;
    const int nInt = n - 1;
    const int first = blockIdx.x * kRefitTile;
    const int last = min(first + kRefitTile, n) - 1;
    auto local_parent = [&](i32 p) -> short {
        const int q = p < 0 ? ~p : p;
        return (q >= first && q <= last) ? (short)(q - first) : kNoParent;
    };
    auto child_ref = [&](int c) -> u32 {
        return c >= nInt ? (kLeafRef | (u32)((c - nInt - first) & 0x7FF)) : (u32)((c - first) & 0x7FF);
    };
    if (threadIdx.x == 0) {
        sListCount = 0u;
        sQCount[0] = 0u;
        sQCount[1] = 0u;
    }
    for (int k = threadIdx.x; k < kRefitTile / 4; k += kRefitThreads) sArrived[k] = 0u;
    u32 listSlot[kRefitIpt];
#pragma unroll
    for (int r = 0; r < kRefitIpt; ++r) {
        const int k = r * kRefitThreads + (int)threadIdx.x;
        const int node = first + k;
        listSlot[r] = ~0u;
        if (node < nInt) {
            sChild[k] = child_ref(children[2 * node]) | (child_ref(children[2 * node + 1]) << 16);
            sParentInt[k] = node == 0 ? kNoParent : local_parent(parents[node]);
            sNodeByte[k] = levels[node] != 0 ? kHigher : 0;
        }
        if (node <= last && n > 1) sParentLeaf[k] = local_parent(parents[node + nInt]);
    }
    __syncthreads();
    // this tile's nodes of higher levels into the straddle list
#pragma unroll
    for (int r = 0; r < kRefitIpt; ++r) {
        const int k = r * kRefitThreads + (int)threadIdx.x;
        if (first + k < nInt && sNodeByte[k] == kHigher) listSlot[r] = atomicAdd(&sListCount, 1u);
    }
    __syncthreads();
    if (threadIdx.x == 0 && sListCount) sListBase = atomicAdd(straddleCount, sListCount);
    int leafIdx[kRefitIpt];
#pragma unroll
    for (int k = 0; k < kRefitIpt; ++k) {
        const int idx = first + k * kRefitThreads + (int)threadIdx.x;
        leafIdx[k] = idx;
        if (idx > last) continue;
        const int prim = ids[idx];
        primIds[idx] = prim;
        V3 lo, hi;
        if constexpr (P::kStaged) {
            if (stage)
                set.record_staged(prim, idx, stage, triVerts, lo, hi);
            else
                set.record(prim, idx, triVerts, lo, hi);
        } else if constexpr (P::kRecords) {
            set.record(prim, idx, triVerts, lo, hi);  // LoD: also lay the primitive out in BVH order
        } else {
            set.box(prim, lo, hi);
        }
        const i32 cnt = C::leaf(set.kind(prim));
        sLeafBox[idx - first][0] = make_float4(lo.x, lo.y, lo.z, __int_as_float(cnt));
        sLeafBox[idx - first][1] = make_float4(hi.x, hi.y, hi.z, C::cost(cnt));
    }
#pragma unroll
    for (int r = 0; r < kRefitIpt; ++r)
        if (listSlot[r] != ~0u) {
            const int node = first + r * kRefitThreads + (int)threadIdx.x;
            straddle[sListBase + listSlot[r]] = ((u32)levels[node] << 29) | (u32)node;
        }
    // arrivals of the leaves: the second one queues the parent
    auto arrive = [&](short p, int q) {
        if (p == kNoParent || sNodeByte[p] == kHigher) return;  // a higher level merges it
        const u32 sh = 8u * (u32)(p & 3);
        if (((atomicAdd(&sArrived[p >> 2], 1u << sh) >> sh) & 0xFFu) == 1u) sQueue[q][atomicAdd(&sQCount[q], 1u)] = p;
    };
    if (n > 1) {
#pragma unroll
        for (int k = 0; k < kRefitIpt; ++k)
            if (leafIdx[k] <= last) arrive(sParentLeaf[leafIdx[k] - first], 0);
    }
    __syncthreads();
    int cur = 0;
    for (int round = 0; round < kMaxDepth; ++round) {
        const u32 items = sQCount[cur];  // at most kRefitTile / 2: each queued node took two children
        if (items == 0) break;
        if (threadIdx.x < items) {
            const short p = sQueue[cur][threadIdx.x];
            const u32 ch = sChild[p];
            const u32 a = ch & 0xFFFFu, b = ch >> 16;
            const float4* lb = (a & kLeafRef) ? sLeafBox[a & 0x7FF] : sIntBox[a];
            const float4* rb = (b & kLeafRef) ? sLeafBox[b & 0x7FF] : sIntBox[b];
            float4 oLo, oHi;
            uint8_t f;
            merge_boxes<C>(lb[0], lb[1], rb[0], rb[1], oLo, oHi, f);
            sIntBox[p][0] = oLo;
            sIntBox[p][1] = oHi;
            sNodeByte[p] = f;
            arrive(sParentInt[p], cur ^ 1);
        }
        __syncthreads();
        if (threadIdx.x == 0) sQCount[cur] = 0u;
        __syncthreads();
        cur ^= 1;
    }
    for (int k = threadIdx.x; k < kRefitTile; k += kRefitThreads) {
        const int leaf = first + k;
        if (leaf <= last) {
            boxes[2 * (size_t)(leaf + nInt)] = sLeafBox[k][0];
            boxes[2 * (size_t)(leaf + nInt) + 1] = sLeafBox[k][1];
        }
        if (leaf < nInt && sNodeByte[k] != kHigher) {
            boxes[2 * (size_t)leaf] = sIntBox[k][0];
            boxes[2 * (size_t)leaf + 1] = sIntBox[k][1];
            flags[leaf] = sNodeByte[k];
        }
    }
}

// Refit, level L >= 1. Block b merges the level-L nodes of its span (node
// index >> shift == b): they are gathered from the straddle list into LDS
// slots with a node -> slot hash. A child or parent of such a node that is
// in the hash is of the same level (ranges nest and levels do not decrease
// towards the root, so every same-level relative is in this span's set);
// every other child is finished (lower level or leaf) and has its box staged
// in LDS. The climb goes in rounds over an LDS queue, as in the tile kernel
// (arrival counters start at the number of finished children; the arrival
// that makes a count 2 queues the node). Boxes and flags go to global memory
// at the end. A span with more than `cap` such nodes climbs through global
// memory instead, with its own nodes' counters in `counters`
// (workgroup-scope ordering).
constexpr int kLevelThreads = 512;
constexpr int kLevelCap = 1024;
constexpr int kLevelHash = 2048;  // power of two, >= 2 * kLevelCap

__device__ __forceinline__ u32 level_hash(int node) { return ((u32)node * 2654435761u) & (kLevelHash - 1); }

template <class C>
__global__ __launch_bounds__(kLevelThreads) void k_refit_level(int n, int level, int shift, int cap,
                                                              const i32* __restrict__ parents,
                                                              const i32* __restrict__ children,
                                                              const uint8_t* __restrict__ levels,
                                                              const u32* __restrict__ straddle,
                                                              const u32* __restrict__ straddleCount, float4* boxes,
                                                              uint8_t* flags, u32* counters) {
    __shared__ i32 sNode[kLevelCap];
    __shared__ short sParentSlot[kLevelCap];
    __shared__ int2 sChild[kLevelCap];  // >= 0: slot of a same-level child; < 0: ~index into sFinal
    __shared__ u32 sCnt[kLevelCap];
    __shared__ uint8_t sFlag[kLevelCap];
    __shared__ float4 sBox[kLevelCap][2];
    __shared__ float4 sFinal[2 * kLevelCap][2];
    __shared__ i32 sHashKey[kLevelHash];
    __shared__ short sHashVal[kLevelHash];
    __shared__ short sQueue[2][kLevelCap];
    __shared__ u32 sQCount[2];
    __shared__ u32 sCount, sFinalCount;
    const int nInt = n - 1;
    const int listCount = (int)*straddleCount;
    if (threadIdx.x == 0) {
        sCount = 0u;
        sFinalCount = 0u;
        sQCount[0] = 0u;
        sQCount[1] = 0u;
    }
    for (int k = threadIdx.x; k < kLevelHash; k += kLevelThreads) sHashKey[k] = -1;
    __syncthreads();
    auto mine = [&](u32 e) { return (int)(e >> 29) == level && (int)((e & 0x1FFFFFFFu) >> shift) == (int)blockIdx.x; };
    // the list in batches of loads issued together (an entry of all ones is
    // level 7: no level's)
    constexpr int kBatch = 8;
    for (int base = 0; base < listCount; base += kBatch * kLevelThreads) {
        u32 es[kBatch];
#pragma unroll
        for (int b = 0; b < kBatch; ++b) {
            const int k = base + b * kLevelThreads + (int)threadIdx.x;
            es[b] = k < listCount ? straddle[k] : ~0u;
        }
#pragma unroll
        for (int b = 0; b < kBatch; ++b) {
            const u32 e = es[b];
            if (!mine(e)) continue;
            const int node = (int)(e & 0x1FFFFFFFu);
            const u32 slot = atomicAdd(&sCount, 1u);
            if ((int)slot >= cap) continue;
            sNode[slot] = node;
            for (u32 h = level_hash(node);; h = (h + 1) & (kLevelHash - 1))
                if (atomicCAS(&sHashKey[h], -1, node) == -1) {
                    sHashVal[h] = (short)slot;
                    break;
                }
        }
    }
    __syncthreads();
    const int count = (int)sCount;
    if (count == 0) return;
    if (count > cap) {
        // global-memory climb over this span's nodes
        auto lower = [&](int c) { return c >= nInt || (int)levels[c] < level; };  // finished before this level
        for (int k = threadIdx.x; k < listCount; k += kLevelThreads) {
            const u32 e = straddle[k];
            if (!mine(e)) continue;
            const int node = (int)(e & 0x1FFFFFFFu);
            counters[node] = (lower(children[2 * node]) ? 1u : 0u) + (lower(children[2 * node + 1]) ? 1u : 0u);
        }
        __threadfence_block();
        __syncthreads();
        for (int k = threadIdx.x; k < listCount; k += kLevelThreads) {
            const u32 e = straddle[k];
            if (!mine(e)) continue;
            int node = (int)(e & 0x1FFFFFFFu);
            // a start node has both children finished before this level (not
            // a counter value, which arrivals of other climbs raise meanwhile)
            if (!lower(children[2 * node]) || !lower(children[2 * node + 1])) continue;
            for (int depth = 0; depth < kMaxDepth; ++depth) {
                merge_node<C>(node, children[2 * node], children[2 * node + 1], boxes, flags);
                if (node == 0) break;
                const int p = parent_of(parents, node);
                if ((int)levels[p] != level) break;
                __threadfence_block();
                if (atomicAdd(&counters[p], 1u) != 1u) break;
                __threadfence_block();
                node = p;
            }
        }
        return;
    }
    // slot of a node of this span's set, -1 for any other node (the probe
    // ends at the first empty entry: the table is at most half full)
    auto slot_of = [&](int node) -> int {
        u32 h = level_hash(node);
        for (int probe = 0; probe < kLevelHash; ++probe, h = (h + 1) & (kLevelHash - 1)) {
            const int k = sHashKey[h];
            if (k == node) return sHashVal[h];
            if (k == -1) return -1;
        }
        return -1;
    };
    for (int s = threadIdx.x; s < count; s += kLevelThreads) {
        const int node = sNode[s];
        const int2 ch = make_int2(children[2 * node], children[2 * node + 1]);
        const int p = node != 0 ? parent_of(parents, node) : -1;
        const int sl = ch.x >= nInt ? -1 : slot_of(ch.x);
        const int sr = ch.y >= nInt ? -1 : slot_of(ch.y);
        int refs[2] = {sl, sr};
        const int cs[2] = {ch.x, ch.y};
        u32 fin = 0;
#pragma unroll
        for (int q = 0; q < 2; ++q)
            if (refs[q] < 0) {
                const u32 f = atomicAdd(&sFinalCount, 1u);
                sFinal[f][0] = boxes[2 * (size_t)cs[q]];
                sFinal[f][1] = boxes[2 * (size_t)cs[q] + 1];
                refs[q] = ~(int)f;
                ++fin;
            }
        sChild[s] = make_int2(refs[0], refs[1]);
        sCnt[s] = fin;
        sParentSlot[s] = (short)(p >= 0 ? slot_of(p) : -1);
        if (fin == 2u) sQueue[0][atomicAdd(&sQCount[0], 1u)] = (short)s;
    }
    __syncthreads();
    int cur = 0;
    for (int round = 0; round < kMaxDepth; ++round) {
        const int items = (int)sQCount[cur];
        if (items == 0) break;
        for (int i = threadIdx.x; i < items; i += kLevelThreads) {
            const int s = sQueue[cur][i];
            const int2 r = sChild[s];
            const float4* lb = r.x >= 0 ? sBox[r.x] : sFinal[~r.x];
            const float4* rb = r.y >= 0 ? sBox[r.y] : sFinal[~r.y];
            float4 oLo, oHi;
            uint8_t f;
            merge_boxes<C>(lb[0], lb[1], rb[0], rb[1], oLo, oHi, f);
            sBox[s][0] = oLo;
            sBox[s][1] = oHi;
            sFlag[s] = f;
            const int ps = sParentSlot[s];
            if (ps >= 0 && atomicAdd(&sCnt[ps], 1u) == 1u) sQueue[cur ^ 1][atomicAdd(&sQCount[cur ^ 1], 1u)] = (short)ps;
        }
        __syncthreads();
        if (threadIdx.x == 0) sQCount[cur] = 0u;
        __syncthreads();
        cur ^= 1;
    }
    for (int s = threadIdx.x; s < count; s += kLevelThreads) {
        const int node = sNode[s];
        boxes[2 * (size_t)node] = sBox[s][0];
        boxes[2 * (size_t)node + 1] = sBox[s][1];
        flags[node] = sFlag[s];
    }
}

// Keep marks of the reference's collapse (marks) and, with marks2, of the
// traversal hierarchy: a node is also kept there when it holds more than
// refineAbove primitives, so no traversal leaf is longer than that. Both mark
// sets are closed under descendants (a removed node's subtree is removed).
// *differ is set when the two sets differ.
template <class C>
__global__ __launch_bounds__(kThreads) void k_mark(int nInt, const i32* __restrict__ parents,
                                                  const uint8_t* __restrict__ flags, i32* __restrict__ marks,
                                                  const float4* __restrict__ boxes, int refineAbove,
                                                  i32* __restrict__ marks2, u32* __restrict__ differ) {
    const int x = blockIdx.x * blockDim.x + threadIdx.x;
    if (x >= nInt) return;
    int keep = 1;
    int y = x;
    for (int depth = 0; y != 0 && depth < kMaxDepth; ++depth) {
        const uint8_t f = flags[y];
        if (f & 1) break;
        if (f & 2) {
            keep = 0;
            break;
        }
        y = parent_of(parents, y);
    }
    marks[x] = keep;
    if (marks2) {
        const int keep2 = (keep || C::sum(__float_as_int(boxes[2 * (size_t)x].w)) > refineAbove) ? 1 : 0;
        marks2[x] = keep2;
        if (keep2 != keep) *differ = 1u;
    }
}

__device__ __forceinline__ bool is_collapsed(const i32* offsets, int nInt, int node) {
    return node > 0 && node < nInt && (offsets[node] - offsets[node - 1]) == 0;
}

// The collapsed BVH's record of node `node` (its box enlarged, its index in
// the collapsed numbering) under the keep scan `offsets`.
template <class C>
__device__ __forceinline__ void collapse_node(int node, int n, const i32* __restrict__ parents,
                                              const float4* __restrict__ boxes, const i32* __restrict__ offsets,
                                              BvhChild* __restrict__ out) {
    const int nInt = n - 1;
    if (is_collapsed(offsets, nInt, node)) return;
    const int nAfter = offsets[nInt - 1];
    int parent = parents[node];
    int off = parent < 0 ? 0 : 1;
    if (parent < 0) parent = ~parent;
    int fin = node;
    if (fin >= nInt) {
        for (int depth = 0; depth < kMaxDepth && is_collapsed(offsets, nInt, parent); ++depth) {
            if (off == 1) return;  // only the left-most leaf writes the collapsed subtree
            fin = parent;
            parent = parents[parent];
            off = parent < 0 ? 0 : 1;
            if (parent < 0) parent = ~parent;
        }
    }
    const float4 a = boxes[2 * (size_t)fin], b = boxes[2 * (size_t)fin + 1];
    const i32 primCount = C::sum(__float_as_int(a.w));
    const int outIdx = (offsets[parent] - 1) * 2;
    const int outNode = (node >= nInt) ? node - nInt + nAfter : offsets[node] - 1;
    const V3 lo = v3(a.x, a.y, a.z), hi = v3(b.x, b.y, b.z);
    const V3 c = (lo + hi) * 0.5f;
    const V3 bl = (lo - c) * 1.0001f + c;
    const V3 bh = (hi - c) * 1.0001f + c;
    BvhChild r;
    r.lo[0] = bl.x; r.lo[1] = bl.y; r.lo[2] = bl.z;
    r.hi[0] = bh.x; r.hi[1] = bh.y; r.hi[2] = bh.z;
    r.index = outNode;
    r.primCount = primCount;
    out[outIdx + off] = r;
}

// The reference's collapsed BVH and, with offsets2, the traversal hierarchy
// (written only when *differ is set) in one pass over the nodes.
template <class C>
__global__ __launch_bounds__(kThreads) void k_collapse(int n, const i32* __restrict__ parents,
                                                      const float4* __restrict__ boxes,
                                                      const i32* __restrict__ offsets, BvhChild* __restrict__ out,
                                                      const i32* __restrict__ offsets2, BvhChild* __restrict__ out2,
                                                      const u32* __restrict__ differ) {
    const int node = blockIdx.x * blockDim.x + threadIdx.x + 1;
    if (node >= 2 * n - 1) return;
    collapse_node<C>(node, n, parents, boxes, offsets, out);
    if (offsets2 && *differ != 0u) collapse_node<C>(node, n, parents, boxes, offsets2, out2);
}

// Zero fill and count hand-off of the build (kernels, not memsets, so that a
// build can sit in a captured stream).
__global__ void k_fill_u32(u32* p, int n, u32 v) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}
__global__ void k_num_internal(const i32* offsets, int n, i32* out) {
    if (threadIdx.x == 0) *out = n > 1 ? offsets[n - 2] : 0;
}
// Internal node counts of the reference BVH and of the traversal hierarchy,
// and whether the two differ, for one copy to the host.
__global__ void k_build_counts(const i32* offsets, const i32* offsets2, const u32* differ, int n, i32* out) {
    if (threadIdx.x == 0) {
        out[0] = offsets[n - 2];
        out[1] = offsets2 ? offsets2[n - 2] : 0;
        out[2] = differ ? (i32)*differ : 0;
    }
}

// World box of each instance (feeds the scene box of the instance level).
__global__ __launch_bounds__(kThreads) void k_inst_boxes(InstPrims set, int n, float* out) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    V3 lo, hi;
    set.box(i, lo, hi);
    float* o = out + 6 * (size_t)i;
    o[0] = lo.x; o[1] = lo.y; o[2] = lo.z;
    o[3] = hi.x; o[4] = hi.y; o[5] = hi.z;
}

inline int blocks_for(long long n) { return (int)((n + kThreads - 1) / kThreads); }

template <class T>
void d2h(hipStream_t s, T* dst, const void* src, size_t count) {
    if (!dst || count == 0) return;
    MFH_HIP(hipMemcpyAsync(dst, src, sizeof(T) * count, hipMemcpyDeviceToHost, s));
    MFH_HIP(hipStreamSynchronize(s));
}

// LDS slots of a refit level block (MFH_LBVH_LEVEL_CAPACITY lowers it, to
// exercise the global-memory climb of fuller spans).
int level_capacity() {
    const char* e = std::getenv("MFH_LBVH_LEVEL_CAPACITY");
    if (!e || !*e) return kLevelCap;
    const int v = std::atoi(e);
    return v < 0 ? 0 : (v > kLevelCap ? kLevelCap : v);
}

void fill_u32(void* p, int n, u32 v, hipStream_t st) {
    hipLaunchKernelGGL(k_fill_u32, dim3(blocks_for(n)), dim3(kThreads), 0, st, static_cast<u32*>(p), n, v);
    MFH_HIP(hipGetLastError());
}

// The traversal hierarchy of a build: the reference's collapsed BVH with
// every collapsed subtree of more than refineAbove primitives kept split
// (leaves of at most refineAbove primitives), over the same sorted primitive
// records. Closest and any hit do not depend on the hierarchy they are found
// through (the closest hit is the lowest (t, instance, primitive)), so the
// traversal kernels use it while the exported BVH stays the reference's.
struct Refinement {
    int refineAbove = 0;  // 0: none
    DevBuf* nodes = nullptr;
    int numInternal = -1;  // -1: equal to the reference (no subtree was split)
};

// Shared orchestration (build_lbvh.cu:586-780). Returns the number of
// internal nodes after the collapse; primIds and nodes receive the result.
// With dNumInternal the count goes to that device int instead and nothing
// waits for the stream (the return value is then -1).
template <class P, class C>
int build_generic(Context& ctx, BuildScratch& s, const P& set, int n, const float* bbox, DevBuf& nodes,
                  DevBuf& primIds, float4* triVerts, StageSink* sink, i32* dNumInternal = nullptr,
                  Refinement* refine = nullptr, float4* stage = nullptr) {
    hipStream_t st = ctx.stream;
    MFH_REQUIRE(n < kMaxPrims, "an LBVH build holds fewer than 2^29 primitives");
    s.keysA.reserve(8 * (size_t)n);
    s.keysB.reserve(8 * (size_t)n);
    s.idsA.reserve(4 * (size_t)n);
    s.idsB.reserve(4 * (size_t)n);
    s.parents.reserve(4 * (size_t)(2 * n - 1));
    s.children.reserve(8 * (size_t)std::max(n - 1, 1));
    s.levels.reserve((size_t)std::max(n - 1, 1));
    s.straddle.reserve(4 * (size_t)n + 4);
    s.boxes.reserve(32 * (size_t)(2 * n - 1));
    s.counters.reserve(4 * (size_t)std::max(n - 1, 1));
    s.flags.reserve((size_t)std::max(n - 1, 1));
    s.offsets.reserve(4 * (size_t)std::max(n - 1, 1));
    primIds.reserve(4 * (size_t)n);
    Box6 bb;
    std::memcpy(bb.v, bbox, sizeof(bb.v));

    u64* keys = s.keysA.as<u64>();
    i32* ids = s.idsA.as<i32>();
    hipLaunchKernelGGL(HIP_KERNEL_NAME(k_morton<P>), dim3(blocks_for(n)), dim3(kThreads), 0, st, set, n, bb, keys,
                       ids, stage);
    MFH_HIP(hipGetLastError());
    if (sink) d2h(st, sink->unsortedKeys, keys, n);
    radix_sort_pairs(keys, s.keysB.as<u64>(), ids, s.idsB.as<i32>(), n, 0, 64, s, st);
    if (sink) {
        d2h(st, sink->sortedKeys, keys, n);
        d2h(st, sink->sortedIds, ids, n);
    }
    i32* parents = s.parents.as<i32>();
    i32* children = s.children.as<i32>();
    float4* boxes = s.boxes.as<float4>();
    uint8_t* levels = s.levels.as<uint8_t>();
    u32* straddle = s.straddle.as<u32>();
    u32* straddleCount = straddle + n;
    fill_u32(straddleCount, 1, 0u, st);
    if (n > 1)
        hipLaunchKernelGGL(k_karras, dim3(blocks_for(n - 1)), dim3(kThreads), 0, st, keys, n, parents, children,
                           levels);
    hipLaunchKernelGGL(HIP_KERNEL_NAME(k_refit_tiles<P, C>), dim3((n + kRefitTile - 1) / kRefitTile),
                       dim3(kRefitThreads), 0, st, set, n, ids, parents, children, levels, boxes,
                       s.flags.as<uint8_t>(), triVerts, stage, primIds.as<i32>(), straddle, straddleCount);
    MFH_HIP(hipGetLastError());
    const int cap = level_capacity();
    int level = 0;
    for (int shift = kTileShift; n > 1 && (1ll << shift) < n;) {
        shift += kGrowthShift;
        ++level;
        const int blocks = (int)((n + (1ll << shift) - 1) >> shift);
        hipLaunchKernelGGL(HIP_KERNEL_NAME(k_refit_level<C>), dim3(blocks), dim3(kLevelThreads), 0, st, n, level,
                           shift, cap, parents, children, levels, straddle, straddleCount, boxes,
                           s.flags.as<uint8_t>(), s.counters.as<u32>());
        MFH_HIP(hipGetLastError());
    }
    if (n == 1) {
        if (dNumInternal) {
            hipLaunchKernelGGL(k_num_internal, dim3(1), dim3(64), 0, st, nullptr, 1, dNumInternal);
            MFH_HIP(hipGetLastError());
            return -1;
        }
        MFH_HIP(hipStreamSynchronize(st));
        return 0;
    }
    i32* offsets = s.offsets.as<i32>();
    const bool refining = refine && refine->refineAbove > 0 && !dNumInternal;
    i32* offsets2 = nullptr;
    u32* differ = nullptr;
    if (refining) {
        s.offsets2.reserve(4 * (size_t)(n - 1) + 4);
        offsets2 = s.offsets2.as<i32>();
        differ = reinterpret_cast<u32*>(offsets2 + (n - 1));
        fill_u32(differ, 1, 0u, st);
    }
    hipLaunchKernelGGL(HIP_KERNEL_NAME(k_mark<C>), dim3(blocks_for(n - 1)), dim3(kThreads), 0, st, n - 1, parents,
                       s.flags.as<uint8_t>(), offsets, boxes, refining ? refine->refineAbove : 0, offsets2, differ);
    // the reference's BVH and the traversal hierarchy (collapsed whether or
    // not it differs, so that the build waits for the device once; its
    // records are written only when it differs)
    if (refining)
        scan_inclusive_i32_pair(offsets, offsets2, n - 1, s, st);
    else
        scan_inclusive_i32(offsets, offsets, n - 1, s, st);
    nodes.reserve(64 * (size_t)(n - 1));
    if (refining) refine->nodes->reserve(64 * (size_t)(n - 1));
    hipLaunchKernelGGL(HIP_KERNEL_NAME(k_collapse<C>), dim3(blocks_for(2 * n - 2)), dim3(kThreads), 0, st, n,
                       parents, boxes, offsets, nodes.as<BvhChild>(), offsets2,
                       refining ? refine->nodes->as<BvhChild>() : nullptr, differ);
    MFH_HIP(hipGetLastError());
    if (dNumInternal) {
        hipLaunchKernelGGL(k_num_internal, dim3(1), dim3(64), 0, st, offsets, n, dNumInternal);
        MFH_HIP(hipGetLastError());
        return -1;
    }
    s.misc.reserve(16);
    i32* dCounts = s.misc.as<i32>();
    hipLaunchKernelGGL(k_build_counts, dim3(1), dim3(64), 0, st, offsets, refining ? offsets2 : nullptr, differ, n,
                       dCounts);
    MFH_HIP(hipGetLastError());
    i32 counts[3] = {0, 0, 0};
    MFH_HIP(hipMemcpyAsync(counts, dCounts, sizeof(counts), hipMemcpyDeviceToHost, st));
    MFH_HIP(hipStreamSynchronize(st));
    const i32 nAfter = counts[0];
    if (refining) refine->numInternal = counts[2] ? counts[1] : -1;
    if (sink) {
        d2h(st, sink->parents, parents, 2 * (size_t)n - 1);
        d2h(st, sink->boxes, boxes, 8 * (size_t)(2 * n - 1));
        d2h(st, sink->offsets, offsets, (size_t)n - 1);
    }
    return nAfter;
}

}  // namespace

void build_lod_bvh(Context& ctx, Lod& lod, const float* bbox, StageSink* sink) {
    const int n = lod.numPrims();
    MFH_REQUIRE(n >= 1, "LoD has no primitives");
    if (bbox) std::memcpy(lod.box, bbox, sizeof(lod.box));
    lod.triVerts.reserve(48 * (size_t)n);
    Refinement refine;
    refine.refineAbove = ctx.traversalLeafLimit;
    refine.nodes = &lod.travNodes;
    if (lod.mixed()) {
        if (lod.numQuads > 0) lod.quadV3.reserve(16 * (size_t)n);
        MixedPrims set{lod.verts.as<float>(), lod.tris.as<u32>(),  lod.quads.as<u32>(), lod.spheres.as<float4>(),
                       lod.numQuads > 0 ? lod.quadV3.as<float4>() : nullptr, lod.numTris, lod.numQuads};
        lod.numInternal = build_generic<MixedPrims, LodCost>(ctx, ctx.scratch, set, n, lod.box, lod.nodes,
                                                             lod.primIds, lod.triVerts.as<float4>(), sink, nullptr,
                                                             &refine);
    } else {
        TriPrims set{lod.verts.as<float>(), lod.tris.as<u32>()};
        ctx.scratch.stage.reserve(48 * (size_t)n);
        lod.numInternal = build_generic<TriPrims, LodCost>(ctx, ctx.scratch, set, n, lod.box, lod.nodes,
                                                           lod.primIds, lod.triVerts.as<float4>(), sink, nullptr,
                                                           &refine, ctx.scratch.stage.as<float4>());
    }
    lod.travNumInternal = refine.numInternal;
    lod.built = true;
}

void build_scene_bvh(Context& ctx, Scene& sc, StageSink* sink) {
    const int n = sc.numInstances;
    MFH_REQUIRE(n >= 1, "scene has no instances");
    InstPrims set{sc.dWorldToInst.as<float>(), sc.dInstLod.as<i32>(), sc.dLods.as<DevLod>()};
    // Scene box: union of the instance world boxes (exact min/max).
    DevBuf& tmp = ctx.scratch.misc;
    tmp.reserve(24 * (size_t)n);
    hipLaunchKernelGGL(k_inst_boxes, dim3(blocks_for(n)), dim3(kThreads), 0, ctx.stream, set, n, tmp.as<float>());
    MFH_HIP(hipGetLastError());
    std::vector<float> ib(6 * (size_t)n);
    d2h(ctx.stream, ib.data(), tmp.ptr, ib.size());
    for (int k = 0; k < 3; ++k) {
        float lo = ib[k], hi = ib[3 + k];
        for (int i = 1; i < n; ++i) {
            const float a = ib[6 * (size_t)i + k], b = ib[6 * (size_t)i + 3 + k];
            lo = (a < lo) ? a : lo;
            hi = (hi < b) ? b : hi;
        }
        sc.box[k] = lo;
        sc.box[3 + k] = hi;
    }
    if (n == 1) {
        sc.tlasNumInternal = 0;
        return;
    }
    sc.tlasNumInternal = build_generic<InstPrims, InstCost>(ctx, ctx.scratch, set, n, sc.box, sc.tlasNodes,
                                                            sc.tlasPrimIds, nullptr, sink);
}

// ------------------------------------------------- build without allocation
// The LoD build over caller-owned device buffers with its scratch in one
// static device array per GPU (no hipMalloc / hipFree anywhere on the path).
// Every build on a GPU shares that GPU's array and leaves its contents behind;
// what must start at zero (arrival counters, queue counts, the index check
// flag) is cleared by fill kernels inside the build. Builds that overlap on
// one GPU (two host threads, or two contexts whose streams run at once) race.

namespace {

constexpr size_t align256(size_t b) { return (b + 255) & ~(size_t)255; }

// Bytes of every scratch buffer of a build over n primitives, in carve order.
constexpr int kNoAllocParts = 17;
struct NoAllocSizes {
    size_t b[kNoAllocParts];
};
constexpr NoAllocSizes noalloc_sizes(long long n) {
    const long long n1 = n > 1 ? n - 1 : 1;
    const long long tiles = (n + kRadixTileKeys - 1) / kRadixTileKeys;  // radix sort tiles
    const long long scanned = n > 256 * tiles ? n : 256 * tiles;        // longest scan: marks or digit counts
    const long long scanBlocks = (scanned + 2047) / 2048;               // scan chunks of 2048 values
    return NoAllocSizes{{(size_t)(8 * n), (size_t)(8 * n), (size_t)(4 * n), (size_t)(4 * n),  // keys, ids A/B
                         (size_t)(4 * (2 * n - 1)),                                            // parents
                         (size_t)(8 * n1), (size_t)n1,                                         // children, levels
                         (size_t)(4 * n + 4),                                                  // straddle list
                         (size_t)(32 * (2 * n - 1)),                                           // node boxes
                         (size_t)(4 * n1), (size_t)n1, (size_t)(4 * n1),                       // counters, flags, offsets
                         (size_t)(4 * 256 * (tiles + 8 * ((tiles + 31) / 32))),               // digit counts + group sums
                         (size_t)(4 * scanBlocks),                                             // scan partials
                         (size_t)(48 * n),                                                     // BVH-order triangles
                         (size_t)16, (size_t)256}};                                            // index check, spare
}
constexpr size_t noalloc_total(long long n) {
    size_t t = 0;
    const NoAllocSizes s = noalloc_sizes(n);
    for (int i = 0; i < kNoAllocParts; ++i) t += align256(s.b[i]);
    return t;
}

constexpr int kNoAllocMaxTris = MFH_NOALLOC_MAX_TRIANGLES;
constexpr size_t kNoAllocBytes = noalloc_total(kNoAllocMaxTris);
static_assert(kNoAllocBytes < ((size_t)64 << 20), "scratch of the allocation-free build above 64 MiB");

// The scratch of the allocation-free build (one copy per GPU, held until the
// process ends). Host code never passes this name to a kernel: it uses the
// per-device address that hipGetSymbolAddress returned for it.
static __device__ __align__(256) unsigned char scratch[kNoAllocBytes];

std::mutex gScratchMutex;
std::vector<void*> gScratchPtr;  // per device index; nullptr until looked up

// 1 in *flag if any of the 3n vertex indices is numVerts or more.
__global__ __launch_bounds__(kThreads) void k_check_indices(const u32* __restrict__ tris, long long count,
                                                           u32 numVerts, u32* flag) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < count && tris[i] >= numVerts) *flag = 1u;
}

}  // namespace

void noalloc_scratch_init(int device) {
    std::lock_guard<std::mutex> lock(gScratchMutex);
    if (device >= (int)gScratchPtr.size()) gScratchPtr.resize(device + 1, nullptr);
    if (gScratchPtr[device]) return;
    void* p = nullptr;
    MFH_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    MFH_REQUIRE(p != nullptr, "no device address for the build scratch");
    gScratchPtr[device] = p;
}

size_t noalloc_scratch_bytes() { return kNoAllocBytes; }

int build_lod_bvh_noalloc(Context& ctx, const float* dVerts, int numVerts, const uint32_t* dTris, int n,
                          const float* bbox, void* dNodes, int32_t* dPrimIds) {
    MFH_REQUIRE(n >= 1, "LoD has no triangles");
    MFH_REQUIRE(n <= kNoAllocMaxTris, "allocation-free build holds at most " + std::to_string(kNoAllocMaxTris) +
                                          " triangles, got " + std::to_string(n));
    MFH_REQUIRE(numVerts >= 1, "LoD has no vertices");
    MFH_REQUIRE(dVerts && dTris && bbox && dPrimIds && (dNodes || n == 1), "null pointer");
    void* base = nullptr;
    {
        std::lock_guard<std::mutex> lock(gScratchMutex);
        if (ctx.device < (int)gScratchPtr.size()) base = gScratchPtr[ctx.device];
    }
    MFH_REQUIRE(base != nullptr, "build scratch of this device not set up (mfh_create does it)");
    // Carve the per-device array in the order and sizes of noalloc_sizes.
    const NoAllocSizes sz = noalloc_sizes(n);
    DevBuf part[kNoAllocParts];
    size_t off = 0;
    for (int i = 0; i < kNoAllocParts; ++i) {
        part[i] = DevBuf::view(static_cast<unsigned char*>(base) + off, sz.b[i]);
        off += align256(sz.b[i]);
    }
    MFH_REQUIRE(off <= kNoAllocBytes, "build scratch layout exceeds the array");
    BuildScratch s;
    s.keysA = part[0]; s.keysB = part[1]; s.idsA = part[2]; s.idsB = part[3];
    s.parents = part[4]; s.children = part[5]; s.levels = part[6];
    s.straddle = part[7]; s.boxes = part[8];
    s.counters = part[9]; s.flags = part[10]; s.offsets = part[11];
    s.hist = part[12]; s.partials = part[13];
    float4* triVerts = part[14].as<float4>();
    u32* badIndex = part[15].as<u32>();

    hipStream_t st = ctx.stream;
    fill_u32(badIndex, 1, 0u, st);
    const long long count = 3ll * n;
    hipLaunchKernelGGL(k_check_indices, dim3((unsigned)((count + kThreads - 1) / kThreads)), dim3(kThreads), 0, st,
                       dTris, count, (u32)numVerts, badIndex);
    MFH_HIP(hipGetLastError());
    u32 bad = 0;
    MFH_HIP(hipMemcpyAsync(&bad, badIndex, 4, hipMemcpyDeviceToHost, st));
    MFH_HIP(hipStreamSynchronize(st));
    MFH_REQUIRE(bad == 0, "triangle index out of range");

    TriPrims set{dVerts, dTris};
    DevBuf nodes = DevBuf::view(dNodes, 64 * (size_t)std::max(n - 1, 1));
    DevBuf primIds = DevBuf::view(dPrimIds, 4 * (size_t)n);
    return build_generic<TriPrims, LodCost>(ctx, s, set, n, bbox, nodes, primIds, triVerts, nullptr);
}


}  // namespace mfh
