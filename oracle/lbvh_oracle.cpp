// Oracle — test infrastructure only.
//
// Sequential CPU restatement of Mufflon's LBVH builder for a triangle LoD and
// for the instance-level (TLAS) hierarchy, following
//   build_lbvh.cu:70-125   Morton codes (63 bit) of the centroids,
//   build_lbvh.cu:662-670  stable sort of (key, primId),
//   build_lbvh.cu:164-245  Karras split hierarchy (parents[], left as ~idx),
//   build_lbvh.cu:326-411  bottom-up refit with the counters[] climb,
//   build_lbvh.cu:417-463  SAH collapse marks,
//   build_lbvh.cu:748-758  inclusive scan of the marks,
//   build_lbvh.cu:503-566  copy into the collapsed 64-byte-node BVH,
//   build_lbvh.cu:586-780  orchestration (actualPrimCount == 1 special case).
// It runs the stages in the reference's serial order (thrust::seq path), which
// is the only deterministic order of the reference (its parallel refit is
// documented as racy at build_lbvh.cu:719-721). clz is exact (the CUDA and MSVC
// semantics of cuda_utils.hpp:182-211), not the log2f approximation of the
// reference's GCC branch.
#include "oracle_math.hpp"

#include <omp.h>

#include <algorithm>
#include <cstring>
#include <numeric>
#include <vector>

using namespace orc;

namespace {

constexpr i32 kTreeHead = 0x10000000;  // build_lbvh.cu:31

inline int clz64(u64 v) { return v == 0 ? 64 : __builtin_clzll(v); }
inline int clz32(u32 v) { return v == 0 ? 32 : __builtin_clz(v); }

// sfcurves.hpp:26-35
inline u64 part_by_two21(u32 x) {
    u64 r = x;
    r = (r | (r << 32)) & 0x001f00000000ffffull;
    r = (r | (r << 16)) & 0x001f0000ff0000ffull;
    r = (r | (r << 8)) & 0x100f00f00f00f00full;
    r = (r | (r << 4)) & 0x10c30c30c30c30c3ull;
    r = (r | (r << 2)) & 0x1249249249249249ull;
    return r;
}

inline float clampf(float v, float lo, float hi) { return mn(mx(v, lo), hi); }

// sdiv of util/types.hpp:124-128: x / (d + 1e-20 * sgn(d)), with sgn(0) = 1
// (Epsilon's sgn is unavailable; a zero sign would make the "safe" division
// divide by zero).
inline float sdiv(float x, float d) { return x / (d + 1e-20f * (d < 0.0f ? -1.0f : 1.0f)); }

// build_lbvh.cu:70-78 (u64 variant) and :81-85 (normalize_position with sdiv).
inline u64 morton_of_point(V3 c, const Box& bb) {
    const V3 span = bb.hi - bb.lo;
    const V3 num = c - bb.lo;
    const V3 n{sdiv(num.x, span.x), sdiv(num.y, span.y), sdiv(num.z, span.z)};
    const u32 qx = (u32)clampf(n.x * 2097152.0f, 0.0f, 2097151.0f);
    const u32 qy = (u32)clampf(n.y * 2097152.0f, 0.0f, 2097151.0f);
    const u32 qz = (u32)clampf(n.z * 2097152.0f, 0.0f, 2097151.0f);
    return part_by_two21(qx) * 4 + part_by_two21(qy) * 2 + part_by_two21(qz);
}

// Primitive sets the builder runs over (the DescType of build_lbvh.cu).
struct TriSet {
    const float* verts;
    const u32* idx;
    i32 n;
    V3 vert(u32 i) const { return {verts[3 * i], verts[3 * i + 1], verts[3 * i + 2]}; }
    // accel_structs_commons.hpp:56-71 (triangle branch)
    V3 centroid(i32 p) const {
        const V3 a = vert(idx[3 * p]), b = vert(idx[3 * p + 1]), c = vert(idx[3 * p + 2]);
        return (a + b + c) / 3.0f;
    }
    Box box(i32 p) const {
        return tri_box(vert(idx[3 * p]), vert(idx[3 * p + 1]), vert(idx[3 * p + 2]));
    }
    bool masked(i32) const { return false; }
    int kind(i32) const { return 0; }
};

// LoD with quads and spheres besides triangles; primitive ids run
// triangles, quads, spheres (descriptors.hpp:33-68). accel_structs_commons.hpp:
// 56-71 (centroids: triangle and quad means, sphere center) and :85-100
// (boxes; a sphere's is center -+ radius).
struct MixedSet {
    const float* verts;
    const u32* tris;
    const u32* quads;
    const float* spheres;  // center xyz, radius
    i32 nt, nq, ns;
    i32 n;
    V3 vert(u32 i) const { return {verts[3 * i], verts[3 * i + 1], verts[3 * i + 2]}; }
    int kind(i32 p) const { return p < nt ? 0 : (p < nt + nq ? 1 : 2); }
    V3 centroid(i32 p) const {
        const int k = kind(p);
        if (k == 0) return (vert(tris[3 * p]) + vert(tris[3 * p + 1]) + vert(tris[3 * p + 2])) / 3.0f;
        if (k == 1) {
            const u32* q = quads + 4 * (size_t)(p - nt);
            return (vert(q[0]) + vert(q[1]) + vert(q[2]) + vert(q[3])) / 4.0f;
        }
        const float* sp = spheres + 4 * (size_t)(p - nt - nq);
        return {sp[0], sp[1], sp[2]};
    }
    Box box(i32 p) const {
        const int k = kind(p);
        if (k == 0) return tri_box(vert(tris[3 * p]), vert(tris[3 * p + 1]), vert(tris[3 * p + 2]));
        if (k == 1) {
            const u32* q = quads + 4 * (size_t)(p - nt);
            return quad_box(vert(q[0]), vert(q[1]), vert(q[2]), vert(q[3]));
        }
        const float* sp = spheres + 4 * (size_t)(p - nt - nq);
        return sphere_box({sp[0], sp[1], sp[2]}, sp[3]);
    }
    bool masked(i32) const { return false; }
};

// Instance set: centroid/box of each instance's object AABB in world space
// (accel_structs_commons.hpp:75-108).
struct InstSet {
    const float* worldToInst;  // 12 floats per instance
    const float* objBoxes;     // 6 floats per LoD
    const i32* lodIdx;
    const u32* present;  // nullptr = all present
    i32 n;
    M34 mat(i32 i) const { return instance_to_world(worldToInst + 12 * (size_t)i); }
    Box obox(i32 i) const {
        const float* b = objBoxes + 6 * lodIdx[i];
        return {{b[0], b[1], b[2]}, {b[3], b[4], b[5]}};
    }
    V3 centroid(i32 p) const {
        const Box b = obox(p);
        const V3 c = (b.lo + b.hi) * 0.5f;
        return xform_point(mat(p), c);
    }
    Box box(i32 p) const { return xform_box(mat(p), obox(p)); }
    bool masked(i32 p) const { return present && !present[p]; }
    int kind(i32) const { return 0; }
};

// Cost model of a triangle LoD (build_lbvh.cu:34-40): Ct = 2, Cp = (1.2, 4, 1).
struct LodCost {
    static constexpr float Ct = 2.0f;
    struct Count {
        i32 x = 0, y = 0, z = 0;
    };
    // get_count (build_lbvh.cu:285-295): one primitive of kind 0 (triangle),
    // 1 (quad) or 2 (sphere)
    static Count leaf(int k) { return {k == 0 ? 1 : 0, k == 1 ? 1 : 0, k == 2 ? 1 : 0}; }
    static i32 encode(const Count& c) {
        return (std::min(c.x, 0x3FF) << 20) | (std::min(c.y, 0x3FF) << 10) | std::min(c.z, 0x3FF);
    }
    static Count extract(i32 v) {
        return {(v & 0x3FF00000) >> 20, (v & 0x000FFC00) >> 10, (v & 0x000003FF)};
    }
    static float cost(const Count& c) {
        return Ct + ((float)c.x * 1.2f + (float)c.y * 4.0f + (float)c.z * 1.0f);
    }
    static Count add(Count a, const Count& b) {
        a.x += b.x;
        a.y += b.y;
        a.z += b.z;
        return a;
    }
    static bool capped(const Count& c) { return c.x >= 1023 || c.y >= 1023 || c.z >= 1023; }
    static i32 sum(const Count& c) { return c.x + c.y + c.z; }
};
// Cost model of the instance level (build_lbvh.cu:41-48): Ct = 1, Cp = 8.5.
struct InstCost {
    static constexpr float Ct = 1.0f;
    struct Count {
        i32 x = 0;
    };
    static Count leaf(int) { return {1}; }
    static i32 encode(const Count& c) { return c.x; }
    static Count extract(i32 v) { return {v}; }
    static float cost(const Count& c) { return Ct + (float)c.x * 8.5f; }
    static Count add(Count a, const Count& b) {
        a.x += b.x;
        return a;
    }
    static bool capped(const Count& c) { return c.x >= 1023; }
    static i32 sum(const Count& c) { return c.x; }
};

inline float i2f(i32 v) {
    float f;
    std::memcpy(&f, &v, 4);
    return f;
}
inline i32 f2i(float f) {
    i32 v;
    std::memcpy(&v, &f, 4);
    return v;
}

// build_lbvh.cu:164-177
inline i32 lcp(const u64* keys, i32 n, i32 i1, i32 i2, u64 k1) {
    if (i2 < 0 || i2 >= n) return -1;
    const u64 k2 = keys[i2];
    if (k1 == k2) return 64 + clz32((u32)(i1 ^ i2));
    return clz64(k1 ^ k2);
}

// build_lbvh.cu:179-245
void karras_node(i32 n, const u64* keys, i32* parents, i32 idx) {
    const u64 k1 = keys[idx];
    const i32 l1 = lcp(keys, n, idx, idx + 1, k1);
    const i32 l2 = lcp(keys, n, idx, idx - 1, k1);
    const i32 d = (l1 - l2) > 0 ? 1 : ((l1 - l2) < 0 ? -1 : 0);
    const i32 minLcp = lcp(keys, n, idx, idx - d, k1);
    i32 lMax = 2;
    while (lcp(keys, n, idx, idx + lMax * d, k1) > minLcp) lMax *= 2;
    i32 l = 0;
    for (i32 t = lMax; t > 1;) {
        t = t / 2;
        if (lcp(keys, n, idx, idx + (l + t) * d, k1) > minLcp) l += t;
    }
    const i32 j = idx + l * d;
    const i32 nodeLcp = lcp(keys, n, idx, j, k1);
    i32 s = 0;
    const i32 maxDiv = 1 << (32 - clz32((u32)l));
    for (i32 div = 2; div <= maxDiv; div *= 2) {
        const i32 t = (l + div - 1) / div;
        if (lcp(keys, n, idx, idx + (s + t) * d, k1) > nodeLcp) s += t;
    }
    const i32 split = idx + s * d + std::min(d, 0);
    const i32 left = (std::min(idx, j) == split) ? split + n - 1 : split;
    const i32 right = (std::max(idx, j) == split + 1) ? split + n : split + 1;
    parents[left] = ~idx;
    parents[right] = idx;
    if (idx == 0) parents[0] = kTreeHead;
}

struct BuildOut {
    std::vector<u64> keys;      // sorted Morton codes
    std::vector<i32> ids;       // sorted primitive ids
    std::vector<i32> parents;   // 2N-1
    std::vector<float> boxes;   // (2N-1) * 8
    std::vector<i32> offsets;   // N-1, inclusive scan of the marks
    std::vector<uint8_t> nodes; // numInternalAfter * 64 bytes
    i32 numInternalAfter = 0;
};

template <class Set, class C>
void build_lbvh(const Set& set, const Box& bb, BuildOut& o, u64* unsortedKeys) {
    const i32 n = set.n;
    if (n == 1) {  // build_lbvh.cu:595-602: no hierarchy
        o.keys.assign(1, set.masked(0) ? ~0ull : morton_of_point(set.centroid(0), bb));
        o.ids.assign(1, 0);
        o.parents.assign(1, 0);
        o.boxes.assign(8, 0.0f);
        o.offsets.clear();
        o.nodes.clear();
        o.numInternalAfter = 0;
        if (unsortedKeys) unsortedKeys[0] = o.keys[0];
        return;
    }
    const i32 nInt = n - 1;
    const i32 nNodes = nInt + n;
    // Morton codes (build_lbvh.cu:96-125, CPU branch :662-668)
    std::vector<std::pair<u64, i32>> kv(n);
    for (i32 i = 0; i < n; ++i) {
        const u64 k = set.masked(i) ? ~0ull : morton_of_point(set.centroid(i), bb);
        kv[i] = {k, i};
        if (unsortedKeys) unsortedKeys[i] = k;
    }
    // thrust::sort_by_key is stable: ties keep ascending primitive id.
    std::stable_sort(kv.begin(), kv.end(),
                     [](const std::pair<u64, i32>& a, const std::pair<u64, i32>& b) {
                         return a.first < b.first;
                     });
    o.keys.resize(n);
    o.ids.resize(n);
    for (i32 i = 0; i < n; ++i) {
        o.keys[i] = kv[i].first;
        o.ids[i] = kv[i].second;
    }
    // Karras hierarchy
    o.parents.assign(nNodes, 0);
    for (i32 i = 0; i < nInt; ++i) karras_node(n, o.keys.data(), o.parents.data(), i);

    // Refit (build_lbvh.cu:326-411), serial thread order.
    o.boxes.assign((size_t)nNodes * 8, 0.0f);
    std::vector<i32> counters(nInt, (i32)0xFFFFFFFF);
    std::vector<i32> marks(nInt, 0);
    float* B = o.boxes.data();
    for (i32 idx = 0; idx < n; ++idx) {
        const i32 prim = o.ids[idx];
        Box cur = set.box(prim);
        float curSA = surface(cur);
        typename C::Count cnt = C::leaf(set.kind(prim));
        float cost = C::cost(cnt);
        const i32 leaf = idx + nInt;
        float* lb = B + (size_t)leaf * 8;
        lb[0] = cur.lo.x; lb[1] = cur.lo.y; lb[2] = cur.lo.z; lb[3] = i2f(C::encode(cnt));
        lb[4] = cur.hi.x; lb[5] = cur.hi.y; lb[6] = cur.hi.z; lb[7] = cost;
        i32 node = o.parents[leaf];
        i32 last = idx;
        bool lastLeft = node < 0;
        if (node < 0) node = ~node;
        i32 other = counters[node];
        counters[node] = leaf;
        while (other != (i32)0xFFFFFFFF) {
            i32 otherNode = lastLeft ? last + 1 : last - 1;
            if (other >= nInt) {
                other -= nInt;
                otherNode += nInt;
            }
            const float* ob = B + (size_t)otherNode * 8;
            const Box oBox{{ob[0], ob[1], ob[2]}, {ob[4], ob[5], ob[6]}};
            cur.lo = vmin(cur.lo, oBox.lo);
            cur.hi = vmax(cur.hi, oBox.hi);
            const float newSA = surface(cur);
            const float otherSA = surface(oBox);
            const typename C::Count oc = C::extract(f2i(ob[3]));
            cost = C::Ct + curSA / newSA * C::cost(cnt) + otherSA / newSA * C::cost(oc);
            cnt = C::add(cnt, oc);
            curSA = newSA;
            float* nb = B + (size_t)node * 8;
            nb[0] = cur.lo.x; nb[1] = cur.lo.y; nb[2] = cur.lo.z; nb[3] = i2f(C::encode(cnt));
            nb[4] = cur.hi.x; nb[5] = cur.hi.y; nb[6] = cur.hi.z; nb[7] = cost;
            last = node;
            node = o.parents[node];
            lastLeft = node < 0;
            if (node < 0) node = ~node;
            if (node == kTreeHead) break;
            other = counters[node];
            counters[node] = idx;
        }
        if (idx < n - 1) marks[idx] = 1;
    }

    // Collapse marks (build_lbvh.cu:417-463), one leaf at a time.
    for (i32 idx = 0; idx < n; ++idx) {
        const i32 leaf = idx + nInt;
        i32 collapse = -1;
        i32 node = o.parents[leaf];
        if (node < 0) node = ~node;
        while (node != 0) {
            const typename C::Count c = C::extract(f2i(B[(size_t)node * 8 + 3]));
            if (C::capped(c)) break;
            const float costNode = B[(size_t)node * 8 + 7];
            if (C::cost(c) < costNode) collapse = node;
            node = o.parents[node];
            if (node < 0) node = ~node;
        }
        if (collapse != -1) {
            i32 m = o.parents[leaf];
            if (m < 0) m = ~m;
            while (m != collapse) {
                marks[m] = 0;
                m = o.parents[m];
                if (m < 0) m = ~m;
            }
            marks[m] = 0;
        }
    }
    // Inclusive scan (build_lbvh.cu:755-757)
    o.offsets.resize(nInt);
    i32 acc = 0;
    for (i32 i = 0; i < nInt; ++i) {
        acc += marks[i];
        o.offsets[i] = acc;
    }
    const i32 nAfter = o.offsets[nInt - 1];
    o.numInternalAfter = nAfter;
    o.nodes.assign((size_t)nAfter * 64, 0);
    auto collapsed = [&](i32 node) {
        return node > 0 && node < nInt && (o.offsets[node] - o.offsets[node - 1]) == 0;
    };
    // Copy into the collapsed BVH (build_lbvh.cu:503-566), serial node order.
    for (i32 node = 1; node < nNodes; ++node) {
        if (collapsed(node)) continue;
        i32 parent = o.parents[node];
        i32 off;
        if (parent < 0) {
            off = 0;
            parent = ~parent;
        } else
            off = 1;
        i32 fin = node;
        bool skip = false;
        if (fin >= nInt) {
            while (collapsed(parent)) {
                if (off == 1) {
                    skip = true;
                    break;
                }
                fin = parent;
                parent = o.parents[parent];
                if (parent < 0) {
                    off = 0;
                    parent = ~parent;
                } else
                    off = 1;
            }
        }
        if (skip) continue;
        const float* fb = B + (size_t)fin * 8;
        const i32 primCount = C::sum(C::extract(f2i(fb[3])));
        const i32 outIdx = (o.offsets[parent] - 1) * 2;
        const i32 outNode = (node >= nInt) ? node - nInt + nAfter : o.offsets[node] - 1;
        const V3 lo{fb[0], fb[1], fb[2]}, hi{fb[4], fb[5], fb[6]};
        const V3 c = (lo + hi) * 0.5f;
        const V3 bl = (lo - c) * 1.0001f + c;
        const V3 bh = (hi - c) * 1.0001f + c;
        uint8_t* dst = o.nodes.data() + (size_t)(outIdx + off) * 32;
        const float f[6] = {bl.x, bl.y, bl.z, bh.x, bh.y, bh.z};
        std::memcpy(dst, f, 24);
        std::memcpy(dst + 24, &outNode, 4);
        std::memcpy(dst + 28, &primCount, 4);
    }
}

// ---------------------------------------------------------------------------
// OpenMP-parallel variant of the same stages: the reference builds LoDs of
// 1000 or more primitives with its parallel path (lod.cpp:26, the OpenMP /
// thrust::par branches of build_lbvh.cu:658-760). Used as the multi-core CPU
// baseline; its output equals build_lbvh's bit for bit:
// * Morton codes, Karras nodes, collapse marks and the collapsed copy are per
//   element and independent (parallel for);
// * the sort is a stable LSD radix sort of (key, id) in 8 passes of 8 bits
//   (per-thread digit counts over contiguous chunks, scattered in thread
//   order), which gives std::stable_sort's order;
// * the refit climbs with one atomic arrival flag per node; the second
//   arriver merges the node in the serial path's operand order (right child
//   as the climbing box, left child as the other) and with the serial path's
//   counts: the right child's carried (unsaturated) count plus the left
//   child's stored (saturated) count, kept per node in a side array.
template <class Set, class C>
void build_lbvh_omp(const Set& set, const Box& bb, BuildOut& o) {
    const i32 n = set.n;
    if (n == 1) {
        build_lbvh<Set, C>(set, bb, o, nullptr);
        return;
    }
    const i32 nInt = n - 1;
    const i32 nNodes = nInt + n;
    std::vector<u64> keys(n), keys2(n);
    std::vector<i32> ids(n), ids2(n);
#pragma omp parallel for schedule(static)
    for (i32 i = 0; i < n; ++i) {
        keys[i] = set.masked(i) ? ~0ull : morton_of_point(set.centroid(i), bb);
        ids[i] = i;
    }
    // stable LSD radix sort
    const int T = std::max(1, omp_get_max_threads());
    std::vector<size_t> cnt((size_t)T * 256);
    for (int shift = 0; shift < 64; shift += 8) {
#pragma omp parallel num_threads(T)
        {
            const int t = omp_get_thread_num();
            const int nt = omp_get_num_threads();
            const int64_t lo = (int64_t)n * t / nt, hi = (int64_t)n * (t + 1) / nt;
            size_t* c = cnt.data() + (size_t)t * 256;
            std::fill(c, c + 256, (size_t)0);
            for (int64_t i = lo; i < hi; ++i) ++c[(keys[i] >> shift) & 255u];
#pragma omp barrier
#pragma omp single
            {
                size_t run = 0;
                for (int d = 0; d < 256; ++d)
                    for (int u = 0; u < nt; ++u) {
                        const size_t v = cnt[(size_t)u * 256 + d];
                        cnt[(size_t)u * 256 + d] = run;
                        run += v;
                    }
            }
            for (int64_t i = lo; i < hi; ++i) {
                const size_t dst = c[(keys[i] >> shift) & 255u]++;
                keys2[dst] = keys[i];
                ids2[dst] = ids[i];
            }
        }
        keys.swap(keys2);
        ids.swap(ids2);
    }
    o.keys = keys;
    o.ids = ids;
    o.parents.assign(nNodes, 0);
#pragma omp parallel for schedule(static)
    for (i32 i = 0; i < nInt; ++i) karras_node(n, o.keys.data(), o.parents.data(), i);
    // explicit children of every internal node
    std::vector<i32> left(nInt), right(nInt);
#pragma omp parallel for schedule(static)
    for (i32 c = 1; c < nNodes; ++c) {
        const i32 p = o.parents[c];
        if (p < 0)
            left[~p] = c;
        else
            right[p] = c;
    }
    // refit
    o.boxes.assign((size_t)nNodes * 8, 0.0f);
    std::vector<typename C::Count> carried(nNodes);
    std::vector<i32> arrived(nInt, 0);
    float* B = o.boxes.data();
#pragma omp parallel for schedule(static)
    for (i32 idx = 0; idx < n; ++idx) {
        const i32 prim = o.ids[idx];
        const Box lb0 = set.box(prim);
        const typename C::Count c0 = C::leaf(set.kind(prim));
        const i32 leaf = idx + nInt;
        float* lb = B + (size_t)leaf * 8;
        lb[0] = lb0.lo.x; lb[1] = lb0.lo.y; lb[2] = lb0.lo.z; lb[3] = i2f(C::encode(c0));
        lb[4] = lb0.hi.x; lb[5] = lb0.hi.y; lb[6] = lb0.hi.z; lb[7] = C::cost(c0);
        carried[leaf] = c0;
        i32 node = o.parents[leaf];
        if (node < 0) node = ~node;
        for (;;) {
            if (__atomic_exchange_n(&arrived[node], 1, __ATOMIC_ACQ_REL) == 0) break;  // first arriver
            const float* rb = B + (size_t)right[node] * 8;
            const float* ob = B + (size_t)left[node] * 8;
            const Box rBox{{rb[0], rb[1], rb[2]}, {rb[4], rb[5], rb[6]}};
            const Box oBox{{ob[0], ob[1], ob[2]}, {ob[4], ob[5], ob[6]}};
            Box cur = rBox;
            cur.lo = vmin(cur.lo, oBox.lo);
            cur.hi = vmax(cur.hi, oBox.hi);
            const float curSA = surface(rBox);
            const float newSA = surface(cur);
            const float otherSA = surface(oBox);
            const typename C::Count rc = carried[right[node]];
            const typename C::Count oc = C::extract(f2i(ob[3]));
            const float cost = C::Ct + curSA / newSA * C::cost(rc) + otherSA / newSA * C::cost(oc);
            const typename C::Count sum = C::add(rc, oc);
            carried[node] = sum;
            float* nb = B + (size_t)node * 8;
            nb[0] = cur.lo.x; nb[1] = cur.lo.y; nb[2] = cur.lo.z; nb[3] = i2f(C::encode(sum));
            nb[4] = cur.hi.x; nb[5] = cur.hi.y; nb[6] = cur.hi.z; nb[7] = cost;
            if (node == 0) break;
            node = o.parents[node];
            if (node < 0) node = ~node;
        }
    }
    // collapse marks, one leaf walk per thread (every write stores 0)
    std::vector<i32> marks(nInt, 1);
#pragma omp parallel for schedule(static)
    for (i32 idx = 0; idx < n; ++idx) {
        const i32 leaf = idx + nInt;
        i32 collapse = -1;
        i32 node = o.parents[leaf];
        if (node < 0) node = ~node;
        while (node != 0) {
            const typename C::Count c = C::extract(f2i(B[(size_t)node * 8 + 3]));
            if (C::capped(c)) break;
            if (C::cost(c) < B[(size_t)node * 8 + 7]) collapse = node;
            node = o.parents[node];
            if (node < 0) node = ~node;
        }
        if (collapse != -1) {
            i32 m = o.parents[leaf];
            if (m < 0) m = ~m;
            for (;;) {
                __atomic_store_n(&marks[m], 0, __ATOMIC_RELAXED);
                if (m == collapse) break;
                m = o.parents[m];
                if (m < 0) m = ~m;
            }
        }
    }
    o.offsets.resize(nInt);
    i32 acc = 0;
    for (i32 i = 0; i < nInt; ++i) {
        acc += marks[i];
        o.offsets[i] = acc;
    }
    const i32 nAfter = o.offsets[nInt - 1];
    o.numInternalAfter = nAfter;
    o.nodes.assign((size_t)nAfter * 64, 0);
    auto collapsed = [&](i32 node) {
        return node > 0 && node < nInt && (o.offsets[node] - o.offsets[node - 1]) == 0;
    };
#pragma omp parallel for schedule(static)
    for (i32 node = 1; node < nNodes; ++node) {
        if (collapsed(node)) continue;
        i32 parent = o.parents[node];
        i32 off = parent < 0 ? 0 : 1;
        if (parent < 0) parent = ~parent;
        i32 fin = node;
        bool skip = false;
        if (fin >= nInt) {
            while (collapsed(parent)) {
                if (off == 1) {
                    skip = true;
                    break;
                }
                fin = parent;
                parent = o.parents[parent];
                off = parent < 0 ? 0 : 1;
                if (parent < 0) parent = ~parent;
            }
        }
        if (skip) continue;
        const float* fb = B + (size_t)fin * 8;
        const i32 primCount = C::sum(C::extract(f2i(fb[3])));
        const i32 outIdx = (o.offsets[parent] - 1) * 2;
        const i32 outNode = (node >= nInt) ? node - nInt + nAfter : o.offsets[node] - 1;
        const V3 lo{fb[0], fb[1], fb[2]}, hi{fb[4], fb[5], fb[6]};
        const V3 c = (lo + hi) * 0.5f;
        const V3 bl = (lo - c) * 1.0001f + c;
        const V3 bh = (hi - c) * 1.0001f + c;
        uint8_t* dst = o.nodes.data() + (size_t)(outIdx + off) * 32;
        const float f[6] = {bl.x, bl.y, bl.z, bh.x, bh.y, bh.z};
        std::memcpy(dst, f, 24);
        std::memcpy(dst + 24, &outNode, 4);
        std::memcpy(dst + 28, &primCount, 4);
    }
}

void copy_out(const BuildOut& o, i32 n, u64* sortedKeys, i32* ids, i32* parents, float* boxes,
              i32* offsets, void* nodes, i32* numInternalAfter) {
    if (sortedKeys) std::memcpy(sortedKeys, o.keys.data(), 8 * (size_t)n);
    if (ids) std::memcpy(ids, o.ids.data(), 4 * (size_t)n);
    if (n > 1) {
        if (parents) std::memcpy(parents, o.parents.data(), 4 * (size_t)(2 * n - 1));
        if (boxes) std::memcpy(boxes, o.boxes.data(), 4 * o.boxes.size());
        if (offsets) std::memcpy(offsets, o.offsets.data(), 4 * (size_t)(n - 1));
        if (nodes) std::memcpy(nodes, o.nodes.data(), o.nodes.size());
    }
    if (numInternalAfter) *numInternalAfter = o.numInternalAfter;
}

}  // namespace

extern "C" {

// Morton code of one point in a box (closed-form checks use this).
uint64_t orc_morton_point(const float p[3], const float bbox[6]) {
    return morton_of_point({p[0], p[1], p[2]}, {{bbox[0], bbox[1], bbox[2]}, {bbox[3], bbox[4], bbox[5]}});
}
uint64_t orc_part_by_two21(uint32_t x) { return part_by_two21(x); }
void orc_part_by_two21_batch(const uint32_t* x, int64_t n, uint64_t* out) {
    for (int64_t i = 0; i < n; ++i) out[i] = part_by_two21(x[i]);
}
void orc_morton_points(const float* p, int64_t n, const float bbox[6], uint64_t* out) {
    const Box bb{{bbox[0], bbox[1], bbox[2]}, {bbox[3], bbox[4], bbox[5]}};
    for (int64_t i = 0; i < n; ++i) out[i] = morton_of_point({p[3 * i], p[3 * i + 1], p[3 * i + 2]}, bb);
}

// Full LBVH build of a triangle LoD. Every output pointer may be null.
// Sizes: unsortedKeys/sortedKeys/ids N, parents 2N-1, boxes (2N-1)*8 floats,
// offsets N-1, nodes (N-1)*64 bytes capacity.
int orc_lbvh_build_tris(const float* verts, const uint32_t* idx, int32_t numTris, const float bbox[6],
                        uint64_t* unsortedKeys, uint64_t* sortedKeys, int32_t* ids, int32_t* parents,
                        float* boxes, int32_t* offsets, void* nodes, int32_t* numInternalAfter) {
    if (numTris <= 0) return 1;
    TriSet set{verts, idx, numTris};
    const Box bb{{bbox[0], bbox[1], bbox[2]}, {bbox[3], bbox[4], bbox[5]}};
    BuildOut o;
    build_lbvh<TriSet, LodCost>(set, bb, o, unsortedKeys);
    copy_out(o, numTris, sortedKeys, ids, parents, boxes, offsets, nodes, numInternalAfter);
    return 0;
}

// The same build with the OpenMP-parallel stages (build_lbvh_omp); equal
// outputs. Runs on omp_get_max_threads() threads.
int orc_lbvh_build_tris_omp(const float* verts, const uint32_t* idx, int32_t numTris, const float bbox[6],
                            uint64_t* sortedKeys, int32_t* ids, int32_t* parents, float* boxes, int32_t* offsets,
                            void* nodes, int32_t* numInternalAfter) {
    if (numTris <= 0) return 1;
    TriSet set{verts, idx, numTris};
    const Box bb{{bbox[0], bbox[1], bbox[2]}, {bbox[3], bbox[4], bbox[5]}};
    BuildOut o;
    build_lbvh_omp<TriSet, LodCost>(set, bb, o);
    copy_out(o, numTris, sortedKeys, ids, parents, boxes, offsets, nodes, numInternalAfter);
    return 0;
}

// The same build over a LoD of triangles, quads (4 vertex indices: p00, p10,
// p11, p01) and spheres (center xyz, radius); bbox: the LoD box.
int orc_lbvh_build_mixed(const float* verts, const uint32_t* tris, int32_t numTris, const uint32_t* quads,
                         int32_t numQuads, const float* spheres, int32_t numSpheres, const float bbox[6],
                         uint64_t* sortedKeys, int32_t* ids, int32_t* parents, float* boxes, int32_t* offsets,
                         void* nodes, int32_t* numInternalAfter) {
    const i32 n = numTris + numQuads + numSpheres;
    if (n <= 0) return 1;
    MixedSet set{verts, tris, quads, spheres, numTris, numQuads, numSpheres, n};
    const Box bb{{bbox[0], bbox[1], bbox[2]}, {bbox[3], bbox[4], bbox[5]}};
    BuildOut o;
    build_lbvh<MixedSet, LodCost>(set, bb, o, nullptr);
    copy_out(o, n, sortedKeys, ids, parents, boxes, offsets, nodes, numInternalAfter);
    return 0;
}

// LBVH over instances (TLAS). worldToInst: 12 floats per instance,
// objBoxes: 6 floats per LoD, sceneBox: world AABB of the scene.
int orc_lbvh_build_instances(const float* worldToInst, const float* objBoxes, const int32_t* lodIdx,
                             int32_t numInst, const float sceneBox[6], uint64_t* sortedKeys,
                             int32_t* ids, int32_t* parents, float* boxes, int32_t* offsets, void* nodes,
                             int32_t* numInternalAfter) {
    if (numInst <= 0) return 1;
    InstSet set{worldToInst, objBoxes, lodIdx, nullptr, numInst};
    const Box bb{{sceneBox[0], sceneBox[1], sceneBox[2]}, {sceneBox[3], sceneBox[4], sceneBox[5]}};
    BuildOut o;
    build_lbvh<InstSet, InstCost>(set, bb, o, nullptr);
    copy_out(o, numInst, sortedKeys, ids, parents, boxes, offsets, nodes, numInternalAfter);
    return 0;
}

// Karras hierarchy alone over given sorted keys (for brute-force checks).
int orc_karras(const uint64_t* sortedKeys, int32_t n, int32_t* parents) {
    if (n < 2) return 1;
    std::fill(parents, parents + (2 * n - 1), 0);
    for (i32 i = 0; i < n - 1; ++i) karras_node(n, sortedKeys, parents, i);
    return 0;
}

// World AABB of every instance and their union (the scene box).
int orc_instance_boxes(const float* worldToInst, const float* objBoxes, const int32_t* lodIdx, int32_t numInst,
                       float* outBoxes, float sceneBox[6]) {
    InstSet set{worldToInst, objBoxes, lodIdx, nullptr, numInst};
    Box u{};
    for (i32 i = 0; i < numInst; ++i) {
        const Box b = set.box(i);
        if (outBoxes) {
            const float f[6] = {b.lo.x, b.lo.y, b.lo.z, b.hi.x, b.hi.y, b.hi.z};
            std::memcpy(outBoxes + 6 * (size_t)i, f, 24);
        }
        if (i == 0)
            u = b;
        else {
            u.lo = vmin(u.lo, b.lo);
            u.hi = vmax(u.hi, b.hi);
        }
    }
    const float f[6] = {u.lo.x, u.lo.y, u.lo.z, u.hi.x, u.hi.y, u.hi.z};
    std::memcpy(sceneBox, f, 24);
    return 0;
}

}  // extern "C"
