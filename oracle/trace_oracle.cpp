// Oracle — test infrastructure only.
//
// CPU restatement of Mufflon's BVH traversal (reference CPU path):
//   intersection.cu:16-24    add_epsilon
//   intersection.cu:57-75    slab test, CPU branch (early outs, no clamp)
//   intersection.cu:181-316  intersects_primitve (triangle, quad, sphere)
//   intersection.cu:321-361  world_to_object_space / object_to_world_space
//   intersection.cu:366-621  first_intersection (stack traversal, two levels)
//   intersection.cu:624-753  any_intersection
// One deliberate change of contract (the reference keeps the first visited of
// two equal-distance hits, intersection.cu:205): a hit with t equal to the
// current closest replaces it when its (instance, primitive) pair is lower,
// so the result is the lexicographic minimum of (t, instance, primitive),
// independent of traversal order. A brute-force all-pairs tracer is provided
// to pin the traversal.
#include "oracle_math.hpp"

#include <cstring>
#include <vector>

using namespace orc;

namespace {

constexpr i32 kSentinel = (i32)0xFFFFFFFF;
constexpr i32 kIgnore = (i32)0xFFFFFFFF;

struct Node {  // lbvh.hpp:13-18, 32 bytes packed
    float lo[3], hi[3];
    i32 index, primCount;
};

}  // namespace

#include "trace_core.hpp"

namespace orc {

struct Lod {
    const Node* bvh;
    const i32* primIds;
    i32 numInternal;
    i32 numPrims;
    i32 nt, nq, ns;
    const float* verts;
    const u32* tris;
    const u32* quads;
    const float* spheres;
    Box box;
};

Lod get_lod(const OrcScene& s, i32 l) {
    Lod o;
    o.bvh = (const Node*)s.nodes + 2 * (size_t)s.lodNodeOffset[l];
    o.numInternal = s.lodNumInternal[l];
    o.nt = s.lodTriOffset[l + 1] - s.lodTriOffset[l];
    o.nq = s.lodQuadOffset ? s.lodQuadOffset[l + 1] - s.lodQuadOffset[l] : 0;
    o.ns = s.lodSphereOffset ? s.lodSphereOffset[l + 1] - s.lodSphereOffset[l] : 0;
    o.numPrims = o.nt + o.nq + o.ns;
    o.primIds = s.primIds + (s.lodPrimOffset ? s.lodPrimOffset[l] : s.lodTriOffset[l]);
    o.verts = s.verts + 3 * (size_t)s.lodVertOffset[l];
    o.tris = s.tris + 3 * (size_t)s.lodTriOffset[l];
    o.quads = o.nq ? s.quads + 4 * (size_t)s.lodQuadOffset[l] : nullptr;
    o.spheres = o.ns ? s.spheres + 4 * (size_t)s.lodSphereOffset[l] : nullptr;
    const float* b = s.lodBoxes + 6 * l;
    o.box = {{b[0], b[1], b[2]}, {b[3], b[4], b[5]}};
    return o;
}

M34 get_mat(const OrcScene& s, i32 i) {
    M34 m;
    std::memcpy(m.m, s.worldToInst + 12 * (size_t)i, 48);
    return m;
}

// intersection.cu:57-75
bool slab(const float lo[3], const float hi[3], const FastRay& r, float tmax, float& cmin) {
    float t0 = lo[0] * r.inv.x - r.odiv.x;
    float t1 = hi[0] * r.inv.x - r.odiv.x;
    cmin = mn(t0, t1);
    float cmax = mx(t0, t1);
    if (cmax < 0.0f || cmin > tmax) return false;
    t0 = lo[1] * r.inv.y - r.odiv.y;
    t1 = hi[1] * r.inv.y - r.odiv.y;
    float min2 = mn(t0, t1), max2 = mx(t0, t1);
    cmin = mx(cmin, min2);
    cmax = mn(cmax, max2);
    if (cmax < 0.0f || cmin > tmax || cmin > cmax) return false;
    t0 = lo[2] * r.inv.z - r.odiv.z;
    t1 = hi[2] * r.inv.z - r.odiv.z;
    min2 = mn(t0, t1);
    max2 = mx(t0, t1);
    cmin = mx(cmin, min2);
    cmax = mn(cmax, max2);
    return (cmax >= 0.0f) && (cmin <= tmax) && (cmin <= cmax);
}
bool slab(const Box& b, const FastRay& r, float tmax, float& cmin) {
    const float lo[3] = {b.lo.x, b.lo.y, b.lo.z}, hi[3] = {b.hi.x, b.hi.y, b.hi.z};
    return slab(lo, hi, r, tmax, cmin);
}

// Strictly closer wins; at equal t the lower (instance, primitive) pair wins.
// Without a previous hit, t must be strictly below tmax (intersection.cu:205).
inline bool better(float t, i32 inst, i32 prim, float hitT, i32 hInst, i32 hPrim) {
    if (t < hitT) return true;
    if (t > hitT || hInst == kIgnore) return false;
    if (inst != hInst) return inst < hInst;
    return prim < hPrim;
}

inline V3 lod_vert(const Lod& L, u32 i) { return {L.verts[3 * i], L.verts[3 * i + 1], L.verts[3 * i + 2]}; }

// intersects_primitve (intersection.cu:181-316) without alpha testing: a
// triangle (accepted with t > 0), a bilinear quad or a sphere (their tests
// return only hits with t > 0). (w0, w1): barycentrics of v0 and v1, or the
// quad's (u, v).
inline bool prim_test(const Lod& L, const FastRay& r, i32 prim, float& t, float& w0, float& w1) {
    if (prim < L.nt) {
        const u32* ix = L.tris + 3 * (size_t)prim;
        return ray_triangle(r, lod_vert(L, ix[0]), lod_vert(L, ix[1]), lod_vert(L, ix[2]), t, w0, w1) && t > 0.0f;
    }
    if (prim < L.nt + L.nq) {
        const u32* ix = L.quads + 4 * (size_t)(prim - L.nt);
        return ray_quad(r, lod_vert(L, ix[0]), lod_vert(L, ix[1]), lod_vert(L, ix[2]), lod_vert(L, ix[3]), t, w0, w1);
    }
    const float* sp = L.spheres + 4 * (size_t)(prim - L.nt - L.nq);
    w0 = w1 = 0.0f;
    return ray_sphere(r, {sp[0], sp[1], sp[2]}, sp[3], t);
}

// world_to_object_space (intersection.cu:321-349)
bool to_object(const OrcScene& s, i32 inst, const FastRay& world, FastRay& cur, float& scale, float& tmax,
               Lod& obj, u64* nodeCount) {
    const M34 M = get_mat(s, inst);
    const V3 d = xform_dir(M, world.d);
    const float sc = len(d);
    const V3 o = xform_point(M, world.o);
    const FastRay fr = make_fast_ray(o, d / sc);
    const float objTMax = tmax * sc;
    const Lod L = get_lod(s, s.instLod[inst]);
    float ot;
    if (nodeCount) ++*nodeCount;
    if (slab(L.box, fr, objTMax, ot)) {
        cur = fr;
        scale = sc;
        tmax = objTMax;
        obj = L;
        return true;
    }
    return false;
}

// first_intersection / any_intersection share this loop (intersection.cu:407-507
// and :666-751). anyHit returns at the first accepted primitive.
bool traverse(const OrcScene& s, const FastRay& fray, float& hitT, Hit& h, bool anyHit, u64* nodeCount,
              u64* primCount_) {
    h.inst = kIgnore;
    h.prim = kIgnore;
    h.w0 = h.w1 = 0.0f;
    FastRay cur = fray;
    float scale = 1.0f;
    bool inObj = false;
    Lod obj{};
    i32 curInst = kIgnore;
    const Node* bvh = (const Node*)s.tlasNodes;
    const i32* bvhPrims = s.tlasPrimIds;
    i32 numInternal = s.tlasNumInternal;

    std::vector<i32> stack;
    stack.reserve(160);
    i32 nodeAddr = 0, primCount = 2, primOffset = 0;

    auto enter_obj = [&]() {
        bvh = obj.bvh;
        bvhPrims = obj.primIds;
        numInternal = obj.numInternal;
        inObj = true;
    };
    auto leave_obj = [&]() {
        cur = fray;
        bvh = (const Node*)s.tlasNodes;
        bvhPrims = s.tlasPrimIds;
        numInternal = s.tlasNumInternal;
        hitT /= scale;
        scale = 1.0f;
        inObj = false;
    };

    if (s.numInstances == 1) {
        if (!to_object(s, 0, fray, cur, scale, hitT, obj, nodeCount))
            primCount = 0;
        else {
            enter_obj();
            curInst = 0;
            if (obj.numPrims == 1) primCount = 1;
        }
    }

    while (!stack.empty() || primCount > 0) {
        if (primCount == 0) {
            nodeAddr = stack.back();
            stack.pop_back();
            if (nodeAddr == kSentinel) {
                if (stack.empty()) break;
                leave_obj();
                nodeAddr = stack.back();
                stack.pop_back();
            }
            if (nodeAddr >= numInternal) {
                primCount = stack.back();
                stack.pop_back();
            } else
                primCount = 2;
            primOffset = 0;
        }
        while (nodeAddr < numInternal && primCount > 0) {
            const Node& L = bvh[2 * nodeAddr];
            const Node& R = bvh[2 * nodeAddr + 1];
            if (nodeCount) ++*nodeCount;
            float c0, c1;
            const bool t0 = slab(L.lo, L.hi, cur, hitT, c0);
            const bool t1 = slab(R.lo, R.hi, cur, hitT, c1);
            if (!t0) c0 = 1e38f;
            if (!t1) c1 = 1e38f;
            if (t0 && t1) {
                const i32 push = (c1 < c0) ? L.index : R.index;
                if (push >= numInternal) stack.push_back((c1 < c0) ? L.primCount : R.primCount);
                stack.push_back(push);
            }
            nodeAddr = (c0 <= c1) ? L.index : R.index;
            if (nodeAddr >= numInternal) {
                primCount = (c0 <= c1) ? L.primCount : R.primCount;
                primOffset = 0;
            } else
                primCount = 2;
            if (!t0 && !t1) primCount = 0;
        }
        if (nodeAddr >= numInternal && primCount > 0) {
            const i32 leafId = nodeAddr - numInternal + primOffset;
            const i32 primId = bvhPrims[leafId];
            ++primOffset;
            --primCount;
            if (!inObj) {
                if (to_object(s, primId, fray, cur, scale, hitT, obj, nodeCount)) {
                    if (obj.numPrims == 1) {
                        float t, w0, w1;
                        if (primCount_) ++*primCount_;
                        if (prim_test(obj, cur, 0, t, w0, w1) && better(t, primId, 0, hitT, h.inst, h.prim)) {
                            hitT = t;
                            h.inst = primId;
                            h.prim = 0;
                            h.w0 = w0;
                            h.w1 = w1;
                            if (anyHit) return true;
                        }
                        // back to the scene level (object_to_world_space)
                        cur = fray;
                        hitT /= scale;
                        scale = 1.0f;
                    } else {
                        if (primCount > 0) {
                            stack.push_back(primCount);
                            stack.push_back(nodeAddr + primOffset);
                        }
                        stack.push_back(kSentinel);
                        enter_obj();
                        nodeAddr = 0;
                        primCount = 2;
                        primOffset = 0;
                        curInst = primId;
                    }
                }
            } else {
                float t, w0, w1;
                if (primCount_) ++*primCount_;
                if (prim_test(obj, cur, primId, t, w0, w1) && better(t, curInst, primId, hitT, h.inst, h.prim)) {
                    hitT = t;
                    h.inst = curInst;
                    h.prim = primId;
                    h.w0 = w0;
                    h.w1 = w1;
                    if (anyHit) return true;
                }
            }
        }
    }
    if (inObj) leave_obj();
    return h.inst != kIgnore;
}

}  // namespace orc

extern "C" {

// Closest hit for rays given as 8 floats (origin xyz, tmax, direction xyz,
// unused). Origins are used as given (callers apply add_epsilon themselves).
// Outputs: t (world), instance, primitive, barycentrics (w0, w1) per ray;
// optional per-ray counts of internal nodes visited and triangles tested.
int orc_trace_closest(const OrcScene* s, const float* rays, int32_t numRays, float* outT, int32_t* outInst,
                      int32_t* outPrim, float* outBary, uint32_t* outNodes, uint32_t* outPrims) {
#pragma omp parallel for schedule(dynamic, 1024)
    for (int32_t i = 0; i < numRays; ++i) {
        const float* r = rays + 8 * (size_t)i;
        const FastRay fr = make_fast_ray({r[0], r[1], r[2]}, {r[4], r[5], r[6]});
        float hitT = r[3];
        Hit h;
        u64 nc = 0, pc = 0;
        traverse(*s, fr, hitT, h, false, &nc, &pc);
        outT[i] = hitT;
        outInst[i] = h.inst;
        outPrim[i] = h.prim;
        if (outBary) {
            outBary[2 * (size_t)i] = h.w0;
            outBary[2 * (size_t)i + 1] = h.w1;
        }
        if (outNodes) outNodes[i] = (uint32_t)nc;
        if (outPrims) outPrims[i] = (uint32_t)pc;
    }
    return 0;
}

// The primitive tests alone (closed-form checks): ray = origin xyz, direction
// xyz (unit). Quad: verts = p00, p10, p11, p01 (12 floats); out = t, u, v.
int orc_ray_quad(const float* ray, const float* verts, float* out) {
    const FastRay r = make_fast_ray({ray[0], ray[1], ray[2]}, {ray[3], ray[4], ray[5]});
    V3 p[4];
    for (int k = 0; k < 4; ++k) p[k] = {verts[3 * k], verts[3 * k + 1], verts[3 * k + 2]};
    return ray_quad(r, p[0], p[1], p[2], p[3], out[0], out[1], out[2]) ? 1 : 0;
}
// Sphere = center xyz, radius; out = t.
int orc_ray_sphere(const float* ray, const float* sphere, float* out) {
    const FastRay r = make_fast_ray({ray[0], ray[1], ray[2]}, {ray[3], ray[4], ray[5]});
    return ray_sphere(r, {sphere[0], sphere[1], sphere[2]}, sphere[3], out[0]) ? 1 : 0;
}
int orc_solve_square_poly(float a, float b, float c, float* out) {
    return solve_square_poly(a, b, c, out[0], out[1]) ? 1 : 0;
}
void orc_perpendicular(const float* v, float* out) {
    const V3 p = perpendicular({v[0], v[1], v[2]});
    out[0] = p.x;
    out[1] = p.y;
    out[2] = p.z;
}
float orc_bilerp(const float* x, float s, float t) { return bilerp(x[0], x[1], x[2], x[3], s, t); }

// Any hit (occlusion) for the same ray records; 1 = occluded.
int orc_trace_any(const OrcScene* s, const float* rays, int32_t numRays, int32_t* outHit) {
#pragma omp parallel for schedule(dynamic, 1024)
    for (int32_t i = 0; i < numRays; ++i) {
        const float* r = rays + 8 * (size_t)i;
        const FastRay fr = make_fast_ray({r[0], r[1], r[2]}, {r[4], r[5], r[6]});
        float hitT = r[3];
        Hit h;
        outHit[i] = traverse(*s, fr, hitT, h, true, nullptr, nullptr) ? 1 : 0;
    }
    return 0;
}

// Brute force over every instance and triangle, with the same transforms:
// the lexicographic minimum of (t, instance, primitive) among hits with
// 0 < t < tmax in the object space of each instance.
int orc_trace_brute(const OrcScene* s, const float* rays, int32_t numRays, float* outT, int32_t* outInst,
                    int32_t* outPrim) {
#pragma omp parallel for schedule(dynamic, 64)
    for (int32_t i = 0; i < numRays; ++i) {
        const float* r = rays + 8 * (size_t)i;
        const FastRay world = make_fast_ray({r[0], r[1], r[2]}, {r[4], r[5], r[6]});
        float bestT = r[3];
        i32 bi = kIgnore, bp = kIgnore;
        for (i32 inst = 0; inst < s->numInstances; ++inst) {
            const M34 M = get_mat(*s, inst);
            const V3 d = xform_dir(M, world.d);
            const float sc = len(d);
            const FastRay fr = make_fast_ray(xform_point(M, world.o), d / sc);
            const Lod L = get_lod(*s, s->instLod[inst]);
            float tmaxObj = bestT * sc;
            for (i32 p = 0; p < L.numPrims; ++p) {
                float t, w0, w1;
                if (prim_test(L, fr, p, t, w0, w1) && better(t, inst, p, tmaxObj, bi, bp)) {
                    tmaxObj = t;
                    bestT = t / sc;
                    bi = inst;
                    bp = p;
                }
            }
        }
        outT[i] = bestT;
        outInst[i] = bi;
        outPrim[i] = bp;
    }
    return 0;
}

}  // extern "C"
