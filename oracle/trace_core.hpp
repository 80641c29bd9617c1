// Oracle — test infrastructure only.
// Shared declarations of the traversal restatement (trace_oracle.cpp) for the
// path-tracer restatement (pt_oracle.cpp).
#pragma once
#include "oracle_math.hpp"

extern "C" {
// Flat scene layout shared with the tests (all arrays host memory).
struct OrcScene {
    int32_t numLods;
    const int32_t* lodVertOffset;   // numLods+1 (in vertices)
    const int32_t* lodTriOffset;    // numLods+1 (in triangles); primIds share it
    const int32_t* lodNodeOffset;   // numLods+1 (in 64-byte internal nodes)
    const int32_t* lodNumInternal;  // numLods
    const float* verts;             // xyz per vertex
    const uint32_t* tris;           // 3 local vertex indices per triangle
    const void* nodes;              // concatenated BVHs
    const int32_t* primIds;         // concatenated, local primitive ids
    const float* lodBoxes;          // 6 per LoD (object AABB)
    int32_t numInstances;
    const float* worldToInst;       // 12 per instance
    const int32_t* instLod;         // LoD of each instance
    const void* tlasNodes;
    int32_t tlasNumInternal;
    const int32_t* tlasPrimIds;
    // LoDs with quads and spheres (all may be null: triangles only). A LoD's
    // primitive ids run triangles, quads, spheres; its sorted primitive ids
    // then start at lodPrimOffset[l] instead of lodTriOffset[l].
    const int32_t* lodQuadOffset;    // numLods+1 (in quads)
    const uint32_t* quads;           // 4 local vertex indices per quad (p00, p10, p11, p01)
    const int32_t* lodSphereOffset;  // numLods+1 (in spheres)
    const float* spheres;            // center xyz, radius
    const int32_t* lodPrimOffset;    // numLods+1 (in primitives)
};
}


namespace orc {

struct Hit {
    float t;
    i32 inst, prim;
    float w0, w1;
};

// Closest (anyHit = false) or any hit along fray up to hitT (updated to the
// world-space hit distance, or the tmax round trip). Counts are optional.
bool traverse(const OrcScene& s, const FastRay& fray, float& hitT, Hit& h, bool anyHit, u64* nodeCount,
              u64* primCount);

}  // namespace orc
