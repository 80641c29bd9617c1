// Host-side internals of the HIP library behind include/mufflon_hip.h.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

namespace mfh {

struct Error : std::runtime_error {
    int code;
    Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

enum : int {
    kOk = 0,
    kInvalidArgument = 1,
    kHipError = 2,
    kOutOfMemory = 3,
    kBadState = 4,
};

#define MFH_HIP(expr)                                                                              \
    do {                                                                                           \
        hipError_t e_ = (expr);                                                                    \
        if (e_ != hipSuccess)                                                                      \
            throw ::mfh::Error(::mfh::kHipError, std::string(#expr) + ": " + hipGetErrorString(e_)); \
    } while (0)

#define MFH_REQUIRE(cond, msg)                                                 \
    do {                                                                       \
        if (!(cond)) throw ::mfh::Error(::mfh::kInvalidArgument, (msg));       \
    } while (0)

// Device buffer owned by a context; grows on demand, never shrinks. A fixed
// buffer is a view of memory owned elsewhere (a caller's buffer, a static
// device array): it never allocates or frees, and asking for more than it
// holds is an error.
struct DevBuf {
    void* ptr = nullptr;
    size_t bytes = 0;
    bool fixed = false;
    static DevBuf view(void* p, size_t n) {
        DevBuf b;
        b.ptr = p;
        b.bytes = n;
        b.fixed = true;
        return b;
    }
    void reserve(size_t n) {
        if (n <= bytes) return;
        if (fixed)
            throw Error(kInvalidArgument, "fixed buffer of " + std::to_string(bytes) + " bytes, " +
                                              std::to_string(n) + " needed");
        if (ptr) MFH_HIP(hipFree(ptr));
        ptr = nullptr;
        bytes = 0;
        if (hipMalloc(&ptr, n) != hipSuccess) {
            (void)hipGetLastError();
            throw Error(kOutOfMemory, "hipMalloc of " + std::to_string(n) + " bytes failed");
        }
        bytes = n;
    }
    void release() {
        if (ptr && !fixed) (void)hipFree(ptr);
        ptr = nullptr;
        bytes = 0;
    }
    template <class T>
    T* as() const {
        return static_cast<T*>(ptr);
    }
};

// Keys per tile of the radix sort's histogram and scatter passes.
constexpr int kRadixTileKeys = 2048;

// Scratch of one LBVH build (sizes for n primitives).
struct BuildScratch {
    DevBuf keysA, keysB, idsA, idsB;  // sort ping-pong
    DevBuf parents;                   // 2n-1 (reference encoding, left child as ~parent)
    DevBuf children;                  // 2(n-1): left, right node of each internal node
    DevBuf levels;                    // n-1 bytes: refit level of each internal node
    DevBuf straddle;                  // n u32: internal nodes of refit level >= 1 (packed), then their count
    DevBuf boxes;                     // (2n-1) * 32 bytes: {lo, countBits}, {hi, cost}
    DevBuf counters;                  // n-1 arrival counters (refit levels past their LDS capacity)
    DevBuf flags;                     // n-1 bytes: bit0 count capped, bit1 leaf cheaper than split
    DevBuf offsets;                   // n-1 inclusive scan of the keep marks
    DevBuf offsets2;                  // n-1 (+1 flag) the same for the traversal hierarchy
    DevBuf hist;                      // radix digit counts
    DevBuf partials;                  // scan partial sums
    DevBuf misc;                      // small values (scene box, counts)
    DevBuf stage;                     // n * 48 bytes: triangle records in primitive order
};

// Host copies of the intermediate stages, filled when a build is asked for them.
struct StageSink {
    uint64_t* unsortedKeys = nullptr;
    uint64_t* sortedKeys = nullptr;
    int32_t* sortedIds = nullptr;
    int32_t* parents = nullptr;
    float* boxes = nullptr;
    int32_t* offsets = nullptr;
};

// Primitive kinds of a LoD, in primitive-id order (descriptors.hpp:33-68):
// triangles [0, numTris), quads [numTris, numTris + numQuads), spheres after.
constexpr int kPrimTriangle = 0;
constexpr int kPrimQuad = 1;
constexpr int kPrimSphere = 2;

struct Lod {
    int32_t numVerts = 0, numTris = 0, numQuads = 0, numSpheres = 0;
    DevBuf verts;     // float xyz per vertex
    DevBuf tris;      // u32 x3 per triangle
    DevBuf quads;     // u32 x4 per quad
    DevBuf spheres;   // float4 per sphere: center xyz, radius
    // host copies and shading attributes (normals, uvs, material per primitive)
    std::vector<float> hostVerts;
    std::vector<uint32_t> hostTris;
    std::vector<uint32_t> hostQuads;
    std::vector<float> hostSpheres;
    std::vector<float> hostUvs;
    std::vector<uint16_t> hostMatIdx;
    bool hasShading = false;
    DevBuf normals, uvs, matIdx;
    float box[6] = {0, 0, 0, 0, 0, 0};
    float coordMax = 0.0f;    // largest |coordinate| of the vertices (bounds every node box of its BVHs)
    bool built = false;
    uint64_t generation = 0;  // incremented by every build
    int32_t numInternal = 0;  // internal nodes of the collapsed BVH
    DevBuf nodes;             // numInternal * 64 bytes
    DevBuf primIds;           // numPrims(), sorted order
    // 48-byte primitive records in sorted order (three float4): a triangle
    // {v0, id}, {v1, 0}, {v2, 0}; a quad {v0, id}, {v1, 1}, {v2, 0} with its
    // fourth vertex in quadV3 at the same position; a sphere {center, id},
    // {radius, 0, 0, 2}, {0}. The kind is the int bits of the second .w.
    DevBuf triVerts;
    DevBuf quadV3;            // float4 per sorted position (LoDs with quads only)
    // traversal hierarchy: the collapsed BVH with its long leaves kept split
    // (travNumInternal = -1: none was, traversal uses nodes)
    DevBuf travNodes;
    int32_t travNumInternal = -1;
    int32_t numPrims() const { return numTris + numQuads + numSpheres; }
    bool mixed() const { return numQuads > 0 || numSpheres > 0; }
};

struct Scene {
    int32_t numInstances = 0;
    std::vector<uint64_t> lodGeneration;  // generation of every LoD when the table was uploaded
    std::vector<int32_t> instLod;
    std::vector<float> worldToInst;
    DevBuf dInstLod, dWorldToInst, dLods;
    DevBuf dLodsRef;  // the LoD table over the reference BVHs (visit counts of the reference order)
    float box[6] = {0, 0, 0, 0, 0, 0};
    int32_t tlasNumInternal = 0;
    DevBuf tlasNodes, tlasPrimIds;
    bool mixed = false;  // an instance's LoD holds quads or spheres
};

struct PathTracer;
struct PtBatch;

// A texture of the context: RGBA32F texels (missing channels are 0), row-major,
// kept on the host too (background flux and summed-area table are host-built).
struct Texture {
    int32_t width = 0, height = 0, channels = 0, sampling = 0;
    std::vector<float> host;  // 4 * width * height
    DevBuf dev;
};

struct Context {
    int device = 0;
    hipStream_t stream = nullptr;
    hipEvent_t evBegin = nullptr, evEnd = nullptr;
    std::vector<Lod*> lods;
    std::vector<Scene*> scenes;
    std::vector<PathTracer*> tracers;
    std::vector<PtBatch*> batches;
    std::vector<Texture*> textures;
    BuildScratch scratch;
    int lastBuilt = -1;
    // LoD builds keep collapsed subtrees of more than this many triangles
    // split in their traversal hierarchy (0: traverse the reference BVH)
    int traversalLeafLimit = 16;
};

// Device view of one LoD, as the traversal kernels read it.
struct DevLod {
    const float4* triVerts;
    const void* nodes;
    const float4* quadV3;  // fourth quad vertices per sorted position (null without quads)
    int32_t numInternal;
    int32_t numPrims;
    float box[6];
    float coordMax;  // largest |coordinate| of the primitives
    int32_t pad;
};

void destroy_tracers(Context& ctx);
void upload(Context& c, DevBuf& b, const void* src, size_t bytes);

// LBVH over a triangle LoD; fills lod.nodes / primIds / triVerts / numInternal.
void build_lod_bvh(Context& ctx, Lod& lod, const float* bbox, StageSink* sink);
// LBVH over the instances of a scene (instance-level hierarchy).
void build_scene_bvh(Context& ctx, Scene& sc, StageSink* sink);
// Device address of the allocation-free build's static scratch array on the
// current device (looked up once per device; the device must be current).
void noalloc_scratch_init(int device);
size_t noalloc_scratch_bytes();
// LBVH of a triangle LoD in caller-owned device memory (dVerts: float xyz,
// dTris: u32 x3), scratch in the static array, result into dNodes (64 bytes
// per internal node, room for n - 1) and dPrimIds (n). Returns the number of
// internal nodes of the collapsed BVH. Allocates nothing.
int build_lod_bvh_noalloc(Context& ctx, const float* dVerts, int numVerts, const uint32_t* dTris, int n,
                          const float* bbox, void* dNodes, int32_t* dPrimIds);

// Stable LSD radix sort of (u64 key, i32 value) pairs; result ends in keysA/idsA.
void radix_sort_pairs(uint64_t* keysA, uint64_t* keysB, int32_t* idsA, int32_t* idsB, int n, int beginBit,
                      int endBit, BuildScratch& s, hipStream_t stream);
// Scratch bytes a radix sort of n pairs over [beginBit, endBit) reserves:
// digit counts (+ group sums) in BuildScratch::hist, scan partials in
// BuildScratch::partials (callers that must not allocate size them up front).
size_t radix_sort_hist_bytes(int n, int beginBit, int endBit);
size_t radix_sort_partials_bytes(int n);
// Inclusive prefix sum of int32 values, in place allowed.
void scan_inclusive_i32(const int32_t* in, int32_t* out, int n, BuildScratch& s, hipStream_t stream);
// The same over two arrays of n values at once, in place.
void scan_inclusive_i32_pair(int32_t* a, int32_t* b, int n, BuildScratch& s, hipStream_t stream);
// Exclusive prefix sum of uint32 values, in place allowed.
void scan_exclusive_u32(const uint32_t* in, uint32_t* out, int n, BuildScratch& s, hipStream_t stream);

struct TraceParams {
    int32_t numInstances;
    const float* worldToInst;
    const int32_t* instLod;
    const DevLod* lods;
    const void* tlasNodes;
    const int32_t* tlasPrimIds;
    int32_t tlasNumInternal;
    float tlasCoordMax;  // largest |coordinate| of the scene box (bounds the instance-level node boxes)
    int32_t mixed;       // an instance's LoD holds quads or spheres (kernels with kMixed)
};
// Traversal over the LoDs' traversal hierarchies, or (reference = true) over
// the reference BVHs, whose visit counts follow the reference's order.
TraceParams trace_params(const Scene& sc, bool reference = false);
void launch_trace_closest(const TraceParams& p, const float* rays, int n, float4* hits, float2* bary, uint2* stats,
                          hipStream_t stream);
void launch_trace_any(const TraceParams& p, const float* rays, int n, int32_t* occluded, hipStream_t stream);
// Re-upload the LoD table of a scene (and rebuild its instance level) when a
// LoD was rebuilt after the scene was created.
void refresh_scene(Context& ctx, Scene& sc);
// Inverse of an affine row-major 3x4 transform as the reference computes it
// on the host (descriptors.hpp:108-129): adjugate of the 3x3 part over its
// determinant, translation -(inverse * t).
void instance_to_world(const float* w, float* o);

}  // namespace mfh
