// Device-side description of a path-tracer scene (the role of Mufflon's
// SceneDescriptor + PtParameters + render buffers for the pt renderer).
#pragma once
#include "mfh_device.hpp"
#include "mfh_internal.hpp"

namespace mfh {

constexpr float kPi = 3.14159265358979323846f;
constexpr float kMaxSceneSize = 1048576.0f;  // types.hpp:26
constexpr float kSceneEps = 1e-4f;          // intersection.cu:16
constexpr float kInfPdf = 4294967296.f;      // AreaPdf::infinite() (util/types.hpp:52-54)
constexpr float kDistSqEps = 1e-10f;        // light_sampling.hpp:78

constexpr int kMatEmissive = 0;  // Materials enum (material_definitions.hpp:25-39)
constexpr int kMatLambert = 1;
constexpr int kMatTorrance = 3;
constexpr int kMatWalter = 4;
constexpr int kMatMicrofacet = 10;
// half-vector-based materials with a roughness (texture); materials with an
// inner medium of their own (Material::compute_medium)
__host__ __device__ constexpr bool mat_has_roughness(int t) {
    return t == kMatTorrance || t == kMatWalter || t == kMatMicrofacet;
}
__host__ __device__ constexpr bool mat_has_inner_medium(int t) { return t == kMatWalter || t == kMatMicrofacet; }
constexpr int kNdfGgx = 0;
constexpr int kNdfBeckmann = 1;
constexpr int kShadowingVCavity = 0;
constexpr int kShadowingSmith = 1;

// Render targets: Radiance, Hit ID, Position, Depth, Normal, UV, Albedo, Lightness
constexpr int kNumTargets = 8;
constexpr int kTargetChannels[kNumTargets] = {3, 2, 3, 1, 3, 2, 3, 1};
constexpr int kTargetRadiance = 0;  // MFH_TARGET_RADIANCE

constexpr uint16_t kLightPoint = 0;
constexpr uint16_t kLightAreaTriangle = 2;
constexpr uint16_t kLightAreaQuad = 3;
constexpr uint16_t kLightAreaSphere = 4;
constexpr uint16_t kLightInternalNode = 0xFFFF;

// mfh_material of include/mufflon_hip.h (128 bytes)
struct PtMaterialRecord {
    int32_t type;
    float albedo[3];
    float radiance[3];
    float scale[3];
    float roughness[2];
    int32_t ndf;
    int32_t shadowing;
    float ior;
    float absorption[3];
    float outerIor;
    float outerAbsorption[3];
    int32_t albedoTexture;
    int32_t roughnessTexture;
    int32_t emissionTexture;
    int32_t reserved[7];
};
static_assert(sizeof(PtMaterialRecord) == 128, "material record");

// A texture as the kernels read it: RGBA32F texels, row-major.
struct DevTexture {
    const float4* texels;
    int32_t width, height, channels, sampling;
    int32_t pad[2];
};

// Background (BackgroundDesc): COLORED (0) or polar ENVMAP (1).
struct DevBackground {
    int32_t type;
    float color[3];
    float scale[3];  // Spectrum scale per channel
    int32_t envmapTexture;
    const float* sat;  // width * height summed-area table (ENVMAP)
};

// Object-space shading attributes of one LoD (primitive ids: triangles,
// quads, spheres).
struct DevPtLod {
    const float* verts;     // xyz per vertex
    const uint32_t* tris;   // 3 per triangle
    const float* normals;   // xyz per vertex
    const float* uvs;       // uv per vertex
    const uint16_t* matIdx; // per primitive
    const uint32_t* quads;  // 4 per quad
    const float4* spheres;  // center xyz, radius
    int32_t numTris, numQuads;
};

struct PtScene {
    TraceParams trace;
    const DevPtLod* ptLods;
    const PtMaterialRecord* materials;
    const DevTexture* textures;
    DevBackground background;
    float sceneBox[6];
    // light tree (Mufflon layout, positional subtree)
    const uint8_t* lightMem;
    float posRootFlux;
    uint32_t posRootType;
    float dirRootFlux;
    float bgFlux;
    const uint64_t* mapKeys;
    const uint32_t* mapCodes;
    int32_t mapCount;
    // pinhole camera (PinholeParams, pinhole.hpp:43-51)
    float camPos[3];
    float tanVFov;
    float camDir[3];
    float camUp[3];
    int32_t width, height;
    // PtParameters (pt_params.hpp:8-14)
    int32_t seed, minPathLength, maxPathLength, neeCount;
    int32_t posGuide;  // "NEE position guide": distance-weighted light tree traversal
    // per-pixel state and targets
    uint64_t* rngs;
    float* cum[kNumTargets];  // running means of the enabled targets (null when disabled)
    float* var[kNumTargets];  // running variances of the targets in varianceMask (null otherwise)
    uint32_t targetMask;
    uint32_t varianceMask;    // subset of targetMask
    uint32_t* nanCounter;
    int32_t iteration;   // index of the iteration being run (0 after a reset)
};

void launch_pt_init_rng(uint64_t* rngs, int num, int seed, hipStream_t stream);
void launch_pt_iterate(const PtScene& S, int leafBatch, hipStream_t stream);
// One iteration of the wireframe renderer into S.cum[MFH_TARGET_RADIANCE]
// (and S.var[...] when recorded), S.iteration the iteration index.
void launch_wireframe(const PtScene& S, const uint32_t* seeds, float lineWidth, int maxTraceDepth,
                      hipStream_t stream);
// One iteration of the light tracer: clears the 5-channel splat image
// (Radiance, Lightness, Density per pixel), traces width * height light
// paths from S.rngs and folds the splats into mean (and var when non-null)
// as iteration S.iteration.
void launch_lighttracer(const PtScene& S, float* splat, float* mean, float* var, int minPathLength,
                        int maxPathLength, bool density, hipStream_t stream);

}  // namespace mfh
