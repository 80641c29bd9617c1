// Host-built light tree of the path tracer, in the byte layout of Mufflon's
// LightSubTree (light_tree.hpp:56-110): internal 32-byte nodes first, then the
// light records (AreaLightTriangle 64 B, AreaLightQuad 96 B, AreaLightSphere
// 32 B, PointLight 32 B, lights.hpp:46-137).
// As in Mufflon, the tree is built once on the host per scene and copied to
// the device verbatim; every device path reads the same bytes.
#pragma once
#include <cstdint>
#include <vector>

namespace mfh {

enum LightType : uint16_t {
    kPointLight = 0,
    kSpotLight = 1,
    kAreaLightTriangle = 2,
    kAreaLightQuad = 3,
    kAreaLightSphere = 4,
    kDirectionalLight = 5,
    kEnvmapLight = 6,
    kInternalNode = 0xFFFF,
};

// One emissive primitive of one instance in world space
// (AreaLightTriangleDesc / AreaLightQuadDesc / AreaLightSphereDesc,
// lights.hpp:77-91, with the primitive it was made from).
constexpr int32_t kAreaTriangle = 0;
constexpr int32_t kAreaQuad = 1;
constexpr int32_t kAreaSphere = 2;
struct HostAreaLight {
    int32_t kind = kAreaTriangle;
    float points[4][3] = {};  // triangle: 3 vertices, quad: 4, sphere: the center in points[0]
    float uv[4][2] = {};
    float radius = 0.0f;      // sphere
    uint16_t material = 0;
    int32_t instance = -1, primitive = -1;
};
using HostAreaTriangle = HostAreaLight;
struct HostPointLight {
    float position[3];
    float intensity[3];
};

// Emission of a material at any uv (the benchmark materials carry constant
// emission: a 1x1 texture times a scale).
struct EmissionLookup {
    virtual ~EmissionLookup() = default;
    virtual void radiance(uint16_t material, float u, float v, float out[3]) const = 0;
};

struct LightTreeHost {
    // positional subtree
    std::vector<uint8_t> mem;
    float posRootFlux = 0.0f;
    uint16_t posRootType = kInternalNode;
    uint32_t posInternalCount = 0;
    uint32_t posLightCount = 0;
    // directional subtree and background (absent in the supported scenes)
    float dirRootFlux = 0.0f;
    float backgroundFlux = 0.0f;  // sum of the background flux channels
    bool sphereLights = false;  // the connect pdf of a hit then depends on the previous vertex
    // primitive (instance << 32 | primitive) -> left-aligned tree path, sorted by key
    std::vector<uint64_t> mapKeys;
    std::vector<uint32_t> mapCodes;
};

// Background of the light tree (BackgroundDesc, background.hpp): COLORED
// (type 0) or a polar ENVMAP (type 1) with its summed-area table.
struct BackgroundHost {
    int32_t type = 0;
    float color[3] = {0.0f, 0.0f, 0.0f};
    float scale[3] = {1.0f, 1.0f, 1.0f};  // Spectrum m_scale (background.hpp:119-123)
    int32_t envmapTexture = -1;
    int32_t width = 0, height = 0;
    std::vector<float> sat;  // width * height, R32F
    float flux[3] = {0.0f, 0.0f, 0.0f};
    float fluxSum = 0.0f;    // ei::sum(flux), the background's weight in the tree
};

// Background::acquire_const (background.cpp:30-66): flux for the scene box and,
// for ENVMAP, create_summed_area_table (texture_sampling.hpp:49-110).
// envRgba: 4 floats per texel (NULL for COLORED).
void build_background(int32_t type, const float color[3], const float scale[3], const float* envRgba, int32_t width,
                      int32_t height, const float sceneBox[6], BackgroundHost& out);

void build_light_tree(const std::vector<HostAreaLight>& area, const std::vector<HostPointLight>& points,
                      const float sceneBox[6], const EmissionLookup& emission, LightTreeHost& out);

}  // namespace mfh
