// Mufflon's core C API (core_interface.h / interface.cpp) for the "Pathtracer"
// renderer on the HIP device: a host-side world (objects, LoDs, instances,
// materials, textures, cameras, lights, scenarios) that is uploaded through
// the mfh_* C-ABI when a scenario is loaded, and the render / target /
// parameter calls of the reference's front-ends (bindings.py, the GUI).
//
// The world lives on the host exactly as the reference's WorldContainer
// keeps it (world_container.cpp); loading a scenario (world_load_scenario,
// or render_iterate after a change, interface.cpp:3611-3684) creates the
// device context on first use, uploads every LoD of an unmasked instance,
// builds the BVHs, the instance level, the light tree and the background,
// and creates the path tracer. Any later change of the world marks the scene
// invalid and the next iteration reloads it.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <filesystem>
#include <fstream>
#include <limits>
#include <map>
#include <memory>
#include <mutex>
#include <set>
#include <stdexcept>
#include <string>
#include <vector>

#include "mfh_internal.hpp"
#include "mufflon_core.h"
#include "mufflon_hip.h"

// The descriptors keep the reference's C layout (core_interface.h:205-318).
static_assert(sizeof(MaterialParams) == 88, "MaterialParams layout");
static_assert(sizeof(LightHdl) == 4, "LightHdl layout");

namespace {

constexpr uint32_t kNoAnimationFrame = 0xFFFFFFFFu;  // Instance::NO_ANIMATION_FRAME
constexpr MatIdx kInvalidMaterial = 0xFFFF;          // INVALID_MATERIAL
constexpr int32_t kInvalidIndex = -1;

// ------------------------------------------------------------------ logging
// delegateLog (interface.cpp:156-166): errors become the DLL error and go to
// the logger callback with their severity.
std::mutex gLogMutex;
void (*gLogCallback)(const char*, int) = nullptr;
int gLogLevel = LOG_INFO;
std::string gLastError;
thread_local std::string tlErrorCopy;

void log_message(int severity, const std::string& msg) {
    std::lock_guard<std::mutex> lock(gLogMutex);
    if (severity >= LOG_ERROR) gLastError = msg;
    if (gLogCallback && severity >= gLogLevel) gLogCallback(msg.c_str(), severity);
}
void log_error(const char* fn, const std::string& msg) { log_message(LOG_ERROR, std::string("[") + fn + "] " + msg); }
void log_warning(const char* fn, const std::string& msg) {
    log_message(LOG_WARNING, std::string("[") + fn + "] " + msg);
}

struct ApiError : std::runtime_error {
    using std::runtime_error::runtime_error;
};
#define CORE_REQUIRE(cond, msg)                    \
    do {                                           \
        if (!(cond)) throw ApiError(msg);          \
    } while (0)

// TRY / CATCH_ALL (interface.cpp:54-83): no exception leaves the API.
template <class R, class F>
R guard(const char* fn, R fail, F&& f) {
    try {
        return f();
    } catch (const std::exception& e) {
        log_error(fn, e.what());
        return fail;
    }
}

void check_mfh(int status, const char* what) {
    if (status != MFH_OK) throw ApiError(std::string(what) + ": " + mfh_last_error());
}

// ------------------------------------------------------------------ world
struct Instance;

struct Texture {
    std::string name;
    int32_t width = 1, height = 1, channels = 4;  // channels of the stored format
    TextureSampling sampling = SAMPLING_NEAREST;
    std::vector<float> texels;                    // width * height * channels
};

struct Lod {
    Instance* owner = nullptr;
    std::vector<float> points, normals, uvs;  // 3, 3, 2 floats per vertex
    std::vector<uint32_t> tris;               // 3 per triangle
    std::vector<uint32_t> quads;              // 4 per quad (added after every triangle)
    std::vector<uint16_t> mat;                // per face: triangles, then quads (scenario material slot)
    std::vector<float> spheres;               // center xyz, radius
    std::vector<uint16_t> sphereMat;          // per sphere
    // the spheres' box (Spheres::m_boundingBox; the polygons' is box)
    float sphereBox[6] = {std::numeric_limits<float>::max(), std::numeric_limits<float>::max(),
                          std::numeric_limits<float>::max(), -std::numeric_limits<float>::max(),
                          -std::numeric_limits<float>::max(), -std::numeric_limits<float>::max()};
    float box[6] = {std::numeric_limits<float>::max(), std::numeric_limits<float>::max(),
                    std::numeric_limits<float>::max(), -std::numeric_limits<float>::max(),
                    -std::numeric_limits<float>::max(), -std::numeric_limits<float>::max()};
    size_t vertices() const { return points.size() / 3; }
    size_t triangles() const { return tris.size() / 3; }
    size_t quad_count() const { return quads.size() / 4; }
    size_t sphere_count() const { return spheres.size() / 4; }
    void grow_box(const float* p) {
        for (int k = 0; k < 3; ++k) {
            box[k] = p[k] < box[k] ? p[k] : box[k];
            box[3 + k] = box[3 + k] < p[k] ? p[k] : box[3 + k];
        }
    }
    void grow_sphere_box(const float* s) {
        for (int k = 0; k < 3; ++k) {
            const float lo = s[k] - s[3], hi = s[k] + s[3];
            sphereBox[k] = lo < sphereBox[k] ? lo : sphereBox[k];
            sphereBox[3 + k] = sphereBox[3 + k] < hi ? hi : sphereBox[3 + k];
        }
    }
    // Lod::get_bounding_box: the polygons' box joined with the spheres'
    void lod_box(float out[6]) const {
        for (int k = 0; k < 3; ++k) {
            out[k] = sphereBox[k] < box[k] ? sphereBox[k] : box[k];
            out[3 + k] = box[3 + k] < sphereBox[3 + k] ? sphereBox[3 + k] : box[3 + k];
        }
    }
};

struct Object {
    std::string name;
    uint32_t flags = 0;
    uint32_t id = 0;
    std::vector<std::unique_ptr<Lod>> lods;  // by level (nullptr: no LoD at that level)
};

struct ObjInstance {
    Object* object = nullptr;
    uint32_t frame = kNoAnimationFrame;
    float worldToInst[12] = {1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0};
};

struct Material {
    std::string name;
    MaterialParams params;
    mfh_material record;        // textures resolved at load
    Texture* textures[3] = {};  // albedo, roughness, emission of the record
};

struct Camera {
    std::string name;
    std::vector<Vec3> position, dir, up;  // per path segment; dir and up orthonormalized
    float nearPlane = 1e-10f, farPlane = 1e10f, vFov = 0.0f, tanVFov = 0.0f;
};

struct PointLight {
    std::string name;
    std::vector<Vec3> position, intensity;  // per path segment
};

struct Background {
    std::string name;
    BackgroundType type = BACKGROUND_MONOCHROME;
    Vec3 color{0, 0, 0};
    Vec3 scale{1, 1, 1};
    Texture* map = nullptr;
};

struct Scenario {
    std::string name;
    uint32_t width = 800, height = 600;
    Camera* camera = nullptr;
    LodLevel globalLod = 0;
    std::vector<std::string> slots;
    std::vector<Material*> assigned;
    std::vector<LightHdl> lights;
    std::set<const Object*> maskedObjects;
    std::set<const ObjInstance*> maskedInstances;
    std::map<const Object*, LodLevel> objectLod;
    std::map<const ObjInstance*, LodLevel> instanceLod;
};

// The "Pathtracer" renderer: parameters (PtParameters, pt_params.hpp:8-14;
// defaults parameter.hpp:327-364) per variation.
struct Variation {
    int32_t seed = 0, minPathLength = 0, maxPathLength = 16, neeCount = 1;
    bool neePositionGuide = false;
};

struct ParamDescription {
    const char* name;
    ParameterType type;
};
constexpr ParamDescription kParams[] = {{"Seed", PARAM_INT},
                                        {"Min. path length", PARAM_INT},
                                        {"Max. path length", PARAM_INT},
                                        {"NEE count", PARAM_INT},
                                        {"NEE position guide", PARAM_BOOL}};
constexpr int kNumParams = 5;

// The "Wireframe" renderer's parameters (WireframeParameters,
// wireframe_params.hpp:10-36) and its single target (BorderTarget, :38-43).
struct WireParams {
    float lineWidth = 1.0f;
    int32_t maxTraceDepth = 1000;
};
constexpr ParamDescription kWireParams[] = {{"Line width", PARAM_FLOAT}, {"Maximum trace depth", PARAM_INT}};
constexpr int kNumWireParams = 2;
constexpr int kRendererPt = 0, kRendererWire = 1, kRendererLt = 2;
// The "Lighttracer" renderer's parameters (LtParameters, lt_params.hpp:8-12):
// the first three of kParams.
constexpr int kNumLtParams = 3;

// PtTargets order (pt_params.hpp:22-26) and the device target ids.
struct TargetName {
    const char* name;
    int32_t device;
    int channels;
};
constexpr TargetName kTargets[] = {{"Radiance", MFH_TARGET_RADIANCE, 3}, {"Position", MFH_TARGET_POSITION, 3},
                                   {"Depth", MFH_TARGET_DEPTH, 1},       {"Normal", MFH_TARGET_NORMAL, 3},
                                   {"UV", MFH_TARGET_UV, 2},             {"Albedo", MFH_TARGET_ALBEDO, 3},
                                   {"Lightness", MFH_TARGET_LIGHTNESS, 1}, {"Hit ID", MFH_TARGET_HIT_ID, 2}};
constexpr int kNumTargetNames = 8;
constexpr TargetName kWireTargets[] = {{"Border", -1, 3}};
// LtTargets (lt_params.hpp:14-17)
constexpr TargetName kLtTargets[] = {{"Radiance", MFH_LT_RADIANCE, 3}, {"Lightness", MFH_LT_LIGHTNESS, 1},
                                     {"Density", MFH_LT_DENSITY, 1}};
constexpr int kLtDensityTarget = 2;

struct TargetTable {
    const TargetName* targets;
    int count;
};
TargetTable target_table(int renderer) {
    if (renderer == kRendererWire) return TargetTable{kWireTargets, 1};
    if (renderer == kRendererLt) return TargetTable{kLtTargets, 3};
    return TargetTable{kTargets, kNumTargetNames};
}

int target_index(int renderer, const char* name) {
    if (!name) return -1;
    const TargetTable tt = target_table(renderer);
    for (int i = 0; i < tt.count; ++i)
        if (std::strcmp(tt.targets[i].name, name) == 0) return i;
    return -1;
}

struct Core {
    std::mutex iterationMutex;
    std::vector<std::unique_ptr<Object>> objects;
    std::vector<std::unique_ptr<ObjInstance>> instances;
    std::vector<std::unique_ptr<Material>> materials;
    std::vector<std::unique_ptr<Texture>> textures;
    std::vector<std::unique_ptr<Camera>> cameras;
    std::vector<std::unique_ptr<PointLight>> pointLights;
    std::vector<std::unique_ptr<Background>> backgrounds;
    std::vector<std::unique_ptr<Scenario>> scenarios;
    uint32_t frameCurrent = 0;
    float worldBox[6] = {0, 0, 0, 0, 0, 0};
    Scenario* current = nullptr;
    bool sceneValid = false;
    // device state of the loaded scene
    mfh_context* ctx = nullptr;
    int32_t pt = -1, scene = -1;
    float sceneBox[6] = {0, 0, 0, 0, 0, 0};
    std::map<const Texture*, int32_t> texIds;  // device texture of every texture larger than 1x1 in use
    const Texture* envMap = nullptr;           // the loaded scene's environment map
    int32_t adjointMode = 0;                   // MFH_ADJOINT_*
    // renderer state
    int variation = -1;  // -1: no renderer enabled
    int renderer = kRendererPt;  // kRendererPt (variations 0, 1) or kRendererWire (variation 0)
    Variation vars[2];
    Variation ltVars;  // the light tracer's own parameters
    WireParams wire;
    uint32_t targetMask = 0, varianceMask = 0;  // bits: index into kTargets
    std::set<std::pair<std::string, bool>> toEnable, toDisable;
    int32_t iteration = -1;  // OutputHandler::m_iteration
    std::vector<float> screen;
    int screenChannels = 0;
    // the scene-file loader's on-demand LoD callback (mufflon_set_lod_loader)
    Boolean (*lodLoader)(void*, ObjectHdl, uint32_t, Boolean) = nullptr;
    Boolean (*matIndexLoader)(void*, uint32_t, uint16_t*, uint32_t*) = nullptr;
    Boolean (*lodMetaLoader)(void*, LodMetadata*, size_t*) = nullptr;
    void* loaderParams = nullptr;

    ~Core() { release_device(); }
    void release_device() {
        if (ctx) (void)mfh_destroy(ctx);
        ctx = nullptr;
        pt = scene = -1;
        texIds.clear();
        envMap = nullptr;
    }
    void invalidate() { sceneValid = false; }
};

Core& core_of(void* h) {
    if (!h) throw ApiError("Invalid mufflon instance (nullptr)");
    return *static_cast<Core*>(h);
}

std::mutex gCoresMutex;
std::set<Core*> gCores;

// The instance that owns a LoD, so that LoD-level calls can invalidate its scene.
std::map<const void*, Core*> gLodOwner;  // LodHdl -> Core

Lod& lod_of(void* h) {
    if (!h) throw ApiError("Invalid LoD handle (nullptr)");
    return *static_cast<Lod*>(h);
}
void touch_lod(const void* h) {
    std::lock_guard<std::mutex> lock(gCoresMutex);
    auto it = gLodOwner.find(h);
    if (it != gLodOwner.end()) it->second->invalidate();
}

void light_check(const LightHdl& h, LightType type, size_t count) {
    if (h.type != (uint32_t)type) throw ApiError("light handle has the wrong type");
    if (h.index >= count) throw ApiError("light handle out of range");
}

LightHdl invalid_light() {
    LightHdl h;
    h.type = 7;
    h.index = 0;
    return h;
}

uint32_t segment(size_t count, uint32_t frame) { return count == 0 ? 0 : std::min<uint32_t>((uint32_t)count - 1, frame); }

// Camera::set_view_dir (camera.hpp:81-97): normalized direction, up made
// orthogonal to it.
Vec3 v3add(Vec3 a, Vec3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
Vec3 v3sub(Vec3 a, Vec3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
Vec3 v3mul(float s, Vec3 a) { return {s * a.x, s * a.y, s * a.z}; }
float v3dot(Vec3 a, Vec3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
Vec3 v3norm(Vec3 a) {
    const float l = std::sqrt(v3dot(a, a));
    return {a.x / l, a.y / l, a.z / l};
}
void view_space(Vec3 dir, Vec3 up, Vec3& outDir, Vec3& outUp) {
    dir = v3norm(dir);
    up = v3norm(up);
    const float zDotY = v3dot(dir, up);
    if (zDotY > 0.999f) throw ApiError("View direction and up-vector are too close to each other");
    up = v3norm(v3sub(up, v3mul(zDotY, dir)));
    outDir = dir;
    outUp = up;
}

// ------------------------------------------------------------------ textures
// Channels of the stored format for a value texture (world_add_texture_value,
// interface.cpp:1938-1980: 1 -> R32F, 2 -> RG32F, 3 and 4 -> RGBA32F).
int value_channels(int num) { return num <= 2 ? num : 4; }

// PFM reader (pfmloader/interface.cpp:75-160): "PF" -> RGBA32F (alpha 0),
// "Pf" -> R32F, rows in file order, byte order from the sign of the scale.
std::unique_ptr<Texture> load_pfm(const std::string& path) {
    std::ifstream f(path, std::ios::binary);
    if (!f) throw ApiError("Could not open texture file '" + path + "'");
    char bands[2];
    f.read(bands, 2);
    if (!f || bands[0] != 'P' || (bands[1] != 'F' && bands[1] != 'f'))
        throw ApiError("Unknown bands description in '" + path + "'");
    int width = 0, height = 0;
    float scale = 0.0f;
    f >> width >> height >> scale;
    if (!f || width <= 0 || height <= 0) throw ApiError("Invalid width/height in '" + path + "'");
    char c = (char)f.get();
    if (c == '\r') c = (char)f.get();
    if (c != '\n') throw ApiError("Expected newline in header of '" + path + "'");
    const bool gray = bands[1] == 'f';
    const int inCh = gray ? 1 : 3;
    std::vector<float> raw((size_t)width * height * inCh);
    f.read(reinterpret_cast<char*>(raw.data()), (std::streamsize)(raw.size() * sizeof(float)));
    if (!f) throw ApiError("Truncated pixel data in '" + path + "'");
    const bool fileLittle = scale < 0.0f;
    const uint16_t probe = 1;
    const bool hostLittle = *reinterpret_cast<const uint8_t*>(&probe) == 1;
    if (fileLittle != hostLittle)
        for (float& v : raw) {
            uint32_t u;
            std::memcpy(&u, &v, 4);
            u = __builtin_bswap32(u);
            std::memcpy(&v, &u, 4);
        }
    auto t = std::make_unique<Texture>();
    t->name = path;
    t->width = width;
    t->height = height;
    t->channels = gray ? 1 : 4;
    t->texels.assign((size_t)width * height * t->channels, 0.0f);
    for (size_t i = 0; i < (size_t)width * height; ++i)
        for (int k = 0; k < inCh; ++k) t->texels[i * t->channels + k] = raw[i * inCh + k];
    return t;
}

// PFM writer (pfmloader/interface.cpp:162-210): one channel as "Pf",
// otherwise the first three channels as "PF", little endian, rows in order.
void store_pfm(const std::string& path, const float* data, int width, int height, int channels) {
    std::ofstream f(path, std::ios::binary);
    if (!f) throw ApiError("Failed to open screenshot file '" + path + "'");
    f << (channels == 1 ? "Pf\n" : "PF\n") << width << " " << height << "\n-1.000000\n";
    for (size_t i = 0; i < (size_t)width * height; ++i) {
        if (channels == 1) {
            f.write(reinterpret_cast<const char*>(data + i), 4);
        } else {
            float px[3] = {0.0f, 0.0f, 0.0f};
            for (int k = 0; k < std::min(channels, 3); ++k) px[k] = data[i * channels + k];
            f.write(reinterpret_cast<const char*>(px), 12);
        }
    }
    if (!f) throw ApiError("Failed to write screenshot file '" + path + "'");
}

// ------------------------------------------------------------------ materials
// convert_material (interface.cpp:1259-1376) onto the device record.
void texture_constant(const Texture* t, float* out, int n) {
    for (int k = 0; k < n; ++k) out[k] = k < t->channels ? t->texels[k] : 0.0f;
}

Material make_material(const char* name, const MaterialParams& p) {
    Material m;
    m.name = name;
    m.params = p;
    std::memset(&m.record, 0, sizeof(m.record));
    mfh_material& r = m.record;
    r.albedoTexture = r.roughnessTexture = r.emissionTexture = -1;
    r.outerRefractionIndex = p.outerMedium.refractionIndex.x;
    r.outerAbsorption[0] = p.outerMedium.absorption.x;
    r.outerAbsorption[1] = p.outerMedium.absorption.y;
    r.outerAbsorption[2] = p.outerMedium.absorption.z;
    r.refractionIndex = r.outerRefractionIndex;
    std::memcpy(r.absorption, r.outerAbsorption, sizeof(r.absorption));
    CORE_REQUIRE(p.alpha == nullptr, "alpha textures are not supported by the HIP path tracer");
    CORE_REQUIRE(p.displacement.map == nullptr, "displacement maps are not supported by the HIP path tracer");
    auto ndf = [](NormalDistFunction n) {
        if (n == NDF_BECKMANN) return (int32_t)MFH_NDF_BECKMANN;
        if (n == NDF_GGX) return (int32_t)MFH_NDF_GGX;
        throw ApiError("the COSINE normal distribution is not supported by the HIP path tracer");
    };
    auto shadowing = [](ShadowingModel s) {
        CORE_REQUIRE(s == SHADOWING_VCAVITY || s == SHADOWING_SMITH, "unknown shadowing model");
        return s == SHADOWING_SMITH ? (int32_t)MFH_SHADOWING_SMITH : (int32_t)MFH_SHADOWING_VCAVITY;
    };
    auto tex = [](TextureHdl h) {
        CORE_REQUIRE(h != nullptr, "material texture handle is null");
        return static_cast<Texture*>(h);
    };
    switch (p.innerType) {
        case MATERIAL_LAMBERT:
            r.type = MFH_MATERIAL_LAMBERT;
            m.textures[0] = tex(p.inner.lambert.albedo);
            break;
        case MATERIAL_TORRANCE:
            r.type = MFH_MATERIAL_TORRANCE;
            m.textures[0] = tex(p.inner.torrance.albedo);
            m.textures[1] = tex(p.inner.torrance.roughness);
            r.ndf = ndf(p.inner.torrance.ndf);
            r.shadowing = shadowing(p.inner.torrance.shadowingModel);
            break;
        case MATERIAL_WALTER:      // interface.cpp:1275-1279
        case MATERIAL_MICROFACET:  // the same parameters (interface.cpp:1280-1284)
            r.type = p.innerType == MATERIAL_WALTER ? MFH_MATERIAL_WALTER : MFH_MATERIAL_MICROFACET;
            m.textures[1] = tex(p.inner.walter.roughness);
            r.ndf = ndf(p.inner.walter.ndf);
            r.shadowing = shadowing(p.inner.walter.shadowingModel);
            r.refractionIndex = p.inner.walter.refractionIndex;
            r.absorption[0] = p.inner.walter.absorption.x;
            r.absorption[1] = p.inner.walter.absorption.y;
            r.absorption[2] = p.inner.walter.absorption.z;
            break;
        case MATERIAL_EMISSIVE:
            r.type = MFH_MATERIAL_EMISSIVE;
            m.textures[2] = tex(p.inner.emissive.radiance);
            CORE_REQUIRE(m.textures[2]->width == 1 && m.textures[2]->height == 1,
                         "emission textures must be 1x1 values for the HIP path tracer");
            r.scale[0] = p.inner.emissive.scale.x;
            r.scale[1] = p.inner.emissive.scale.y;
            r.scale[2] = p.inner.emissive.scale.z;
            break;
        case MATERIAL_ORENNAYAR:
        case MATERIAL_BLEND:
        case MATERIAL_FRESNEL:
            throw ApiError("material type " + std::to_string((int)p.innerType) +
                           " is not supported by the HIP path tracer (LAMBERT, TORRANCE, WALTER, MICROFACET and "
                           "EMISSIVE are)");
        default:
            throw ApiError("Invalid material params: type unknown.");
    }
    return m;
}

// Texture slots of a material record: constants for 1x1 textures, device
// ids (from `ids`) otherwise.
void resolve_textures(Material& m, const std::map<const Texture*, int32_t>& ids) {
    mfh_material& r = m.record;
    if (const Texture* t = m.textures[0]) {
        if (t->width == 1 && t->height == 1)
            texture_constant(t, r.albedo, 3);
        else
            r.albedoTexture = ids.at(t);
    }
    if (const Texture* t = m.textures[1]) {
        if (t->width == 1 && t->height == 1) {
            r.roughness[0] = t->texels[0];
            r.roughness[1] = t->channels >= 2 ? t->texels[1] : t->texels[0];
        } else {
            r.roughnessTexture = ids.at(t);
        }
    }
    if (const Texture* t = m.textures[2]) texture_constant(t, r.radiance, 3);
}

// ------------------------------------------------------------------ loading
// Scenario::get_effective_lod (scenario.cpp:131-154): instance LoD, else
// object LoD, else the global one, clamped to the object's last LoD slot.
LodLevel lod_level(const Scenario& s, const ObjInstance& inst) {
    LodLevel level = s.globalLod;
    auto ii = s.instanceLod.find(&inst);
    auto oi = s.objectLod.find(inst.object);
    if (ii != s.instanceLod.end()) level = ii->second;
    else if (oi != s.objectLod.end()) level = oi->second;
    const size_t slots = inst.object->lods.size();
    return slots > 0 ? std::min<LodLevel>(level, (LodLevel)(slots - 1)) : level;
}

bool instance_active(const Core& c, const Scenario& s, const ObjInstance& inst) {
    if (inst.frame != kNoAnimationFrame && inst.frame != c.frameCurrent) return false;
    if (s.maskedInstances.count(&inst) || s.maskedObjects.count(inst.object)) return false;
    return true;
}

mfh_pt_params device_params(const Variation& v) {
    mfh_pt_params p;
    std::memset(&p, 0, sizeof(p));
    p.seed = v.seed;
    p.minPathLength = v.minPathLength;
    p.maxPathLength = v.maxPathLength;
    p.neeCount = v.neeCount;
    p.neePositionGuide = v.neePositionGuide ? 1 : 0;
    return p;
}

mfh_pinhole device_camera(const Core& c, const Scenario& s) {
    const Camera& cam = *s.camera;
    const uint32_t k = segment(cam.position.size(), c.frameCurrent);
    mfh_pinhole p;
    std::memset(&p, 0, sizeof(p));
    p.position[0] = cam.position[k].x;
    p.position[1] = cam.position[k].y;
    p.position[2] = cam.position[k].z;
    p.tanVFov = cam.tanVFov;
    p.viewDir[0] = cam.dir[k].x;
    p.viewDir[1] = cam.dir[k].y;
    p.viewDir[2] = cam.dir[k].z;
    p.up[0] = cam.up[k].x;
    p.up[1] = cam.up[k].y;
    p.up[2] = cam.up[k].z;
    p.width = (int32_t)s.width;
    p.height = (int32_t)s.height;
    return p;
}

// The wireframe's Border target: a fresh running mean (with variance when
// requested); the renderer's seed generator keeps running.
void wire_reset(Core& c) {
    if (c.pt < 0) return;
    check_mfh(mfh_pt_wireframe_reset(c.ctx, c.pt, (c.varianceMask & 1u) ? 1 : 0), "reset");
}

// The light tracer's targets: fresh running means (with variances when any
// target records them) and re-seeded generators.
void lt_reset(Core& c) {
    if (c.pt < 0) return;
    check_mfh(mfh_pt_lighttracer_reset(c.ctx, c.pt, c.varianceMask ? 1 : 0), "reset");
}

void apply_targets(Core& c) {
    if (c.pt < 0 || c.renderer != kRendererPt) return;
    for (int i = 0; i < kNumTargetNames; ++i) {
        const bool on = (c.targetMask >> i) & 1u, var = (c.varianceMask >> i) & 1u;
        check_mfh(mfh_pt_enable_target(c.ctx, c.pt, kTargets[i].device, on ? 1 : 0), "enable target");
        check_mfh(mfh_pt_enable_variance(c.ctx, c.pt, kTargets[i].device, var ? 1 : 0), "enable variance");
    }
}

void apply_mode(Core& c) {
    if (c.pt < 0 || c.variation < 0 || c.renderer != kRendererPt) return;
    check_mfh(mfh_pt_set_mode(c.ctx, c.pt, c.variation == 1 ? MFH_PT_WAVEFRONT : MFH_PT_MEGAKERNEL,
                              c.variation == 1 ? 1 : 0),
              "set mode");
}

// WorldContainer::load_scene + RendererBase::load_scene / pre_reset: the
// whole scenario onto the device.
void load_scene(Core& c, Scenario& s) {
    CORE_REQUIRE(s.camera != nullptr, "scenario '" + s.name + "' has no camera");
    // LoDs that active instances need and that are not loaded yet come from the
    // scene-file loader (WorldContainer::load_lod, world_container.cpp:684-693),
    // before anything touches the device
    if (c.lodLoader) {
        for (const auto& ip : c.instances) {
            if (!instance_active(c, s, *ip)) continue;
            Object& obj = *ip->object;
            const LodLevel level = lod_level(s, *ip);
            if (level < obj.lods.size() && obj.lods[level]) continue;
            CORE_REQUIRE(c.lodLoader(c.loaderParams, &obj, level, false),
                         "the scene loader failed to load LoD " + std::to_string(level) + " of object '" + obj.name +
                             "'");
        }
    }
    c.release_device();
    check_mfh(mfh_create(0, &c.ctx), "device context");
    // textures used by the materials and the background
    std::map<const Texture*, int32_t> ids;
    auto upload_texture = [&](const Texture* t) {
        if (!t || (t->width == 1 && t->height == 1) || ids.count(t)) return;
        int32_t id = -1;
        check_mfh(mfh_texture_create(c.ctx, t->width, t->height, t->channels, (int32_t)t->sampling, t->texels.data(),
                                     &id),
                  "texture upload");
        ids[t] = id;
    };
    std::vector<mfh_material> mats;
    for (size_t i = 0; i < s.slots.size(); ++i) {
        Material* m = i < s.assigned.size() ? s.assigned[i] : nullptr;
        CORE_REQUIRE(m != nullptr, "material slot '" + s.slots[i] + "' of scenario '" + s.name +
                                       "' has no material assigned");
        for (const Texture* t : m->textures) upload_texture(t);
        resolve_textures(*m, ids);
        mats.push_back(m->record);
    }
    CORE_REQUIRE(!mats.empty(), "scenario '" + s.name + "' declares no material slots");
    // geometry: one device LoD per (object, level) in use
    std::map<std::pair<const Object*, LodLevel>, int32_t> lodIds;
    std::vector<int32_t> instLod;
    std::vector<float> w2i;
    for (const auto& ip : c.instances) {
        const ObjInstance& inst = *ip;
        if (!instance_active(c, s, inst)) continue;
        const LodLevel level = lod_level(s, inst);
        const Object& obj = *inst.object;
        CORE_REQUIRE(level < obj.lods.size() && obj.lods[level], "object '" + obj.name + "' has no LoD " +
                                                                     std::to_string(level));
        const auto key = std::make_pair(&obj, level);
        auto it = lodIds.find(key);
        if (it == lodIds.end()) {
            const Lod& L = *obj.lods[level];
            CORE_REQUIRE(!L.mat.empty() || !L.sphereMat.empty(), "LoD " + std::to_string(level) + " of object '" +
                                                                     obj.name + "' has no primitives");
            // material per primitive: triangles, quads, spheres (descriptors.hpp:43, 64)
            std::vector<uint16_t> primMat(L.mat);
            primMat.insert(primMat.end(), L.sphereMat.begin(), L.sphereMat.end());
            for (uint16_t mi : primMat)
                CORE_REQUIRE(mi < mats.size(), "object '" + obj.name + "' uses material slot " + std::to_string(mi) +
                                                   " which the scenario does not declare");
            int32_t id = -1;
            if (L.quads.empty() && L.spheres.empty())
                check_mfh(mfh_lod_create(c.ctx, L.points.data(), (int32_t)L.vertices(), L.tris.data(),
                                         (int32_t)L.triangles(), &id),
                          "LoD upload");
            else
                check_mfh(mfh_lod_create_mixed(c.ctx, L.points.data(), (int32_t)L.vertices(), L.tris.data(),
                                               (int32_t)L.triangles(), L.quads.data(), (int32_t)L.quad_count(),
                                               L.spheres.data(), (int32_t)L.sphere_count(), &id),
                          "LoD upload");
            check_mfh(mfh_lod_set_shading(c.ctx, id, L.normals.data(), L.uvs.data(), primMat.data()), "LoD shading");
            float box[6];
            L.lod_box(box);
            check_mfh(mfh_lbvh_build(c.ctx, id, box), "LBVH build");
            it = lodIds.emplace(key, id).first;
        }
        instLod.push_back(it->second);
        w2i.insert(w2i.end(), inst.worldToInst, inst.worldToInst + 12);
    }
    CORE_REQUIRE(!instLod.empty(), "scenario '" + s.name + "' has no instances");
    check_mfh(mfh_scene_create(c.ctx, instLod.data(), w2i.data(), (int32_t)instLod.size(), &c.scene), "scene");
    int32_t numInternal = 0;
    check_mfh(mfh_scene_get(c.ctx, c.scene, c.sceneBox, &numInternal, nullptr, nullptr), "scene box");
    // lights of the scenario
    std::vector<mfh_point_light> points;
    const Background* bg = nullptr;
    for (const LightHdl& h : s.lights) {
        if (h.type == LIGHT_POINT) {
            const PointLight& L = *c.pointLights.at(h.index);
            const uint32_t k = segment(L.position.size(), c.frameCurrent);
            mfh_point_light p;
            p.position[0] = L.position[k].x;
            p.position[1] = L.position[k].y;
            p.position[2] = L.position[k].z;
            p.intensity[0] = L.intensity[k].x;
            p.intensity[1] = L.intensity[k].y;
            p.intensity[2] = L.intensity[k].z;
            points.push_back(p);
        } else if (h.type == LIGHT_ENVMAP) {
            bg = c.backgrounds.at(h.index).get();
        }
    }
    const mfh_pinhole cam = device_camera(c, s);
    const Variation& v = (c.renderer == kRendererLt && c.variation >= 0) ? c.ltVars
                                                                        : c.vars[c.variation >= 0 ? c.variation : 0];
    const mfh_pt_params params = device_params(v);
    check_mfh(mfh_pt_create(c.ctx, c.scene, mats.data(), (int32_t)mats.size(), points.empty() ? nullptr : points.data(),
                            (int32_t)points.size(), &cam, &params, &c.pt),
              "path tracer");
    if (bg) {
        mfh_background b;
        std::memset(&b, 0, sizeof(b));
        b.envmapTexture = -1;
        b.scale[0] = bg->scale.x;
        b.scale[1] = bg->scale.y;
        b.scale[2] = bg->scale.z;
        if (bg->type == BACKGROUND_ENVMAP) {
            CORE_REQUIRE(bg->map != nullptr, "environment light '" + bg->name + "' has no map");
            upload_texture(bg->map);
            b.type = MFH_BACKGROUND_ENVMAP;
            CORE_REQUIRE(ids.count(bg->map), "environment maps must be larger than 1x1");
            b.envmapTexture = ids.at(bg->map);
        } else {
            b.type = MFH_BACKGROUND_COLORED;
            b.color[0] = bg->color.x;
            b.color[1] = bg->color.y;
            b.color[2] = bg->color.z;
        }
        check_mfh(mfh_pt_set_background(c.ctx, c.pt, &b), "background");
    }
    c.texIds = ids;
    c.envMap = (bg && bg->type == BACKGROUND_ENVMAP) ? bg->map : nullptr;
    check_mfh(mfh_pt_set_adjoint_mode(c.ctx, c.pt, c.adjointMode), "adjoint mode");
    apply_targets(c);
    apply_mode(c);
    check_mfh(mfh_pt_reset(c.ctx, c.pt), "reset");
    if (c.renderer == kRendererWire) wire_reset(c);
    if (c.renderer == kRendererLt) lt_reset(c);
    c.iteration = -1;
    c.screen.clear();
    c.sceneValid = true;
}

// WorldContainer::finalize_world sanity (world_container.cpp:31-56).
const char* world_sanity(const Core& c) {
    if (c.instances.empty()) return "No instances";
    if (c.objects.empty()) return "No objects";
    if (c.pointLights.empty() && c.backgrounds.empty()) {
        bool emitters = false;
        for (const auto& m : c.materials) emitters = emitters || m->params.innerType == MATERIAL_EMISSIVE;
        if (!emitters) return "No lights or emitters";
    }
    if (c.cameras.empty()) return "No camera";
    return "";
}

// Scenario name tags of screenshots (interface.cpp:168-195).
std::string replace_tag(std::string s, const std::string& from, const std::string& to) {
    const size_t at = s.find(from);
    if (at != std::string::npos) s.replace(at, from.size(), to);
    return s;
}

uint64_t cycles() {
#if defined(__x86_64__)
    return __builtin_ia32_rdtsc();
#else
    return 0;
#endif
}
uint64_t micros() {
    return (uint64_t)std::chrono::duration_cast<std::chrono::microseconds>(
               std::chrono::steady_clock::now().time_since_epoch())
        .count();
}

Variation& current_variation(Core& c) {
    CORE_REQUIRE(c.variation >= 0, "Currently, no renderer is set.");
    return c.renderer == kRendererLt ? c.ltVars : c.vars[c.variation];
}

void push_params(Core& c) {
    if (c.pt < 0 || !c.sceneValid) return;
    if (c.renderer == kRendererWire) {  // a parameter change resets the output
        wire_reset(c);
        c.iteration = -1;
        return;
    }
    if (c.renderer == kRendererLt) {
        const mfh_pt_params p = device_params(c.ltVars);
        check_mfh(mfh_pt_set_params(c.ctx, c.pt, &p), "parameters");
        lt_reset(c);
        c.iteration = -1;
        return;
    }
    const mfh_pt_params p = device_params(c.vars[c.variation]);
    check_mfh(mfh_pt_set_params(c.ctx, c.pt, &p), "parameters");
    c.iteration = -1;
}

// Looks a parameter up; unknown names are logged and ignored
// (ParameterHandler::get, parameter.hpp:242-254).
int find_param(const Core& c, const char* fn, const char* name, ParameterType expected) {
    const bool wire = c.renderer == kRendererWire;
    const ParamDescription* params = wire ? kWireParams : kParams;
    const int count = wire ? kNumWireParams : (c.renderer == kRendererLt ? kNumLtParams : kNumParams);
    for (int i = 0; i < count; ++i)
        if (name && std::strcmp(params[i].name, name) == 0) {
            if (params[i].type != expected) {
                log_error(fn, std::string("Parameter '") + name + "' has another type");
                return -2;
            }
            return i;
        }
    log_warning(fn, std::string("Trying to access an unknown parameter '") + (name ? name : "") + "'.");
    return -1;
}
int32_t* int_param(Variation& v, int i) {
    switch (i) {
        case 0: return &v.seed;
        case 1: return &v.minPathLength;
        case 2: return &v.maxPathLength;
        case 3: return &v.neeCount;
        default: return nullptr;
    }
}

}  // namespace

using mfh::instance_to_world;

// Helpers of the API functions below.
namespace {

FaceHdl add_triangle(void* h, UVec3 v, MatIdx idx) {
    Lod& L = lod_of(h);
    const uint32_t n = (uint32_t)L.vertices();
    CORE_REQUIRE(v.x < n && v.y < n && v.z < n, "triangle vertex index out of range");
    // faces keep triangles before quads (polygon.cpp:200, "to keep the order implicitly")
    CORE_REQUIRE(L.quads.empty(), "triangles must be added before the LoD's first quad");
    L.tris.insert(L.tris.end(), {v.x, v.y, v.z});
    L.mat.push_back(idx);
    touch_lod(h);
    return (FaceHdl)(L.mat.size() - 1);
}

FaceHdl add_quad(void* h, UVec4 v, MatIdx idx) {
    Lod& L = lod_of(h);
    const uint32_t n = (uint32_t)L.vertices();
    CORE_REQUIRE(v.x < n && v.y < n && v.z < n && v.w < n, "quad vertex index out of range");
    L.quads.insert(L.quads.end(), {v.x, v.y, v.z, v.w});
    L.mat.push_back(idx);
    touch_lod(h);
    return (FaceHdl)(L.mat.size() - 1);
}

SphereHdl add_sphere(void* h, Vec3 p, float radius, MatIdx idx) {
    Lod& L = lod_of(h);
    CORE_REQUIRE(radius >= 0.0f, "negative sphere radius");
    const float s[4] = {p.x, p.y, p.z, radius};
    L.spheres.insert(L.spheres.end(), s, s + 4);
    L.sphereMat.push_back(idx);
    L.grow_sphere_box(s);
    touch_lod(h);
    return (SphereHdl)(L.sphere_count() - 1);
}

// Bytes of a bulk stream: an array, or the next `bytes` of a FILE
// (BulkLoader, core_interface.h:322-333). Returns the elements read.
size_t read_bulk(const BulkLoader* b, size_t elemBytes, size_t count, std::vector<uint8_t>& out) {
    out.assign(elemBytes * count, 0);
    if (b->type == BULK_ARRAY) {
        CORE_REQUIRE(b->descriptor.bytes != nullptr || count == 0, "bulk array is null");
        if (count) std::memcpy(out.data(), b->descriptor.bytes, out.size());
        return count;
    }
    CORE_REQUIRE(b->type == BULK_FILE && b->descriptor.file != nullptr, "invalid bulk loader");
    const size_t n = std::fread(out.data(), elemBytes, count, b->descriptor.file);
    out.resize(n * elemBytes);
    return n;
}

template <class F>
Boolean camera_vec(const char* fn, ConstCameraHdl cam, Vec3* out, uint32_t pathIndex, F&& pick) {
    return guard(fn, (Boolean) false, [&]() -> Boolean {
        CORE_REQUIRE(cam != nullptr && out != nullptr, "Invalid camera handle or output pointer (nullptr)");
        const Camera& c = *static_cast<const Camera*>(cam);
        CORE_REQUIRE(pathIndex < c.position.size(), "path index out of range");
        *out = pick(c)[pathIndex];
        return true;
    });
}

void camera_changed(const Camera* cam) {
    std::lock_guard<std::mutex> lock(gCoresMutex);
    for (Core* c : gCores)
        if (c->current && c->current->camera == cam) c->invalidate();
}

// A change of a light of the loaded scenario reloads it (on_light_changed).
void light_changed(Core& c, const LightHdl& h) {
    if (!c.current) return;
    for (const LightHdl& l : c.current->lights)
        if (l.type == h.type && l.index == h.index) c.invalidate();
}
Vec3& point_field(Core& c, LightHdl hdl, uint32_t frame, bool position) {
    light_check(hdl, LIGHT_POINT, c.pointLights.size());
    PointLight& L = *c.pointLights[hdl.index];
    CORE_REQUIRE(frame < L.position.size(), "frame out of the light's path segments");
    return position ? L.position[frame] : L.intensity[frame];
}

Background& background_of(Core& c, LightHdl hdl) {
    light_check(hdl, LIGHT_ENVMAP, c.backgrounds.size());
    return *c.backgrounds[hdl.index];
}

Scenario& scenario_of(void* h) {
    if (!h) throw ApiError("Invalid scenario handle (nullptr)");
    return *static_cast<Scenario*>(h);
}
void scenario_changed(const Scenario* s) {
    std::lock_guard<std::mutex> lock(gCoresMutex);
    for (Core* c : gCores)
        if (c->current == s) c->invalidate();
}

}  // namespace

extern "C" {

// ======================================================================= DLL
const char* core_get_dll_error() {
    std::lock_guard<std::mutex> lock(gLogMutex);
    tlErrorCopy = gLastError;
    return tlErrorCopy.c_str();
}

Boolean core_set_logger(void (*logCallback)(const char*, int)) {
    std::lock_guard<std::mutex> lock(gLogMutex);
    gLogCallback = logCallback;
    return true;
}

Boolean core_set_log_level(LogLevel level) {
    if ((int)level < LOG_PEDANTIC || (int)level > LOG_FATAL_ERROR) {
        log_error(__func__, "Invalid log level");
        return false;
    }
    std::lock_guard<std::mutex> lock(gLogMutex);
    gLogLevel = (int)level;
    return true;
}

MufflonInstanceHdl mufflon_initialize() {
    return guard(__func__, (MufflonInstanceHdl) nullptr, [&]() -> MufflonInstanceHdl {
        auto* c = new Core();
        std::lock_guard<std::mutex> lock(gCoresMutex);
        gCores.insert(c);
        return c;
    });
}

void mufflon_destroy(MufflonInstanceHdl instHdl) {
    guard(__func__, 0, [&] {
        Core* c = &core_of(instHdl);
        {
            std::lock_guard<std::mutex> lock(gCoresMutex);
            for (auto it = gLodOwner.begin(); it != gLodOwner.end();) it = it->second == c ? gLodOwner.erase(it) : ++it;
            gCores.erase(c);
        }
        delete c;
        return 0;
    });
}

Boolean mufflon_is_cuda_available() {
    int n = 0;
    return hipGetDeviceCount(&n) == hipSuccess && n > 0;
}

int32_t mufflon_get_cuda_device_index() { return mufflon_is_cuda_available() ? 0 : -1; }

Boolean mufflon_set_lod_loader(MufflonInstanceHdl instHdl, Boolean (*func)(void*, ObjectHdl, uint32_t, Boolean),
                               Boolean (*objFunc)(void*, uint32_t, uint16_t*, uint32_t*),
                               Boolean (*metaFunc)(void*, LodMetadata*, size_t*), void* userParams) {
    return guard(__func__, (Boolean) false, [&]() -> Boolean {
        Core& c = core_of(instHdl);
        CORE_REQUIRE(func != nullptr, "Invalid LoD loader function (nullptr)");
        c.lodLoader = func;
        c.matIndexLoader = objFunc;
        c.lodMetaLoader = metaFunc;
        c.loaderParams = userParams;
        return true;
    });
}

Boolean mufflon_get_target_image(MufflonInstanceHdl instHdl, const char* name, Boolean variance, const float** ptr) {
    return guard(__func__, (Boolean) false, [&]() -> Boolean {
        Core& c = core_of(instHdl);
        CORE_REQUIRE(c.variation >= 0, "Invalid current renderer (nullptr)");
        std::lock_guard<std::mutex> lock(c.iterationMutex);
        if (c.pt < 0) {
            if (ptr) *ptr = nullptr;
            return true;
        }
        const int t = target_index(c.renderer, name);
        CORE_REQUIRE(t >= 0, std::string("Cannot find render target '") + (name ? name : "") + "'");
        const uint32_t bit = 1u << t;
        CORE_REQUIRE(variance ? (c.varianceMask & bit) : (c.targetMask & bit),
                     std::string("Render target '") + name + (variance ? "' does not record variance" : "' is not enabled"));
        const Scenario& s = *c.current;
        const TargetName& tn = target_table(c.renderer).targets[t];
        c.screenChannels = tn.channels;
        c.screen.resize((size_t)s.width * s.height * c.screenChannels);
        if (c.renderer == kRendererWire)
            check_mfh(mfh_pt_wireframe_get(c.ctx, c.pt, variance ? 1 : 0, c.screen.data()), "target read-back");
        else if (c.renderer == kRendererLt)
            check_mfh(mfh_pt_lighttracer_get(c.ctx, c.pt, tn.device, variance ? 1 : 0, c.screen.data()),
                      "target read-back");
        else
            check_mfh(variance ? mfh_pt_get_variance(c.ctx, c.pt, tn.device, c.screen.data())
                               : mfh_pt_get_target(c.ctx, c.pt, tn.device, c.screen.data()),
                      "target read-back");
        if (ptr) *ptr = c.screen.data();
        return true;
    });
}

Boolean mufflon_get_target_image_num_channels(MufflonInstanceHdl instHdl, int* numChannels) {
    return guard(__func__, (Boolean) false, [&]() -> Boolean {
        Core& c = core_of(instHdl);
        CORE_REQUIRE(!c.screen.empty(), "Invalid screen texture (nullptr)");
        CORE_REQUIRE(numChannels != nullptr, "Invalid channel count pointer (nullptr)");
        *numChannels = c.screenChannels;
        return true;
    });
}

// ===================================================================== world
void world_clear_all(MufflonInstanceHdl instHdl) {
    guard(__func__, 0, [&] {
        Core& c = core_of(instHdl);
        std::lock_guard<std::mutex> lock(c.iterationMutex);
        {
            std::lock_guard<std::mutex> g(gCoresMutex);
            for (auto it = gLodOwner.begin(); it != gLodOwner.end();)
                it = it->second == &c ? gLodOwner.erase(it) : ++it;
        }
        c.release_device();
        c.objects.clear();
        c.instances.clear();
        c.materials.clear();
        c.textures.clear();
        c.cameras.clear();
        c.pointLights.clear();
        c.backgrounds.clear();
        c.scenarios.clear();
        c.current = nullptr;
        c.sceneValid = false;
        c.iteration = -1;
        c.screen.clear();
        return 0;
    });
}

Boolean world_finalize(MufflonInstanceHdl instHdl, const Vec3 min, const Vec3 max, const char** msg) {
    return guard(__func__, (Boolean) false, [&]() -> Boolean {
        Core& c = core_of(instHdl);
        c.worldBox[0] = min.x;
        c.worldBox[1] = min.y;
        c.worldBox[2] = min.z;
        c.worldBox[3] = max.x;
        c.worldBox[4] = max.y;
        c.worldBox[5] = max.z;
        const char* m = world_sanity(c);
        if (msg) *msg = m;
        return m[0] == '\0';
    });
}

void world_reserve_objects_instances(MufflonInstanceHdl instHdl, const uint32_t objects, const uint32_t instances) {
    guard(__func__, 0, [&] {
        Core& c = core_of(instHdl);
        CORE_REQUIRE(c.objects.empty() && c.instances.empty(),
                     "This method may only be called on a clean world state!");
        c.objects.reserve(objects);
        c.instances.reserve(instances);
        return 0;
    });
}

Boolean world_set_frame_current(MufflonInstanceHdl instHdl, const uint32_t animationFrame) {
    return guard(__func__, (Boolean) false, [&]() -> Boolean {
        Core& c = core_of(instHdl);
        if (c.frameCurrent != animationFrame) {
            c.frameCurrent = animationFrame;
            c.invalidate();
        }
        return true;
    });
}

Boolean world_get_frame_current(MufflonInstanceHdl instHdl, uint32_t* animationFrame) {
    return guard(__func__, (Boolean) false, [&]() -> Boolean {
        Core& c = core_of(instHdl);
        CORE_REQUIRE(animationFrame != nullptr, "Invalid frame pointer (nullptr)");
        *animationFrame = c.frameCurrent;
        return true;
    });
}

// ================================================================== polygons
Boolean polygon_reserve(LodHdl lvlDtl, size_t vertices, size_t tris, size_t quads) {
    return guard(__func__, (Boolean) false, [&]() -> Boolean {
        Lod& L = lod_of(lvlDtl);
        L.points.reserve(3 * vertices);
        L.normals.reserve(3 * vertices);
        L.uvs.reserve(2 * vertices);
        L.tris.reserve(3 * tris);
        L.quads.reserve(4 * quads);
        L.mat.reserve(tris + quads);
        return true;
    });
}

VertexHdl polygon_add_vertex(LodHdl lvlDtl, Vec3 point, Vec3 normal, Vec2 uv) {
    return guard(__func__, (VertexHdl)kInvalidIndex, [&]() -> VertexHdl {
        Lod& L = lod_of(lvlDtl);
        const float p[3] = {point.x, point.y, point.z};
        L.points.insert(L.points.end(), p, p + 3);
        L.normals.insert(L.normals.end(), {normal.x, normal.y, normal.z});
        L.uvs.insert(L.uvs.end(), {uv.x, uv.y});
        L.grow_box(p);
        touch_lod(lvlDtl);
        return (VertexHdl)(L.vertices() - 1);
    });
}

FaceHdl polygon_add_triangle(LodHdl lvlDtl, UVec3 vertices) {
    return guard(__func__, (FaceHdl)kInvalidIndex, [&] { return add_triangle(lvlDtl, vertices, 0); });
}

FaceHdl polygon_add_triangle_material(LodHdl lvlDtl, UVec3 vertices, MatIdx idx) {
    return guard(__func__, (FaceHdl)kInvalidIndex, [&] { return add_triangle(lvlDtl, vertices, idx); });
}

FaceHdl polygon_add_quad(LodHdl lvlDtl, UVec4 vertices) {
    return guard(__func__, (FaceHdl)kInvalidIndex, [&] { return add_quad(lvlDtl, vertices, 0); });
}

FaceHdl polygon_add_quad_material(LodHdl lvlDtl, UVec4 vertices, MatIdx idx) {
    return guard(__func__, (FaceHdl)kInvalidIndex, [&] { return add_quad(lvlDtl, vertices, idx); });
}

VertexHdl polygon_add_vertex_bulk(LodHdl lvlDtl, size_t count, const BulkLoader* points, const BulkLoader* normals,
                                  const BulkLoader* uvs, const AABB* aabb, size_t* pointsRead, size_t* normalsRead,
                                  size_t* uvsRead) {
    return guard(__func__, (VertexHdl)kInvalidIndex, [&]() -> VertexHdl {
        Lod& L = lod_of(lvlDtl);
        CORE_REQUIRE(points != nullptr, "Invalid points stream descriptor (nullptr)");
        CORE_REQUIRE(uvs != nullptr, "Invalid UV coordinates stream descriptor (nullptr)");
        const size_t start = L.vertices();
        std::vector<uint8_t> p, n, u;
        const size_t np = read_bulk(points, 12, count, p);
        const size_t nn = normals ? read_bulk(normals, 12, count, n) : 0;
        const size_t nu = read_bulk(uvs, 8, count, u);
        // Polygons::add_bulk (polygon.cpp:274-334): every attribute gets
        // `count` entries; what a stream did not deliver stays zero.
        L.points.resize(3 * (start + count), 0.0f);
        L.normals.resize(3 * (start + count), 0.0f);
        L.uvs.resize(2 * (start + count), 0.0f);
        std::memcpy(L.points.data() + 3 * start, p.data(), 12 * np);
        if (nn) std::memcpy(L.normals.data() + 3 * start, n.data(), 12 * nn);
        std::memcpy(L.uvs.data() + 2 * start, u.data(), 8 * nu);
        if (aabb) {
            const float lo[3] = {aabb->min.x, aabb->min.y, aabb->min.z};
            const float hi[3] = {aabb->max.x, aabb->max.y, aabb->max.z};
            L.grow_box(lo);
            L.grow_box(hi);
        } else {
            for (size_t i = start; i < start + np; ++i) L.grow_box(L.points.data() + 3 * i);
        }
        if (pointsRead) *pointsRead = np;
        if (normalsRead) *normalsRead = nn;
        if (uvsRead) *uvsRead = nu;
        touch_lod(lvlDtl);
        return (VertexHdl)start;
    });
}

Boolean polygon_set_vertex_normal(LodHdl lvlDtl, VertexHdl vertex, Vec3 normal) {
    return guard(__func__, (Boolean) false, [&]() -> Boolean {
        Lod& L = lod_of(lvlDtl);
        CORE_REQUIRE(vertex >= 0 && (size_t)vertex < L.vertices(), "vertex index out of range");
        float* n = L.normals.data() + 3 * (size_t)vertex;
        n[0] = normal.x;
        n[1] = normal.y;
        n[2] = normal.z;
        touch_lod(lvlDtl);
        return true;
    });
}

Boolean polygon_set_vertex_uv(LodHdl lvlDtl, VertexHdl vertex, Vec2 uv) {
    return guard(__func__, (Boolean) false, [&]() -> Boolean {
        Lod& L = lod_of(lvlDtl);
        CORE_REQUIRE(vertex >= 0 && (size_t)vertex < L.vertices(), "vertex index out of range");
        L.uvs[2 * (size_t)vertex] = uv.x;
        L.uvs[2 * (size_t)vertex + 1] = uv.y;
        touch_lod(lvlDtl);
        return true;
    });
}

Boolean polygon_set_material_idx(LodHdl lvlDtl, FaceHdl face, MatIdx idx) {
    return guard(__func__, (Boolean) false, [&]() -> Boolean {
        Lod& L = lod_of(lvlDtl);
        CORE_REQUIRE(face >= 0 && (size_t)face < L.mat.size(), "face index out of range");
        L.mat[(size_t)face] = idx;
        touch_lod(lvlDtl);
        return true;
    });
}

size_t polygon_set_material_idx_bulk(LodHdl lvlDtl, FaceHdl startFace, size_t count, const BulkLoader* stream) {
    return guard(__func__, (size_t)0, [&]() -> size_t {
        Lod& L = lod_of(lvlDtl);
        CORE_REQUIRE(stream != nullptr, "Invalid material index stream (nullptr)");
        CORE_REQUIRE(startFace >= 0 && (size_t)startFace + count <= L.mat.size(), "face range out of bounds");
        std::vector<uint8_t> b;
        const size_t n = read_bulk(stream, 2, count, b);
        std::memcpy(L.mat.data() + startFace, b.data(), 2 * n);
        touch_lod(lvlDtl);
        return n;
    });
}

size_t polygon_get_vertex_count(LodHdl lvlDtl) {
    return guard(__func__, (size_t)0, [&] { return lod_of(lvlDtl).vertices(); });
}
size_t polygon_get_face_count(LodHdl lvlDtl) {
    return guard(__func__, (size_t)0, [&] { return lod_of(lvlDtl).mat.size(); });
}
size_t polygon_get_triangle_count(LodHdl lvlDtl) {
    return guard(__func__, (size_t)0, [&] { return lod_of(lvlDtl).triangles(); });
}
size_t polygon_get_quad_count(LodHdl lvlDtl) {
    return guard(__func__, (size_t)0, [&] { return lod_of(lvlDtl).quad_count(); });
}

Boolean polygon_get_bounding_box(LodHdl lvlDtl, Vec3* min, Vec3* max) {
    return guard(__func__, (Boolean) false, [&]() -> Boolean {
        const Lod& L = lod_of(lvlDtl);
        if (min) *min = {L.box[0], L.box[1], L.box[2]};
        if (max) *max = {L.box[3], L.box[4], L.box[5]};
        return true;
    });
}

// ================================================================== spheres
Boolean spheres_reserve(LodHdl lvlDtl, size_t count) {
    return guard(__func__, (Boolean) false, [&]() -> Boolean {
        Lod& L = lod_of(lvlDtl);
        L.spheres.reserve(4 * count);
        L.sphereMat.reserve(count);
        return true;
    });
}

SphereHdl spheres_add_sphere(LodHdl lvlDtl, Vec3 point, float radius) {
    return guard(__func__, (SphereHdl)kInvalidIndex, [&] { return add_sphere(lvlDtl, point, radius, 0); });
}

SphereHdl spheres_add_sphere_material(LodHdl lvlDtl, Vec3 point, float radius, MatIdx idx) {
    return guard(__func__, (SphereHdl)kInvalidIndex, [&] { return add_sphere(lvlDtl, point, radius, idx); });
}

SphereHdl spheres_add_sphere_bulk(LodHdl lvlDtl, size_t count, const BulkLoader* stream, const AABB* aabb,
                                  size_t* readSpheres) {
    return guard(__func__, (SphereHdl)kInvalidIndex, [&]() -> SphereHdl {
        Lod& L = lod_of(lvlDtl);
        CORE_REQUIRE(stream != nullptr, "Invalid sphere stream descriptor (nullptr)");
        const size_t start = L.sphere_count();
        std::vector<uint8_t> b;
        const size_t n = read_bulk(stream, 16, count, b);
        // Spheres::add_bulk (sphere.cpp:87-115): the spheres read; the box
        // grows by the given AABB, else by every sphere read
        L.spheres.resize(4 * (start + n));
        L.sphereMat.resize(start + n, 0);
        if (n) std::memcpy(L.spheres.data() + 4 * start, b.data(), 16 * n);
        if (aabb) {
            const float lo[4] = {aabb->min.x, aabb->min.y, aabb->min.z, 0.0f};
            const float hi[4] = {aabb->max.x, aabb->max.y, aabb->max.z, 0.0f};
            L.grow_sphere_box(lo);
            L.grow_sphere_box(hi);
        } else {
            for (size_t i = start; i < start + n; ++i) L.grow_sphere_box(L.spheres.data() + 4 * i);
        }
        if (readSpheres) *readSpheres = n;
        touch_lod(lvlDtl);
        return (SphereHdl)start;
    });
}

Boolean spheres_set_material_idx(LodHdl lvlDtl, SphereHdl sphere, MatIdx idx) {
    return guard(__func__, (Boolean) false, [&]() -> Boolean {
        Lod& L = lod_of(lvlDtl);
        CORE_REQUIRE(sphere >= 0 && (size_t)sphere < L.sphere_count(), "sphere index out of range");
        L.sphereMat[(size_t)sphere] = idx;
        touch_lod(lvlDtl);
        return true;
    });
}

size_t spheres_set_material_idx_bulk(LodHdl lvlDtl, SphereHdl startSphere, size_t count, const BulkLoader* stream) {
    return guard(__func__, (size_t)0, [&]() -> size_t {
        Lod& L = lod_of(lvlDtl);
        CORE_REQUIRE(stream != nullptr, "Invalid material index stream (nullptr)");
        CORE_REQUIRE(startSphere >= 0 && (size_t)startSphere + count <= L.sphere_count(),
                     "sphere range out of bounds");
        std::vector<uint8_t> b;
        const size_t n = read_bulk(stream, 2, count, b);
        std::memcpy(L.sphereMat.data() + startSphere, b.data(), 2 * n);
        touch_lod(lvlDtl);
        return n;
    });
}

size_t spheres_get_sphere_count(LodHdl lvlDtl) {
    return guard(__func__, (size_t)0, [&] { return lod_of(lvlDtl).sphere_count(); });
}

Boolean spheres_get_bounding_box(LodHdl lvlDtl, Vec3* min, Vec3* max) {
    return guard(__func__, (Boolean) false, [&]() -> Boolean {
        const Lod& L = lod_of(lvlDtl);
        if (min) *min = {L.sphereBox[0], L.sphereBox[1], L.sphereBox[2]};
        if (max) *max = {L.sphereBox[3], L.sphereBox[4], L.sphereBox[5]};
        return true;
    });
}

// ===================================================== objects and instances
ObjectHdl world_create_object(MufflonInstanceHdl instHdl, const char* name, ObjectFlags flags) {
    return guard(__func__, (ObjectHdl) nullptr, [&]() -> ObjectHdl {
        Core& c = core_of(instHdl);
        CORE_REQUIRE(name != nullptr, "Invalid object name (nullptr)");
        for (const auto& o : c.objects) CORE_REQUIRE(o->name != name, std::string("object '") + name + "' exists");
        auto o = std::make_unique<Object>();
        o->name = name;
        o->flags = (uint32_t)flags;
        o->id = (uint32_t)c.objects.size();
        c.objects.push_back(std::move(o));
        c.invalidate();
        return c.objects.back().get();
    });
}

ObjectHdl world_get_object(MufflonInstanceHdl instHdl, const char* name) {
    return guard(__func__, (ObjectHdl) nullptr, [&]() -> ObjectHdl {
        Core& c = core_of(instHdl);
        CORE_REQUIRE(name != nullptr, "Invalid object name (nullptr)");
        for (const auto& o : c.objects)
            if (o->name == name) return o.get();
        throw ApiError(std::string("Could not find object '") + name + "'");
    });
}

const char* world_get_object_name(ObjectHdl obj) {
    return guard(__func__, (const char*)nullptr, [&]() -> const char* {
        CORE_REQUIRE(obj != nullptr, "Invalid object handle (nullptr)");
        return static_cast<Object*>(obj)->name.c_str();
    });
}

Boolean object_has_lod(ConstObjectHdl obj, LodLevel level) {
    return guard(__func__, (Boolean) false, [&]() -> Boolean {
        CORE_REQUIRE(obj != nullptr, "Invalid object handle (nullptr)");
        const Object& o = *static_cast<const Object*>(obj);
        return level < o.lods.size() && o.lods[level] != nullptr;
    });
}

Boolean object_allocate_lod_slots(ObjectHdl obj, LodLevel slots) {
    return guard(__func__, (Boolean) false, [&]() -> Boolean {
        CORE_REQUIRE(obj != nullptr, "Invalid object handle (nullptr)");
        Object& o = *static_cast<Object*>(obj);
        if (o.lods.size() < slots) o.lods.resize(slots);
        return true;
    });
}

LodHdl object_add_lod(ObjectHdl obj, LodLevel level, Boolean asReduced) {
    (void)asReduced;
    return guard(__func__, (LodHdl) nullptr, [&]() -> LodHdl {
        CORE_REQUIRE(obj != nullptr, "Invalid object handle (nullptr)");
        Object& o = *static_cast<Object*>(obj);
        if (o.lods.size() <= level) o.lods.resize(level + 1);
        if (!o.lods[level]) o.lods[level] = std::make_unique<Lod>();
        Lod* L = o.lods[level].get();
        // the owning instance is the one whose objects hold this object
        std::lock_guard<std::mutex> lock(gCoresMutex);
        for (Core* c : gCores)
            for (const auto& op : c->objects)
                if (op.get() == &o) {
                    gLodOwner[L] = c;
                    c->invalidate();
                }
        return L;
    });
}

Boolean object_get_id(ObjectHdl obj, uint32_t* id) {
    return guard(__func__, (Boolean) false, [&]() -> Boolean {
        CORE_REQUIRE(obj != nullptr && id != nullptr, "Invalid object handle or id pointer (nullptr)");
        *id = static_cast<Object*>(obj)->id;
        return true;
    });
}

InstanceHdl world_create_instance(MufflonInstanceHdl instHdl, ObjectHdl obj, const uint32_t animationFrame) {
    return guard(__func__, (InstanceHdl) nullptr, [&]() -> InstanceHdl {
        Core& c = core_of(instHdl);
        CORE_REQUIRE(obj != nullptr, "Invalid object handle");
        auto i = std::make_unique<ObjInstance>();
        i->object = static_cast<Object*>(obj);
        i->frame = animationFrame;
        c.instances.push_back(std::move(i));
        c.invalidate();
        return c.instances.back().get();
    });
}

uint32_t world_get_instance_count(MufflonInstanceHdl instHdl, const uint32_t frame) {
    return guard(__func__, (uint32_t)0, [&]() -> uint32_t {
        Core& c = core_of(instHdl);
        uint32_t n = 0;
        for (const auto& i : c.instances) n += i->frame == frame ? 1 : 0;
        return n;
    });
}

InstanceHdl world_get_instance_by_index(MufflonInstanceHdl instHdl, uint32_t index, const uint32_t animationFrame) {
    return guard(__func__, (InstanceHdl) nullptr, [&]() -> InstanceHdl {
        Core& c = core_of(instHdl);
        uint32_t n = 0;
        for (const auto& i : c.instances)
            if (i->frame == animationFrame && n++ == index) return i.get();
        throw ApiError("Instance index '" + std::to_string(index) + "' out of bounds (" + std::to_string(n) + ")");
    });
}

Boolean instance_set_transformation_matrix(MufflonInstanceHdl instHdl, InstanceHdl inst, const Mat3x4* mat,
                                           const Boolean isWorldToInst) {
    return guard(__func__, (Boolean) false, [&]() -> Boolean {
        Core& c = core_of(instHdl);
        CORE_REQUIRE(inst != nullptr, "Invalid instance handle (nullptr)");
        CORE_REQUIRE(mat != nullptr, "Invalid transformation matrix (nullptr)");
        ObjInstance& i = *static_cast<ObjInstance*>(inst);
        if (isWorldToInst)
            std::memcpy(i.worldToInst, mat->v, sizeof(i.worldToInst));
        else  // set_instance_to_world_transformation inverts (world_container.cpp:240-245)
            instance_to_world(mat->v, i.worldToInst);
        c.invalidate();
        return true;
    });
}

Boolean instance_get_transformation_matrix(MufflonInstanceHdl instHdl, InstanceHdl inst, Mat3x4* mat) {
    return guard(__func__, (Boolean) false, [&]() -> Boolean {
        core_of(instHdl);
        CORE_REQUIRE(inst != nullptr, "Invalid instance handle (nullptr)");
        if (mat) instance_to_world(static_cast<ObjInstance*>(inst)->worldToInst, mat->v);
        return true;
    });
}

// ==================================================== materials and textures
MaterialHdl world_add_material(MufflonInstanceHdl instHdl, const char* name, const MaterialParams* mat) {
    return guard(__func__, (MaterialHdl) nullptr, [&]() -> MaterialHdl {
        Core& c = core_of(instHdl);
        CORE_REQUIRE(name != nullptr, "Invalid material name (nullptr)");
        CORE_REQUIRE(mat != nullptr, "Invalid material parameters (nullptr)");
        CORE_REQUIRE((int)mat->innerType >= 0 && mat->innerType < MATERIAL_NUM,
                     "Invalid material params: type unknown.");
        c.materials.push_back(std::make_unique<Material>(make_material(name, *mat)));
        c.invalidate();
        return c.materials.back().get();
    });
}

IndexType world_get_material_count(MufflonInstanceHdl instHdl) {
    return guard(__func__, (IndexType)0, [&] { return (IndexType)core_of(instHdl).materials.size(); });
}

MaterialHdl world_get_material(MufflonInstanceHdl instHdl, IndexType index) {
    return guard(__func__, (MaterialHdl) nullptr, [&]() -> MaterialHdl {
        Core& c = core_of(instHdl);
        CORE_REQUIRE(index >= 0 && (size_t)index < c.materials.size(), "material index out of range");
        return c.materials[(size_t)index].get();
    });
}

const char* world_get_material_name(MaterialHdl material) {
    return guard(__func__, (const char*)nullptr, [&]() -> const char* {
        CORE_REQUIRE(material != nullptr, "Invalid material handle (nullptr)");
        return static_cast<Material*>(material)->name.c_str();
    });
}

TextureHdl world_add_texture_value(MufflonInstanceHdl instHdl, const float* value, int num, TextureSampling sampling) {
    return guard(__func__, (TextureHdl) nullptr, [&]() -> TextureHdl {
        Core& c = core_of(instHdl);
        CORE_REQUIRE(value != nullptr, "Invalid texture value (nullptr)");
        CORE_REQUIRE(num >= 1 && num <= 4, "Invalid number of channels (" + std::to_string(num) + ")");
        // the name of a value texture is its values (interface.cpp:1943-1953)
        std::string name = std::to_string(value[0]);
        for (int i = 1; i < num; ++i) name += " " + std::to_string(value[i]);
        for (const auto& t : c.textures)
            if (t->name == name) return t.get();
        auto t = std::make_unique<Texture>();
        t->name = name;
        t->channels = value_channels(num);
        t->sampling = sampling;
        t->texels.assign(t->channels, 0.0f);
        for (int i = 0; i < num; ++i) t->texels[i] = value[i];
        c.textures.push_back(std::move(t));
        return c.textures.back().get();
    });
}

TextureHdl world_add_texture(MufflonInstanceHdl instHdl, const char* path, TextureSampling sampling, MipmapType type,
                             TextureCallback callback, void* userParams) {
    (void)type;
    return guard(__func__, (TextureHdl) nullptr, [&]() -> TextureHdl {
        Core& c = core_of(instHdl);
        CORE_REQUIRE(path != nullptr, "Invalid texture path (nullptr)");
        for (const auto& t : c.textures)
            if (t->name == path) return t.get();
        const std::string ext = std::filesystem::path(path).extension().string();
        CORE_REQUIRE(ext == ".pfm" || ext == ".PFM",
                     std::string("No texture loader for '") + path + "' (PFM files are supported)");
        std::unique_ptr<Texture> t = load_pfm(path);
        t->sampling = sampling;
        if (callback) {
            const TextureFormat format = t->channels == 1 ? FORMAT_R32F : FORMAT_RGBA32F;
            for (int y = 0; y < t->height; ++y)
                for (int x = 0; x < t->width; ++x) {
                    float* px = t->texels.data() + ((size_t)y * t->width + x) * t->channels;
                    Vec4 v{px[0], t->channels > 1 ? px[1] : 0.0f, t->channels > 2 ? px[2] : 0.0f,
                           t->channels > 3 ? px[3] : 0.0f};
                    v = callback((uint32_t)x, (uint32_t)y, 0u, format, v, userParams);
                    const float out[4] = {v.x, v.y, v.z, v.w};
                    for (int k = 0; k < t->channels; ++k) px[k] = out[k];
                }
        }
        c.textures.push_back(std::move(t));
        return c.textures.back().get();
    });
}

TextureHdl world_get_texture(MufflonInstanceHdl instHdl, const char* path) {
    return guard(__func__, (TextureHdl) nullptr, [&]() -> TextureHdl {
        Core& c = core_of(instHdl);
        CORE_REQUIRE(path != nullptr, "Invalid texture path (nullptr)");
        for (const auto& t : c.textures)
            if (t->name == path) return t.get();
        throw ApiError(std::string("Could not find texture ") + path);
    });
}

const char* world_get_texture_name(TextureHdl hdl) {
    return guard(__func__, (const char*)nullptr, [&]() -> const char* {
        CORE_REQUIRE(hdl != nullptr, "Invalid texture handle (nullptr)");
        return static_cast<Texture*>(hdl)->name.c_str();
    });
}

Boolean world_get_texture_size(TextureHdl hdl, IVec2* size) {
    return guard(__func__, (Boolean) false, [&]() -> Boolean {
        CORE_REQUIRE(hdl != nullptr && size != nullptr, "Invalid texture handle or size pointer (nullptr)");
        const Texture& t = *static_cast<Texture*>(hdl);
        *size = {t.width, t.height};
        return true;
    });
}

// =================================================================== cameras
CameraHdl world_add_pinhole_camera(MufflonInstanceHdl instHdl, const char* name, const Vec3* position, const Vec3* dir,
                                   const Vec3* up, const uint32_t pathCount, float near, float far, float vFov) {
    return guard(__func__, (CameraHdl) nullptr, [&]() -> CameraHdl {
        Core& c = core_of(instHdl);
        CORE_REQUIRE(name != nullptr, "Invalid camera name (nullptr)");
        CORE_REQUIRE(position && dir && up && pathCount > 0, "Invalid camera path");
        auto cam = std::make_unique<Camera>();
        cam->name = name;
        cam->position.assign(position, position + pathCount);
        cam->dir.resize(pathCount);
        cam->up.resize(pathCount);
        for (uint32_t i = 0; i < pathCount; ++i) view_space(dir[i], up[i], cam->dir[i], cam->up[i]);
        cam->nearPlane = near;
        cam->farPlane = far;
        cam->vFov = vFov;
        cam->tanVFov = std::tan(vFov / 2.f);  // Pinhole::Pinhole (pinhole.hpp:17-23)
        c.cameras.push_back(std::move(cam));
        return c.cameras.back().get();
    });
}

CameraHdl world_get_camera(MufflonInstanceHdl instHdl, const char* name) {
    return guard(__func__, (CameraHdl) nullptr, [&]() -> CameraHdl {
        Core& c = core_of(instHdl);
        CORE_REQUIRE(name != nullptr, "Invalid camera name (nullptr)");
        for (const auto& cam : c.cameras)
            if (cam->name == name) return cam.get();
        throw ApiError(std::string("Could not find camera '") + name + "'");
    });
}

size_t world_get_camera_count(MufflonInstanceHdl instHdl) {
    return guard(__func__, (size_t)0, [&] { return core_of(instHdl).cameras.size(); });
}

CameraHdl world_get_camera_by_index(MufflonInstanceHdl instHdl, size_t index) {
    return guard(__func__, (CameraHdl) nullptr, [&]() -> CameraHdl {
        Core& c = core_of(instHdl);
        CORE_REQUIRE(index < c.cameras.size(), "camera index out of range");
        return c.cameras[index].get();
    });
}

const char* world_get_camera_name(ConstCameraHdl cam) {
    return guard(__func__, (const char*)nullptr, [&]() -> const char* {
        CORE_REQUIRE(cam != nullptr, "Invalid camera handle (nullptr)");
        return static_cast<const Camera*>(cam)->name.c_str();
    });
}

Boolean world_get_camera_position(ConstCameraHdl cam, Vec3* pos, const uint32_t pathIndex) {
    return camera_vec(__func__, cam, pos, pathIndex, [](const Camera& c) -> const std::vector<Vec3>& {
        return c.position;
    });
}
Boolean world_get_camera_direction(ConstCameraHdl cam, Vec3* dir, const uint32_t pathIndex) {
    return camera_vec(__func__, cam, dir, pathIndex, [](const Camera& c) -> const std::vector<Vec3>& { return c.dir; });
}
Boolean world_get_camera_up(ConstCameraHdl cam, Vec3* up, const uint32_t pathIndex) {
    return camera_vec(__func__, cam, up, pathIndex, [](const Camera& c) -> const std::vector<Vec3>& { return c.up; });
}

Boolean world_set_camera_position(CameraHdl cam, Vec3 pos, const uint32_t pathIndex) {
    return guard(__func__, (Boolean) false, [&]() -> Boolean {
        CORE_REQUIRE(cam != nullptr, "Invalid camera handle (nullptr)");
        Camera& c = *static_cast<Camera*>(cam);
        CORE_REQUIRE(pathIndex < c.position.size(), "path index out of range");
        c.position[pathIndex] = pos;
        camera_changed(&c);
        return true;
    });
}

Boolean world_set_camera_direction(CameraHdl cam, Vec3 dir, Vec3 up, const uint32_t pathIndex) {
    return guard(__func__, (Boolean) false, [&]() -> Boolean {
        CORE_REQUIRE(cam != nullptr, "Invalid camera handle (nullptr)");
        Camera& c = *static_cast<Camera*>(cam);
        CORE_REQUIRE(pathIndex < c.position.size(), "path index out of range");
        view_space(dir, up, c.dir[pathIndex], c.up[pathIndex]);
        camera_changed(&c);
        return true;
    });
}

Boolean world_get_pinhole_camera_fov(ConstCameraHdl cam, float* vFov) {
    return guard(__func__, (Boolean) false, [&]() -> Boolean {
        CORE_REQUIRE(cam != nullptr && vFov != nullptr, "Invalid camera handle or output pointer (nullptr)");
        *vFov = static_cast<const Camera*>(cam)->vFov;
        return true;
    });
}

Boolean world_set_pinhole_camera_fov(CameraHdl cam, float vFov) {
    return guard(__func__, (Boolean) false, [&]() -> Boolean {
        CORE_REQUIRE(cam != nullptr, "Invalid camera handle (nullptr)");
        Camera& c = *static_cast<Camera*>(cam);
        c.vFov = vFov;
        c.tanVFov = std::tan(vFov / 2);  // Pinhole::set_vertical_fov (pinhole.hpp:26-29)
        camera_changed(&c);
        return true;
    });
}

// ==================================================================== lights
LightHdl world_add_light(MufflonInstanceHdl instHdl, const char* name, LightType type, const uint32_t count) {
    return guard(__func__, invalid_light(), [&]() -> LightHdl {
        Core& c = core_of(instHdl);
        CORE_REQUIRE(name != nullptr, "Invalid light name (nullptr)");
        if (type == LIGHT_ENVMAP)
            throw ApiError("This function is not suited for adding background lights; please use "
                           "'world_add_background_light' instead");
        CORE_REQUIRE(type == LIGHT_POINT, "spot and directional lights are not supported by the HIP path tracer");
        CORE_REQUIRE(count > 0, "a light needs at least one path segment");
        auto l = std::make_unique<PointLight>();
        l->name = name;
        l->position.assign(count, Vec3{0, 0, 0});
        l->intensity.assign(count, Vec3{0, 0, 0});
        c.pointLights.push_back(std::move(l));
        LightHdl h;
        h.type = LIGHT_POINT;
        h.index = (uint32_t)c.pointLights.size() - 1;
        return h;
    });
}

LightHdl world_add_background_light(MufflonInstanceHdl instHdl, const char* name, BackgroundType type) {
    return guard(__func__, invalid_light(), [&]() -> LightHdl {
        Core& c = core_of(instHdl);
        CORE_REQUIRE(name != nullptr, "Invalid light name (nullptr)");
        CORE_REQUIRE(type == BACKGROUND_MONOCHROME || type == BACKGROUND_ENVMAP,
                     type == BACKGROUND_SKY_HOSEK ? "the Hosek sky is not supported by the HIP path tracer"
                                                  : "Unknown background type");
        auto b = std::make_unique<Background>();
        b->name = name;
        b->type = type;
        c.backgrounds.push_back(std::move(b));
        LightHdl h;
        h.type = LIGHT_ENVMAP;
        h.index = (uint32_t)c.backgrounds.size() - 1;
        return h;
    });
}

Boolean world_find_light(MufflonInstanceHdl instHdl, const char* name, LightHdl* hdl) {
    return guard(__func__, (Boolean) false, [&]() -> Boolean {
        Core& c = core_of(instHdl);
        CORE_REQUIRE(name != nullptr && hdl != nullptr, "Invalid light name or handle pointer (nullptr)");
        for (size_t i = 0; i < c.pointLights.size(); ++i)
            if (c.pointLights[i]->name == name) {
                hdl->type = LIGHT_POINT;
                hdl->index = (uint32_t)i;
                return true;
            }
        for (size_t i = 0; i < c.backgrounds.size(); ++i)
            if (c.backgrounds[i]->name == name) {
                hdl->type = LIGHT_ENVMAP;
                hdl->index = (uint32_t)i;
                return true;
            }
        return false;
    });
}

size_t world_get_point_light_count(MufflonInstanceHdl instHdl) {
    return guard(__func__, (size_t)0, [&] { return core_of(instHdl).pointLights.size(); });
}
size_t world_get_env_light_count(MufflonInstanceHdl instHdl) {
    return guard(__func__, (size_t)0, [&] { return core_of(instHdl).backgrounds.size(); });
}

Boolean world_get_point_light_position(MufflonInstanceHdl instHdl, LightHdl hdl, Vec3* pos, const uint32_t frame) {
    return guard(__func__, (Boolean) false, [&]() -> Boolean {
        CORE_REQUIRE(pos != nullptr, "Invalid output pointer (nullptr)");
        *pos = point_field(core_of(instHdl), hdl, frame, true);
        return true;
    });
}
Boolean world_get_point_light_intensity(MufflonInstanceHdl instHdl, LightHdl hdl, Vec3* intensity,
                                        const uint32_t frame) {
    return guard(__func__, (Boolean) false, [&]() -> Boolean {
        CORE_REQUIRE(intensity != nullptr, "Invalid output pointer (nullptr)");
        *intensity = point_field(core_of(instHdl), hdl, frame, false);
        return true;
    });
}
Boolean world_set_point_light_position(MufflonInstanceHdl instHdl, LightHdl hdl, Vec3 pos, const uint32_t frame) {
    return guard(__func__, (Boolean) false, [&]() -> Boolean {
        Core& c = core_of(instHdl);
        point_field(c, hdl, frame, true) = pos;
        light_changed(c, hdl);
        return true;
    });
}
Boolean world_set_point_light_intensity(MufflonInstanceHdl instHdl, LightHdl hdl, Vec3 intensity,
                                        const uint32_t frame) {
    return guard(__func__, (Boolean) false, [&]() -> Boolean {
        Core& c = core_of(instHdl);
        point_field(c, hdl, frame, false) = intensity;
        light_changed(c, hdl);
        return true;
    });
}

Boolean world_set_env_light_map(MufflonInstanceHdl instHdl, LightHdl hdl, TextureHdl tex) {
    return guard(__func__, (Boolean) false, [&]() -> Boolean {
        Core& c = core_of(instHdl);
        Background& b = background_of(c, hdl);
        CORE_REQUIRE(b.type == BACKGROUND_ENVMAP, "The background is not an environment-mapped light");
        CORE_REQUIRE(tex != nullptr, "Invalid texture handle (nullptr)");
        b.map = static_cast<Texture*>(tex);
        light_changed(c, hdl);
        return true;
    });
}

Boolean world_set_env_light_color(MufflonInstanceHdl instHdl, LightHdl hdl, Vec3 color) {
    return guard(__func__, (Boolean) false, [&]() -> Boolean {
        Core& c = core_of(instHdl);
        Background& b = background_of(c, hdl);
        CORE_REQUIRE(b.type == BACKGROUND_MONOCHROME, "The background is not a monochrom light");
        b.color = color;
        light_changed(c, hdl);
        return true;
    });
}

Boolean world_get_env_light_color(MufflonInstanceHdl instHdl, LightHdl hdl, Vec3* color) {
    return guard(__func__, (Boolean) false, [&]() -> Boolean {
        Core& c = core_of(instHdl);
        Background& b = background_of(c, hdl);
        CORE_REQUIRE(b.type == BACKGROUND_MONOCHROME, "The background is not a monochrom light");
        CORE_REQUIRE(color != nullptr, "Invalid output pointer (nullptr)");
        *color = b.color;
        return true;
    });
}

Boolean world_set_env_light_scale(MufflonInstanceHdl instHdl, LightHdl hdl, Vec3 color) {
    return guard(__func__, (Boolean) false, [&]() -> Boolean {
        Core& c = core_of(instHdl);
        std::lock_guard<std::mutex> lock(c.iterationMutex);
        background_of(c, hdl).scale = color;
        light_changed(c, hdl);
        return true;
    });
}

Boolean world_get_env_light_scale(MufflonInstanceHdl instHdl, LightHdl hdl, Vec3* color) {
    return guard(__func__, (Boolean) false, [&]() -> Boolean {
        Core& c = core_of(instHdl);
        CORE_REQUIRE(color != nullptr, "Invalid output pointer (nullptr)");
        *color = background_of(c, hdl).scale;
        return true;
    });
}

// ================================================================= scenarios

void world_reserve_scenarios(MufflonInstanceHdl instHdl, const uint32_t scenarios) {
    guard(__func__, 0, [&] {
        core_of(instHdl).scenarios.reserve(scenarios);
        return 0;
    });
}

ScenarioHdl world_create_scenario(MufflonInstanceHdl instHdl, const char* name) {
    return guard(__func__, (ScenarioHdl) nullptr, [&]() -> ScenarioHdl {
        Core& c = core_of(instHdl);
        CORE_REQUIRE(name != nullptr, "Invalid scenario name (nullptr)");
        auto s = std::make_unique<Scenario>();
        s->name = name;
        c.scenarios.push_back(std::move(s));
        return c.scenarios.back().get();
    });
}

ScenarioHdl world_find_scenario(MufflonInstanceHdl instHdl, const char* name) {
    return guard(__func__, (ScenarioHdl) nullptr, [&]() -> ScenarioHdl {
        Core& c = core_of(instHdl);
        CORE_REQUIRE(name != nullptr, "Invalid scenario name (nullptr)");
        for (const auto& s : c.scenarios)
            if (s->name == name) return s.get();
        throw ApiError(std::string("Could not find scenario '") + name + "'");
    });
}

uint32_t world_get_scenario_count(MufflonInstanceHdl instHdl) {
    return guard(__func__, (uint32_t)0, [&] { return (uint32_t)core_of(instHdl).scenarios.size(); });
}

ScenarioHdl world_get_scenario_by_index(MufflonInstanceHdl instHdl, uint32_t index) {
    return guard(__func__, (ScenarioHdl) nullptr, [&]() -> ScenarioHdl {
        Core& c = core_of(instHdl);
        CORE_REQUIRE(index < c.scenarios.size(), "scenario index out of range");
        return c.scenarios[index].get();
    });
}

ConstScenarioHdl world_get_current_scenario(MufflonInstanceHdl instHdl) {
    return guard(__func__, (ConstScenarioHdl) nullptr, [&] { return (ConstScenarioHdl)core_of(instHdl).current; });
}

Boolean world_finalize_scenario(MufflonInstanceHdl instHdl, ScenarioHdl scenario, const char** msg) {
    return guard(__func__, (Boolean) false, [&]() -> Boolean {
        Core& c = core_of(instHdl);
        Scenario& s = scenario_of(scenario);
        // WorldContainer::finalize_scenario (world_container.cpp:59-82)
        bool objects = false;
        for (const auto& i : c.instances)
            objects = objects || !(s.maskedInstances.count(i.get()) || s.maskedObjects.count(i->object));
        const char* m = !objects ? "No objects" : (s.camera == nullptr ? "No camera" : "");
        if (msg) *msg = m;
        return m[0] == '\0';
    });
}

SceneHdl world_load_scenario(MufflonInstanceHdl instHdl, ScenarioHdl scenario) {
    return guard(__func__, (SceneHdl) nullptr, [&]() -> SceneHdl {
        Core& c = core_of(instHdl);
        Scenario& s = scenario_of(scenario);
        std::lock_guard<std::mutex> lock(c.iterationMutex);
        c.current = &s;
        c.invalidate();
        c.screen.clear();
        load_scene(c, s);
        return &s;
    });
}

SceneHdl world_get_current_scene(MufflonInstanceHdl instHdl) {
    return guard(__func__, (SceneHdl) nullptr, [&]() -> SceneHdl { return core_of(instHdl).current; });
}

Boolean scene_get_bounding_box(SceneHdl scene, Vec3* min, Vec3* max) {
    return guard(__func__, (Boolean) false, [&]() -> Boolean {
        CORE_REQUIRE(scene != nullptr, "Invalid scene handle (nullptr)");
        std::lock_guard<std::mutex> lock(gCoresMutex);
        for (Core* c : gCores)
            if (c->current == scene && c->sceneValid) {
                if (min) *min = {c->sceneBox[0], c->sceneBox[1], c->sceneBox[2]};
                if (max) *max = {c->sceneBox[3], c->sceneBox[4], c->sceneBox[5]};
                return true;
            }
        throw ApiError("the scene is not loaded");
    });
}

const char* scenario_get_name(ScenarioHdl scenario) {
    return guard(__func__, (const char*)nullptr, [&] { return scenario_of(scenario).name.c_str(); });
}

LodLevel scenario_get_global_lod_level(ScenarioHdl scenario) {
    return guard(__func__, (LodLevel)0, [&] { return scenario_of(scenario).globalLod; });
}

Boolean scenario_set_global_lod_level(ScenarioHdl scenario, LodLevel level) {
    return guard(__func__, (Boolean) false, [&]() -> Boolean {
        scenario_of(scenario).globalLod = level;
        scenario_changed(static_cast<Scenario*>(scenario));
        return true;
    });
}

Boolean scenario_get_resolution(ScenarioHdl scenario, uint32_t* width, uint32_t* height) {
    return guard(__func__, (Boolean) false, [&]() -> Boolean {
        const Scenario& s = scenario_of(scenario);
        if (width) *width = s.width;
        if (height) *height = s.height;
        return true;
    });
}

Boolean scenario_set_resolution(ScenarioHdl scenario, uint32_t width, uint32_t height) {
    return guard(__func__, (Boolean) false, [&]() -> Boolean {
        Scenario& s = scenario_of(scenario);
        CORE_REQUIRE(width > 0 && height > 0 && width < 65536 && height < 65536, "Invalid resolution");
        s.width = width;
        s.height = height;
        scenario_changed(&s);
        return true;
    });
}

CameraHdl scenario_get_camera(ScenarioHdl scenario) {
    return guard(__func__, (CameraHdl) nullptr, [&] { return (CameraHdl)scenario_of(scenario).camera; });
}

Boolean scenario_set_camera(MufflonInstanceHdl instHdl, ScenarioHdl scenario, CameraHdl cam) {
    return guard(__func__, (Boolean) false, [&]() -> Boolean {
        core_of(instHdl);
        Scenario& s = scenario_of(scenario);
        CORE_REQUIRE(cam != nullptr, "Invalid camera handle (nullptr)");
        s.camera = static_cast<Camera*>(cam);
        scenario_changed(&s);
        return true;
    });
}

Boolean scenario_is_object_masked(ScenarioHdl scenario, ObjectHdl obj) {
    return guard(__func__, (Boolean) false, [&]() -> Boolean {
        return scenario_of(scenario).maskedObjects.count(static_cast<Object*>(obj)) > 0;
    });
}

Boolean scenario_mask_object(ScenarioHdl scenario, ObjectHdl obj) {
    return guard(__func__, (Boolean) false, [&]() -> Boolean {
        CORE_REQUIRE(obj != nullptr, "Invalid object handle (nullptr)");
        scenario_of(scenario).maskedObjects.insert(static_cast<Object*>(obj));
        scenario_changed(static_cast<Scenario*>(scenario));
        return true;
    });
}

Boolean scenario_mask_instance(ScenarioHdl scenario, InstanceHdl inst) {
    return guard(__func__, (Boolean) false, [&]() -> Boolean {
        CORE_REQUIRE(inst != nullptr, "Invalid instance handle (nullptr)");
        scenario_of(scenario).maskedInstances.insert(static_cast<ObjInstance*>(inst));
        scenario_changed(static_cast<Scenario*>(scenario));
        return true;
    });
}

LodLevel scenario_get_object_lod(ScenarioHdl scenario, ObjectHdl obj) {
    return guard(__func__, (LodLevel)0, [&]() -> LodLevel {
        const Scenario& s = scenario_of(scenario);
        auto it = s.objectLod.find(static_cast<Object*>(obj));
        return it == s.objectLod.end() ? s.globalLod : it->second;
    });
}

Boolean scenario_set_object_lod(ScenarioHdl scenario, ObjectHdl obj, LodLevel level) {
    return guard(__func__, (Boolean) false, [&]() -> Boolean {
        CORE_REQUIRE(obj != nullptr, "Invalid object handle (nullptr)");
        scenario_of(scenario).objectLod[static_cast<Object*>(obj)] = level;
        scenario_changed(static_cast<Scenario*>(scenario));
        return true;
    });
}

Boolean scenario_set_instance_lod(ScenarioHdl scenario, InstanceHdl inst, LodLevel level) {
    return guard(__func__, (Boolean) false, [&]() -> Boolean {
        CORE_REQUIRE(inst != nullptr, "Invalid instance handle (nullptr)");
        scenario_of(scenario).instanceLod[static_cast<ObjInstance*>(inst)] = level;
        scenario_changed(static_cast<Scenario*>(scenario));
        return true;
    });
}

// Capacity hints of the reference's hash maps; nothing to reserve here.
void scenario_reserve_custom_object_properties(ScenarioHdl scenario, size_t objects) {
    (void)objects;
    guard(__func__, 0, [&] {
        (void)scenario_of(scenario);
        return 0;
    });
}

void scenario_reserve_custom_instance_properties(ScenarioHdl scenario, size_t instances) {
    (void)instances;
    guard(__func__, 0, [&] {
        (void)scenario_of(scenario);
        return 0;
    });
}

IndexType scenario_get_point_light_count(ScenarioHdl scenario) {
    return guard(__func__, (IndexType)0, [&]() -> IndexType {
        IndexType n = 0;
        for (const LightHdl& h : scenario_of(scenario).lights) n += h.type == LIGHT_POINT ? 1 : 0;
        return n;
    });
}

Boolean scenario_has_envmap_light(MufflonInstanceHdl instHdl, ScenarioHdl scenario) {
    return guard(__func__, (Boolean) false, [&]() -> Boolean {
        Core& c = core_of(instHdl);
        for (const LightHdl& h : scenario_of(scenario).lights)
            if (h.type == LIGHT_ENVMAP && c.backgrounds.at(h.index)->type == BACKGROUND_ENVMAP) return true;
        return false;
    });
}

Boolean scenario_add_light(MufflonInstanceHdl instHdl, ScenarioHdl scenario, LightHdl hdl) {
    return guard(__func__, (Boolean) false, [&]() -> Boolean {
        Core& c = core_of(instHdl);
        Scenario& s = scenario_of(scenario);
        if (hdl.type == LIGHT_POINT)
            light_check(hdl, LIGHT_POINT, c.pointLights.size());
        else
            light_check(hdl, LIGHT_ENVMAP, c.backgrounds.size());
        for (const LightHdl& l : s.lights) {
            if (l.type == hdl.type && l.index == hdl.index) return true;
            // one background per scenario (Scenario::set_background)
            CORE_REQUIRE(!(l.type == LIGHT_ENVMAP && hdl.type == LIGHT_ENVMAP),
                         "the scenario already has a background light");
        }
        s.lights.push_back(hdl);
        scenario_changed(&s);
        return true;
    });
}

Boolean scenario_remove_light(MufflonInstanceHdl instHdl, ScenarioHdl scenario, LightHdl hdl) {
    return guard(__func__, (Boolean) false, [&]() -> Boolean {
        core_of(instHdl);
        Scenario& s = scenario_of(scenario);
        const auto it = std::find_if(s.lights.begin(), s.lights.end(), [&](const LightHdl& l) {
            return l.type == hdl.type && l.index == hdl.index;
        });
        CORE_REQUIRE(it != s.lights.end(), "the light is not part of the scenario");
        s.lights.erase(it);
        scenario_changed(&s);
        return true;
    });
}

void scenario_reserve_material_slots(ScenarioHdl scenario, size_t count) {
    guard(__func__, 0, [&] {
        Scenario& s = scenario_of(scenario);
        s.slots.reserve(count);
        s.assigned.reserve(count);
        return 0;
    });
}

MatIdx scenario_declare_material_slot(ScenarioHdl scenario, const char* name, size_t nameLength) {
    return guard(__func__, kInvalidMaterial, [&]() -> MatIdx {
        Scenario& s = scenario_of(scenario);
        CORE_REQUIRE(name != nullptr, "Invalid material name (nullptr)");
        const std::string n(name, std::min(nameLength, std::strlen(name)));
        for (size_t i = 0; i < s.slots.size(); ++i)
            if (s.slots[i] == n) return (MatIdx)i;
        CORE_REQUIRE(s.slots.size() < kInvalidMaterial, "too many material slots");
        s.slots.push_back(n);
        s.assigned.push_back(nullptr);
        scenario_changed(&s);
        return (MatIdx)(s.slots.size() - 1);
    });
}

MatIdx scenario_get_material_slot(ScenarioHdl scenario, const char* name, size_t nameLength) {
    return guard(__func__, kInvalidMaterial, [&]() -> MatIdx {
        const Scenario& s = scenario_of(scenario);
        CORE_REQUIRE(name != nullptr, "Invalid material name (nullptr)");
        const std::string n(name, std::min(nameLength, std::strlen(name)));
        for (size_t i = 0; i < s.slots.size(); ++i)
            if (s.slots[i] == n) return (MatIdx)i;
        throw ApiError("Could not find material slot '" + n + "'");
    });
}

const char* scenario_get_material_slot_name(ScenarioHdl scenario, MatIdx slot) {
    return guard(__func__, (const char*)nullptr, [&]() -> const char* {
        const Scenario& s = scenario_of(scenario);
        CORE_REQUIRE(slot < s.slots.size(), "material slot out of range");
        return s.slots[slot].c_str();
    });
}

size_t scenario_get_material_slot_count(ScenarioHdl scenario) {
    return guard(__func__, (size_t)0, [&] { return scenario_of(scenario).slots.size(); });
}

MaterialHdl scenario_get_assigned_material(ScenarioHdl scenario, MatIdx index) {
    return guard(__func__, (MaterialHdl) nullptr, [&]() -> MaterialHdl {
        const Scenario& s = scenario_of(scenario);
        CORE_REQUIRE(index < s.assigned.size(), "Invalid material index");
        return s.assigned[index];
    });
}

Boolean scenario_assign_material(ScenarioHdl scenario, MatIdx index, MaterialHdl handle) {
    return guard(__func__, (Boolean) false, [&]() -> Boolean {
        Scenario& s = scenario_of(scenario);
        CORE_REQUIRE(handle != nullptr, "Invalid material handle (nullptr)");
        CORE_REQUIRE(index < s.slots.size(), "Invalid material index");
        s.assigned[index] = static_cast<Material*>(handle);
        scenario_changed(&s);
        return true;
    });
}

// ================================================================= rendering
// Renderers: 0 "Pathtracer"/"PT" (variations 0 megakernel, 1 wavefront),
// 1 "Wireframe"/"WF", 2 "Lighttracer"/"LT" (one variation each), all on
// DEVICE_CUDA.
constexpr int kNumRenderers = 3;
constexpr const char* kRendererNames[kNumRenderers][2] = {
    {"Pathtracer", "PT"}, {"Wireframe", "WF"}, {"Lighttracer", "LT"}};  // gpu_pt.hpp, gpu_wireframe.hpp, gpu_lt.hpp
constexpr uint32_t kRendererVariations[kNumRenderers] = {2u, 1u, 1u};

uint32_t render_get_renderer_count(MufflonInstanceHdl instHdl) {
    return guard(__func__, (uint32_t)0, [&] {
        core_of(instHdl);
        return (uint32_t)kNumRenderers;
    });
}

uint32_t render_get_renderer_variations(MufflonInstanceHdl instHdl, uint32_t index) {
    return guard(__func__, (uint32_t)0, [&]() -> uint32_t {
        core_of(instHdl);
        CORE_REQUIRE(index < (uint32_t)kNumRenderers, "renderer index out of bounds");
        return kRendererVariations[index];
    });
}

const char* render_get_renderer_name(MufflonInstanceHdl instHdl, uint32_t index) {
    return guard(__func__, (const char*)nullptr, [&]() -> const char* {
        core_of(instHdl);
        CORE_REQUIRE(index < (uint32_t)kNumRenderers, "renderer index out of bounds");
        return kRendererNames[index][0];
    });
}

const char* render_get_renderer_short_name(MufflonInstanceHdl instHdl, uint32_t index) {
    return guard(__func__, (const char*)nullptr, [&]() -> const char* {
        core_of(instHdl);
        CORE_REQUIRE(index < (uint32_t)kNumRenderers, "renderer index out of bounds");
        return kRendererNames[index][1];
    });
}

RenderDevice render_get_renderer_devices(MufflonInstanceHdl instHdl, uint32_t index, uint32_t variation) {
    return guard(__func__, DEVICE_NONE, [&]() -> RenderDevice {
        core_of(instHdl);
        CORE_REQUIRE(index < (uint32_t)kNumRenderers && variation < kRendererVariations[index],
                     "renderer index out of bounds");
        return DEVICE_CUDA;
    });
}

Boolean render_enable_renderer(MufflonInstanceHdl instHdl, uint32_t index, uint32_t variation) {
    return guard(__func__, (Boolean) false, [&]() -> Boolean {
        Core& c = core_of(instHdl);
        CORE_REQUIRE(index < (uint32_t)kNumRenderers && variation < kRendererVariations[index],
                     "renderer index out of bounds");
        std::lock_guard<std::mutex> lock(c.iterationMutex);
        c.renderer = (int)index;
        c.variation = (int)variation;
        // a fresh output handler: no targets, iteration -1 (interface.cpp:3587-3609)
        c.targetMask = c.varianceMask = 0;
        c.toEnable.clear();
        c.toDisable.clear();
        c.iteration = -1;
        c.screen.clear();
        if (c.current) {
            if (!c.sceneValid || c.pt < 0) {
                load_scene(c, *c.current);
            } else if (c.renderer == kRendererWire) {
                wire_reset(c);
            } else if (c.renderer == kRendererLt) {
                push_params(c);  // the light tracer's parameters, then a fresh output
            } else {
                apply_targets(c);
                apply_mode(c);
                push_params(c);
                check_mfh(mfh_pt_reset(c.ctx, c.pt), "reset");
            }
        }
        return true;
    });
}

Boolean render_iterate(MufflonInstanceHdl instHdl, ProcessTime* iterateTime, ProcessTime* preTime,
                       ProcessTime* postTime) {
    return guard(__func__, (Boolean) false, [&]() -> Boolean {
        Core& c = core_of(instHdl);
        std::lock_guard<std::mutex> lock(c.iterationMutex);
        CORE_REQUIRE(c.variation >= 0, "No renderer is currently set");
        uint64_t t0 = micros(), k0 = cycles();
        // requested target changes (interface.cpp:3625-3636)
        const uint32_t oldMask = c.targetMask, oldVar = c.varianceMask;
        for (const auto& t : c.toDisable) {
            const int i = target_index(c.renderer, t.first.c_str());
            CORE_REQUIRE(i >= 0, "Cannot find render target '" + t.first + "'");
            c.varianceMask &= ~(1u << i);
            if (!t.second) c.targetMask &= ~(1u << i);
        }
        for (const auto& t : c.toEnable) {
            const int i = target_index(c.renderer, t.first.c_str());
            CORE_REQUIRE(i >= 0, "Cannot find render target '" + t.first + "'");
            c.targetMask |= 1u << i;
            if (t.second) c.varianceMask |= 1u << i;
        }
        c.toEnable.clear();
        c.toDisable.clear();
        CORE_REQUIRE(c.current != nullptr, "Failed to load scenario");
        if (!c.sceneValid || c.pt < 0) {
            load_scene(c, *c.current);
        } else if (oldMask != c.targetMask || oldVar != c.varianceMask) {
            if (c.renderer == kRendererWire)
                wire_reset(c);
            else if (c.renderer == kRendererLt)
                lt_reset(c);
            else
                apply_targets(c);  // resets
            c.iteration = -1;
        }
        if (preTime) {
            preTime->cycles = cycles() - k0;
            preTime->microseconds = micros() - t0;
        }
        t0 = micros();
        k0 = cycles();
        if (c.renderer == kRendererWire)
            check_mfh(mfh_pt_wireframe_iterate(c.ctx, c.pt, 1, c.wire.lineWidth, c.wire.maxTraceDepth), "iterate");
        else if (c.renderer == kRendererLt)
            check_mfh(mfh_pt_lighttracer_iterate(c.ctx, c.pt, 1, (c.targetMask >> kLtDensityTarget) & 1u),
                      "iterate");
        else
            check_mfh(mfh_pt_iterate(c.ctx, c.pt, 1), "iterate");
        check_mfh(mfh_synchronize(c.ctx), "iterate");
        ++c.iteration;
        if (iterateTime) {
            iterateTime->cycles = cycles() - k0;
            iterateTime->microseconds = micros() - t0;
        }
        if (postTime) postTime->cycles = postTime->microseconds = 0;
        return true;
    });
}

uint32_t render_get_current_iteration(ConstMufflonInstanceHdl instHdl) {
    return guard(__func__, (uint32_t)0, [&]() -> uint32_t {
        const Core& c = core_of(const_cast<void*>(instHdl));
        return c.variation < 0 ? 0u : (uint32_t)(c.iteration + 1);
    });
}

Boolean render_reset(MufflonInstanceHdl instHdl) {
    return guard(__func__, (Boolean) false, [&]() -> Boolean {
        Core& c = core_of(instHdl);
        std::lock_guard<std::mutex> lock(c.iterationMutex);
        if (c.pt >= 0 && c.sceneValid) {
            if (c.renderer == kRendererWire)
                wire_reset(c);
            else if (c.renderer == kRendererLt)
                lt_reset(c);
            else
                check_mfh(mfh_pt_reset(c.ctx, c.pt), "reset");
        }
        c.iteration = -1;
        return true;
    });
}

// ---- differentiable rendering (no counterpart in the reference)

Boolean render_set_texture_data(MufflonInstanceHdl instHdl, TextureHdl tex, const float* texels) {
    return guard(__func__, (Boolean) false, [&]() -> Boolean {
        Core& c = core_of(instHdl);
        std::lock_guard<std::mutex> lock(c.iterationMutex);
        CORE_REQUIRE(tex != nullptr && texels != nullptr, "Invalid texture handle or texel pointer (nullptr)");
        Texture* t = static_cast<Texture*>(tex);
        bool owned = false;
        for (const auto& p : c.textures) owned = owned || p.get() == t;
        CORE_REQUIRE(owned, "texture handle of another instance");
        std::copy(texels, texels + t->texels.size(), t->texels.begin());
        if (!c.sceneValid || c.pt < 0) return true;  // uploaded with the next scene load
        auto it = c.texIds.find(t);
        const bool inRecord = t->width == 1 && t->height == 1;  // constants inside material records
        if (inRecord || t == c.envMap) {
            // material constants and the environment map's flux / sampling
            // tables are built at load: reload the scene
            c.invalidate();
        } else if (it != c.texIds.end()) {
            check_mfh(mfh_texture_update(c.ctx, it->second, t->texels.data()), "texture update");
            if (c.renderer == kRendererWire)
                wire_reset(c);
            else if (c.renderer == kRendererLt)
                lt_reset(c);
            else
                check_mfh(mfh_pt_reset(c.ctx, c.pt), "reset");
        }
        c.iteration = -1;
        return true;
    });
}

Boolean render_set_emission_scale(MufflonInstanceHdl instHdl, MatIdx slot, Vec3 scale) {
    return guard(__func__, (Boolean) false, [&]() -> Boolean {
        Core& c = core_of(instHdl);
        std::lock_guard<std::mutex> lock(c.iterationMutex);
        CORE_REQUIRE(c.current != nullptr, "No scenario is loaded");
        CORE_REQUIRE(slot < c.current->assigned.size() && c.current->assigned[slot] != nullptr,
                     "material slot " + std::to_string(slot) + " has no material assigned");
        Material& m = *c.current->assigned[slot];
        CORE_REQUIRE(m.params.innerType == MATERIAL_EMISSIVE, "material '" + m.name + "' is not EMISSIVE");
        m.params.inner.emissive.scale = scale;
        m.record.scale[0] = scale.x;
        m.record.scale[1] = scale.y;
        m.record.scale[2] = scale.z;
        if (c.sceneValid && c.pt >= 0) {
            // every slot of the loaded scenario that holds this material
            for (size_t i = 0; i < c.current->assigned.size(); ++i)
                if (c.current->assigned[i] == &m)
                    check_mfh(mfh_pt_set_material(c.ctx, c.pt, (int32_t)i, &m.record), "material update");
            if (c.renderer == kRendererWire) wire_reset(c);
            if (c.renderer == kRendererLt) lt_reset(c);
        }
        c.iteration = -1;
        return true;
    });
}

Boolean render_set_adjoint_mode(MufflonInstanceHdl instHdl, int32_t mode) {
    return guard(__func__, (Boolean) false, [&]() -> Boolean {
        Core& c = core_of(instHdl);
        std::lock_guard<std::mutex> lock(c.iterationMutex);
        CORE_REQUIRE(mode == MFH_ADJOINT_EXPECTED || mode == MFH_ADJOINT_FIXED_DECISIONS, "unknown adjoint mode");
        c.adjointMode = mode;
        if (c.sceneValid && c.pt >= 0) check_mfh(mfh_pt_set_adjoint_mode(c.ctx, c.pt, mode), "adjoint mode");
        return true;
    });
}

namespace {
void require_adjoint_state(const Core& c) {
    CORE_REQUIRE(c.variation >= 0 && c.renderer == kRendererPt, "the Pathtracer renderer is not enabled");
    CORE_REQUIRE(c.sceneValid && c.pt >= 0, "no scene is loaded on the device");
}
}  // namespace

Boolean render_adjoint(MufflonInstanceHdl instHdl, const float* radianceAdjoint) {
    return guard(__func__, (Boolean) false, [&]() -> Boolean {
        Core& c = core_of(instHdl);
        std::lock_guard<std::mutex> lock(c.iterationMutex);
        require_adjoint_state(c);
        CORE_REQUIRE(radianceAdjoint != nullptr, "Invalid adjoint image (nullptr)");
        CORE_REQUIRE(c.iteration >= 0, "no iteration was rendered since the last reset");
        check_mfh(mfh_pt_adjoint(c.ctx, c.pt, c.iteration + 1, radianceAdjoint), "adjoint");
        check_mfh(mfh_synchronize(c.ctx), "adjoint");
        return true;
    });
}

Boolean render_get_texture_gradient(MufflonInstanceHdl instHdl, TextureHdl tex, float* gradient) {
    return guard(__func__, (Boolean) false, [&]() -> Boolean {
        Core& c = core_of(instHdl);
        std::lock_guard<std::mutex> lock(c.iterationMutex);
        require_adjoint_state(c);
        CORE_REQUIRE(tex != nullptr && gradient != nullptr, "Invalid texture handle or output (nullptr)");
        auto it = c.texIds.find(static_cast<const Texture*>(tex));
        CORE_REQUIRE(it != c.texIds.end(), "the texture is not a device texture of the loaded scene");
        check_mfh(mfh_pt_get_texture_gradient(c.ctx, c.pt, it->second, gradient), "texture gradient");
        return true;
    });
}

Boolean render_get_emission_gradient(MufflonInstanceHdl instHdl, MatIdx slot, Vec3* gradient) {
    return guard(__func__, (Boolean) false, [&]() -> Boolean {
        Core& c = core_of(instHdl);
        std::lock_guard<std::mutex> lock(c.iterationMutex);
        require_adjoint_state(c);
        CORE_REQUIRE(gradient != nullptr, "Invalid output (nullptr)");
        float g[3] = {0, 0, 0};
        check_mfh(mfh_pt_get_emission_gradient(c.ctx, c.pt, (int32_t)slot, g), "emission gradient");
        gradient->x = g[0];
        gradient->y = g[1];
        gradient->z = g[2];
        return true;
    });
}

Boolean render_save_screenshot(MufflonInstanceHdl instHdl, const char* filename, const char* targetName,
                               Boolean variance) {
    return guard(__func__, (Boolean) false, [&]() -> Boolean {
        Core& c = core_of(instHdl);
        CORE_REQUIRE(c.variation >= 0, "No renderer is currently set");
        CORE_REQUIRE(filename != nullptr && targetName != nullptr, "Invalid file or target name (nullptr)");
        const float* data = nullptr;
        CORE_REQUIRE(mufflon_get_target_image(instHdl, targetName, variance, &data) && data,
                     "no image of target '" + std::string(targetName) + "'");
        namespace fs = std::filesystem;
        fs::path file = fs::u8path(filename);
        if (file.extension() == "" || file.extension() == ".") file.replace_extension(".pfm");
        CORE_REQUIRE(file.extension() == ".pfm", "No exporter could save screenshot of file type '" +
                                                     file.extension().string() + "'");
        if (!file.is_absolute()) file = fs::absolute(file);
        std::string name = file.stem().u8string();
        name = replace_tag(name, "#scenario", c.current ? c.current->name : "");
        name = replace_tag(name, "#renderer", kRendererNames[c.renderer][0]);
        name = replace_tag(name, "#shortrenderer", kRendererNames[c.renderer][1]);
        name = replace_tag(name, "#iteration", std::to_string(c.iteration + 1));
        name = replace_tag(name, "#frame", std::to_string(c.frameCurrent));
        name = replace_tag(name, "#target", variance ? std::string(targetName) + "(Variance)" : std::string(targetName));
        name = replace_tag(name, "#devices", "CUDA");
        name = replace_tag(name, "#camera", c.current && c.current->camera ? c.current->camera->name : "");
        file = file.parent_path() / fs::u8path(name + file.extension().u8string());
        if (!fs::exists(file.parent_path())) fs::create_directories(file.parent_path());
        const Scenario& s = *c.current;
        store_pfm(file.u8string(), data, (int)s.width, (int)s.height, c.screenChannels);
        return true;
    });
}

uint32_t render_get_render_target_count(MufflonInstanceHdl instHdl) {
    return guard(__func__, (uint32_t)0, [&]() -> uint32_t {
        const Core& c = core_of(instHdl);
        return c.variation < 0 ? 0u : (uint32_t)target_table(c.renderer).count;
    });
}

const char* render_get_render_target_name(MufflonInstanceHdl instHdl, uint32_t index) {
    return guard(__func__, (const char*)nullptr, [&]() -> const char* {
        const Core& c = core_of(instHdl);
        if (c.variation < 0) return "";
        const TargetTable tt = target_table(c.renderer);
        CORE_REQUIRE(index < (uint32_t)tt.count, "render target index out of range");
        return tt.targets[index].name;
    });
}

Boolean render_enable_render_target(MufflonInstanceHdl instHdl, const char* target, Boolean variance) {
    return guard(__func__, (Boolean) false, [&]() -> Boolean {
        Core& c = core_of(instHdl);
        CORE_REQUIRE(target != nullptr, "Invalid target name (nullptr)");
        c.toDisable.erase({target, variance != 0});
        c.toEnable.insert({target, variance != 0});
        return true;
    });
}

Boolean render_disable_render_target(MufflonInstanceHdl instHdl, const char* target, Boolean variance) {
    return guard(__func__, (Boolean) false, [&]() -> Boolean {
        Core& c = core_of(instHdl);
        CORE_REQUIRE(target != nullptr, "Invalid target name (nullptr)");
        c.toEnable.erase({target, variance != 0});
        c.toDisable.insert({target, variance != 0});
        return true;
    });
}

Boolean render_is_render_target_enabled(MufflonInstanceHdl instHdl, const char* name, Boolean variance) {
    return guard(__func__, (Boolean) false, [&]() -> Boolean {
        Core& c = core_of(instHdl);
        CORE_REQUIRE(c.variation >= 0, "No renderer is enabled yet");
        CORE_REQUIRE(name != nullptr, "Invalid target name (nullptr)");
        const std::pair<std::string, bool> key{name, variance != 0};
        if (c.toEnable.count(key)) return true;
        if (c.toDisable.count(key)) return false;
        const int i = target_index(c.renderer, name);
        CORE_REQUIRE(i >= 0, std::string("Cannot find render target '") + name + "'");
        return ((variance ? c.varianceMask : c.targetMask) >> i) & 1u;
    });
}

Boolean render_is_render_target_required(MufflonInstanceHdl instHdl, const char* name, Boolean variance) {
    (void)variance;
    return guard(__func__, (Boolean) false, [&]() -> Boolean {
        Core& c = core_of(instHdl);
        CORE_REQUIRE(c.variation >= 0, "No renderer is enabled yet");
        CORE_REQUIRE(target_index(c.renderer, name) >= 0,
                     std::string("Cannot find render target '") + (name ? name : "") + "'");
        return false;  // neither PtTargets nor WireframeTargets declares a REQUIRED target
    });
}

// ================================================================ parameters
// Parameters of the enabled renderer: the path tracer's per variation, the
// wireframe's "Line width" and "Maximum trace depth". Unknown names are
// logged and ignored (ParameterHandler, parameter.hpp:242-254); a change
// resets the render (on_renderer_parameter_changed).
uint32_t renderer_get_num_parameters(MufflonInstanceHdl instHdl) {
    return guard(__func__, (uint32_t)0, [&]() -> uint32_t {
        Core& c = core_of(instHdl);
        current_variation(c);
        return (uint32_t)(c.renderer == kRendererWire ? kNumWireParams
                                                      : (c.renderer == kRendererLt ? kNumLtParams : kNumParams));
    });
}

const char* renderer_get_parameter_desc(MufflonInstanceHdl instHdl, uint32_t idx, ParameterType* type) {
    return guard(__func__, (const char*)nullptr, [&]() -> const char* {
        Core& c = core_of(instHdl);
        current_variation(c);
        const bool wire = c.renderer == kRendererWire;
        CORE_REQUIRE(idx < (uint32_t)(wire ? kNumWireParams : (c.renderer == kRendererLt ? kNumLtParams : kNumParams)),
                     "parameter index out of range");
        const ParamDescription& d = wire ? kWireParams[idx] : kParams[idx];
        if (type) *type = d.type;
        return d.name;
    });
}

Boolean renderer_set_parameter_int(MufflonInstanceHdl instHdl, const char* name, int32_t value) {
    return guard(__func__, (Boolean) false, [&]() -> Boolean {
        Core& c = core_of(instHdl);
        std::lock_guard<std::mutex> lock(c.iterationMutex);
        Variation& v = current_variation(c);
        const int i = find_param(c, __func__, name, PARAM_INT);
        if (c.renderer == kRendererWire) {
            if (i == 1) c.wire.maxTraceDepth = value;
        } else if (i >= 0) {
            *int_param(v, i) = value;
        }
        push_params(c);
        return true;
    });
}

Boolean renderer_get_parameter_int(MufflonInstanceHdl instHdl, const char* name, int32_t* value) {
    return guard(__func__, (Boolean) false, [&]() -> Boolean {
        Core& c = core_of(instHdl);
        Variation& v = current_variation(c);
        CORE_REQUIRE(value != nullptr, "Invalid value pointer (nullptr)");
        const int i = find_param(c, __func__, name, PARAM_INT);
        if (c.renderer == kRendererWire)
            *value = i == 1 ? c.wire.maxTraceDepth : 0;
        else
            *value = i >= 0 ? *int_param(v, i) : 0;
        return true;
    });
}

Boolean renderer_set_parameter_float(MufflonInstanceHdl instHdl, const char* name, float value) {
    return guard(__func__, (Boolean) false, [&]() -> Boolean {
        Core& c = core_of(instHdl);
        std::lock_guard<std::mutex> lock(c.iterationMutex);
        current_variation(c);
        const int i = find_param(c, __func__, name, PARAM_FLOAT);  // the path tracer has none
        if (c.renderer == kRendererWire && i == 0) c.wire.lineWidth = value;
        push_params(c);
        return true;
    });
}

Boolean renderer_get_parameter_float(MufflonInstanceHdl instHdl, const char* name, float* value) {
    return guard(__func__, (Boolean) false, [&]() -> Boolean {
        Core& c = core_of(instHdl);
        current_variation(c);
        CORE_REQUIRE(value != nullptr, "Invalid value pointer (nullptr)");
        const int i = find_param(c, __func__, name, PARAM_FLOAT);
        *value = (c.renderer == kRendererWire && i == 0) ? c.wire.lineWidth : 0.0f;
        return true;
    });
}

Boolean renderer_set_parameter_bool(MufflonInstanceHdl instHdl, const char* name, Boolean value) {
    return guard(__func__, (Boolean) false, [&]() -> Boolean {
        Core& c = core_of(instHdl);
        std::lock_guard<std::mutex> lock(c.iterationMutex);
        Variation& v = current_variation(c);
        const int i = find_param(c, __func__, name, PARAM_BOOL);
        if (c.renderer == kRendererPt && i == 4) v.neePositionGuide = value != 0;
        push_params(c);
        return true;
    });
}

Boolean renderer_get_parameter_bool(MufflonInstanceHdl instHdl, const char* name, Boolean* value) {
    return guard(__func__, (Boolean) false, [&]() -> Boolean {
        Core& c = core_of(instHdl);
        Variation& v = current_variation(c);
        CORE_REQUIRE(value != nullptr, "Invalid value pointer (nullptr)");
        const int i = find_param(c, __func__, name, PARAM_BOOL);
        *value = c.renderer == kRendererPt && i == 4 && v.neePositionGuide;
        return true;
    });
}

Boolean renderer_set_parameter_enum(MufflonInstanceHdl instHdl, const char* name, int value) {
    (void)value;
    return guard(__func__, (Boolean) false, [&]() -> Boolean {
        Core& c = core_of(instHdl);
        std::lock_guard<std::mutex> lock(c.iterationMutex);
        current_variation(c);
        find_param(c, __func__, name, PARAM_ENUM);  // neither renderer has an enum parameter
        push_params(c);
        return true;
    });
}

Boolean renderer_get_parameter_enum(MufflonInstanceHdl instHdl, const char* name, int* value) {
    return guard(__func__, (Boolean) false, [&]() -> Boolean {
        Core& c = core_of(instHdl);
        current_variation(c);
        CORE_REQUIRE(value != nullptr, "Invalid value pointer (nullptr)");
        find_param(c, __func__, name, PARAM_ENUM);
        *value = 0;
        return true;
    });
}

void profiling_enable() {}
void profiling_disable() {}

}  // extern "C"
