/*
 * mufflon_core.h — the subset of Mufflon's core C API (core_interface.h,
 * implemented in interface.cpp) that builds a world procedurally and renders
 * it with the "Pathtracer" ("PT") renderer on the MI355X device of
 * mufflon_hip.h. Names, argument meaning, value types and return conventions
 * follow the reference API so that a ctypes caller written against it
 * (bindings.py:322-537) runs unchanged against this library:
 *   - functions return Boolean false / a null handle / 0 on failure, never
 *     throw; the message is available from core_get_dll_error() and is sent
 *     to the core_set_logger callback (interface.cpp:54-83, :156-166);
 *   - handles are opaque pointers owned by the instance; mufflon_destroy
 *     frees everything, including the device context;
 *   - render_iterate, render_enable_renderer, render_reset,
 *     world_load_scenario and mufflon_get_target_image are serialized by one
 *     mutex per instance (interface.cpp:99-100).
 * The device is created on the first scene load, so the world-building calls
 * run on a host without a GPU.
 *
 * Not supported by the HIP path (the call fails with a message): alpha and
 * displacement maps, spot and directional lights, focus cameras, the Hosek
 * sky, and the ORENNAYAR, BLEND and FRESNEL materials.
 */
#ifndef MUFFLON_CORE_H
#define MUFFLON_CORE_H

#include <stddef.h>
#include <stdint.h>
#include <stdio.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef uint32_t Boolean;

typedef struct { float x, y; } Vec2;
typedef struct { float x, y, z; } Vec3;
typedef struct { float x, y, z, w; } Vec4;
typedef struct { uint32_t x, y, z; } UVec3;
typedef struct { uint32_t x, y, z, w; } UVec4;
typedef struct { int32_t x, y; } IVec2;
typedef struct { float v[12]; } Mat3x4; /* row-major 3x4 */
typedef struct { Vec3 min; Vec3 max; } AABB;
typedef struct { uint64_t cycles; uint64_t microseconds; } ProcessTime;

/* Enumerations with the reference's values (core_interface.h:84-230). */
typedef enum { LIGHT_POINT, LIGHT_SPOT, LIGHT_DIRECTIONAL, LIGHT_ENVMAP, LIGHT_COUNT } LightType;
typedef enum { BACKGROUND_MONOCHROME, BACKGROUND_ENVMAP, BACKGROUND_SKY_HOSEK, BACKGROUND_COUNT } BackgroundType;
typedef enum { SAMPLING_NEAREST, SAMPLING_LINEAR } TextureSampling;
typedef enum { SHADOWING_VCAVITY, SHADOWING_SMITH } ShadowingModel;
typedef enum { NDF_BECKMANN, NDF_GGX, NDF_COSINE } NormalDistFunction;
typedef enum {
    MATERIAL_EMISSIVE,
    MATERIAL_LAMBERT,
    MATERIAL_ORENNAYAR,
    MATERIAL_TORRANCE,
    MATERIAL_WALTER,
    MATERIAL_BLEND,
    MATERIAL_FRESNEL,
    MATERIAL_MICROFACET,
    MATERIAL_NUM
} MaterialParamType;
typedef enum { MIPMAP_NONE, MIPMAP_AVG, MIPMAP_MIN, MIPMAP_MAX } MipmapType;
typedef enum { LOG_PEDANTIC, LOG_INFO, LOG_WARNING, LOG_ERROR, LOG_FATAL_ERROR } LogLevel;
typedef enum { NONE = 0, OBJ_EMISSIVE = 1 } ObjectFlags;
typedef enum { PARAM_INT, PARAM_FLOAT, PARAM_BOOL, PARAM_ENUM } ParameterType;
typedef enum { BULK_FILE, BULK_ARRAY } BulkType;
typedef enum { DEVICE_NONE = 0u, DEVICE_CPU = 1u, DEVICE_CUDA = 2u, DEVICE_OPENGL = 4u } RenderDevice;
/* Texture formats (texture_data.h:10-30). */
typedef enum {
    FORMAT_R8U, FORMAT_RG8U, FORMAT_RGBA8U, FORMAT_R16U, FORMAT_RG16U, FORMAT_RGBA16U,
    FORMAT_R16F, FORMAT_RG16F, FORMAT_RGBA16F, FORMAT_R32F, FORMAT_RG32F, FORMAT_RGBA32F,
    FORMAT_RGB8U, FORMAT_RGB16U, FORMAT_RGB16F, FORMAT_RGB32F, FORMAT_NUM
} TextureFormat;

typedef struct {
    BulkType type;
    union {
        FILE* file;
        const char* bytes;
    } descriptor;
} BulkLoader;

/* Sizes of one LoD as stored in a scene file (core_interface.h:77-83). */
typedef struct {
    uint32_t vertices;
    uint32_t triangles;
    uint32_t quads;
    uint32_t edges;
    uint32_t spheres;
} LodMetadata;

/* A light: 3-bit type (LightType) and 29-bit index (core_interface.h:205-208);
 * type 7 marks an invalid handle. */
typedef struct {
    uint32_t type : 3;
    uint32_t index : 29;
} LightHdl;

typedef int32_t IndexType;
typedef IndexType VertexHdl;
typedef IndexType FaceHdl;
typedef IndexType SphereHdl;
typedef uint32_t LodLevel;
typedef uint16_t MatIdx;
typedef void* ObjectHdl;
typedef void* ConstObjectHdl;
typedef void* LodHdl;
typedef void* InstanceHdl;
typedef void* ScenarioHdl;
typedef const void* ConstScenarioHdl;
typedef void* SceneHdl;
typedef void* MaterialHdl;
typedef void* CameraHdl;
typedef const void* ConstCameraHdl;
typedef void* TextureHdl;
typedef void* MufflonInstanceHdl;
typedef const void* ConstMufflonInstanceHdl;

typedef Vec4 (*TextureCallback)(uint32_t x, uint32_t y, uint32_t layer, TextureFormat format, Vec4 value,
                                void* userParams);

/* Material descriptor of world_add_material (core_interface.h:236-318); the
 * layout (88 bytes on LP64) is the reference's. */
typedef struct { Vec2 refractionIndex; Vec3 absorption; } Medium;
typedef struct { TextureHdl albedo; } LambertParams;
typedef struct { TextureHdl roughness; ShadowingModel shadowingModel; NormalDistFunction ndf; TextureHdl albedo; } TorranceParams;
typedef struct {
    TextureHdl roughness;
    ShadowingModel shadowingModel;
    NormalDistFunction ndf;
    Vec3 absorption;
    float refractionIndex;
} WalterParams; /* WALTER and MICROFACET */
typedef struct { TextureHdl radiance; Vec3 scale; } EmissiveParams;
typedef struct { TextureHdl albedo; float roughness; } OrennayarParams;
struct MaterialParamsStruct;
typedef struct { float factor; struct MaterialParamsStruct* mat; } BlendLayer;
typedef struct { BlendLayer a; BlendLayer b; } BlendParams;
typedef struct { Vec2 refractionIndex; struct MaterialParamsStruct* a; struct MaterialParamsStruct* b; } FresnelParams;
typedef struct { TextureHdl map; TextureHdl maxMips; float bias; float scale; } MaterialParamsDisplacement;
typedef struct MaterialParamsStruct {
    Medium outerMedium;
    MaterialParamType innerType;
    TextureHdl alpha;
    MaterialParamsDisplacement displacement;
    union {
        LambertParams lambert;
        TorranceParams torrance;
        WalterParams walter;
        EmissiveParams emissive;
        OrennayarParams orennayar;
        BlendParams blend;
        FresnelParams fresnel;
    } inner;
} MaterialParams;

/* ---- DLL interface (interface.cpp:199-240, 4295-4500) */
const char* core_get_dll_error(void);
Boolean core_set_logger(void (*logCallback)(const char*, int));
Boolean core_set_log_level(LogLevel level);
MufflonInstanceHdl mufflon_initialize(void);
void mufflon_destroy(MufflonInstanceHdl instHdl);
/* A HIP device is present (the reference asks for a CUDA device with unified
 * addressing, interface.cpp:4470-4484). */
Boolean mufflon_is_cuda_available(void);
int32_t mufflon_get_cuda_device_index(void);
/* The scene-file loader's callbacks (core_interface.h:355-357): `func` loads
 * LoD `level` of an object on demand (world_load_scenario calls it for every
 * LoD an active instance needs that is not loaded yet), `objFunc` reads an
 * object's unique material indices, `metaFunc` the LoD sizes. */
Boolean mufflon_set_lod_loader(MufflonInstanceHdl instHdl, Boolean (*func)(void*, ObjectHdl, uint32_t, Boolean),
                               Boolean (*objFunc)(void*, uint32_t, uint16_t*, uint32_t*),
                               Boolean (*metaFunc)(void*, LodMetadata*, size_t*), void* userParams);
/* Running mean (or variance) of a target of the current render; the buffer
 * (width * height * channels floats, row-major) is owned by the instance and
 * valid until the next call (interface.cpp:241-262). */
Boolean mufflon_get_target_image(MufflonInstanceHdl instHdl, const char* name, Boolean variance, const float** ptr);
Boolean mufflon_get_target_image_num_channels(MufflonInstanceHdl instHdl, int* numChannels);

/* ---- World (interface.cpp:1775-1800, world_container.cpp:31-82) */
void world_clear_all(MufflonInstanceHdl instHdl);
Boolean world_finalize(MufflonInstanceHdl instHdl, const Vec3 min, const Vec3 max, const char** msg);
void world_reserve_objects_instances(MufflonInstanceHdl instHdl, const uint32_t objects, const uint32_t instances);
Boolean world_set_frame_current(MufflonInstanceHdl instHdl, const uint32_t animationFrame);
Boolean world_get_frame_current(MufflonInstanceHdl instHdl, uint32_t* animationFrame);

/* ---- Polygons (interface.cpp:316-560) */
Boolean polygon_reserve(LodHdl lvlDtl, size_t vertices, size_t tris, size_t quads);
VertexHdl polygon_add_vertex(LodHdl lvlDtl, Vec3 point, Vec3 normal, Vec2 uv);
FaceHdl polygon_add_triangle(LodHdl lvlDtl, UVec3 vertices);
FaceHdl polygon_add_triangle_material(LodHdl lvlDtl, UVec3 vertices, MatIdx idx);
FaceHdl polygon_add_quad(LodHdl lvlDtl, UVec4 vertices);
FaceHdl polygon_add_quad_material(LodHdl lvlDtl, UVec4 vertices, MatIdx idx);
VertexHdl polygon_add_vertex_bulk(LodHdl lvlDtl, size_t count, const BulkLoader* points, const BulkLoader* normals,
                                  const BulkLoader* uvs, const AABB* aabb, size_t* pointsRead, size_t* normalsRead,
                                  size_t* uvsRead);
Boolean polygon_set_vertex_normal(LodHdl lvlDtl, VertexHdl vertex, Vec3 normal);
Boolean polygon_set_vertex_uv(LodHdl lvlDtl, VertexHdl vertex, Vec2 uv);
Boolean polygon_set_material_idx(LodHdl lvlDtl, FaceHdl face, MatIdx idx);
size_t polygon_set_material_idx_bulk(LodHdl lvlDtl, FaceHdl startFace, size_t count, const BulkLoader* stream);
size_t polygon_get_vertex_count(LodHdl lvlDtl);
size_t polygon_get_face_count(LodHdl lvlDtl);
size_t polygon_get_triangle_count(LodHdl lvlDtl);
size_t polygon_get_quad_count(LodHdl lvlDtl);
Boolean polygon_get_bounding_box(LodHdl lvlDtl, Vec3* min, Vec3* max);

/* ---- Spheres (interface.cpp:754-940): a sphere stream holds center xyz and
 * radius per sphere (16 bytes, ei::Sphere). Sphere ids follow the LoD's
 * triangles and quads in its primitive order. */
Boolean spheres_reserve(LodHdl lvlDtl, size_t count);
SphereHdl spheres_add_sphere(LodHdl lvlDtl, Vec3 point, float radius);
SphereHdl spheres_add_sphere_material(LodHdl lvlDtl, Vec3 point, float radius, MatIdx idx);
SphereHdl spheres_add_sphere_bulk(LodHdl lvlDtl, size_t count, const BulkLoader* stream, const AABB* aabb,
                                  size_t* readSpheres);
Boolean spheres_set_material_idx(LodHdl lvlDtl, SphereHdl sphere, MatIdx idx);
size_t spheres_set_material_idx_bulk(LodHdl lvlDtl, SphereHdl startSphere, size_t count, const BulkLoader* stream);
size_t spheres_get_sphere_count(LodHdl lvlDtl);
Boolean spheres_get_bounding_box(LodHdl lvlDtl, Vec3* min, Vec3* max);

/* ---- Objects and instances (interface.cpp:880-1060) */
ObjectHdl world_create_object(MufflonInstanceHdl instHdl, const char* name, ObjectFlags flags);
ObjectHdl world_get_object(MufflonInstanceHdl instHdl, const char* name);
const char* world_get_object_name(ObjectHdl obj);
Boolean object_has_lod(ConstObjectHdl obj, LodLevel level);
Boolean object_allocate_lod_slots(ObjectHdl obj, LodLevel slots);
LodHdl object_add_lod(ObjectHdl obj, LodLevel level, Boolean asReduced);
Boolean object_get_id(ObjectHdl obj, uint32_t* id);
InstanceHdl world_create_instance(MufflonInstanceHdl instHdl, ObjectHdl obj, const uint32_t animationFrame);
uint32_t world_get_instance_count(MufflonInstanceHdl instHdl, const uint32_t frame);
InstanceHdl world_get_instance_by_index(MufflonInstanceHdl instHdl, uint32_t index, const uint32_t animationFrame);
Boolean instance_set_transformation_matrix(MufflonInstanceHdl instHdl, InstanceHdl inst, const Mat3x4* mat,
                                           const Boolean isWorldToInst);
Boolean instance_get_transformation_matrix(MufflonInstanceHdl instHdl, InstanceHdl inst, Mat3x4* mat);

/* ---- Materials and textures (interface.cpp:1259-1490, 1900-1990) */
MaterialHdl world_add_material(MufflonInstanceHdl instHdl, const char* name, const MaterialParams* mat);
IndexType world_get_material_count(MufflonInstanceHdl instHdl);
MaterialHdl world_get_material(MufflonInstanceHdl instHdl, IndexType index);
const char* world_get_material_name(MaterialHdl material);
TextureHdl world_add_texture_value(MufflonInstanceHdl instHdl, const float* value, int num, TextureSampling sampling);
/* Loads a PFM file (the reference's pfmloader plugin, pfmloader/interface.cpp:75-160);
 * callback, if given, may rewrite every texel. */
TextureHdl world_add_texture(MufflonInstanceHdl instHdl, const char* path, TextureSampling sampling, MipmapType type,
                             TextureCallback callback, void* userParams);
TextureHdl world_get_texture(MufflonInstanceHdl instHdl, const char* path);
const char* world_get_texture_name(TextureHdl hdl);
Boolean world_get_texture_size(TextureHdl hdl, IVec2* size);

/* ---- Cameras (interface.cpp:1492-1520, camera.hpp:44-97) */
CameraHdl world_add_pinhole_camera(MufflonInstanceHdl instHdl, const char* name, const Vec3* position, const Vec3* dir,
                                   const Vec3* up, const uint32_t pathCount, float near, float far, float vFov);
CameraHdl world_get_camera(MufflonInstanceHdl instHdl, const char* name);
size_t world_get_camera_count(MufflonInstanceHdl instHdl);
CameraHdl world_get_camera_by_index(MufflonInstanceHdl instHdl, size_t index);
const char* world_get_camera_name(ConstCameraHdl cam);
Boolean world_get_camera_position(ConstCameraHdl cam, Vec3* pos, const uint32_t pathIndex);
Boolean world_get_camera_direction(ConstCameraHdl cam, Vec3* dir, const uint32_t pathIndex);
Boolean world_get_camera_up(ConstCameraHdl cam, Vec3* up, const uint32_t pathIndex);
Boolean world_set_camera_position(CameraHdl cam, Vec3 pos, const uint32_t pathIndex);
Boolean world_set_camera_direction(CameraHdl cam, Vec3 dir, Vec3 up, const uint32_t pathIndex);
Boolean world_get_pinhole_camera_fov(ConstCameraHdl cam, float* vFov);
Boolean world_set_pinhole_camera_fov(CameraHdl cam, float vFov);

/* ---- Lights (interface.cpp:1544-1600, 3023-3530) */
LightHdl world_add_light(MufflonInstanceHdl instHdl, const char* name, LightType type, const uint32_t count);
LightHdl world_add_background_light(MufflonInstanceHdl instHdl, const char* name, BackgroundType type);
Boolean world_find_light(MufflonInstanceHdl instHdl, const char* name, LightHdl* hdl);
size_t world_get_point_light_count(MufflonInstanceHdl instHdl);
size_t world_get_env_light_count(MufflonInstanceHdl instHdl);
Boolean world_get_point_light_position(MufflonInstanceHdl instHdl, LightHdl hdl, Vec3* pos, const uint32_t frame);
Boolean world_get_point_light_intensity(MufflonInstanceHdl instHdl, LightHdl hdl, Vec3* intensity,
                                        const uint32_t frame);
Boolean world_set_point_light_position(MufflonInstanceHdl instHdl, LightHdl hdl, Vec3 pos, const uint32_t frame);
Boolean world_set_point_light_intensity(MufflonInstanceHdl instHdl, LightHdl hdl, Vec3 intensity,
                                        const uint32_t frame);
Boolean world_set_env_light_map(MufflonInstanceHdl instHdl, LightHdl hdl, TextureHdl tex);
Boolean world_set_env_light_color(MufflonInstanceHdl instHdl, LightHdl hdl, Vec3 color);
Boolean world_get_env_light_color(MufflonInstanceHdl instHdl, LightHdl hdl, Vec3* color);
Boolean world_set_env_light_scale(MufflonInstanceHdl instHdl, LightHdl hdl, Vec3 color);
Boolean world_get_env_light_scale(MufflonInstanceHdl instHdl, LightHdl hdl, Vec3* color);

/* ---- Scenarios (interface.cpp:2600-2900) */
void world_reserve_scenarios(MufflonInstanceHdl instHdl, const uint32_t scenarios);
ScenarioHdl world_create_scenario(MufflonInstanceHdl instHdl, const char* name);
ScenarioHdl world_find_scenario(MufflonInstanceHdl instHdl, const char* name);
uint32_t world_get_scenario_count(MufflonInstanceHdl instHdl);
ScenarioHdl world_get_scenario_by_index(MufflonInstanceHdl instHdl, uint32_t index);
ConstScenarioHdl world_get_current_scenario(MufflonInstanceHdl instHdl);
Boolean world_finalize_scenario(MufflonInstanceHdl instHdl, ScenarioHdl scenario, const char** msg);
/* Uploads the scenario to the device (creating the device context on first
 * use): LoD BVHs, instance level, materials, textures, light tree,
 * background and camera. */
SceneHdl world_load_scenario(MufflonInstanceHdl instHdl, ScenarioHdl scenario);
SceneHdl world_get_current_scene(MufflonInstanceHdl instHdl);
Boolean scene_get_bounding_box(SceneHdl scene, Vec3* min, Vec3* max);
const char* scenario_get_name(ScenarioHdl scenario);
LodLevel scenario_get_global_lod_level(ScenarioHdl scenario);
Boolean scenario_set_global_lod_level(ScenarioHdl scenario, LodLevel level);
Boolean scenario_get_resolution(ScenarioHdl scenario, uint32_t* width, uint32_t* height);
Boolean scenario_set_resolution(ScenarioHdl scenario, uint32_t width, uint32_t height);
CameraHdl scenario_get_camera(ScenarioHdl scenario);
Boolean scenario_set_camera(MufflonInstanceHdl instHdl, ScenarioHdl scenario, CameraHdl cam);
Boolean scenario_is_object_masked(ScenarioHdl scenario, ObjectHdl obj);
Boolean scenario_mask_object(ScenarioHdl scenario, ObjectHdl obj);
Boolean scenario_mask_instance(ScenarioHdl scenario, InstanceHdl inst);
LodLevel scenario_get_object_lod(ScenarioHdl scenario, ObjectHdl obj);
Boolean scenario_set_object_lod(ScenarioHdl scenario, ObjectHdl obj, LodLevel level);
Boolean scenario_set_instance_lod(ScenarioHdl scenario, InstanceHdl inst, LodLevel level);
void scenario_reserve_custom_object_properties(ScenarioHdl scenario, size_t objects);
void scenario_reserve_custom_instance_properties(ScenarioHdl scenario, size_t instances);
IndexType scenario_get_point_light_count(ScenarioHdl scenario);
Boolean scenario_has_envmap_light(MufflonInstanceHdl instHdl, ScenarioHdl scenario);
Boolean scenario_add_light(MufflonInstanceHdl instHdl, ScenarioHdl scenario, LightHdl hdl);
Boolean scenario_remove_light(MufflonInstanceHdl instHdl, ScenarioHdl scenario, LightHdl hdl);
void scenario_reserve_material_slots(ScenarioHdl scenario, size_t count);
MatIdx scenario_declare_material_slot(ScenarioHdl scenario, const char* name, size_t nameLength);
MatIdx scenario_get_material_slot(ScenarioHdl scenario, const char* name, size_t nameLength);
const char* scenario_get_material_slot_name(ScenarioHdl scenario, MatIdx slot);
size_t scenario_get_material_slot_count(ScenarioHdl scenario);
MaterialHdl scenario_get_assigned_material(ScenarioHdl scenario, MatIdx index);
Boolean scenario_assign_material(ScenarioHdl scenario, MatIdx index, MaterialHdl handle);

/* ---- Rendering (interface.cpp:3540-3990). One renderer, "Pathtracer" /
 * "PT", with two variations on DEVICE_CUDA (the GPU device mask):
 * 0 = megakernel, 1 = wavefront with HIP graph replay. Targets: "Radiance",
 * "Position", "Depth", "Normal", "UV", "Albedo", "Lightness", "Hit ID"
 * (PtTargets, pt_params.hpp:16-26). */
uint32_t render_get_renderer_count(MufflonInstanceHdl instHdl);
uint32_t render_get_renderer_variations(MufflonInstanceHdl instHdl, uint32_t index);
const char* render_get_renderer_name(MufflonInstanceHdl instHdl, uint32_t index);
const char* render_get_renderer_short_name(MufflonInstanceHdl instHdl, uint32_t index);
RenderDevice render_get_renderer_devices(MufflonInstanceHdl instHdl, uint32_t index, uint32_t variation);
Boolean render_enable_renderer(MufflonInstanceHdl instHdl, uint32_t index, uint32_t variation);
/* One iteration, synchronous (the times are host times around the device work). */
Boolean render_iterate(MufflonInstanceHdl instHdl, ProcessTime* iterateTime, ProcessTime* preTime,
                       ProcessTime* postTime);
uint32_t render_get_current_iteration(ConstMufflonInstanceHdl instHdl);
Boolean render_reset(MufflonInstanceHdl instHdl);
/* Writes a PFM (pfmloader/interface.cpp:162-210) after the reference's
 * file-name tag replacement (interface.cpp:168-195). */
Boolean render_save_screenshot(MufflonInstanceHdl instHdl, const char* filename, const char* targetName,
                               Boolean variance);
uint32_t render_get_render_target_count(MufflonInstanceHdl instHdl);
const char* render_get_render_target_name(MufflonInstanceHdl instHdl, uint32_t index);
Boolean render_enable_render_target(MufflonInstanceHdl instHdl, const char* target, Boolean variance);
Boolean render_disable_render_target(MufflonInstanceHdl instHdl, const char* target, Boolean variance);
Boolean render_is_render_target_enabled(MufflonInstanceHdl instHdl, const char* name, Boolean variance);
Boolean render_is_render_target_required(MufflonInstanceHdl instHdl, const char* name, Boolean variance);

/* ---- Renderer parameters (interface.cpp:3920-4100; PtParameters,
 * pt_params.hpp:8-14): "Seed" (int), "Min. path length" (int),
 * "Max. path length" (int), "NEE count" (int), "NEE position guide" (bool).
 * Unknown names are logged and ignored (parameter.hpp:242-247). */
uint32_t renderer_get_num_parameters(MufflonInstanceHdl instHdl);
const char* renderer_get_parameter_desc(MufflonInstanceHdl instHdl, uint32_t idx, ParameterType* type);
Boolean renderer_set_parameter_int(MufflonInstanceHdl instHdl, const char* name, int32_t value);
Boolean renderer_get_parameter_int(MufflonInstanceHdl instHdl, const char* name, int32_t* value);
Boolean renderer_set_parameter_float(MufflonInstanceHdl instHdl, const char* name, float value);
Boolean renderer_get_parameter_float(MufflonInstanceHdl instHdl, const char* name, float* value);
Boolean renderer_set_parameter_bool(MufflonInstanceHdl instHdl, const char* name, Boolean value);
Boolean renderer_get_parameter_bool(MufflonInstanceHdl instHdl, const char* name, Boolean* value);
Boolean renderer_set_parameter_enum(MufflonInstanceHdl instHdl, const char* name, int value);
Boolean renderer_get_parameter_enum(MufflonInstanceHdl instHdl, const char* name, int* value);

/* ---- Differentiable rendering. No counterpart in the reference (it has no
 * differentiable renderer): the adjoint (path replay) of the "Pathtracer"
 * renderer made reachable for callers of this API, such as the ctypes
 * binding of bindings.py:373-386, which binds core_interface.h. */
/* Replace the texels of a texture (width * height * channels floats of its
 * stored format, channels as world_get_texture_size's texture was created:
 * 1 for R32F, 4 for RGBA32F). Resets the renderer (the scene is reloaded at
 * the next iteration if the texture is a 1x1 value or the environment map). */
Boolean render_set_texture_data(MufflonInstanceHdl instHdl, TextureHdl tex, const float* texels);
/* Set the radiance scale of the EMISSIVE material assigned to a material slot
 * of the current scenario (MaterialParams' emissive.scale); resets. */
Boolean render_set_emission_scale(MufflonInstanceHdl instHdl, MatIdx slot, Vec3 scale);
/* Derivative form of render_adjoint: 0 (default) an unbiased estimate of the
 * derivative of the expected image; 1 the derivative of this realization
 * with every random decision held fixed. */
Boolean render_set_adjoint_mode(MufflonInstanceHdl instHdl, int32_t mode);
/* Given radianceAdjoint = dLoss/dRadiance (height x width x 3 floats, the
 * layout of mufflon_get_target_image) of the current Radiance target, the
 * mean of render_get_current_iteration() iterations since the last reset,
 * replay those paths and accumulate the gradients read below. Synchronous. */
Boolean render_adjoint(MufflonInstanceHdl instHdl, const float* radianceAdjoint);
/* dLoss/d(texel) of an albedo or roughness texture larger than 1x1 used by
 * the loaded scene: width * height * channels floats, the texel layout. */
Boolean render_get_texture_gradient(MufflonInstanceHdl instHdl, TextureHdl tex, float* gradient);
/* dLoss/d(emission scale) of the material in a slot of the current scenario. */
Boolean render_get_emission_gradient(MufflonInstanceHdl instHdl, MatIdx slot, Vec3* gradient);

/* ---- Profiling (interface.cpp:4180-4260): accepted and ignored; no core
 * profiler is kept. */
void profiling_enable(void);
void profiling_disable(void);

#ifdef __cplusplus
}
#endif

#endif /* MUFFLON_CORE_H */
