/*
 * mufflon_hip.h — C-ABI of the MI355X (gfx950) device of Mufflon's GPU
 * ray-tracing hot path: LBVH build, closest/any-hit traversal.
 *
 * Conventions (as Mufflon's core API, core_interface.h / interface.cpp:54-83):
 * every function returns an int status, 0 on success; on failure the message
 * is available from mfh_last_error() (thread-local) and no exception crosses
 * the boundary. All device memory is owned by the context and freed by
 * mfh_destroy. Pointers named d* are device pointers (from mfh_alloc or any
 * HIP allocation on the context's device); all others are host pointers.
 * Calls on one context are not thread-safe; callers serialize them, as the
 * core API does with its iteration mutex (interface.cpp:99-100).
 */
#ifndef MUFFLON_HIP_H
#define MUFFLON_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define MFH_OK 0
#define MFH_ERROR_INVALID_ARGUMENT 1
#define MFH_ERROR_HIP 2
#define MFH_ERROR_OUT_OF_MEMORY 3
#define MFH_ERROR_STATE 4

typedef struct mfh_context mfh_context;

/* Message of the last failed call on this thread ("" if none). */
const char* mfh_last_error(void);

/* Context on one HIP device with its own stream.
 * Replaces the device selection of mufflon_initialize (interface.cpp:4295-4425). */
int mfh_create(int32_t deviceIndex, mfh_context** outCtx);
int mfh_destroy(mfh_context* ctx);
/* Wait for all work queued on the context's stream. */
int mfh_synchronize(mfh_context* ctx);

/* Device memory helpers (the role of Mufflon's unique_device_ptr / copy()). */
int mfh_alloc(mfh_context* ctx, size_t bytes, void** outDevPtr);
int mfh_free(mfh_context* ctx, void* devPtr);
int mfh_upload(mfh_context* ctx, void* dDst, const void* src, size_t bytes);
int mfh_download(mfh_context* ctx, void* dst, const void* dSrc, size_t bytes);

/* Triangle LoD in object space: vertices as float xyz, triangles as three
 * uint32 vertex indices. The layout of PolygonsDescriptor's vertices /
 * vertexIndices for a triangle-only mesh (descriptors.hpp:33-57). */
int mfh_lod_create(mfh_context* ctx, const float* vertices, int32_t numVertices, const uint32_t* triIndices,
                   int32_t numTriangles, int32_t* outLodId);

/* LoD of triangles, quads and spheres (PolygonsDescriptor + SpheresDescriptor,
 * descriptors.hpp:33-68): quads as four uint32 vertex indices (p00, p10,
 * p11, p01 of the bilinear patch, intersection.cu:99-179), spheres as four
 * floats (center xyz, radius). Primitive ids run triangles, quads, spheres.
 * Any count may be 0 but not all; vertices may be NULL without polygons. The
 * LoD box joins the vertices' box with the spheres' boxes
 * (Lod::get_bounding_box). Replaces the polygon and sphere uploads of
 * Lod::get_descriptor (lod.cpp:16-20). */
int mfh_lod_create_mixed(mfh_context* ctx, const float* vertices, int32_t numVertices, const uint32_t* triIndices,
                         int32_t numTriangles, const uint32_t* quadIndices, int32_t numQuads, const float* spheres,
                         int32_t numSpheres, int32_t* outLodId);

/* Build the LBVH of a LoD on the device. bbox = {min xyz, max xyz} of the
 * LoD (Lod::get_bounding_box); NULL takes the box computed at creation.
 * Replaces LBVHBuilder::build(LodDescriptor&, const ei::Box&, bool)
 * (build_lbvh.cu:782-789, lod.cpp:33). */
int mfh_lbvh_build(mfh_context* ctx, int32_t lodId, const float* bbox);

/* Traversal hierarchy of later LoD builds (no counterpart in the
 * reference): every collapsed subtree of the reference's BVH with more than
 * `limit` triangles stays split in the hierarchy the trace and path-tracing
 * kernels walk, so that no traversal leaf holds more than `limit` triangles
 * (default 16; 0 walks the reference BVH itself). Hits do not change (the
 * closest hit is the lowest (t, instance, primitive) whatever the order of
 * the tests); mfh_lbvh_get and mfh_trace_closest_stats keep the reference
 * BVH and its visiting order. */
int mfh_set_traversal_leaf_limit(mfh_context* ctx, int32_t limit);


/* Result of the last build of a LoD: numInternalNodes of the collapsed BVH
 * (LBVH::numInternalNodes, lbvh.hpp:23) and, when the pointers are non-NULL,
 * the nodes (numInternalNodes * 64 bytes, the BvhNode layout of lbvh.hpp:13-18)
 * and the primitive ids in BVH order (one int32 per primitive).
 * Replaces LBVHBuilder::acquire_const (lbvh.hpp:70-86). */
int mfh_lbvh_get(mfh_context* ctx, int32_t lodId, int32_t* outNumInternal, void* outNodes, int32_t* outPrimIds);

/* Largest LoD of mfh_lbvh_build_noalloc, in triangles. */
#define MFH_NOALLOC_MAX_TRIANGLES 262144

/* LBVH build that allocates nothing. No counterpart in the reference: it is
 * for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc). The same
 * build as mfh_lbvh_build + mfh_lbvh_get over a triangle LoD held in the
 * caller's device memory: dVertices (numVertices float xyz), dTriIndices
 * (numTriangles x 3 uint32), bbox = {min xyz, max xyz} on the host
 * (required). Writes the collapsed BVH into dNodes (room for
 * 64 * (numTriangles - 1) bytes; may be NULL for one triangle), the primitive
 * ids in BVH order into dPrimIds (numTriangles int32) and the number of
 * internal nodes into *outNumInternal; returns after the build has finished.
 * Its scratch is one static device array per GPU (41.7 MiB, held until the
 * process ends), so at most MFH_NOALLOC_MAX_TRIANGLES triangles; larger LoDs
 * and out-of-range vertex indices are rejected with
 * MFH_ERROR_INVALID_ARGUMENT. Every call on one GPU shares that array:
 * calls that overlap on one GPU (other host threads, other contexts) race. */
int mfh_lbvh_build_noalloc(mfh_context* ctx, const float* dVertices, int32_t numVertices,
                           const uint32_t* dTriIndices, int32_t numTriangles, const float* bbox,
                           int32_t* outNumInternal, void* dNodes, int32_t* dPrimIds);

/* Build a LoD and copy every intermediate stage to the host (NULL skips one):
 * Morton codes in primitive order and sorted, sorted ids (N each), parents
 * (2N-1, Mufflon's encoding: left child stores ~parent), node boxes
 * ((2N-1)*8 floats: min xyz, count bits, max xyz, SAH cost) and the inclusive
 * scan of the keep marks (N-1). Synchronizes after every stage; for tests. */
int mfh_lbvh_build_stages(mfh_context* ctx, int32_t lodId, const float* bbox, uint64_t* unsortedKeys,
                          uint64_t* sortedKeys, int32_t* sortedIds, int32_t* parents, float* boxes,
                          int32_t* offsets);

/* Scene of instances over built LoDs. worldToInstance: 12 floats per instance
 * (row-major Mat3x4). With more than one instance the instance-level LBVH is
 * built (LBVHBuilder::build(SceneDescriptor&), build_lbvh.cu:796-806).
 * The LoD BVHs must be built first. A LoD rebuilt later is picked up by the
 * next trace call on the scene (its table and instance level are refreshed). */
int mfh_scene_create(mfh_context* ctx, const int32_t* instLod, const float* worldToInstance, int32_t numInstances,
                     int32_t* outSceneId);
/* Scene box (6 floats) and the instance-level BVH (numInternal, nodes, ids). */
int mfh_scene_get(mfh_context* ctx, int32_t sceneId, float* outBox, int32_t* outNumInternal, void* outNodes,
                  int32_t* outPrimIds);

/* Rays: 8 floats each {origin xyz, tmax, direction xyz (normalized), unused};
 * origins are used as given (offsetting from a surface is the caller's).
 * Closest hit (first_intersection, intersection.cu:366-621): dHits gets
 * 4 x 32 bit per ray {t (float), instance, primitive, status (float: 0 ok,
 * 1 stack overflow)}, instance = primitive = -1 on a miss, in which case t is
 * tmax after the object-space round trip. dBary (may be NULL) gets the
 * barycentric weights (w0, w1) of v0 and v1. Asynchronous on the stream. */
int mfh_trace_closest(mfh_context* ctx, int32_t sceneId, const float* dRays, int32_t numRays, void* dHits,
                      float* dBary);
/* Closest hit as above that also writes, per ray, 2 x uint32 {internal nodes
 * visited (an instance's object-box test counts as one), triangles tested}:
 * the traversal work that the algorithmic byte count is made of. */
int mfh_trace_closest_stats(mfh_context* ctx, int32_t sceneId, const float* dRays, int32_t numRays, void* dHits,
                            float* dBary, uint32_t* dStats);
/* Any hit (any_intersection, intersection.cu:624-753): 1 occluded, 0 free,
 * -1 stack overflow. Asynchronous on the stream. */
int mfh_trace_any(mfh_context* ctx, int32_t sceneId, const float* dRays, int32_t numRays, int32_t* dOccluded);

/* ------------------------------------------------------------------------
 * Path tracer (the GPU device of Mufflon's "Pathtracer"/"PT" renderer,
 * gpu_pt.cpp / gpu_pt.cu / pt_common.hpp).
 * ------------------------------------------------------------------------ */

/* Material record (128 bytes). type follows Mufflon's Materials enum
 * (material_definitions.hpp:25-46); the world_add_material descriptor it
 * replaces is core_interface.h's MaterialParams (inner/outer medium, layer
 * parameters). Texture slots hold a texture id of mfh_texture_create or -1
 * for the constant in the record (Mufflon's 1x1 value textures,
 * world_add_texture_value). The inner medium of WALTER and MICROFACET is
 * {refractionIndex, absorption}; every other type has inner = outer medium
 * (Material::compute_medium, material.cpp:54-60). */
#define MFH_MATERIAL_EMISSIVE 0
#define MFH_MATERIAL_LAMBERT 1
#define MFH_MATERIAL_TORRANCE 3
#define MFH_MATERIAL_WALTER 4 /* rough refraction (microfacet_refractive.hpp); inner medium as MICROFACET */
#define MFH_MATERIAL_MICROFACET 10
#define MFH_NDF_GGX 0
#define MFH_NDF_BECKMANN 1
#define MFH_SHADOWING_VCAVITY 0
#define MFH_SHADOWING_SMITH 1
typedef struct mfh_material {
    int32_t type;
    float albedo[3];             /* LAMBERT, TORRANCE: albedo (ALBEDO texture) */
    float radiance[3];           /* EMISSIVE: radiance (EMISSION texture) */
    float scale[3];              /* EMISSIVE: radiance scale (MatEmissive::NonTexParams) */
    float roughness[2];          /* TORRANCE, WALTER, MICROFACET: alpha (ROUGHNESS texture) */
    int32_t ndf;                 /* TORRANCE, WALTER, MICROFACET: MFH_NDF_* */
    int32_t shadowing;           /* TORRANCE, WALTER, MICROFACET: MFH_SHADOWING_* */
    float refractionIndex;       /* WALTER, MICROFACET: inner medium */
    float absorption[3];         /* WALTER, MICROFACET: inner medium absorption per unit length */
    float outerRefractionIndex;  /* medium on the normal side, 1 = vacuum */
    float outerAbsorption[3];
    int32_t albedoTexture;       /* -1 or a texture id */
    int32_t roughnessTexture;    /* -1 or a texture id (1 channel: isotropic, 2+: x/y) */
    int32_t emissionTexture;     /* must be -1 (emission textures are not supported) */
    int32_t reserved[7];
} mfh_material;

/* Textures (world_add_texture / world_add_texture_value, interface.cpp:1938-2100)
 * with the CPU texture's sampling (cputexture.cpp:372-411): NEAREST reads
 * texel floor(uv * size); LINEAR blends 4 texels with weights frac(uv * size)
 * (no half-texel offset); both wrap x and y by a non-negative modulo. */
#define MFH_SAMPLING_NEAREST 0
#define MFH_SAMPLING_LINEAR 1
/* texels: width * height * channels floats, row-major, 1 <= channels <= 4. */
int mfh_texture_create(mfh_context* ctx, int32_t width, int32_t height, int32_t channels, int32_t samplingMode,
                       const float* texels, int32_t* outTextureId);

/* Replace the texels of a texture (same size and channels). Path tracers
 * read the new values from their next iteration; reset them to restart. */
int mfh_texture_update(mfh_context* ctx, int32_t textureId, const float* texels);

/* Background (BackgroundDesc, background.hpp; world_set_env_light...):
 * COLORED = constant radiance color * scale; ENVMAP = polar (equirectangular)
 * map texture * scale, importance sampled through a summed-area table. */
#define MFH_BACKGROUND_COLORED 0
#define MFH_BACKGROUND_ENVMAP 1
typedef struct mfh_background {
    int32_t type;
    float color[3];
    float scale[3];        /* per channel (Background::set_scale, background.hpp:119-123) */
    int32_t envmapTexture; /* ENVMAP: texture id (RGB, polar) */
} mfh_background;

/* Point light (lights.hpp:46-50): position and intensity (W/sr). */
typedef struct mfh_point_light {
    float position[3];
    float intensity[3];
} mfh_point_light;

/* Pinhole camera (PinholeParams, pinhole.hpp:43-51); tanVFov = tan(vFov / 2). */
typedef struct mfh_pinhole {
    float position[3];
    float tanVFov;
    float viewDir[3];
    float up[3];
    int32_t width;
    int32_t height;
} mfh_pinhole;

/* PtParameters (pt_params.hpp:8-14, parameter.hpp:327-366). */
typedef struct mfh_pt_params {
    int32_t seed;          /* "Seed", default 0 */
    int32_t minPathLength; /* "Min. path length", default 0 */
    int32_t maxPathLength; /* "Max. path length", default 16 */
    int32_t neeCount;      /* "NEE count", default 1 */
    int32_t neePositionGuide; /* "NEE position guide", default 0 (false): light tree
                                 traversal by flux over squared distance to the
                                 shading point (guide_flux_pos,
                                 light_tree_sampling.hpp:64-71) */
} mfh_pt_params;

/* Render targets of the pt renderer (PtTargets, pt_params.hpp:16-26;
 * render_targets.hpp:13-74), running means per pixel, row-major. Radiance
 * and Hit ID are enabled when a path tracer is created. */
#define MFH_TARGET_RADIANCE 0  /* 3 floats per pixel */
#define MFH_TARGET_HIT_ID 1    /* 2 floats per pixel: instance, primitive */
#define MFH_TARGET_POSITION 2  /* 3 floats, guide-weighted */
#define MFH_TARGET_DEPTH 3     /* 1 float */
#define MFH_TARGET_NORMAL 4    /* 3 floats (shading normal) */
#define MFH_TARGET_UV 5        /* 2 floats */
#define MFH_TARGET_ALBEDO 6    /* 3 floats */
#define MFH_TARGET_LIGHTNESS 7 /* 1 float */
#define MFH_TARGET_COUNT 8

/* Light tree as built on the host (light_tree.cpp:340-447), Mufflon layout. */
typedef struct mfh_light_tree_info {
    float posRootFlux;
    uint32_t posRootType;
    uint32_t posInternalCount;
    uint32_t posLightCount;
    uint32_t memBytes;
    int32_t mapCount;
    float dirRootFlux;
    float backgroundFlux;
} mfh_light_tree_info;

/* Emissive triangle in world space for the light tree (AreaLightTriangleDesc,
 * lights.hpp:77-81, plus the primitive it was made from). */
typedef struct mfh_area_light {
    float points[9];
    float uv[6];
    int32_t material;
    int32_t instance;
    int32_t primitive;
} mfh_area_light;

/* Host-only light-tree build (no device needed): area lights and point
 * lights in the order Mufflon gathers them, the scene box (for the Hilbert
 * ordering) and the materials (for the emitted flux). Capacities: outMem at
 * least 64 * (numAreaLights + numPointLights) * 2 bytes, outMap* at least
 * numAreaLights entries. */
int mfh_light_tree_build(const mfh_area_light* areaLights, int32_t numAreaLights, const mfh_point_light* pointLights,
                         int32_t numPointLights, const float* sceneBox, const mfh_material* materials,
                         int32_t numMaterials, mfh_light_tree_info* outInfo, void* outMem, size_t memCapacity,
                         uint64_t* outMapKeys, uint32_t* outMapCodes, int32_t mapCapacity);

/* Emissive triangle, quad or sphere in world space for the light tree
 * (AreaLightTriangleDesc / AreaLightQuadDesc / AreaLightSphereDesc,
 * lights.hpp:77-91, plus the primitive it was made from). kind 0: points
 * 0..2 and uv 0..2 of a triangle; 1: points 0..3 and uv 0..3 of a quad (its
 * vertices in patch order p00, p10, p11, p01); 2: a sphere, center in
 * points 0..2, and radius. */
typedef struct mfh_area_light_ex {
    int32_t kind;
    float points[12];
    float uv[8];
    float radius;
    int32_t material;
    int32_t instance;
    int32_t primitive;
} mfh_area_light_ex;

/* mfh_light_tree_build over triangles, quads and spheres. Records:
 * AreaLightQuad 96 bytes, AreaLightSphere 32 bytes (lights.hpp:108-137); a
 * quad's flux samples its area density at 128 golden-ratio points seeded with
 * the bits of its parallelogram area (the reference seeds with the record's
 * memory address, lights.cpp:46). Capacities: outMem at least
 * 96 * (numAreaLights + numPointLights) * 2 bytes. */
int mfh_light_tree_build_ex(const mfh_area_light_ex* areaLights, int32_t numAreaLights,
                            const mfh_point_light* pointLights, int32_t numPointLights, const float* sceneBox,
                            const mfh_material* materials, int32_t numMaterials, mfh_light_tree_info* outInfo,
                            void* outMem, size_t memCapacity, uint64_t* outMapKeys, uint32_t* outMapCodes,
                            int32_t mapCapacity);

/* Shading attributes of a LoD: per-vertex normals (xyz) and uvs, and a
 * material index per primitive (PolygonsDescriptor normals/uvs/matIndices,
 * then SpheresDescriptor matIndices); normals and uvs may be NULL for a LoD
 * without vertices. */
int mfh_lod_set_shading(mfh_context* ctx, int32_t lodId, const float* normals, const float* uvs,
                        const uint16_t* materialIndices);

/* Path tracer over a scene. Emissive triangles of every instance become area
 * lights (world_container.cpp:869-990); point lights are added; the light tree
 * is built on the host and uploaded. Starts reset. Replaces
 * render_enable_renderer + the scene-descriptor upload of
 * RendererBase::pre_iteration (renderer_base.hpp:41-61). */
int mfh_pt_create(mfh_context* ctx, int32_t sceneId, const mfh_material* materials, int32_t numMaterials,
                  const mfh_point_light* pointLights, int32_t numPointLights, const mfh_pinhole* camera,
                  const mfh_pt_params* params, int32_t* outPtId);
/* Replace the parameters (renderer_set_parameter_*, interface.cpp; a change
 * resets the render as IRenderer::on_renderer_parameter_changed does,
 * renderer.hpp:91-143). */
int mfh_pt_set_params(mfh_context* ctx, int32_t ptId, const mfh_pt_params* params);
/* Replace the camera, resolution included (scenario_set_camera /
 * scenario_set_resolution / world_set_camera_*; resets the render). */
int mfh_pt_set_camera(mfh_context* ctx, int32_t ptId, const mfh_pinhole* camera);
/* Re-seed the per-pixel RNGs and clear the targets (render_reset,
 * GpuPathTracer::post_reset, gpu_pt.cpp:33-38). */
int mfh_pt_reset(mfh_context* ctx, int32_t ptId);
/* Run iterations (render_iterate, interface.cpp:3611-3684); asynchronous. */
int mfh_pt_iterate(mfh_context* ctx, int32_t ptId, int32_t iterations);
/* Enable or disable a render target (render_enable_render_target /
 * render_disable_render_target, interface.cpp:3876-3940); resets the path tracer. */
int mfh_pt_enable_target(mfh_context* ctx, int32_t ptId, int32_t target, int32_t enable);
/* Record the running variance of a target (render_enable_render_target with
 * variance = true, OutputHandler::enable_render_target, output_handler.hpp:391-405;
 * enabling also enables the target). Disabling drops only the variance
 * (disable_render_target with variance = true, :407-430). Resets the path
 * tracer. The variance follows UpdateIter (output_handler_variance.hpp:7-31):
 * diff = x - mean; mean += diff / max(1, i + 1);
 * var += (diff * (x - mean) - var) / max(1, i) for iteration index i. */
int mfh_pt_enable_variance(mfh_context* ctx, int32_t ptId, int32_t target, int32_t enable);
/* Copy a target's running variance (same layout as mfh_pt_get_target;
 * mufflon_get_target_image with variance = true, interface.cpp:241-262). */
int mfh_pt_get_variance(mfh_context* ctx, int32_t ptId, int32_t target, float* out);
/* Execution form of an iteration. MEGAKERNEL: one thread runs a pixel's
 * whole path (GpuPathTracer's sample_pt kernel, gpu_pt.cu:54-138).
 * WAVEFRONT: persistent kernels drain ray queues stage by stage (shade,
 * shadow rays, extension rays) with wave-level compaction; with useGraph
 * each iteration is replayed from a captured HIP graph. Both forms produce
 * identical results. No counterpart in the reference (its GPU device is the
 * megakernel only). */
#define MFH_PT_MEGAKERNEL 0
#define MFH_PT_WAVEFRONT 1
int mfh_pt_set_mode(mfh_context* ctx, int32_t ptId, int32_t mode, int32_t useGraph);
/* Launch shape of the wavefront trace stages: block size x LDS stack depth
 * (64x8, 64x16, 128x8, 128x16, 256x8, 256x16, 256x24), the largest queue
 * chunk a wave grabs at once (multiple of 64; smaller for short queues), and
 * the number of leading bounces whose extension and shadow rays are traced
 * in sorted order (origin cell, direction octant). */
typedef struct mfh_wavefront_tuning {
    int32_t traceBlock;
    int32_t stackDepth;
    int32_t queueChunk;
    int32_t sortBounces;
    int32_t refillBudget; /* 0: chunked queue walk; n > 0: lanes whose ray ended take the next ray
                             every n traversal steps once a quarter of the wave is idle */
    int32_t leafBatch;    /* 1 or 8: triangles of a BVH leaf loaded ahead of their tests (long
                             leaves); also used by the megakernel form; the refill kernels use 1 */
    int32_t tailAfter;    /* 0: one launch per stage and bounce; k > 0: after the extension rays of
                             bounce k one launch finishes every remaining path (megakernel-style,
                             same bits), so the last bounces of a few long paths do not cost three
                             launches each */
} mfh_wavefront_tuning;
int mfh_pt_set_tuning(mfh_context* ctx, int32_t ptId, const mfh_wavefront_tuning* tuning);
int mfh_pt_get_tuning(mfh_context* ctx, int32_t ptId, mfh_wavefront_tuning* outTuning);
/* Times `iterations` wavefront iterations for every tuning candidate (block x
 * stack depth x leaf batch 1/8 x {chunk 256, chunk 1024, refill 16, refill 32}
 * x sorted bounces 0/1/2), then the tail bounce (never, 1, 2, 3, 4, 6) on the
 * fastest of those, keeps the fastest and resets the path tracer. */
int mfh_pt_autotune(mfh_context* ctx, int32_t ptId, int32_t iterations, mfh_wavefront_tuning* outBest,
                    float* outMsPerIteration);
/* Batches: several path tracers (scenes of any size, resolution, path
 * length) iterated together by the wavefront stages, one launch per stage
 * over the paths of all scenes (optionally replayed from a HIP graph). The
 * batch resets its path tracers on creation and on mfh_batch_reset; their
 * targets are read with mfh_pt_get_target. No counterpart in the reference. */
int mfh_batch_create(mfh_context* ctx, const int32_t* ptIds, int32_t count, int32_t useGraph, int32_t* outBatchId);
int mfh_batch_reset(mfh_context* ctx, int32_t batchId);
int mfh_batch_iterate(mfh_context* ctx, int32_t batchId, int32_t iterations);
int mfh_batch_set_tuning(mfh_context* ctx, int32_t batchId, const mfh_wavefront_tuning* tuning);
int mfh_batch_autotune(mfh_context* ctx, int32_t batchId, int32_t iterations, mfh_wavefront_tuning* outBest,
                       float* outMsPerIteration);
/* Running mean of an enabled target over the completed iterations (mufflon_get_target_image). */
int mfh_pt_get_target(mfh_context* ctx, int32_t ptId, int32_t target, float* out);
/* Index of the last completed iteration, -1 after a reset (render_get_current_iteration). */
int mfh_pt_get_iteration(mfh_context* ctx, int32_t ptId, int32_t* outIteration);
/* Number of NaN contributions dropped since the last reset. */
int mfh_pt_get_nan_count(mfh_context* ctx, int32_t ptId, uint32_t* outCount);
/* The host-built light tree: info, then (when non-NULL) the memory blob
 * (info.memBytes), and the primitive -> tree-path map (info.mapCount entries,
 * keys = instance << 32 | primitive, sorted). */
int mfh_pt_get_light_tree(mfh_context* ctx, int32_t ptId, mfh_light_tree_info* outInfo, void* outMem,
                          uint64_t* outMapKeys, uint32_t* outMapCodes);

/* Host-only background preparation (no device needed), shared by the device
 * upload and any CPU renderer: the flux the light tree weighs the background
 * with (Background::acquire_const, background.cpp:30-66, and its
 * compute_envmap_flux / compute_constant_flux, :78-135) into outFlux[3], and
 * for ENVMAP the summed-area table of create_summed_area_table
 * (texture_sampling.hpp:49-110) into outSat (width * height floats).
 * envTexels as in mfh_texture_create (NULL for COLORED). */
int mfh_background_build(const mfh_background* bg, const float* envTexels, int32_t width, int32_t height,
                         int32_t channels, const float* sceneBox, float* outFlux, float* outSat);
/* Sets the background of a path tracer (world_set_env_light +
 * the light-tree background of LightTreeBuilder, light_tree.cpp:592-597) and
 * resets it. The default background is COLORED black. */
int mfh_pt_set_background(mfh_context* ctx, int32_t ptId, const mfh_background* bg);

/* Wireframe renderer over a path tracer's scene, camera and background: the
 * GPU device of Mufflon's "Wireframe"/"WF" renderer (GpuWireframe,
 * gpu_wireframe.cpp:18-50, gpu_wireframe.cu:19-51, wireframe_common.hpp:94-181).
 * Each iteration draws one seed per pixel from the renderer's own PCG64
 * (default-constructed when the path tracer is created; resets do not
 * re-seed it), jitters the camera ray as the path tracer does, and walks the
 * ray through the scene (at most maxTraceDepth hits): a hit whose closest
 * triangle edge lies within lineWidth pixel footprints of the camera ray
 * scores the border colour 1, a walk that leaves the scene scores the
 * background radiance if positive. The per-iteration values go into the
 * running mean of the Border target (3 floats per pixel, row-major) and,
 * when recorded, its variance (the output handler's recursion).
 * _reset clears the Border target (call it after changing the camera);
 * _get copies the mean (variance = 0) or the variance (1) to host memory. */
int mfh_pt_wireframe_reset(mfh_context* ctx, int32_t ptId, int32_t recordVariance);
int mfh_pt_wireframe_iterate(mfh_context* ctx, int32_t ptId, int32_t iterations, float lineWidth,
                             int32_t maxTraceDepth);
int mfh_pt_wireframe_get(mfh_context* ctx, int32_t ptId, int32_t variance, float* hostOut);

/* Light tracer over a path tracer's scene, camera and lights: the GPU device
 * of Mufflon's "Lighttracer"/"LT" renderer (GpuLightTracer, gpu_lt.cpp:25-37,
 * gpu_lt.cu:19-101, lt_sample of lt_common.hpp:17-78). An iteration traces
 * width * height light paths: path i emits from the light tree with index i
 * (emit, light_tree_sampling.hpp:81-187), walks on adjoint BSDF samples up
 * to the path tracer's maximum path length, and every vertex between its
 * minimum and maximum length is connected to the pinhole camera and added to
 * the pixel it projects to (float atomics: sums agree with a sequential
 * evaluation to rounding). Targets (LtTargets, lt_params.hpp:13-17):
 * MFH_LT_RADIANCE (3 floats per pixel), MFH_LT_LIGHTNESS (1), MFH_LT_DENSITY
 * (1, recorded when `density` is set: it also tests the visibility of
 * connections that carry no radiance). Seeded from the path tracer's seed as
 * the reference seeds it. Positional lights only: a background that emits is
 * refused. _reset clears the targets and re-seeds; _get copies the running
 * mean (variance = 0) or variance (1) of one target to host memory. */
#define MFH_LT_RADIANCE 0
#define MFH_LT_LIGHTNESS 1
#define MFH_LT_DENSITY 2
int mfh_pt_lighttracer_reset(mfh_context* ctx, int32_t ptId, int32_t recordVariance);
int mfh_pt_lighttracer_iterate(mfh_context* ctx, int32_t ptId, int32_t iterations, int32_t density);
int mfh_pt_lighttracer_get(mfh_context* ctx, int32_t ptId, int32_t target, int32_t variance, float* hostOut);

/* Adjoint (path replay) of the Radiance target. No counterpart in the
 * reference (it has no differentiable renderer). For the image of
 * `iterations` iterations from a reset, and radianceAdjoint = dLoss/dRadiance
 * (host, 3 floats per pixel, row-major), replays the same paths and
 * accumulates dLoss/d(texel) of every albedo texture (LAMBERT, TORRANCE),
 * every roughness texture (TORRANCE, WALTER, MICROFACET; one channel drives both
 * roughnesses, two channels x and y) and dLoss/d(emission scale) of every
 * material. By default (MFH_ADJOINT_EXPECTED) the result is an unbiased
 * estimate of the derivative of the expected image: each contribution times
 * the d ln of the BSDF values and emission it contains and of its MIS
 * weight (sampling pdfs and Russian-roulette probabilities need no term).
 * Maximum path length 32. Asynchronous; read with the getters below. */
int mfh_pt_adjoint(mfh_context* ctx, int32_t ptId, int32_t iterations, const float* radianceAdjoint);
/* Which derivative mfh_pt_adjoint forms (no counterpart in the reference):
 * MFH_ADJOINT_EXPECTED (0, default) the derivative of the expected image;
 * MFH_ADJOINT_FIXED_DECISIONS (1) the derivative of this realization of the
 * estimator with every random decision and sampled direction held fixed,
 * which central differences with common random numbers reproduce where no
 * decision flips (for roughness: of renders replaying the sampled
 * directions); biased as an estimate of the derivative of the expectation. */
#define MFH_ADJOINT_EXPECTED 0
#define MFH_ADJOINT_FIXED_DECISIONS 1
int mfh_pt_set_adjoint_mode(mfh_context* ctx, int32_t ptId, int32_t mode);
/* Gradient of an albedo or roughness texture: width * height * channels floats. */
int mfh_pt_get_texture_gradient(mfh_context* ctx, int32_t ptId, int32_t textureId, float* out);
/* Gradient with respect to a material's emission scale: 3 floats. */
int mfh_pt_get_emission_gradient(mfh_context* ctx, int32_t ptId, int32_t materialIndex, float* out);
/* Replace one material of a path tracer (world_get_material + material
 * edits); the light tree is rebuilt and the path tracer reset. */
int mfh_pt_set_material(mfh_context* ctx, int32_t ptId, int32_t materialIndex, const mfh_material* material);

/* Event timing on the context's stream: begin records an event, end records
 * a second one, waits for it and returns the elapsed milliseconds. */
int mfh_timer_begin(mfh_context* ctx);
int mfh_timer_end(mfh_context* ctx, float* outMs);

#ifdef __cplusplus
}
#endif

#endif /* MUFFLON_HIP_H */
