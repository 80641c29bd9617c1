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

/* Build the LBVH of a LoD on the device. bbox = {min xyz, max xyz} of the
 * LoD (Lod::get_bounding_box); NULL computes it from the vertices.
 * Replaces LBVHBuilder::build(LodDescriptor&, const ei::Box&, bool)
 * (build_lbvh.cu:782-789, lod.cpp:33). */
int mfh_lbvh_build(mfh_context* ctx, int32_t lodId, const float* bbox);

/* Result of the last build of a LoD: numInternalNodes of the collapsed BVH
 * (LBVH::numInternalNodes, lbvh.hpp:23) and, when the pointers are non-NULL,
 * the nodes (numInternalNodes * 64 bytes, the BvhNode layout of lbvh.hpp:13-18)
 * and the primitive ids in BVH order (numTriangles int32).
 * Replaces LBVHBuilder::acquire_const (lbvh.hpp:70-86). */
int mfh_lbvh_get(mfh_context* ctx, int32_t lodId, int32_t* outNumInternal, void* outNodes, int32_t* outPrimIds);

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

/* Event timing on the context's stream: begin records an event, end records
 * a second one, waits for it and returns the elapsed milliseconds. */
int mfh_timer_begin(mfh_context* ctx);
int mfh_timer_end(mfh_context* ctx, float* outMs);

#ifdef __cplusplus
}
#endif

#endif /* MUFFLON_HIP_H */
