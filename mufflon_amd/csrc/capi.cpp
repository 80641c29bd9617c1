// extern "C" boundary of the HIP library (include/mufflon_hip.h).
// Every entry point converts exceptions into a status code plus a
// thread-local message, the convention of Mufflon's core API
// (interface.cpp:54-83, core_get_dll_error).
#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>

#include "capi_util.hpp"

using namespace mfh;

std::string& mfh::last_error() {
    thread_local std::string msg;
    return msg;
}

namespace {

void vertex_box(const float* v, int32_t n, float* box) {
    for (int k = 0; k < 3; ++k) {
        float lo = v[k], hi = v[k];
        for (int32_t i = 1; i < n; ++i) {
            const float x = v[3 * (size_t)i + k];
            lo = (x < lo) ? x : lo;
            hi = (hi < x) ? x : hi;
        }
        box[k] = lo;
        box[3 + k] = hi;
    }
}

}  // namespace

void mfh::upload(Context& c, DevBuf& b, const void* src, size_t bytes) {
    b.reserve(bytes);
    MFH_HIP(hipMemcpyAsync(b.ptr, src, bytes, hipMemcpyHostToDevice, c.stream));
}

namespace mfh {

void refresh_scene(Context& c, Scene& s) {
    bool stale = s.lodGeneration.size() != c.lods.size();
    for (size_t i = 0; !stale && i < c.lods.size(); ++i)
        stale = c.lods[i] && c.lods[i]->generation != s.lodGeneration[i];
    if (!stale) return;
    std::vector<DevLod> lods(c.lods.size()), ref(c.lods.size());
    s.lodGeneration.assign(c.lods.size(), 0);
    for (size_t i = 0; i < c.lods.size(); ++i) {
        std::memset(&lods[i], 0, sizeof(DevLod));
        const Lod* l = c.lods[i];
        if (!l || !l->built) {
            ref[i] = lods[i];
            continue;
        }
        s.lodGeneration[i] = l->generation;
        lods[i].triVerts = l->triVerts.as<float4>();
        lods[i].nodes = l->nodes.ptr;
        lods[i].numInternal = l->numInternal;
        lods[i].numPrims = l->numPrims();
        lods[i].quadV3 = l->numQuads > 0 ? l->quadV3.as<float4>() : nullptr;
        std::memcpy(lods[i].box, l->box, sizeof(l->box));
        lods[i].coordMax = l->coordMax;
        ref[i] = lods[i];
        if (l->travNumInternal >= 0) {
            lods[i].nodes = l->travNodes.ptr;
            lods[i].numInternal = l->travNumInternal;
        }
    }
    s.mixed = false;
    for (int32_t li : s.instLod) {
        MFH_REQUIRE(c.lods[li]->built, "instance refers to a LoD whose BVH is not built");
        if (c.lods[li]->mixed()) s.mixed = true;
    }
    upload(c, s.dLods, lods.data(), sizeof(DevLod) * lods.size());
    upload(c, s.dLodsRef, ref.data(), sizeof(DevLod) * ref.size());
    build_scene_bvh(c, s, nullptr);
    MFH_HIP(hipStreamSynchronize(c.stream));
}

}  // namespace mfh

extern "C" {

const char* mfh_last_error(void) { return last_error().c_str(); }

int mfh_create(int32_t deviceIndex, mfh_context** outCtx) {
    return guarded([&] {
        MFH_REQUIRE(outCtx != nullptr, "null output pointer");
        *outCtx = nullptr;
        int count = 0;
        MFH_HIP(hipGetDeviceCount(&count));
        MFH_REQUIRE(deviceIndex >= 0 && deviceIndex < count, "device index out of range");
        auto* ctx = new mfh_context();
        ctx->c.device = deviceIndex;
        try {
            MFH_HIP(hipSetDevice(deviceIndex));
            MFH_HIP(hipStreamCreateWithFlags(&ctx->c.stream, hipStreamNonBlocking));
            MFH_HIP(hipEventCreate(&ctx->c.evBegin));
            MFH_HIP(hipEventCreate(&ctx->c.evEnd));
            noalloc_scratch_init(deviceIndex);
        } catch (...) {
            delete ctx;
            throw;
        }
        *outCtx = ctx;
    });
}

int mfh_destroy(mfh_context* ctx) {
    return guarded([&] {
        if (!ctx) return;
        Context& c = ctx->c;
        (void)hipSetDevice(c.device);
        if (c.stream) (void)hipStreamSynchronize(c.stream);
        for (Lod* l : c.lods) {
            if (!l) continue;
            l->verts.release(); l->tris.release(); l->nodes.release(); l->primIds.release(); l->triVerts.release();
            l->quads.release(); l->spheres.release(); l->quadV3.release();
            l->travNodes.release();
            l->normals.release(); l->uvs.release(); l->matIdx.release();
            delete l;
        }
        for (Scene* s : c.scenes) {
            if (!s) continue;
            s->dInstLod.release(); s->dWorldToInst.release(); s->dLods.release(); s->dLodsRef.release();
            s->tlasNodes.release(); s->tlasPrimIds.release();
            delete s;
        }
        destroy_tracers(c);
        for (Texture* t : c.textures) {
            if (!t) continue;
            t->dev.release();
            delete t;
        }
        BuildScratch& b = c.scratch;
        for (DevBuf* d : {&b.keysA, &b.keysB, &b.idsA, &b.idsB, &b.parents, &b.children, &b.levels, &b.straddle,
                          &b.boxes, &b.counters, &b.flags, &b.offsets, &b.offsets2, &b.hist, &b.partials,
                          &b.misc, &b.stage})
            d->release();
        if (c.evBegin) (void)hipEventDestroy(c.evBegin);
        if (c.evEnd) (void)hipEventDestroy(c.evEnd);
        if (c.stream) (void)hipStreamDestroy(c.stream);
        delete ctx;
    });
}

int mfh_synchronize(mfh_context* ctx) {
    return guarded([&] { MFH_HIP(hipStreamSynchronize(ctx_of(ctx).stream)); });
}

int mfh_alloc(mfh_context* ctx, size_t bytes, void** outDevPtr) {
    return guarded([&] {
        ctx_of(ctx);
        MFH_REQUIRE(outDevPtr != nullptr, "null output pointer");
        *outDevPtr = nullptr;
        if (hipMalloc(outDevPtr, bytes ? bytes : 1) != hipSuccess) {
            (void)hipGetLastError();
            throw Error(kOutOfMemory, "hipMalloc of " + std::to_string(bytes) + " bytes failed");
        }
    });
}

int mfh_free(mfh_context* ctx, void* devPtr) {
    return guarded([&] {
        Context& c = ctx_of(ctx);
        MFH_HIP(hipStreamSynchronize(c.stream));
        if (devPtr) MFH_HIP(hipFree(devPtr));
    });
}

int mfh_upload(mfh_context* ctx, void* dDst, const void* src, size_t bytes) {
    return guarded([&] {
        Context& c = ctx_of(ctx);
        MFH_REQUIRE(dDst && src, "null pointer");
        MFH_HIP(hipMemcpyAsync(dDst, src, bytes, hipMemcpyHostToDevice, c.stream));
        MFH_HIP(hipStreamSynchronize(c.stream));
    });
}

int mfh_download(mfh_context* ctx, void* dst, const void* dSrc, size_t bytes) {
    return guarded([&] {
        Context& c = ctx_of(ctx);
        MFH_REQUIRE(dst && dSrc, "null pointer");
        MFH_HIP(hipMemcpyAsync(dst, dSrc, bytes, hipMemcpyDeviceToHost, c.stream));
        MFH_HIP(hipStreamSynchronize(c.stream));
    });
}

namespace mfh {
namespace {

// A LoD of triangles, quads and spheres (any of them may be absent, not all).
// Box: the vertices' box joined with every sphere's center -+ radius
// (Lod::get_bounding_box: polygons and spheres).
int32_t create_lod(Context& c, const float* vertices, int32_t numVertices, const uint32_t* triIndices,
                   int32_t numTriangles, const uint32_t* quadIndices, int32_t numQuads, const float* spheres,
                   int32_t numSpheres) {
    MFH_REQUIRE(numVertices >= 0 && numTriangles >= 0 && numQuads >= 0 && numSpheres >= 0, "negative count");
    MFH_REQUIRE((numTriangles == 0 || triIndices) && (numQuads == 0 || quadIndices) && (numSpheres == 0 || spheres) &&
                    (numVertices == 0 || vertices),
                "null pointer");
    const long long prims = (long long)numTriangles + numQuads + numSpheres;
    MFH_REQUIRE(prims > 0, "empty LoD");
    MFH_REQUIRE(prims * 2 - 1 < (1ll << 31), "too many primitives");
    MFH_REQUIRE(numTriangles + numQuads == 0 || numVertices > 0, "polygons without vertices");
    for (int64_t i = 0; i < 3 * (int64_t)numTriangles; ++i)
        MFH_REQUIRE(triIndices[i] < (uint32_t)numVertices, "triangle index out of range");
    for (int64_t i = 0; i < 4 * (int64_t)numQuads; ++i)
        MFH_REQUIRE(quadIndices[i] < (uint32_t)numVertices, "quad index out of range");
    for (int32_t i = 0; i < numSpheres; ++i)
        MFH_REQUIRE(spheres[4 * (size_t)i + 3] >= 0.0f, "negative sphere radius");
    auto* l = new Lod();
    try {
        l->numVerts = numVertices;
        l->numTris = numTriangles;
        l->numQuads = numQuads;
        l->numSpheres = numSpheres;
        if (numVertices > 0) {
            l->hostVerts.assign(vertices, vertices + 3 * (size_t)numVertices);
            upload(c, l->verts, vertices, 12 * (size_t)numVertices);
        }
        if (numTriangles > 0) {
            l->hostTris.assign(triIndices, triIndices + 3 * (size_t)numTriangles);
            upload(c, l->tris, triIndices, 12 * (size_t)numTriangles);
        }
        if (numQuads > 0) {
            l->hostQuads.assign(quadIndices, quadIndices + 4 * (size_t)numQuads);
            upload(c, l->quads, quadIndices, 16 * (size_t)numQuads);
        }
        if (numSpheres > 0) {
            l->hostSpheres.assign(spheres, spheres + 4 * (size_t)numSpheres);
            upload(c, l->spheres, spheres, 16 * (size_t)numSpheres);
        }
        bool first = true;
        if (numVertices > 0) {
            vertex_box(vertices, numVertices, l->box);
            first = false;
        }
        for (int32_t i = 0; i < numSpheres; ++i) {
            const float* sp = spheres + 4 * (size_t)i;
            for (int k = 0; k < 3; ++k) {
                const float lo = sp[k] - sp[3], hi = sp[k] + sp[3];
                l->box[k] = (first || lo < l->box[k]) ? lo : l->box[k];
                l->box[3 + k] = (first || l->box[3 + k] < hi) ? hi : l->box[3 + k];
            }
            first = false;
        }
        for (int k = 0; k < 6; ++k) l->coordMax = std::max(l->coordMax, std::fabs(l->box[k]));
        MFH_HIP(hipStreamSynchronize(c.stream));
    } catch (...) {
        l->verts.release();
        l->tris.release();
        l->quads.release();
        l->spheres.release();
        delete l;
        throw;
    }
    c.lods.push_back(l);
    return (int32_t)c.lods.size() - 1;
}

}  // namespace
}  // namespace mfh

int mfh_lod_create(mfh_context* ctx, const float* vertices, int32_t numVertices, const uint32_t* triIndices,
                   int32_t numTriangles, int32_t* outLodId) {
    return guarded([&] {
        Context& c = ctx_of(ctx);
        MFH_REQUIRE(vertices && triIndices && outLodId, "null pointer");
        MFH_REQUIRE(numVertices > 0 && numTriangles > 0, "empty LoD");
        *outLodId = create_lod(c, vertices, numVertices, triIndices, numTriangles, nullptr, 0, nullptr, 0);
    });
}

int mfh_lod_create_mixed(mfh_context* ctx, const float* vertices, int32_t numVertices, const uint32_t* triIndices,
                         int32_t numTriangles, const uint32_t* quadIndices, int32_t numQuads, const float* spheres,
                         int32_t numSpheres, int32_t* outLodId) {
    return guarded([&] {
        Context& c = ctx_of(ctx);
        MFH_REQUIRE(outLodId, "null pointer");
        *outLodId = create_lod(c, vertices, numVertices, triIndices, numTriangles, quadIndices, numQuads, spheres,
                               numSpheres);
    });
}

int mfh_lbvh_build(mfh_context* ctx, int32_t lodId, const float* bbox) {
    return guarded([&] {
        Context& c = ctx_of(ctx);
        Lod& l = lod_of(c, lodId);
        build_lod_bvh(c, l, bbox, nullptr);
        ++l.generation;
        c.lastBuilt = lodId;
    });
}

int mfh_set_traversal_leaf_limit(mfh_context* ctx, int32_t limit) {
    return guarded([&] {
        Context& c = ctx_of(ctx);
        MFH_REQUIRE(limit >= 0 && limit < 1023, "traversal leaf limit out of range (0 .. 1022)");
        c.traversalLeafLimit = limit;
    });
}

int mfh_lbvh_build_stages(mfh_context* ctx, int32_t lodId, const float* bbox, uint64_t* unsortedKeys,
                          uint64_t* sortedKeys, int32_t* sortedIds, int32_t* parents, float* boxes,
                          int32_t* offsets) {
    return guarded([&] {
        Context& c = ctx_of(ctx);
        Lod& l = lod_of(c, lodId);
        StageSink sink{unsortedKeys, sortedKeys, sortedIds, parents, boxes, offsets};
        build_lod_bvh(c, l, bbox, &sink);
        ++l.generation;
        c.lastBuilt = lodId;
    });
}

int mfh_lbvh_get(mfh_context* ctx, int32_t lodId, int32_t* outNumInternal, void* outNodes, int32_t* outPrimIds) {
    return guarded([&] {
        Context& c = ctx_of(ctx);
        Lod& l = lod_of(c, lodId);
        MFH_REQUIRE(l.built, "LoD BVH not built");
        MFH_HIP(hipStreamSynchronize(c.stream));
        if (outNumInternal) *outNumInternal = l.numInternal;
        if (outNodes && l.numInternal > 0)
            MFH_HIP(hipMemcpy(outNodes, l.nodes.ptr, 64 * (size_t)l.numInternal, hipMemcpyDeviceToHost));
        if (outPrimIds)
            MFH_HIP(hipMemcpy(outPrimIds, l.primIds.ptr, 4 * (size_t)l.numPrims(), hipMemcpyDeviceToHost));
    });
}

int mfh_lbvh_build_noalloc(mfh_context* ctx, const float* dVertices, int32_t numVertices,
                           const uint32_t* dTriIndices, int32_t numTriangles, const float* bbox,
                           int32_t* outNumInternal, void* dNodes, int32_t* dPrimIds) {
    return guarded([&] {
        Context& c = ctx_of(ctx);
        MFH_REQUIRE(outNumInternal != nullptr, "null output pointer");
        *outNumInternal = build_lod_bvh_noalloc(c, dVertices, numVertices, dTriIndices, numTriangles, bbox, dNodes,
                                                dPrimIds);
    });
}

int mfh_scene_create(mfh_context* ctx, const int32_t* instLod, const float* worldToInstance, int32_t numInstances,
                     int32_t* outSceneId) {
    return guarded([&] {
        Context& c = ctx_of(ctx);
        MFH_REQUIRE(instLod && worldToInstance && outSceneId, "null pointer");
        MFH_REQUIRE(numInstances > 0, "scene without instances");
        for (int32_t i = 0; i < numInstances; ++i) {
            const int32_t li = instLod[i];
            MFH_REQUIRE(li >= 0 && li < (int32_t)c.lods.size() && c.lods[li], "instance refers to an unknown LoD");
            MFH_REQUIRE(c.lods[li]->built, "instance refers to a LoD whose BVH is not built");
        }
        auto* s = new Scene();
        try {
            s->numInstances = numInstances;
            s->instLod.assign(instLod, instLod + numInstances);
            s->worldToInst.assign(worldToInstance, worldToInstance + 12 * (size_t)numInstances);
            upload(c, s->dInstLod, instLod, 4 * (size_t)numInstances);
            upload(c, s->dWorldToInst, worldToInstance, 48 * (size_t)numInstances);
            refresh_scene(c, *s);
            MFH_HIP(hipStreamSynchronize(c.stream));
        } catch (...) {
            s->dInstLod.release(); s->dWorldToInst.release(); s->dLods.release(); s->dLodsRef.release();
            s->tlasNodes.release(); s->tlasPrimIds.release();
            delete s;
            throw;
        }
        c.scenes.push_back(s);
        *outSceneId = (int32_t)c.scenes.size() - 1;
    });
}

int mfh_scene_get(mfh_context* ctx, int32_t sceneId, float* outBox, int32_t* outNumInternal, void* outNodes,
                  int32_t* outPrimIds) {
    return guarded([&] {
        Context& c = ctx_of(ctx);
        Scene& s = scene_of(c, sceneId);
        MFH_HIP(hipStreamSynchronize(c.stream));
        if (outBox) std::memcpy(outBox, s.box, sizeof(s.box));
        if (outNumInternal) *outNumInternal = s.tlasNumInternal;
        if (outNodes && s.tlasNumInternal > 0)
            MFH_HIP(hipMemcpy(outNodes, s.tlasNodes.ptr, 64 * (size_t)s.tlasNumInternal, hipMemcpyDeviceToHost));
        if (outPrimIds && s.numInstances > 1)
            MFH_HIP(hipMemcpy(outPrimIds, s.tlasPrimIds.ptr, 4 * (size_t)s.numInstances, hipMemcpyDeviceToHost));
    });
}

int mfh_trace_closest(mfh_context* ctx, int32_t sceneId, const float* dRays, int32_t numRays, void* dHits,
                      float* dBary) {
    return guarded([&] {
        Context& c = ctx_of(ctx);
        Scene& s = scene_of(c, sceneId);
        MFH_REQUIRE(numRays >= 0, "negative ray count");
        if (numRays == 0) return;
        MFH_REQUIRE(dRays && dHits, "null pointer");
        refresh_scene(c, s);
        launch_trace_closest(trace_params(s), dRays, numRays, static_cast<float4*>(dHits),
                             reinterpret_cast<float2*>(dBary), nullptr, c.stream);
    });
}

int mfh_trace_closest_stats(mfh_context* ctx, int32_t sceneId, const float* dRays, int32_t numRays, void* dHits,
                            float* dBary, uint32_t* dStats) {
    return guarded([&] {
        Context& c = ctx_of(ctx);
        Scene& s = scene_of(c, sceneId);
        MFH_REQUIRE(numRays >= 0, "negative ray count");
        if (numRays == 0) return;
        MFH_REQUIRE(dRays && dHits && dStats, "null pointer");
        refresh_scene(c, s);
        launch_trace_closest(trace_params(s, true), dRays, numRays, static_cast<float4*>(dHits),
                             reinterpret_cast<float2*>(dBary), reinterpret_cast<uint2*>(dStats), c.stream);
    });
}

int mfh_trace_any(mfh_context* ctx, int32_t sceneId, const float* dRays, int32_t numRays, int32_t* dOccluded) {
    return guarded([&] {
        Context& c = ctx_of(ctx);
        Scene& s = scene_of(c, sceneId);
        MFH_REQUIRE(numRays >= 0, "negative ray count");
        if (numRays == 0) return;
        MFH_REQUIRE(dRays && dOccluded, "null pointer");
        refresh_scene(c, s);
        launch_trace_any(trace_params(s), dRays, numRays, dOccluded, c.stream);
    });
}

int mfh_timer_begin(mfh_context* ctx) {
    return guarded([&] {
        Context& c = ctx_of(ctx);
        MFH_HIP(hipEventRecord(c.evBegin, c.stream));
    });
}

int mfh_timer_end(mfh_context* ctx, float* outMs) {
    return guarded([&] {
        Context& c = ctx_of(ctx);
        MFH_REQUIRE(outMs != nullptr, "null output pointer");
        MFH_HIP(hipEventRecord(c.evEnd, c.stream));
        MFH_HIP(hipEventSynchronize(c.evEnd));
        MFH_HIP(hipEventElapsedTime(outMs, c.evBegin, c.evEnd));
    });
}

}  // extern "C"
