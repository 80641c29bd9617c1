// Host side of the path tracer: scene-descriptor assembly (area lights from
// emissive triangles, light tree, materials, camera), per-pixel state, and
// the extern "C" entry points of include/mufflon_hip.h for the pt renderer.
#include <algorithm>
#include <cstring>

#include "capi_util.hpp"
#include "light_tree.hpp"
#include "pt_device.hpp"
#include "pt_adjoint.hpp"
#include "pt_wavefront.hpp"

namespace mfh {

struct PathTracer {
    int32_t scene = -1;
    std::vector<PtMaterialRecord> materials;
    mfh_pinhole camera{};
    mfh_pt_params params{};
    LightTreeHost tree;
    BackgroundHost background;
    std::vector<HostPointLight> pointLights;
    DevBuf dMaterials, dPtLods, dLightMem, dMapKeys, dMapCodes, dTextures, dSat;
    DevBuf dRngs, dNan;
    DevBuf dTargets[kNumTargets];
    DevBuf dVariance[kNumTargets];
    uint32_t targetMask = (1u << MFH_TARGET_RADIANCE) | (1u << MFH_TARGET_HIT_ID);
    uint32_t varianceMask = 0;  // subset of targetMask
    int32_t iteration = -1;
    uint64_t sceneGeneration = ~0ull;
    // wavefront form
    int32_t mode = MFH_PT_MEGAKERNEL;
    bool useGraph = false;
    WfTuning tune;
    DevBuf wfCounters, wfTp, wfAcc, wfRay[2][4], wfRayP[2], wfSh[3], wfShPath, wfSortKeys[2], wfSortIds[2];
    // digit counts and scan partials of this tracer's queue sort: its own,
    // never the context's build scratch, which an LBVH build may reallocate
    // under a captured graph
    BuildScratch wfSortScratch;
    int wfPaths = 0, wfMaxLen = -1, wfNee = -1;
    hipGraph_t graph = nullptr;
    hipGraphExec_t graphExec = nullptr;
    std::vector<uint8_t> graphKey;  // bytes of (S, W, tune, sort scratch) the graph was captured with
    // adjoint
    DevBuf adjRngs, adjDelta, adjEmGrad, adjTexGradPtrs;
    int32_t adjMode = MFH_ADJOINT_EXPECTED;
    std::vector<DevBuf> adjTexGrad;  // per texture id
    bool adjValid = false;
    // wireframe renderer over the same scene and camera (GpuWireframe,
    // gpu_wireframe.cpp:18-50): Border running mean (and variance), the
    // per-pixel seeds of an iteration and the renderer's generator, a
    // default-constructed PCG64 that no reset re-seeds
    DevBuf wireBorder, wireVar, wireSeeds;
    std::vector<uint32_t> wireSeedHost;
    uint64_t wireRng = 0x2f0ae9bc6431af73ull;
    int32_t wireIteration = -1;
    bool wireVariance = false;
    // light tracer over the same scene (GpuLightTracer, gpu_lt.cpp:25-37): its
    // generators, the splat image of an iteration and the running means (and
    // variances) of Radiance, Lightness, Density, 5 floats per pixel
    DevBuf ltRngs, ltSplat, ltMean, ltVar;
    int32_t ltIteration = -1;
    bool ltVariance = false;
};

void drop_graph(PathTracer& t) {
    if (t.graphExec) (void)hipGraphExecDestroy(t.graphExec);
    if (t.graph) (void)hipGraphDestroy(t.graph);
    t.graphExec = nullptr;
    t.graph = nullptr;
    t.graphKey.clear();
}

void release_wavefront(PathTracer& t) {
    drop_graph(t);
    for (DevBuf* b : {&t.wfCounters, &t.wfTp, &t.wfAcc, &t.wfSh[0], &t.wfSh[1], &t.wfSh[2], &t.wfShPath,
                      &t.wfSortKeys[0], &t.wfSortKeys[1], &t.wfSortIds[0], &t.wfSortIds[1],
                      &t.wfSortScratch.hist, &t.wfSortScratch.partials})
        b->release();
    for (auto& r : t.wfRay)
        for (DevBuf& b : r) b.release();
    for (DevBuf& b : t.wfRayP) b.release();
    t.wfPaths = 0;
    t.wfMaxLen = t.wfNee = -1;
}

// Trace-stage forms the autotuners try: chunked queue walks with two chunk
// sizes, and lane refill every 16 or 32 traversal steps.
constexpr int kTraceModes = 4;
constexpr int kTraceModeChunk[kTraceModes] = {256, 1024, 256, 256};
constexpr int kTraceModeRefill[kTraceModes] = {0, 0, 16, 32};

// A batch of path tracers iterated together by the wavefront stages.
struct PtBatch {
    std::vector<int32_t> tracers;
    std::vector<int32_t> offsets;  // first global path of every tracer, then the total
    bool useGraph = false;
    WfTuning tune;
    int32_t iteration = -1;
    DevBuf dScenes, dOffsets;
    DevBuf counters, tp, acc, ray[2][4], rayP[2], sh[3], shPath, sortKeys[2], sortIds[2];
    int paths = 0, maxLen = 0, nee = 0;
    hipGraph_t graph = nullptr;
    hipGraphExec_t graphExec = nullptr;
    std::vector<uint8_t> graphKey;
};

void drop_batch_graph(PtBatch& b) {
    if (b.graphExec) (void)hipGraphExecDestroy(b.graphExec);
    if (b.graph) (void)hipGraphDestroy(b.graph);
    b.graphExec = nullptr;
    b.graph = nullptr;
    b.graphKey.clear();
}

void destroy_batch(PtBatch* b) {
    drop_batch_graph(*b);
    for (DevBuf* d : {&b->dScenes, &b->dOffsets, &b->counters, &b->tp, &b->acc, &b->sh[0], &b->sh[1], &b->sh[2],
                      &b->shPath, &b->sortKeys[0], &b->sortKeys[1], &b->sortIds[0], &b->sortIds[1], &b->rayP[0],
                      &b->rayP[1]})
        d->release();
    for (auto& r : b->ray)
        for (DevBuf& d : r) d.release();
    delete b;
}

void destroy_tracers(Context& c) {
    for (PtBatch* b : c.batches)
        if (b) destroy_batch(b);
    c.batches.clear();
    for (PathTracer* t : c.tracers) {
        if (!t) continue;
        for (DevBuf* b : {&t->dMaterials, &t->dPtLods, &t->dLightMem, &t->dMapKeys, &t->dMapCodes, &t->dTextures,
                          &t->dSat, &t->dRngs, &t->dNan})
            b->release();
        for (DevBuf& b : t->dTargets) b.release();
        for (DevBuf& b : t->dVariance) b.release();
        release_wavefront(*t);
        for (DevBuf* b : {&t->adjRngs, &t->adjDelta, &t->adjEmGrad, &t->adjTexGradPtrs}) b->release();
        for (DevBuf& b : t->adjTexGrad) b.release();
        delete t;
    }
    c.tracers.clear();
}

// descriptors.hpp:108-129 (host copy of the device inversion)
void instance_to_world(const float* w, float* o) {
    const float r00 = w[0], r01 = w[1], r02 = w[2];
    const float r10 = w[4], r11 = w[5], r12 = w[6];
    const float r20 = w[8], r21 = w[9], r22 = w[10];
    const float m00 = r11 * r22 - r21 * r12;
    const float m01 = r21 * r02 - r01 * r22;
    const float m02 = r01 * r12 - r11 * r02;
    const float m10 = r20 * r12 - r10 * r22;
    const float m11 = r00 * r22 - r20 * r02;
    const float m12 = r10 * r02 - r00 * r12;
    const float m20 = r10 * r21 - r20 * r11;
    const float m21 = r20 * r01 - r00 * r21;
    const float m22 = r00 * r11 - r10 * r01;
    const float det = r00 * m00 + r01 * m10 + r02 * m20;
    const float s = 1.0f / det;
    const float a[9] = {m00, m01, m02, m10, m11, m12, m20, m21, m22};
    for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c) o[4 * r + c] = s * a[3 * r + c];
    const float t[3] = {w[3], w[7], w[11]};
    for (int r = 0; r < 3; ++r) o[4 * r + 3] = -(o[4 * r] * t[0] + o[4 * r + 1] * t[1] + o[4 * r + 2] * t[2]);
}

namespace {

struct MaterialEmission : EmissionLookup {
    const std::vector<PtMaterialRecord>& mats;
    explicit MaterialEmission(const std::vector<PtMaterialRecord>& m) : mats(m) {}
    void radiance(uint16_t material, float, float, float out[3]) const override {
        const PtMaterialRecord& m = mats[material];
        for (int k = 0; k < 3; ++k) out[k] = m.type == kMatEmissive ? m.radiance[k] * m.scale[k] : 0.0f;
    }
};

void validate_material(const Context& c, const mfh_material& m) {
    const int32_t t = m.type;
    MFH_REQUIRE(t == MFH_MATERIAL_EMISSIVE || t == MFH_MATERIAL_LAMBERT || t == MFH_MATERIAL_TORRANCE ||
                    t == MFH_MATERIAL_WALTER || t == MFH_MATERIAL_MICROFACET,
                "unsupported material type " + std::to_string(t));
    auto tex_ok = [&](int32_t id) {
        return id == -1 || (id >= 0 && id < (int32_t)c.textures.size() && c.textures[id] != nullptr);
    };
    MFH_REQUIRE(tex_ok(m.albedoTexture) && tex_ok(m.roughnessTexture), "unknown texture id in a material");
    MFH_REQUIRE(m.emissionTexture == -1, "emission textures are not supported");
    if (mat_has_roughness(t)) {
        MFH_REQUIRE(m.ndf == MFH_NDF_GGX || m.ndf == MFH_NDF_BECKMANN, "unknown NDF");
        MFH_REQUIRE(m.shadowing == MFH_SHADOWING_VCAVITY || m.shadowing == MFH_SHADOWING_SMITH,
                    "unknown shadowing model");
        MFH_REQUIRE(!(m.ndf == MFH_NDF_BECKMANN && m.shadowing == MFH_SHADOWING_SMITH),
                    "Smith shadowing with the Beckmann NDF is not supported");
    }
    if (mat_has_inner_medium(t)) MFH_REQUIRE(m.refractionIndex > 0.0f, "refraction index must be positive");
    MFH_REQUIRE(m.outerRefractionIndex > 0.0f, "outer refraction index must be positive");
}

PathTracer& tracer_of(Context& c, int32_t id) {
    MFH_REQUIRE(id >= 0 && id < (int32_t)c.tracers.size() && c.tracers[id], "unknown path tracer id");
    return *c.tracers[id];
}

// Area lights of every emissive triangle, quad and sphere, per LoD and
// instance, in world space (world_container.cpp:895-960): quads keep their
// vertex order (v1 and v3 swap under a mirroring transform), a sphere's
// radius scales with the length of the transform's first column
// (Instance::extract_scale, instance.hpp:27-33).
std::vector<HostAreaLight> gather_area_lights(const Context& c, const Scene& sc,
                                              const std::vector<PtMaterialRecord>& mats) {
    std::vector<HostAreaLight> out;
    for (size_t lodId = 0; lodId < c.lods.size(); ++lodId) {
        const Lod* l = c.lods[lodId];
        if (!l || !l->hasShading) continue;
        bool emissive = false;
        for (uint16_t m : l->hostMatIdx) emissive |= mats[m].type == kMatEmissive;
        if (!emissive) continue;
        for (int32_t inst = 0; inst < sc.numInstances; ++inst) {
            if (sc.instLod[inst] != (int32_t)lodId) continue;
            float M[12];
            instance_to_world(sc.worldToInst.data() + 12 * (size_t)inst, M);
            const float det = M[0] * (M[5] * M[10] - M[9] * M[6]) + M[1] * (M[8] * M[6] - M[4] * M[10]) +
                              M[2] * (M[4] * M[9] - M[8] * M[5]);
            const bool mirroring = det < 0.0f;
            for (int32_t p = 0; p < l->numTris; ++p) {
                const uint16_t mi = l->hostMatIdx[p];
                if (mats[mi].type != kMatEmissive) continue;
                HostAreaTriangle t;
                t.material = mi;
                t.instance = inst;
                t.primitive = p;
                for (int k = 0; k < 3; ++k) {
                    const uint32_t vi = l->hostTris[3 * (size_t)p + k];
                    const float* v = l->hostVerts.data() + 3 * (size_t)vi;
                    for (int r = 0; r < 3; ++r)
                        t.points[k][r] = M[4 * r] * v[0] + M[4 * r + 1] * v[1] + M[4 * r + 2] * v[2] + M[4 * r + 3];
                    t.uv[k][0] = l->hostUvs[2 * (size_t)vi];
                    t.uv[k][1] = l->hostUvs[2 * (size_t)vi + 1];
                }
                if (mirroring) {
                    std::swap(t.points[1], t.points[2]);
                    std::swap(t.uv[1], t.uv[2]);
                }
                out.push_back(t);
            }
            for (int32_t q = 0; q < l->numQuads; ++q) {
                const int32_t p = l->numTris + q;
                const uint16_t mi = l->hostMatIdx[p];
                if (mats[mi].type != kMatEmissive) continue;
                HostAreaLight a;
                a.kind = kAreaQuad;
                a.material = mi;
                a.instance = inst;
                a.primitive = p;
                for (int k = 0; k < 4; ++k) {
                    const uint32_t vi = l->hostQuads[4 * (size_t)q + k];
                    const float* v = l->hostVerts.data() + 3 * (size_t)vi;
                    for (int r = 0; r < 3; ++r)
                        a.points[k][r] = M[4 * r] * v[0] + M[4 * r + 1] * v[1] + M[4 * r + 2] * v[2] + M[4 * r + 3];
                    a.uv[k][0] = l->hostUvs[2 * (size_t)vi];
                    a.uv[k][1] = l->hostUvs[2 * (size_t)vi + 1];
                }
                if (mirroring) {
                    std::swap(a.points[1], a.points[3]);
                    std::swap(a.uv[1], a.uv[3]);
                }
                out.push_back(a);
            }
            const float scale = std::sqrt(M[0] * M[0] + M[4] * M[4] + M[8] * M[8]);
            for (int32_t k = 0; k < l->numSpheres; ++k) {
                const int32_t p = l->numTris + l->numQuads + k;
                const uint16_t mi = l->hostMatIdx[p];
                if (mats[mi].type != kMatEmissive) continue;
                const float* sp = l->hostSpheres.data() + 4 * (size_t)k;
                HostAreaLight a;
                a.kind = kAreaSphere;
                a.material = mi;
                a.instance = inst;
                a.primitive = p;
                for (int r = 0; r < 3; ++r)
                    a.points[0][r] = M[4 * r] * sp[0] + M[4 * r + 1] * sp[1] + M[4 * r + 2] * sp[2] + M[4 * r + 3];
                a.radius = scale * sp[3];
                out.push_back(a);
            }
        }
    }
    return out;
}

PtScene device_scene(Context& c, PathTracer& t) {
    Scene& sc = scene_of(c, t.scene);
    refresh_scene(c, sc);
    // per-LoD shading table (pointers may move when a LoD is rebuilt)
    std::vector<DevPtLod> lods(c.lods.size());
    for (size_t i = 0; i < c.lods.size(); ++i) {
        std::memset(&lods[i], 0, sizeof(DevPtLod));
        const Lod* l = c.lods[i];
        if (!l) continue;
        lods[i].verts = l->verts.as<float>();
        lods[i].tris = l->tris.as<uint32_t>();
        lods[i].normals = l->normals.as<float>();
        lods[i].uvs = l->uvs.as<float>();
        lods[i].matIdx = l->matIdx.as<uint16_t>();
        lods[i].quads = l->quads.as<uint32_t>();
        lods[i].spheres = l->spheres.as<float4>();
        lods[i].numTris = l->numTris;
        lods[i].numQuads = l->numQuads;
    }
    upload(c, t.dPtLods, lods.data(), sizeof(DevPtLod) * lods.size());
    std::vector<DevTexture> texs(std::max<size_t>(c.textures.size(), 1));
    std::memset(texs.data(), 0, sizeof(DevTexture) * texs.size());
    for (size_t i = 0; i < c.textures.size(); ++i) {
        const Texture* x = c.textures[i];
        texs[i].texels = x->dev.as<float4>();
        texs[i].width = x->width;
        texs[i].height = x->height;
        texs[i].channels = x->channels;
        texs[i].sampling = x->sampling;
    }
    upload(c, t.dTextures, texs.data(), sizeof(DevTexture) * texs.size());
    PtScene S;
    std::memset(&S, 0, sizeof(S));
    S.trace = trace_params(sc);
    S.ptLods = t.dPtLods.as<DevPtLod>();
    S.materials = t.dMaterials.as<PtMaterialRecord>();
    S.textures = t.dTextures.as<DevTexture>();
    S.background.type = t.background.type;
    for (int k = 0; k < 3; ++k) S.background.color[k] = t.background.color[k];
    for (int k = 0; k < 3; ++k) S.background.scale[k] = t.background.scale[k];
    S.background.envmapTexture = t.background.envmapTexture;
    S.background.sat = t.dSat.as<float>();
    std::memcpy(S.sceneBox, sc.box, sizeof(S.sceneBox));
    S.lightMem = t.dLightMem.as<uint8_t>();
    S.posRootFlux = t.tree.posRootFlux;
    S.posRootType = t.tree.posRootType;
    S.dirRootFlux = t.tree.dirRootFlux;
    S.bgFlux = t.background.fluxSum;
    S.mapKeys = t.dMapKeys.as<uint64_t>();
    S.mapCodes = t.dMapCodes.as<uint32_t>();
    S.mapCount = (int32_t)t.tree.mapKeys.size();
    std::memcpy(S.camPos, t.camera.position, 12);
    S.tanVFov = t.camera.tanVFov;
    std::memcpy(S.camDir, t.camera.viewDir, 12);
    std::memcpy(S.camUp, t.camera.up, 12);
    S.width = t.camera.width;
    S.height = t.camera.height;
    S.seed = t.params.seed;
    S.minPathLength = t.params.minPathLength;
    S.maxPathLength = t.params.maxPathLength;
    S.neeCount = t.params.neeCount;
    S.posGuide = t.params.neePositionGuide ? 1 : 0;
    S.rngs = t.dRngs.as<uint64_t>();
    for (int k = 0; k < kNumTargets; ++k) {
        S.cum[k] = (t.targetMask >> k) & 1u ? t.dTargets[k].as<float>() : nullptr;
        S.var[k] = (t.varianceMask >> k) & 1u ? t.dVariance[k].as<float>() : nullptr;
    }
    S.targetMask = t.targetMask;
    S.varianceMask = t.varianceMask;
    S.nanCounter = t.dNan.as<uint32_t>();
    return S;
}

// PCG64::next (rng.hpp:83-90) on the host, for the wireframe seeds.
uint64_t pcg64_next(uint64_t& state) {
    uint64_t x = state;
    const uint64_t count = state >> 59;
    state = state * 6364136223846793005ull + 1442695040888963407ull;
    x ^= x >> (5 + count);
    x *= 12605985483714917081ull;
    return x ^ (x >> 43);
}

void reset(Context& c, PathTracer& t) {
    const int num = t.camera.width * t.camera.height;
    launch_pt_init_rng(t.dRngs.as<uint64_t>(), num, t.params.seed * (num + 1), c.stream);
    for (int k = 0; k < kNumTargets; ++k) {
        if (!((t.targetMask >> k) & 1u)) continue;
        t.dTargets[k].reserve(4 * (size_t)kTargetChannels[k] * num);
        MFH_HIP(hipMemsetAsync(t.dTargets[k].ptr, 0, 4 * (size_t)kTargetChannels[k] * num, c.stream));
        if (!((t.varianceMask >> k) & 1u)) continue;
        t.dVariance[k].reserve(4 * (size_t)kTargetChannels[k] * num);
        MFH_HIP(hipMemsetAsync(t.dVariance[k].ptr, 0, 4 * (size_t)kTargetChannels[k] * num, c.stream));
    }
    MFH_HIP(hipMemsetAsync(t.dNan.ptr, 0, 4, c.stream));
    t.iteration = -1;
}

int num_cus(const Context& c) {
    int n = 0;
    MFH_HIP(hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, c.device));
    return std::max(n, 1);
}

// Buffers of the wavefront form for the current resolution, path length and
// NEE count; the counter block holds queue lengths, shadow counts, work
// counters and the device iteration counter.
WfState wavefront_state(Context& c, PathTracer& t) {
    const int N = t.camera.width * t.camera.height;
    const int L = std::max(t.params.maxPathLength, 1);
    const int nee = std::max(t.params.neeCount, 1);
    MFH_REQUIRE(t.params.neeCount <= 32, "the wavefront form supports at most 32 NEE samples per vertex");
    const int numCounters = (L + 1) + L + 3 * L;
    if (N != t.wfPaths || L != t.wfMaxLen || nee != t.wfNee) {
        drop_graph(t);
        t.wfCounters.reserve(4 * (size_t)(numCounters + 1));
        t.wfTp.reserve(16 * (size_t)N);
        t.wfAcc.reserve(4 * 18 * (size_t)N);
        for (auto& r : t.wfRay)
            for (DevBuf& b : r) b.reserve(16 * (size_t)N);
        for (DevBuf& b : t.wfSh) b.reserve(16 * (size_t)N * nee);
        t.wfShPath.reserve(8 * (size_t)N);
        for (int b = 0; b < 2; ++b) {
            t.wfSortKeys[b].reserve(8 * (size_t)N);
            t.wfSortIds[b].reserve(4 * (size_t)N);
        }
        t.wfPaths = N;
        t.wfMaxLen = L;
        t.wfNee = nee;
    }
    WfState W;
    std::memset(&W, 0, sizeof(W));
    W.numPaths = N;
    uint32_t slots = 0;
    int slot = 0;
    for (int k = 0; k < kNumTargets; ++k) {
        if ((t.targetMask >> k) & 1u)
            for (int j = 0; j < kTargetChannels[k]; ++j) slots |= 1u << (slot + j);
        slot += kTargetChannels[k];
    }
    W.slotMask = slots;
    W.chunk = t.tune.chunk;
    W.numCounters = numCounters;
    W.qBase = 0;
    W.shBase = L + 1;
    W.workBase = 2 * L + 1;
    W.counters = t.wfCounters.as<int>();
    W.iteration = t.wfCounters.as<int>() + numCounters;
    W.tp = t.wfTp.as<float4>();
    W.acc = t.wfAcc.as<float>();
    for (int b = 0; b < 2; ++b) {
        W.rayO[b] = t.wfRay[b][0].as<float4>();
        W.rayD[b] = t.wfRay[b][1].as<float4>();
        W.rayM[b] = t.wfRay[b][2].as<float4>();
        W.hit[b] = t.wfRay[b][3].as<float4>();
        // the previous vertex of each path: the light-tree guide and sphere
        // lights' connect pdfs depend on it
        if (t.params.neePositionGuide || t.tree.sphereLights) {
            t.wfRayP[b].reserve(16 * (size_t)N);
            W.rayP[b] = t.wfRayP[b].as<float4>();
        }
    }
    W.shO = t.wfSh[0].as<float4>();
    W.shD = t.wfSh[1].as<float4>();
    W.shC = t.wfSh[2].as<float4>();
    W.shPath = t.wfShPath.as<int2>();
    W.neeStride = nee;
    W.scenes = nullptr;
    W.sceneOffset = nullptr;
    W.numScenes = 1;
    for (int b = 0; b < 2; ++b) {
        W.sortKeys[b] = t.wfSortKeys[b].as<uint64_t>();
        W.sortIds[b] = t.wfSortIds[b].as<int32_t>();
    }
    // radix-sort scratch of the tracer sized up front (no allocation inside a captured graph)
    t.wfSortScratch.hist.reserve(radix_sort_hist_bytes(N, 0, 32));
    t.wfSortScratch.partials.reserve(radix_sort_partials_bytes(N));
    if (t.tune.traceGrid <= 0 || t.tune.shadeGrid <= 0) {
        const int cus = num_cus(c);
        t.tune.traceGrid = 0;
        t.tune.shadeGrid = wf_shade_grid(cus);
    }
    if (t.tune.tailGrid <= 0) t.tune.tailGrid = wf_tail_grid(num_cus(c));
    return W;
}

// Runs `iterations` wavefront iterations, replaying a captured graph when enabled.
void run_wavefront(Context& c, PathTracer& t, PtScene& S, int32_t iterations) {
    WfState W = wavefront_state(c, t);
    if (t.tune.traceGrid <= 0) t.tune.traceGrid = wf_grid(S, t.tune, num_cus(c));
    // the device iteration counter continues from the host's
    const int32_t it = t.iteration;
    MFH_HIP(hipMemcpyAsync(W.iteration, &it, 4, hipMemcpyHostToDevice, c.stream));
    if (!t.useGraph) {
        for (int32_t k = 0; k < iterations; ++k) {
            S.iteration = ++t.iteration;
            launch_pt_wavefront(S, W, t.tune, t.wfSortScratch, c.stream);
        }
        MFH_HIP(hipStreamSynchronize(c.stream));  // the host copy of the iteration counter is read above
        return;
    }
    PtScene Sk = S;
    Sk.iteration = 0;  // the graph reads the device counter
    // every device address the captured kernels read is in the key, the
    // sort scratch included, so a reallocated buffer forces a new capture
    const void* sortScratch[2] = {t.wfSortScratch.hist.ptr, t.wfSortScratch.partials.ptr};
    std::vector<uint8_t> key(sizeof(Sk) + sizeof(W) + sizeof(WfTuning) + sizeof(sortScratch));
    uint8_t* kp = key.data();
    std::memcpy(kp, &Sk, sizeof(Sk));
    std::memcpy(kp += sizeof(Sk), &W, sizeof(W));
    std::memcpy(kp += sizeof(W), &t.tune, sizeof(WfTuning));
    std::memcpy(kp += sizeof(WfTuning), sortScratch, sizeof(sortScratch));
    if (!t.graphExec || key != t.graphKey) {
        drop_graph(t);
        MFH_HIP(hipStreamSynchronize(c.stream));
        MFH_HIP(hipStreamBeginCapture(c.stream, hipStreamCaptureModeThreadLocal));
        try {
            launch_pt_wavefront(Sk, W, t.tune, t.wfSortScratch, c.stream);
        } catch (...) {
            hipGraph_t g = nullptr;
            (void)hipStreamEndCapture(c.stream, &g);
            if (g) (void)hipGraphDestroy(g);
            throw;
        }
        MFH_HIP(hipStreamEndCapture(c.stream, &t.graph));
        MFH_HIP(hipGraphInstantiate(&t.graphExec, t.graph, nullptr, nullptr, 0));
        t.graphKey = key;
    }
    for (int32_t k = 0; k < iterations; ++k) {
        ++t.iteration;
        MFH_HIP(hipGraphLaunch(t.graphExec, c.stream));
    }
    MFH_HIP(hipStreamSynchronize(c.stream));
}

}  // namespace
}  // namespace mfh

using namespace mfh;

extern "C" {

int mfh_lod_set_shading(mfh_context* ctx, int32_t lodId, const float* normals, const float* uvs,
                        const uint16_t* materialIndices) {
    return guarded([&] {
        Context& c = ctx_of(ctx);
        Lod& l = lod_of(c, lodId);
        MFH_REQUIRE(materialIndices && ((normals && uvs) || l.numVerts == 0), "null pointer");
        if (l.numVerts > 0) {
            l.hostUvs.assign(uvs, uvs + 2 * (size_t)l.numVerts);
            upload(c, l.normals, normals, 12 * (size_t)l.numVerts);
            upload(c, l.uvs, uvs, 8 * (size_t)l.numVerts);
        }
        l.hostMatIdx.assign(materialIndices, materialIndices + (size_t)l.numPrims());
        upload(c, l.matIdx, materialIndices, 2 * (size_t)l.numPrims());
        MFH_HIP(hipStreamSynchronize(c.stream));
        l.hasShading = true;
    });
}

int mfh_pt_create(mfh_context* ctx, int32_t sceneId, const mfh_material* materials, int32_t numMaterials,
                  const mfh_point_light* pointLights, int32_t numPointLights, const mfh_pinhole* camera,
                  const mfh_pt_params* params, int32_t* outPtId) {
    return guarded([&] {
        Context& c = ctx_of(ctx);
        Scene& sc = scene_of(c, sceneId);
        MFH_REQUIRE(materials && numMaterials > 0 && camera && params && outPtId, "null pointer");
        MFH_REQUIRE(numPointLights == 0 || pointLights, "null point lights");
        MFH_REQUIRE(camera->width > 0 && camera->height > 0 && camera->width < 65536 && camera->height < 65536,
                    "camera resolution out of range");
        MFH_REQUIRE(params->maxPathLength >= 0 && params->neeCount >= 0, "invalid path tracer parameters");
        for (int32_t i = 0; i < numMaterials; ++i) validate_material(c, materials[i]);
        for (int32_t li : sc.instLod) {
            const Lod* l = c.lods[li];
            MFH_REQUIRE(l->hasShading, "a LoD of the scene has no shading attributes (mfh_lod_set_shading)");
            for (uint16_t m : l->hostMatIdx) MFH_REQUIRE(m < numMaterials, "material index out of range");
        }
        refresh_scene(c, sc);
        auto* t = new PathTracer();
        try {
            t->scene = sceneId;
            t->materials.resize(numMaterials);
            std::memcpy(t->materials.data(), materials, sizeof(mfh_material) * (size_t)numMaterials);
            t->camera = *camera;
            t->params = *params;
            std::vector<HostPointLight> pl(numPointLights);
            for (int32_t i = 0; i < numPointLights; ++i) {
                std::memcpy(pl[i].position, pointLights[i].position, 12);
                std::memcpy(pl[i].intensity, pointLights[i].intensity, 12);
            }
            const std::vector<HostAreaLight> area = gather_area_lights(c, sc, t->materials);
            t->pointLights = pl;
            build_light_tree(area, pl, sc.box, MaterialEmission(t->materials), t->tree);
            const float black[3] = {0.0f, 0.0f, 0.0f}, one[3] = {1.0f, 1.0f, 1.0f};
            build_background(MFH_BACKGROUND_COLORED, black, one, nullptr, 0, 0, sc.box, t->background);
            t->tree.backgroundFlux = t->background.fluxSum;
            t->dSat.reserve(16);
            upload(c, t->dMaterials, t->materials.data(), sizeof(PtMaterialRecord) * t->materials.size());
            upload(c, t->dLightMem, t->tree.mem.data(), std::max<size_t>(t->tree.mem.size(), 16));
            const size_t nm = t->tree.mapKeys.size();
            t->dMapKeys.reserve(8 * std::max<size_t>(nm, 1));
            t->dMapCodes.reserve(4 * std::max<size_t>(nm, 1));
            if (nm) {
                upload(c, t->dMapKeys, t->tree.mapKeys.data(), 8 * nm);
                upload(c, t->dMapCodes, t->tree.mapCodes.data(), 4 * nm);
            }
            const size_t num = (size_t)camera->width * camera->height;
            t->dRngs.reserve(8 * num);
            t->dNan.reserve(4);
            reset(c, *t);
            MFH_HIP(hipStreamSynchronize(c.stream));
        } catch (...) {
            for (DevBuf* b : {&t->dMaterials, &t->dPtLods, &t->dLightMem, &t->dMapKeys, &t->dMapCodes, &t->dTextures,
                              &t->dSat, &t->dRngs, &t->dNan})
                b->release();
            for (DevBuf& b : t->dTargets) b.release();
            for (DevBuf& b : t->dVariance) b.release();
            delete t;
            throw;
        }
        c.tracers.push_back(t);
        *outPtId = (int32_t)c.tracers.size() - 1;
    });
}

int mfh_pt_set_params(mfh_context* ctx, int32_t ptId, const mfh_pt_params* params) {
    return guarded([&] {
        Context& c = ctx_of(ctx);
        PathTracer& t = tracer_of(c, ptId);
        MFH_REQUIRE(params != nullptr, "null pointer");
        MFH_REQUIRE(params->maxPathLength >= 0 && params->neeCount >= 0, "invalid path tracer parameters");
        MFH_HIP(hipStreamSynchronize(c.stream));
        t.params = *params;
        drop_graph(t);
        reset(c, t);
    });
}

int mfh_pt_set_camera(mfh_context* ctx, int32_t ptId, const mfh_pinhole* camera) {
    return guarded([&] {
        Context& c = ctx_of(ctx);
        PathTracer& t = tracer_of(c, ptId);
        MFH_REQUIRE(camera != nullptr, "null pointer");
        MFH_REQUIRE(camera->width > 0 && camera->height > 0 && camera->width < 65536 && camera->height < 65536,
                    "camera resolution out of range");
        MFH_HIP(hipStreamSynchronize(c.stream));
        t.camera = *camera;
        t.dRngs.reserve(8 * (size_t)camera->width * camera->height);
        t.adjValid = false;
        drop_graph(t);
        reset(c, t);
    });
}

int mfh_pt_reset(mfh_context* ctx, int32_t ptId) {
    return guarded([&] {
        Context& c = ctx_of(ctx);
        reset(c, tracer_of(c, ptId));
    });
}

int mfh_pt_iterate(mfh_context* ctx, int32_t ptId, int32_t iterations) {
    return guarded([&] {
        Context& c = ctx_of(ctx);
        PathTracer& t = tracer_of(c, ptId);
        MFH_REQUIRE(iterations >= 0, "negative iteration count");
        MFH_REQUIRE(t.tree.posLightCount > 0 || t.background.fluxSum > 0.0f, "scene without lights");
        PtScene S = device_scene(c, t);
        if (t.mode == MFH_PT_WAVEFRONT) {
            run_wavefront(c, t, S, iterations);
            return;
        }
        for (int32_t k = 0; k < iterations; ++k) {
            S.iteration = ++t.iteration;
            launch_pt_iterate(S, kWfTraceShapes[t.tune.traceVariant].leafBatch, c.stream);
        }
    });
}

int mfh_pt_wireframe_reset(mfh_context* ctx, int32_t ptId, int32_t recordVariance) {
    return guarded([&] {
        Context& c = ctx_of(ctx);
        PathTracer& t = tracer_of(c, ptId);
        t.wireIteration = -1;
        t.wireVariance = recordVariance != 0;
        const size_t num = (size_t)t.camera.width * t.camera.height;
        t.wireBorder.reserve(12 * num);
        MFH_HIP(hipMemsetAsync(t.wireBorder.ptr, 0, 12 * num, c.stream));
        if (t.wireVariance) {
            t.wireVar.reserve(12 * num);
            MFH_HIP(hipMemsetAsync(t.wireVar.ptr, 0, 12 * num, c.stream));
        } else {
            t.wireVar.release();
        }
    });
}

int mfh_pt_wireframe_iterate(mfh_context* ctx, int32_t ptId, int32_t iterations, float lineWidth,
                             int32_t maxTraceDepth) {
    return guarded([&] {
        Context& c = ctx_of(ctx);
        PathTracer& t = tracer_of(c, ptId);
        MFH_REQUIRE(iterations >= 0, "negative iteration count");
        MFH_REQUIRE(maxTraceDepth >= 0, "negative maximum trace depth");
        const size_t num = (size_t)t.camera.width * t.camera.height;
        if (t.wireBorder.bytes < 12 * num || (t.wireVariance && t.wireVar.bytes < 12 * num)) {
            // first use, or the resolution grew: a fresh Border target
            t.wireIteration = -1;
            t.wireBorder.reserve(12 * num);
            MFH_HIP(hipMemsetAsync(t.wireBorder.ptr, 0, 12 * num, c.stream));
            if (t.wireVariance) {
                t.wireVar.reserve(12 * num);
                MFH_HIP(hipMemsetAsync(t.wireVar.ptr, 0, 12 * num, c.stream));
            }
        }
        PtScene S = device_scene(c, t);
        for (int k = 0; k < kNumTargets; ++k) S.cum[k] = S.var[k] = nullptr;
        S.cum[MFH_TARGET_RADIANCE] = t.wireBorder.as<float>();
        S.var[MFH_TARGET_RADIANCE] = t.wireVariance ? t.wireVar.as<float>() : nullptr;
        S.targetMask = 1u << MFH_TARGET_RADIANCE;
        S.varianceMask = t.wireVariance ? S.targetMask : 0u;
        t.wireSeedHost.resize(num);
        t.wireSeeds.reserve(4 * num);
        for (int32_t it = 0; it < iterations; ++it) {
            // GpuWireframe::iterate (gpu_wireframe.cpp:33-48): one seed per pixel
            for (size_t i = 0; i < num; ++i) t.wireSeedHost[i] = (uint32_t)pcg64_next(t.wireRng);
            // the copy is ordered on the stream before the kernel that reads it;
            // the host vector is reused only after the stream has drained
            MFH_HIP(hipMemcpyAsync(t.wireSeeds.ptr, t.wireSeedHost.data(), 4 * num, hipMemcpyHostToDevice, c.stream));
            S.iteration = ++t.wireIteration;
            launch_wireframe(S, t.wireSeeds.as<uint32_t>(), lineWidth, maxTraceDepth, c.stream);
            MFH_HIP(hipStreamSynchronize(c.stream));
        }
    });
}

int mfh_pt_wireframe_get(mfh_context* ctx, int32_t ptId, int32_t variance, float* out) {
    return guarded([&] {
        Context& c = ctx_of(ctx);
        PathTracer& t = tracer_of(c, ptId);
        MFH_REQUIRE(out != nullptr, "null output pointer");
        MFH_REQUIRE(!variance || t.wireVariance, "the Border variance is not recorded");
        const size_t num = (size_t)t.camera.width * t.camera.height;
        MFH_REQUIRE(t.wireBorder.bytes >= 12 * num, "the wireframe renderer has not run");
        MFH_HIP(hipStreamSynchronize(c.stream));
        MFH_HIP(hipMemcpy(out, (variance ? t.wireVar : t.wireBorder).ptr, 12 * num, hipMemcpyDeviceToHost));
    });
}

namespace {
constexpr int kLtChannels = 5;
void lighttracer_reset(Context& c, PathTracer& t, bool recordVariance) {
    const size_t num = (size_t)t.camera.width * t.camera.height;
    t.ltIteration = -1;
    t.ltVariance = recordVariance;
    t.ltMean.reserve(4 * kLtChannels * num);
    MFH_HIP(hipMemsetAsync(t.ltMean.ptr, 0, 4 * kLtChannels * num, c.stream));
    if (recordVariance) {
        t.ltVar.reserve(4 * kLtChannels * num);
        MFH_HIP(hipMemsetAsync(t.ltVar.ptr, 0, 4 * kLtChannels * num, c.stream));
    } else {
        t.ltVar.release();
    }
    t.ltSplat.reserve(4 * kLtChannels * num);
    // GpuLightTracer::post_reset (gpu_lt.cpp:31-36): generator i seeded with
    // i + seed * (pixels + 1), as the path tracer's
    t.ltRngs.reserve(8 * num);
    launch_pt_init_rng(t.ltRngs.as<uint64_t>(), (int)num, t.params.seed * ((int)num + 1), c.stream);
}
}  // namespace

int mfh_pt_lighttracer_reset(mfh_context* ctx, int32_t ptId, int32_t recordVariance) {
    return guarded([&] {
        Context& c = ctx_of(ctx);
        lighttracer_reset(c, tracer_of(c, ptId), recordVariance != 0);
    });
}

int mfh_pt_lighttracer_iterate(mfh_context* ctx, int32_t ptId, int32_t iterations, int32_t density) {
    return guarded([&] {
        Context& c = ctx_of(ctx);
        PathTracer& t = tracer_of(c, ptId);
        MFH_REQUIRE(iterations >= 0, "negative iteration count");
        MFH_REQUIRE(t.tree.posLightCount > 0, "the light tracer needs a point or area light");
        MFH_REQUIRE(t.tree.posRootFlux > 0.0f, "the light tracer's lights emit nothing");
        MFH_REQUIRE(!(t.background.fluxSum > 0.0f),
                    "the light tracer emits from point and area lights only; the background emits");
        const size_t num = (size_t)t.camera.width * t.camera.height;
        if (t.ltMean.bytes < 4 * kLtChannels * num || t.ltRngs.bytes < 8 * num ||
            (t.ltVariance && t.ltVar.bytes < 4 * kLtChannels * num))
            lighttracer_reset(c, t, t.ltVariance);  // first use, or the resolution grew
        PtScene S = device_scene(c, t);
        S.rngs = t.ltRngs.as<uint64_t>();
        for (int32_t it = 0; it < iterations; ++it) {
            S.iteration = ++t.ltIteration;
            launch_lighttracer(S, t.ltSplat.as<float>(), t.ltMean.as<float>(),
                               t.ltVariance ? t.ltVar.as<float>() : nullptr, t.params.minPathLength,
                               t.params.maxPathLength, density != 0, c.stream);
        }
        MFH_HIP(hipStreamSynchronize(c.stream));
    });
}

int mfh_pt_lighttracer_get(mfh_context* ctx, int32_t ptId, int32_t target, int32_t variance, float* out) {
    return guarded([&] {
        Context& c = ctx_of(ctx);
        PathTracer& t = tracer_of(c, ptId);
        MFH_REQUIRE(out != nullptr, "null output pointer");
        MFH_REQUIRE(target >= MFH_LT_RADIANCE && target <= MFH_LT_DENSITY, "unknown light tracer target");
        MFH_REQUIRE(!variance || t.ltVariance, "the light tracer's variance is not recorded");
        const size_t num = (size_t)t.camera.width * t.camera.height;
        MFH_REQUIRE(t.ltIteration >= 0 && t.ltMean.bytes >= 4 * kLtChannels * num, "the light tracer has not run");
        std::vector<float> all((size_t)kLtChannels * num);
        MFH_HIP(hipStreamSynchronize(c.stream));
        MFH_HIP(hipMemcpy(all.data(), (variance ? t.ltVar : t.ltMean).ptr, 4 * all.size(), hipMemcpyDeviceToHost));
        const int first = target == MFH_LT_RADIANCE ? 0 : (target == MFH_LT_LIGHTNESS ? 3 : 4);
        const int nc = target == MFH_LT_RADIANCE ? 3 : 1;
        for (size_t i = 0; i < num; ++i)
            for (int k = 0; k < nc; ++k) out[nc * i + k] = all[kLtChannels * i + first + k];
    });
}

int mfh_pt_adjoint(mfh_context* ctx, int32_t ptId, int32_t iterations, const float* radianceAdjoint) {
    return guarded([&] {
        Context& c = ctx_of(ctx);
        PathTracer& t = tracer_of(c, ptId);
        MFH_REQUIRE(radianceAdjoint != nullptr, "null adjoint image");
        MFH_REQUIRE(iterations > 0, "the adjoint needs at least one iteration");
        MFH_REQUIRE(t.params.maxPathLength <= kAdjMaxVertices,
                    "the adjoint replays paths of at most " + std::to_string(kAdjMaxVertices) + " vertices");
        MFH_REQUIRE(t.tree.posLightCount > 0 || t.background.fluxSum > 0.0f, "scene without lights");
        PtScene S = device_scene(c, t);
        const int num = t.camera.width * t.camera.height;
        t.adjRngs.reserve(8 * (size_t)num);
        upload(c, t.adjDelta, radianceAdjoint, 12 * (size_t)num);
        const size_t nm = t.materials.size();
        t.adjEmGrad.reserve(12 * nm);
        MFH_HIP(hipMemsetAsync(t.adjEmGrad.ptr, 0, 12 * nm, c.stream));
        // gradient buffers for every texture used as an albedo or a roughness
        t.adjTexGrad.resize(c.textures.size());
        std::vector<float*> ptrs(std::max<size_t>(c.textures.size(), 1), nullptr);
        auto want = [&](int32_t id) {
            if (id < 0 || ptrs[id]) return;
            const Texture& x = *c.textures[id];
            DevBuf& g = t.adjTexGrad[id];
            g.reserve(16 * (size_t)x.width * x.height);
            ptrs[id] = g.as<float>();
        };
        for (const PtMaterialRecord& m : t.materials) {
            if (m.type == kMatLambert || m.type == kMatTorrance) want(m.albedoTexture);
            if (mat_has_roughness(m.type)) want(m.roughnessTexture);
        }
        for (size_t k = 0; k < c.textures.size(); ++k)
            if (ptrs[k]) MFH_HIP(hipMemsetAsync(ptrs[k], 0, 16 * (size_t)c.textures[k]->width * c.textures[k]->height,
                                                c.stream));
        upload(c, t.adjTexGradPtrs, ptrs.data(), sizeof(float*) * ptrs.size());
        AdjointState A;
        A.delta = t.adjDelta.as<float>();
        A.iterations = iterations;
        A.rngs = t.adjRngs.as<uint64_t>();
        A.texGrad = t.adjTexGradPtrs.as<float* const>();
        A.emissionGrad = t.adjEmGrad.as<float>();
        A.fixedDecisions = t.adjMode == MFH_ADJOINT_FIXED_DECISIONS ? 1 : 0;
        launch_pt_adjoint(S, A, t.params.seed * (num + 1), c.stream);
        t.adjValid = true;
    });
}

int mfh_pt_set_adjoint_mode(mfh_context* ctx, int32_t ptId, int32_t mode) {
    return guarded([&] {
        Context& c = ctx_of(ctx);
        PathTracer& t = tracer_of(c, ptId);
        MFH_REQUIRE(mode == MFH_ADJOINT_EXPECTED || mode == MFH_ADJOINT_FIXED_DECISIONS, "unknown adjoint mode");
        t.adjMode = mode;
    });
}

int mfh_pt_get_texture_gradient(mfh_context* ctx, int32_t ptId, int32_t textureId, float* out) {
    return guarded([&] {
        Context& c = ctx_of(ctx);
        PathTracer& t = tracer_of(c, ptId);
        MFH_REQUIRE(out != nullptr, "null output");
        MFH_REQUIRE(t.adjValid, "no adjoint has been run");
        MFH_REQUIRE(textureId >= 0 && textureId < (int32_t)t.adjTexGrad.size() && t.adjTexGrad[textureId].ptr,
                    "the texture has no gradient (not an albedo or roughness texture of this path tracer)");
        const Texture& x = *c.textures[textureId];
        const size_t n = (size_t)x.width * x.height;
        std::vector<float> rgba(4 * n);
        MFH_HIP(hipStreamSynchronize(c.stream));
        MFH_HIP(hipMemcpy(rgba.data(), t.adjTexGrad[textureId].ptr, 16 * n, hipMemcpyDeviceToHost));
        for (size_t i = 0; i < n; ++i)
            for (int k = 0; k < x.channels; ++k) out[(size_t)x.channels * i + k] = rgba[4 * i + k];
    });
}

int mfh_pt_get_emission_gradient(mfh_context* ctx, int32_t ptId, int32_t materialIndex, float* out) {
    return guarded([&] {
        Context& c = ctx_of(ctx);
        PathTracer& t = tracer_of(c, ptId);
        MFH_REQUIRE(out != nullptr, "null output");
        MFH_REQUIRE(t.adjValid, "no adjoint has been run");
        MFH_REQUIRE(materialIndex >= 0 && materialIndex < (int32_t)t.materials.size(), "material index out of range");
        MFH_HIP(hipStreamSynchronize(c.stream));
        MFH_HIP(hipMemcpy(out, t.adjEmGrad.as<float>() + 3 * (size_t)materialIndex, 12, hipMemcpyDeviceToHost));
    });
}

int mfh_pt_set_material(mfh_context* ctx, int32_t ptId, int32_t materialIndex, const mfh_material* material) {
    return guarded([&] {
        Context& c = ctx_of(ctx);
        PathTracer& t = tracer_of(c, ptId);
        MFH_REQUIRE(material != nullptr, "null material");
        MFH_REQUIRE(materialIndex >= 0 && materialIndex < (int32_t)t.materials.size(), "material index out of range");
        validate_material(c, *material);
        MFH_HIP(hipStreamSynchronize(c.stream));
        std::memcpy(&t.materials[materialIndex], material, sizeof(mfh_material));
        Scene& sc = scene_of(c, t.scene);
        std::vector<HostPointLight> pl = t.pointLights;
        const std::vector<HostAreaLight> area = gather_area_lights(c, sc, t.materials);
        build_light_tree(area, pl, sc.box, MaterialEmission(t.materials), t.tree);
        t.tree.backgroundFlux = t.background.fluxSum;
        upload(c, t.dMaterials, t.materials.data(), sizeof(PtMaterialRecord) * t.materials.size());
        upload(c, t.dLightMem, t.tree.mem.data(), std::max<size_t>(t.tree.mem.size(), 16));
        const size_t nm = t.tree.mapKeys.size();
        if (nm) {
            upload(c, t.dMapKeys, t.tree.mapKeys.data(), 8 * nm);
            upload(c, t.dMapCodes, t.tree.mapCodes.data(), 4 * nm);
        }
        drop_graph(t);
        reset(c, t);
    });
}

namespace mfh {
namespace {

PtBatch& batch_of(Context& c, int32_t id) {
    MFH_REQUIRE(id >= 0 && id < (int32_t)c.batches.size() && c.batches[id], "unknown batch id");
    return *c.batches[id];
}

// Device state of a batch: per-scene descriptors, path offsets, and the
// wavefront buffers over the paths of all scenes.
WfState batch_state(Context& c, PtBatch& b, std::vector<PtScene>& scenes) {
    scenes.clear();
    int32_t maxLen = 1, nee = 1;
    uint32_t slotMask = 0;
    for (int32_t id : b.tracers) {
        PathTracer& t = tracer_of(c, id);
        MFH_REQUIRE(t.tree.posLightCount > 0 || t.background.fluxSum > 0.0f, "a scene of the batch has no lights");
        scenes.push_back(device_scene(c, t));
        maxLen = std::max(maxLen, t.params.maxPathLength);
        nee = std::max(nee, t.params.neeCount);
        int slot = 0;
        for (int k = 0; k < kNumTargets; ++k) {
            if ((t.targetMask >> k) & 1u)
                for (int j = 0; j < kTargetChannels[k]; ++j) slotMask |= 1u << (slot + j);
            slot += kTargetChannels[k];
        }
    }
    upload(c, b.dScenes, scenes.data(), sizeof(PtScene) * scenes.size());
    const int N = b.offsets.back();
    const int numCounters = (maxLen + 1) + maxLen + 3 * maxLen;
    if (N != b.paths || maxLen != b.maxLen || nee != b.nee) {
        drop_batch_graph(b);
        b.counters.reserve(4 * (size_t)(numCounters + 1));
        b.tp.reserve(16 * (size_t)N);
        b.acc.reserve(4 * 18 * (size_t)N);
        for (auto& r : b.ray)
            for (DevBuf& d : r) d.reserve(16 * (size_t)N);
        for (DevBuf& d : b.sh) d.reserve(16 * (size_t)N * nee);
        b.shPath.reserve(8 * (size_t)N);
        b.paths = N;
        b.maxLen = maxLen;
        b.nee = nee;
    }
    WfState W;
    std::memset(&W, 0, sizeof(W));
    W.numPaths = N;
    W.slotMask = slotMask;
    W.chunk = b.tune.chunk;
    W.numCounters = numCounters;
    W.qBase = 0;
    W.shBase = maxLen + 1;
    W.workBase = 2 * maxLen + 1;
    W.counters = b.counters.as<int>();
    W.iteration = b.counters.as<int>() + numCounters;
    W.tp = b.tp.as<float4>();
    W.acc = b.acc.as<float>();
    bool posGuide = false;
    for (const PtScene& S : scenes) posGuide = posGuide || S.posGuide;
    for (int k = 0; k < 2; ++k) {
        W.rayO[k] = b.ray[k][0].as<float4>();
        W.rayD[k] = b.ray[k][1].as<float4>();
        W.rayM[k] = b.ray[k][2].as<float4>();
        W.hit[k] = b.ray[k][3].as<float4>();
        if (posGuide) {
            b.rayP[k].reserve(16 * (size_t)N);
            W.rayP[k] = b.rayP[k].as<float4>();
        }
    }
    W.shO = b.sh[0].as<float4>();
    W.shD = b.sh[1].as<float4>();
    W.shC = b.sh[2].as<float4>();
    W.shPath = b.shPath.as<int2>();
    W.neeStride = nee;
    W.scenes = b.dScenes.as<PtScene>();
    W.sceneOffset = b.dOffsets.as<int>();
    W.numScenes = (int)b.tracers.size();
    if (b.tune.traceGrid <= 0) b.tune.traceGrid = wf_grid_batch(b.tune, num_cus(c));
    if (b.tune.shadeGrid <= 0) b.tune.shadeGrid = wf_shade_grid(num_cus(c));
    if (b.tune.tailGrid <= 0) b.tune.tailGrid = wf_tail_grid(num_cus(c));
    return W;
}

void run_batch(Context& c, PtBatch& b, int32_t iterations) {
    std::vector<PtScene> scenes;
    WfState W = batch_state(c, b, scenes);
    const int32_t it = b.iteration;
    MFH_HIP(hipMemcpyAsync(W.iteration, &it, 4, hipMemcpyHostToDevice, c.stream));
    if (!b.useGraph) {
        for (int32_t k = 0; k < iterations; ++k) launch_pt_wavefront_batch(W, b.tune, b.maxLen, c.stream);
    } else {
        std::vector<uint8_t> key(sizeof(W) + sizeof(WfTuning) + sizeof(PtScene) * scenes.size());
        std::memcpy(key.data(), &W, sizeof(W));
        std::memcpy(key.data() + sizeof(W), &b.tune, sizeof(WfTuning));
        std::memcpy(key.data() + sizeof(W) + sizeof(WfTuning), scenes.data(), sizeof(PtScene) * scenes.size());
        if (!b.graphExec || key != b.graphKey) {
            drop_batch_graph(b);
            MFH_HIP(hipStreamSynchronize(c.stream));
            MFH_HIP(hipStreamBeginCapture(c.stream, hipStreamCaptureModeThreadLocal));
            try {
                launch_pt_wavefront_batch(W, b.tune, b.maxLen, c.stream);
            } catch (...) {
                hipGraph_t g = nullptr;
                (void)hipStreamEndCapture(c.stream, &g);
                if (g) (void)hipGraphDestroy(g);
                throw;
            }
            MFH_HIP(hipStreamEndCapture(c.stream, &b.graph));
            MFH_HIP(hipGraphInstantiate(&b.graphExec, b.graph, nullptr, nullptr, 0));
            b.graphKey = key;
        }
        for (int32_t k = 0; k < iterations; ++k) MFH_HIP(hipGraphLaunch(b.graphExec, c.stream));
    }
    b.iteration += iterations;
    for (int32_t id : b.tracers) tracer_of(c, id).iteration = b.iteration;
    MFH_HIP(hipStreamSynchronize(c.stream));
}

}  // namespace
}  // namespace mfh

int mfh_batch_create(mfh_context* ctx, const int32_t* ptIds, int32_t count, int32_t useGraph, int32_t* outBatchId) {
    return guarded([&] {
        Context& c = ctx_of(ctx);
        MFH_REQUIRE(ptIds && outBatchId && count > 0, "a batch needs at least one path tracer");
        auto* b = new PtBatch();
        try {
            int64_t total = 0;
            for (int32_t i = 0; i < count; ++i) {
                PathTracer& t = tracer_of(c, ptIds[i]);
                MFH_REQUIRE(t.params.neeCount <= 32, "the wavefront form supports at most 32 NEE samples per vertex");
                MFH_REQUIRE(!c.scenes[t.scene]->mixed,
                            "a batch holds scenes of triangles only (scene of path tracer " + std::to_string(ptIds[i]) +
                                " has quads or spheres)");
                b->tracers.push_back(ptIds[i]);
                b->offsets.push_back((int32_t)total);
                total += (int64_t)t.camera.width * t.camera.height;
            }
            MFH_REQUIRE(total < (int64_t)1 << 30, "too many paths in one batch");
            b->offsets.push_back((int32_t)total);
            b->useGraph = useGraph != 0;
            upload(c, b->dOffsets, b->offsets.data(), 4 * b->offsets.size());
            for (int32_t id : b->tracers) reset(c, tracer_of(c, id));
            MFH_HIP(hipStreamSynchronize(c.stream));
        } catch (...) {
            destroy_batch(b);
            throw;
        }
        c.batches.push_back(b);
        *outBatchId = (int32_t)c.batches.size() - 1;
    });
}

int mfh_batch_reset(mfh_context* ctx, int32_t batchId) {
    return guarded([&] {
        Context& c = ctx_of(ctx);
        PtBatch& b = batch_of(c, batchId);
        for (int32_t id : b.tracers) reset(c, tracer_of(c, id));
        b.iteration = -1;
    });
}

int mfh_batch_iterate(mfh_context* ctx, int32_t batchId, int32_t iterations) {
    return guarded([&] {
        Context& c = ctx_of(ctx);
        PtBatch& b = batch_of(c, batchId);
        MFH_REQUIRE(iterations >= 0, "negative iteration count");
        for (int32_t id : b.tracers)
            MFH_REQUIRE(tracer_of(c, id).iteration == b.iteration,
                        "a path tracer of the batch was iterated or reset on its own; reset the batch");
        run_batch(c, b, iterations);
    });
}

namespace mfh {
namespace {

// Trace variant of a tuning record (block, LDS stack depth, leaf batch); -1 if none matches.
int trace_variant_of(const mfh_wavefront_tuning& tuning) {
    for (int k = 0; k < kWfNumTraceVariants; ++k)
        if (kWfTraceShapes[k].block == tuning.traceBlock && kWfTraceShapes[k].lds == tuning.stackDepth &&
            kWfTraceShapes[k].leafBatch == tuning.leafBatch)
            return k;
    return -1;
}

void tuning_record(const WfTuning& tune, mfh_wavefront_tuning* out) {
    out->traceBlock = kWfTraceShapes[tune.traceVariant].block;
    out->stackDepth = kWfTraceShapes[tune.traceVariant].lds;
    out->queueChunk = tune.chunk;
    out->sortBounces = tune.sortBounces;
    out->refillBudget = tune.refill;
    out->leafBatch = kWfTraceShapes[tune.traceVariant].leafBatch;
    out->tailAfter = tune.tailAfter;
}

// Bounces after which the tail launch finishes the paths (0: never), tried by the autotuners.
constexpr int kTailChoices[] = {0, 1, 2, 3, 4, 6};

}  // namespace
}  // namespace mfh

int mfh_batch_set_tuning(mfh_context* ctx, int32_t batchId, const mfh_wavefront_tuning* tuning) {
    return guarded([&] {
        Context& c = ctx_of(ctx);
        PtBatch& b = batch_of(c, batchId);
        MFH_REQUIRE(tuning != nullptr, "null tuning");
        const int v = trace_variant_of(*tuning);
        MFH_REQUIRE(v >= 0, "unsupported trace block size / LDS stack depth / leaf batch combination");
        MFH_REQUIRE(tuning->queueChunk >= 64 && tuning->queueChunk % 64 == 0 && tuning->queueChunk <= 65536,
                    "queue chunk must be a multiple of 64 in [64, 65536]");
        MFH_REQUIRE(tuning->sortBounces == 0, "batches do not sort rays");
        MFH_REQUIRE(tuning->refillBudget >= 0 && tuning->refillBudget <= 4096, "refillBudget out of range");
        MFH_REQUIRE(tuning->tailAfter >= 0 && tuning->tailAfter <= 64, "tailAfter out of range");
        b.tune.traceVariant = v;
        b.tune.chunk = tuning->queueChunk;
        b.tune.refill = tuning->refillBudget;
        b.tune.tailAfter = tuning->tailAfter;
        b.tune.traceGrid = 0;
        drop_batch_graph(b);
    });
}

int mfh_batch_autotune(mfh_context* ctx, int32_t batchId, int32_t iterations, mfh_wavefront_tuning* outBest,
                       float* outMsPerIteration) {
    return guarded([&] {
        Context& c = ctx_of(ctx);
        PtBatch& b = batch_of(c, batchId);
        MFH_REQUIRE(iterations > 0, "autotune needs at least one iteration per candidate");
        float best = 3.4e38f;
        WfTuning bestTune = b.tune;
        for (int v = 0; v < kWfNumTraceVariants; ++v) {
            for (int m = 0; m < kTraceModes; ++m) {
                if (kTraceModeRefill[m] > 0 && kWfTraceShapes[v].leafBatch != 1) continue;  // same kernels as leaf 1
                b.tune.traceVariant = v;
                b.tune.chunk = kTraceModeChunk[m];
                b.tune.refill = kTraceModeRefill[m];
                b.tune.traceGrid = 0;
                drop_batch_graph(b);
                run_batch(c, b, 1);
                MFH_HIP(hipEventRecord(c.evBegin, c.stream));
                run_batch(c, b, iterations);
                MFH_HIP(hipEventRecord(c.evEnd, c.stream));
                MFH_HIP(hipEventSynchronize(c.evEnd));
                float ms = 0.0f;
                MFH_HIP(hipEventElapsedTime(&ms, c.evBegin, c.evEnd));
                ms /= (float)iterations;
                if (ms < best) {
                    best = ms;
                    bestTune = b.tune;
                }
            }
        }
        // then the bounce of the tail launch, on the best trace shape
        for (int tail : kTailChoices) {
            b.tune = bestTune;
            b.tune.tailAfter = tail;
            drop_batch_graph(b);
            run_batch(c, b, 1);
            MFH_HIP(hipEventRecord(c.evBegin, c.stream));
            run_batch(c, b, iterations);
            MFH_HIP(hipEventRecord(c.evEnd, c.stream));
            MFH_HIP(hipEventSynchronize(c.evEnd));
            float ms = 0.0f;
            MFH_HIP(hipEventElapsedTime(&ms, c.evBegin, c.evEnd));
            ms /= (float)iterations;
            if (ms < best) {
                best = ms;
                bestTune = b.tune;
            }
        }
        b.tune = bestTune;
        drop_batch_graph(b);
        for (int32_t id : b.tracers) reset(c, tracer_of(c, id));
        b.iteration = -1;
        if (outBest) tuning_record(b.tune, outBest);
        if (outMsPerIteration) *outMsPerIteration = best;
    });
}

int mfh_pt_set_mode(mfh_context* ctx, int32_t ptId, int32_t mode, int32_t useGraph) {
    return guarded([&] {
        Context& c = ctx_of(ctx);
        PathTracer& t = tracer_of(c, ptId);
        MFH_REQUIRE(mode == MFH_PT_MEGAKERNEL || mode == MFH_PT_WAVEFRONT, "unknown path tracer mode");
        MFH_REQUIRE(!(mode == MFH_PT_WAVEFRONT && t.params.neeCount > 32),
                    "the wavefront form supports at most 32 NEE samples per vertex");
        MFH_HIP(hipStreamSynchronize(c.stream));
        t.mode = mode;
        t.useGraph = useGraph != 0;
        if (mode != MFH_PT_WAVEFRONT) release_wavefront(t);
    });
}

int mfh_pt_set_tuning(mfh_context* ctx, int32_t ptId, const mfh_wavefront_tuning* tuning) {
    return guarded([&] {
        Context& c = ctx_of(ctx);
        PathTracer& t = tracer_of(c, ptId);
        MFH_REQUIRE(tuning != nullptr, "null tuning");
        const int v = trace_variant_of(*tuning);
        MFH_REQUIRE(v >= 0, "unsupported trace block size / LDS stack depth / leaf batch combination");
        MFH_REQUIRE(tuning->queueChunk >= 64 && tuning->queueChunk % 64 == 0 && tuning->queueChunk <= 65536,
                    "queue chunk must be a multiple of 64 in [64, 65536]");
        MFH_HIP(hipStreamSynchronize(c.stream));
        MFH_REQUIRE(tuning->sortBounces >= 0 && tuning->sortBounces <= 64, "sortBounces out of range");
        MFH_REQUIRE(tuning->refillBudget >= 0 && tuning->refillBudget <= 4096, "refillBudget out of range");
        MFH_REQUIRE(tuning->tailAfter >= 0 && tuning->tailAfter <= 64, "tailAfter out of range");
        t.tune.traceVariant = v;
        t.tune.chunk = tuning->queueChunk;
        t.tune.sortBounces = tuning->sortBounces;
        t.tune.refill = tuning->refillBudget;
        t.tune.tailAfter = tuning->tailAfter;
        t.tune.traceGrid = 0;
        drop_graph(t);
    });
}

int mfh_pt_get_tuning(mfh_context* ctx, int32_t ptId, mfh_wavefront_tuning* out) {
    return guarded([&] {
        Context& c = ctx_of(ctx);
        PathTracer& t = tracer_of(c, ptId);
        MFH_REQUIRE(out != nullptr, "null output");
        tuning_record(t.tune, out);
    });
}

int mfh_pt_autotune(mfh_context* ctx, int32_t ptId, int32_t iterations, mfh_wavefront_tuning* outBest,
                    float* outMsPerIteration) {
    return guarded([&] {
        Context& c = ctx_of(ctx);
        PathTracer& t = tracer_of(c, ptId);
        MFH_REQUIRE(iterations > 0, "autotune needs at least one iteration per candidate");
        MFH_REQUIRE(t.mode == MFH_PT_WAVEFRONT, "autotune applies to the wavefront form");
        MFH_REQUIRE(t.tree.posLightCount > 0 || t.background.fluxSum > 0.0f, "scene without lights");
        PtScene S = device_scene(c, t);
        const int sorts[3] = {0, 1, 2};
        float best = 3.4e38f;
        WfTuning bestTune = t.tune;
        for (int v = 0; v < kWfNumTraceVariants; ++v) {
            for (int cs = 0; cs < 3 * kTraceModes; ++cs) {
                const int m = cs % kTraceModes;
                if (kTraceModeRefill[m] > 0 && kWfTraceShapes[v].leafBatch != 1) continue;  // same kernels as leaf 1
                t.tune.traceVariant = v;
                t.tune.chunk = kTraceModeChunk[m];
                t.tune.refill = kTraceModeRefill[m];
                t.tune.sortBounces = sorts[cs / kTraceModes];
                t.tune.traceGrid = 0;
                drop_graph(t);
                run_wavefront(c, t, S, 1);  // warm-up (and graph capture)
                MFH_HIP(hipEventRecord(c.evBegin, c.stream));
                run_wavefront(c, t, S, iterations);
                MFH_HIP(hipEventRecord(c.evEnd, c.stream));
                MFH_HIP(hipEventSynchronize(c.evEnd));
                float ms = 0.0f;
                MFH_HIP(hipEventElapsedTime(&ms, c.evBegin, c.evEnd));
                ms /= (float)iterations;
                if (ms < best) {
                    best = ms;
                    bestTune = t.tune;
                }
            }
        }
        // then the bounce of the tail launch, on the best trace shape
        for (int tail : kTailChoices) {
            t.tune = bestTune;
            t.tune.tailAfter = tail;
            drop_graph(t);
            run_wavefront(c, t, S, 1);
            MFH_HIP(hipEventRecord(c.evBegin, c.stream));
            run_wavefront(c, t, S, iterations);
            MFH_HIP(hipEventRecord(c.evEnd, c.stream));
            MFH_HIP(hipEventSynchronize(c.evEnd));
            float ms = 0.0f;
            MFH_HIP(hipEventElapsedTime(&ms, c.evBegin, c.evEnd));
            ms /= (float)iterations;
            if (ms < best) {
                best = ms;
                bestTune = t.tune;
            }
        }
        t.tune = bestTune;
        drop_graph(t);
        reset(c, t);
        if (outBest) tuning_record(t.tune, outBest);
        if (outMsPerIteration) *outMsPerIteration = best;
    });
}

int mfh_pt_get_target(mfh_context* ctx, int32_t ptId, int32_t target, float* out) {
    return guarded([&] {
        Context& c = ctx_of(ctx);
        PathTracer& t = tracer_of(c, ptId);
        MFH_REQUIRE(out != nullptr, "null output pointer");
        const size_t num = (size_t)t.camera.width * t.camera.height;
        MFH_HIP(hipStreamSynchronize(c.stream));
        MFH_REQUIRE(target >= 0 && target < kNumTargets, "unknown render target " + std::to_string(target));
        MFH_REQUIRE((t.targetMask >> target) & 1u, "render target " + std::to_string(target) + " is not enabled");
        MFH_HIP(hipMemcpy(out, t.dTargets[target].ptr, 4 * (size_t)kTargetChannels[target] * num,
                          hipMemcpyDeviceToHost));
    });
}

namespace {

// OutputHandler::enable_render_target / disable_render_target
// (output_handler.hpp:391-430): enabling with variance records both,
// disabling always drops the variance and, unless only the variance is
// disabled, the target. A change resets the render
// (IRenderer::on_render_target_changed, interface.cpp:3627-3628).
void set_targets(mfh::Context& c, mfh::PathTracer& t, uint32_t mask, uint32_t varMask) {
    varMask &= mask;
    if (mask == t.targetMask && varMask == t.varianceMask) return;
    MFH_HIP(hipStreamSynchronize(c.stream));
    for (int k = 0; k < kNumTargets; ++k) {
        if (!((mask >> k) & 1u)) t.dTargets[k].release();
        if (!((varMask >> k) & 1u)) t.dVariance[k].release();
    }
    t.targetMask = mask;
    t.varianceMask = varMask;
    drop_graph(t);
    reset(c, t);
}

}  // namespace

int mfh_pt_enable_target(mfh_context* ctx, int32_t ptId, int32_t target, int32_t enable) {
    return guarded([&] {
        Context& c = ctx_of(ctx);
        PathTracer& t = tracer_of(c, ptId);
        MFH_REQUIRE(target >= 0 && target < kNumTargets, "unknown render target " + std::to_string(target));
        const uint32_t bit = 1u << target;
        if (enable)
            set_targets(c, t, t.targetMask | bit, t.varianceMask);
        else
            set_targets(c, t, t.targetMask & ~bit, t.varianceMask & ~bit);
    });
}

int mfh_pt_enable_variance(mfh_context* ctx, int32_t ptId, int32_t target, int32_t enable) {
    return guarded([&] {
        Context& c = ctx_of(ctx);
        PathTracer& t = tracer_of(c, ptId);
        MFH_REQUIRE(target >= 0 && target < kNumTargets, "unknown render target " + std::to_string(target));
        const uint32_t bit = 1u << target;
        if (enable)
            set_targets(c, t, t.targetMask | bit, t.varianceMask | bit);
        else
            set_targets(c, t, t.targetMask, t.varianceMask & ~bit);
    });
}

int mfh_pt_get_variance(mfh_context* ctx, int32_t ptId, int32_t target, float* out) {
    return guarded([&] {
        Context& c = ctx_of(ctx);
        PathTracer& t = tracer_of(c, ptId);
        MFH_REQUIRE(out != nullptr, "null output pointer");
        MFH_REQUIRE(target >= 0 && target < kNumTargets, "unknown render target " + std::to_string(target));
        MFH_REQUIRE((t.varianceMask >> target) & 1u,
                    "render target " + std::to_string(target) + " does not record variance");
        const size_t num = (size_t)t.camera.width * t.camera.height;
        MFH_HIP(hipStreamSynchronize(c.stream));
        MFH_HIP(hipMemcpy(out, t.dVariance[target].ptr, 4 * (size_t)kTargetChannels[target] * num,
                          hipMemcpyDeviceToHost));
    });
}

int mfh_pt_get_iteration(mfh_context* ctx, int32_t ptId, int32_t* outIteration) {
    return guarded([&] {
        Context& c = ctx_of(ctx);
        MFH_REQUIRE(outIteration != nullptr, "null output pointer");
        *outIteration = tracer_of(c, ptId).iteration;
    });
}

int mfh_pt_get_nan_count(mfh_context* ctx, int32_t ptId, uint32_t* outCount) {
    return guarded([&] {
        Context& c = ctx_of(ctx);
        PathTracer& t = tracer_of(c, ptId);
        MFH_REQUIRE(outCount != nullptr, "null output pointer");
        MFH_HIP(hipStreamSynchronize(c.stream));
        MFH_HIP(hipMemcpy(outCount, t.dNan.ptr, 4, hipMemcpyDeviceToHost));
    });
}

int mfh_pt_get_light_tree(mfh_context* ctx, int32_t ptId, mfh_light_tree_info* outInfo, void* outMem,
                          uint64_t* outMapKeys, uint32_t* outMapCodes) {
    return guarded([&] {
        Context& c = ctx_of(ctx);
        PathTracer& t = tracer_of(c, ptId);
        MFH_REQUIRE(outInfo != nullptr, "null output pointer");
        outInfo->posRootFlux = t.tree.posRootFlux;
        outInfo->posRootType = t.tree.posRootType;
        outInfo->posInternalCount = t.tree.posInternalCount;
        outInfo->posLightCount = t.tree.posLightCount;
        outInfo->memBytes = (uint32_t)t.tree.mem.size();
        outInfo->mapCount = (int32_t)t.tree.mapKeys.size();
        outInfo->dirRootFlux = t.tree.dirRootFlux;
        outInfo->backgroundFlux = t.tree.backgroundFlux;
        if (outMem && !t.tree.mem.empty()) std::memcpy(outMem, t.tree.mem.data(), t.tree.mem.size());
        if (outMapKeys) std::memcpy(outMapKeys, t.tree.mapKeys.data(), 8 * t.tree.mapKeys.size());
        if (outMapCodes) std::memcpy(outMapCodes, t.tree.mapCodes.data(), 4 * t.tree.mapCodes.size());
    });
}

int mfh_texture_create(mfh_context* ctx, int32_t width, int32_t height, int32_t channels, int32_t samplingMode,
                       const float* texels, int32_t* outTextureId) {
    return guarded([&] {
        Context& c = ctx_of(ctx);
        MFH_REQUIRE(texels && outTextureId, "null pointer");
        MFH_REQUIRE(width > 0 && height > 0 && (int64_t)width * height <= (int64_t)1 << 28, "texture size out of range");
        MFH_REQUIRE(channels >= 1 && channels <= 4, "texture channels must be 1..4");
        MFH_REQUIRE(samplingMode == MFH_SAMPLING_NEAREST || samplingMode == MFH_SAMPLING_LINEAR, "unknown sampling mode");
        auto* t = new Texture();
        try {
            t->width = width;
            t->height = height;
            t->channels = channels;
            t->sampling = samplingMode;
            const size_t n = (size_t)width * height;
            t->host.assign(4 * n, 0.0f);
            for (size_t i = 0; i < n; ++i)
                for (int k = 0; k < channels; ++k) t->host[4 * i + k] = texels[(size_t)channels * i + k];
            upload(c, t->dev, t->host.data(), 16 * n);
            MFH_HIP(hipStreamSynchronize(c.stream));
        } catch (...) {
            t->dev.release();
            delete t;
            throw;
        }
        c.textures.push_back(t);
        *outTextureId = (int32_t)c.textures.size() - 1;
    });
}

int mfh_texture_update(mfh_context* ctx, int32_t textureId, const float* texels) {
    return guarded([&] {
        Context& c = ctx_of(ctx);
        MFH_REQUIRE(texels != nullptr, "null texels");
        MFH_REQUIRE(textureId >= 0 && textureId < (int32_t)c.textures.size() && c.textures[textureId],
                    "unknown texture id");
        Texture& t = *c.textures[textureId];
        const size_t n = (size_t)t.width * t.height;
        for (size_t i = 0; i < n; ++i)
            for (int k = 0; k < t.channels; ++k) t.host[4 * i + k] = texels[(size_t)t.channels * i + k];
        MFH_HIP(hipStreamSynchronize(c.stream));
        MFH_HIP(hipMemcpy(t.dev.ptr, t.host.data(), 16 * n, hipMemcpyHostToDevice));
    });
}

int mfh_background_build(const mfh_background* bg, const float* envTexels, int32_t width, int32_t height,
                         int32_t channels, const float* sceneBox, float* outFlux, float* outSat) {
    return guarded([&] {
        MFH_REQUIRE(bg && sceneBox && outFlux, "null pointer");
        MFH_REQUIRE(bg->type == MFH_BACKGROUND_COLORED || bg->type == MFH_BACKGROUND_ENVMAP, "unknown background type");
        BackgroundHost out;
        if (bg->type == MFH_BACKGROUND_ENVMAP) {
            MFH_REQUIRE(envTexels && outSat, "an environment map needs texels and a SAT output");
            MFH_REQUIRE(width > 0 && height > 0 && channels >= 1 && channels <= 4, "invalid environment map");
            std::vector<float> rgba(4 * (size_t)width * height, 0.0f);
            for (size_t i = 0; i < (size_t)width * height; ++i)
                for (int k = 0; k < channels; ++k) rgba[4 * i + k] = envTexels[(size_t)channels * i + k];
            build_background(bg->type, bg->color, bg->scale, rgba.data(), width, height, sceneBox, out);
            std::memcpy(outSat, out.sat.data(), 4 * out.sat.size());
        } else {
            build_background(bg->type, bg->color, bg->scale, nullptr, 0, 0, sceneBox, out);
        }
        for (int k = 0; k < 3; ++k) outFlux[k] = out.flux[k];
    });
}

int mfh_pt_set_background(mfh_context* ctx, int32_t ptId, const mfh_background* bg) {
    return guarded([&] {
        Context& c = ctx_of(ctx);
        PathTracer& t = tracer_of(c, ptId);
        MFH_REQUIRE(bg != nullptr, "null background");
        MFH_REQUIRE(bg->type == MFH_BACKGROUND_COLORED || bg->type == MFH_BACKGROUND_ENVMAP, "unknown background type");
        Scene& sc = scene_of(c, t.scene);
        refresh_scene(c, sc);
        BackgroundHost out;
        if (bg->type == MFH_BACKGROUND_ENVMAP) {
            const int32_t id = bg->envmapTexture;
            MFH_REQUIRE(id >= 0 && id < (int32_t)c.textures.size() && c.textures[id], "unknown environment map texture");
            const Texture& x = *c.textures[id];
            build_background(bg->type, bg->color, bg->scale, x.host.data(), x.width, x.height, sc.box, out);
            out.envmapTexture = id;
            upload(c, t.dSat, out.sat.data(), 4 * out.sat.size());
        } else {
            build_background(bg->type, bg->color, bg->scale, nullptr, 0, 0, sc.box, out);
        }
        t.background = out;
        t.tree.backgroundFlux = out.fluxSum;
        drop_graph(t);
        reset(c, t);
    });
}

namespace {
int light_tree_build(const std::vector<HostAreaLight>& area, const mfh_point_light* pointLights,
                     int32_t numPointLights, const float* sceneBox, const mfh_material* materials,
                     int32_t numMaterials, mfh_light_tree_info* outInfo, void* outMem, size_t memCapacity,
                     uint64_t* outMapKeys, uint32_t* outMapCodes, int32_t mapCapacity);
}  // namespace

int mfh_light_tree_build(const mfh_area_light* areaLights, int32_t numAreaLights, const mfh_point_light* pointLights,
                         int32_t numPointLights, const float* sceneBox, const mfh_material* materials,
                         int32_t numMaterials, mfh_light_tree_info* outInfo, void* outMem, size_t memCapacity,
                         uint64_t* outMapKeys, uint32_t* outMapCodes, int32_t mapCapacity) {
    return guarded([&] {
        MFH_REQUIRE(numAreaLights >= 0, "negative light count");
        MFH_REQUIRE(numAreaLights == 0 || areaLights, "null area lights");
        std::vector<HostAreaLight> area(numAreaLights);
        for (int32_t i = 0; i < numAreaLights; ++i) {
            const mfh_area_light& a = areaLights[i];
            MFH_REQUIRE(a.material >= 0 && a.material < numMaterials, "material index out of range");
            for (int k = 0; k < 3; ++k) {
                for (int r = 0; r < 3; ++r) area[i].points[k][r] = a.points[3 * k + r];
                area[i].uv[k][0] = a.uv[2 * k];
                area[i].uv[k][1] = a.uv[2 * k + 1];
            }
            area[i].material = (uint16_t)a.material;
            area[i].instance = a.instance;
            area[i].primitive = a.primitive;
        }
        light_tree_build(area, pointLights, numPointLights, sceneBox, materials, numMaterials, outInfo, outMem,
                         memCapacity, outMapKeys, outMapCodes, mapCapacity);
    });
}

int mfh_light_tree_build_ex(const mfh_area_light_ex* areaLights, int32_t numAreaLights,
                            const mfh_point_light* pointLights, int32_t numPointLights, const float* sceneBox,
                            const mfh_material* materials, int32_t numMaterials, mfh_light_tree_info* outInfo,
                            void* outMem, size_t memCapacity, uint64_t* outMapKeys, uint32_t* outMapCodes,
                            int32_t mapCapacity) {
    return guarded([&] {
        MFH_REQUIRE(numAreaLights >= 0, "negative light count");
        MFH_REQUIRE(numAreaLights == 0 || areaLights, "null area lights");
        std::vector<HostAreaLight> area(numAreaLights);
        for (int32_t i = 0; i < numAreaLights; ++i) {
            const mfh_area_light_ex& a = areaLights[i];
            MFH_REQUIRE(a.kind >= kAreaTriangle && a.kind <= kAreaSphere, "unknown area light kind");
            MFH_REQUIRE(a.material >= 0 && a.material < numMaterials, "material index out of range");
            area[i].kind = a.kind;
            for (int k = 0; k < 4; ++k) {
                for (int r = 0; r < 3; ++r) area[i].points[k][r] = a.points[3 * k + r];
                area[i].uv[k][0] = a.uv[2 * k];
                area[i].uv[k][1] = a.uv[2 * k + 1];
            }
            area[i].radius = a.radius;
            area[i].material = (uint16_t)a.material;
            area[i].instance = a.instance;
            area[i].primitive = a.primitive;
        }
        light_tree_build(area, pointLights, numPointLights, sceneBox, materials, numMaterials, outInfo, outMem,
                         memCapacity, outMapKeys, outMapCodes, mapCapacity);
    });
}

}  // extern "C"

namespace {
int light_tree_build(const std::vector<HostAreaLight>& area, const mfh_point_light* pointLights,
                     int32_t numPointLights, const float* sceneBox, const mfh_material* materials,
                     int32_t numMaterials, mfh_light_tree_info* outInfo, void* outMem, size_t memCapacity,
                     uint64_t* outMapKeys, uint32_t* outMapCodes, int32_t mapCapacity) {
    {
        MFH_REQUIRE(sceneBox && materials && outInfo, "null pointer");
        MFH_REQUIRE(numPointLights >= 0, "negative light count");
        MFH_REQUIRE(numPointLights == 0 || pointLights, "null point lights");
        std::vector<PtMaterialRecord> mats(numMaterials);
        std::memcpy(mats.data(), materials, sizeof(mfh_material) * (size_t)numMaterials);
        std::vector<HostPointLight> pl(numPointLights);
        for (int32_t i = 0; i < numPointLights; ++i) {
            std::memcpy(pl[i].position, pointLights[i].position, 12);
            std::memcpy(pl[i].intensity, pointLights[i].intensity, 12);
        }
        LightTreeHost tree;
        build_light_tree(area, pl, sceneBox, MaterialEmission(mats), tree);
        MFH_REQUIRE(tree.mem.size() <= memCapacity, "light tree memory capacity too small");
        MFH_REQUIRE((int32_t)tree.mapKeys.size() <= mapCapacity, "light tree map capacity too small");
        outInfo->posRootFlux = tree.posRootFlux;
        outInfo->posRootType = tree.posRootType;
        outInfo->posInternalCount = tree.posInternalCount;
        outInfo->posLightCount = tree.posLightCount;
        outInfo->memBytes = (uint32_t)tree.mem.size();
        outInfo->mapCount = (int32_t)tree.mapKeys.size();
        outInfo->dirRootFlux = tree.dirRootFlux;
        outInfo->backgroundFlux = tree.backgroundFlux;
        if (outMem && !tree.mem.empty()) std::memcpy(outMem, tree.mem.data(), tree.mem.size());
        if (outMapKeys && !tree.mapKeys.empty()) std::memcpy(outMapKeys, tree.mapKeys.data(), 8 * tree.mapKeys.size());
        if (outMapCodes && !tree.mapCodes.empty())
            std::memcpy(outMapCodes, tree.mapCodes.data(), 4 * tree.mapCodes.size());
    }
    return 0;
}
}  // namespace
