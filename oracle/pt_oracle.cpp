// Oracle — test infrastructure only.
//
// CPU restatement of Mufflon's pure path tracer iteration for the supported
// scene content (triangles; LAMBERT and EMISSIVE materials with constant
// textures; area-triangle and point lights through the light tree; pinhole
// camera; vacuum), following
//   pt_common.hpp:75-156       pt_sample (RNG consumption order, NEE, MIS)
//   random_walk.hpp:38-112     walk (sample, Russian roulette, first hit)
//   path_util.hpp:117-748      PathVertex (incident encoding, evaluate, sample,
//                              get_emission, create_surface / create_void)
//   pt_common.hpp:27-68        PtVertexExt (incidentPdf for MIS)
//   pinhole.hpp:43-84          pinholecam_sample_ray
//   material_sampling.hpp:68-237, lambert.hpp, emissive.hpp
//   light_tree_sampling.hpp:195-329 (connect), :340-412 (light_pdf)
//   light_sampling.hpp:118-140,308-375 (area triangle, point connect_light)
//   intersection.cu:366-621 (hit attributes), tangent_space.hpp:10-60
//   rng.hpp:12-93, sampling.hpp:23-46,175-200 (PCG64, samplers, percentage_of)
//   render_target.hpp:62-76,189-195 (contribute, NaN check)
//   output_handler_variance.hpp:7-31 (running mean), cpu_pt.cpp:18-45 (seeds)
// Every float expression keeps the reference's operand order; the build uses
// -ffp-contract=off. Epsilon functions (normalize, sgn, clamp) are this
// project's definitions: normalize(v) = v / len(v), sgn(x) = x < 0 ? -1 : 1.
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

#include "trace_core.hpp"

using namespace orc;

namespace {

constexpr float kPi = 3.14159265358979323846f;
constexpr float kMaxSceneSize = 1048576.0f;  // types.hpp:26
constexpr float kSceneEps = 1e-4f;          // intersection.cu:16
constexpr float kInfPdf = 4294967296.f;      // AreaPdf/AngularPdf::infinite()
constexpr float kDistSqEps = 1e-10f;        // light_sampling.hpp:78
constexpr uint16_t kInternalNode = 0xFFFF;

inline float sgnf(float x) { return x < 0.0f ? -1.0f : 1.0f; }
inline float clampf(float v, float lo, float hi) { return mn(mx(v, lo), hi); }
inline V3 normalize(V3 v) { return v / len(v); }

struct Pcg64 {  // rng.hpp:70-93
    u64 s;
    static u64 scramble(u32 x) {
        u64 z = (u64)x + 0x9e3779b97f4a7c15ull;
        z = (z ^ (z >> 30u)) * 0xbf58476d1ce4e5b9ull;
        z = (z ^ (z >> 27u)) * 0x94d049bb133111ebull;
        return z ^ (z >> 31u);
    }
    explicit Pcg64(u32 seed) {
        s = scramble(seed);
        next();
        s += seed;
        next();
    }
    u64 next() {
        u64 x = s;
        const u64 count = s >> 59ull;
        s = s * 6364136223846793005ull + 1442695040888963407ull;
        x ^= x >> (5 + count);
        x *= 12605985483714917081ull;
        return x ^ (x >> 43);
    }
};

inline float sample_uniform(u32 u) { return (float)u / 4294967810.0f; }
inline u64 percentage_of(u64 num, float p) {
    const u64 pfix = (u64)(p * 2147483648.0f);
    return ((pfix * (num & 0x7fffffff)) >> 31) + (pfix * (num >> 31));
}

}  // namespace

extern "C" {
struct OrcMaterial {  // mfh_material, 64 bytes
    int32_t type;
    float albedo[3];
    float radiance[3];
    float scale;
    int32_t reserved[8];
};
struct OrcPtScene {
    OrcScene geo;
    const float* normals;     // 3 per vertex, concatenated like geo.verts
    const float* uvs;         // 2 per vertex
    const uint16_t* matIdx;   // 1 per triangle, concatenated like geo.tris
    const OrcMaterial* materials;
    int32_t numMaterials;
    const uint8_t* lightMem;  // positional light subtree (Mufflon layout)
    float posRootFlux;
    uint32_t posRootType;
    uint32_t posInternalCount;
    uint32_t posLightCount;
    float dirRootFlux;
    float bgFlux;
    const uint64_t* mapKeys;
    const uint32_t* mapCodes;
    int32_t mapCount;
    float sceneBox[6];
    float camPos[3];
    float tanVFov;
    float camDir[3];
    float camUp[3];
    int32_t width, height;
};
}

namespace {

enum Mat : int32_t { kEmissive = 0, kLambert = 1 };

struct TS {  // TangentSpace
    V3 shadingN{0, 0, 0}, geoN{0, 0, 0}, shadingTX{0, 0, 0}, shadingTY{0, 0, 0};
};

struct Vertex {
    bool isCamera = true, isVoid = false;
    V3 position{0, 0, 0};
    V3 incEnc{0, 0, 0};  // (x, y, dist * sgn(z)) for surfaces, raw dir for void
    TS ts;
    i32 inst = -1, prim = -1;
    float sp0 = 0, sp1 = 0;  // surface parameters (barycentrics)
    const OrcMaterial* mat = nullptr;
    float incidentPdf = 1.0f;  // PtVertexExt
    V3 incident_dir() const {
        const float z = sgnf(incEnc.z) * std::sqrt(clampf(1.0f - incEnc.x * incEnc.x - incEnc.y * incEnc.y, 0.0f, 1.0f));
        return {incEnc.x, incEnc.y, z};
    }
    V3 geo_normal() const { return (isCamera || isVoid) ? V3{0, 0, 0} : ts.geoN; }
};

struct Ctx {
    const OrcPtScene& S;
    V3 vert(i32 lodBase, u32 i) const {
        const float* v = S.geo.verts + 3 * ((size_t)lodBase + i);
        return {v[0], v[1], v[2]};
    }
};

M34 mat_of(const OrcScene& g, i32 inst) {
    M34 m;
    std::memcpy(m.m, g.worldToInst + 12 * (size_t)inst, 48);
    return m;
}

// intersection.cu:16-24
inline void add_epsilon(V3& o, V3 dir, V3 geoN) {
    const V3 off = geoN * kSceneEps;
    if (dot(geoN, dir) >= 0.0f)
        o = o + off;
    else
        o = o - off;
}

struct HitInfo {
    bool hit;
    float t;
    i32 inst, prim;
    float b0, b1;
    V3 geoN, tangentX;
    float u, v;
};

// first_intersection (intersection.cu:366-621), triangle meshes.
HitInfo first_intersection(const OrcPtScene& S, V3& origin, V3 dir, V3 geoNormal, float tmax) {
    add_epsilon(origin, dir, geoNormal);
    const FastRay fr = make_fast_ray(origin, dir);
    float hitT = tmax;
    Hit h;
    traverse(S.geo, fr, hitT, h, false, nullptr, nullptr);
    HitInfo r;
    r.t = hitT;
    r.inst = h.inst;
    r.prim = h.prim;
    r.hit = h.inst != -1;
    if (!r.hit) return r;
    r.b0 = h.w0;
    r.b1 = h.w1;
    const OrcScene& g = S.geo;
    const i32 lod = g.instLod[h.inst];
    const i32 vb = g.lodVertOffset[lod], tb = g.lodTriOffset[lod];
    const u32* ix = g.tris + 3 * ((size_t)tb + h.prim);
    V3 v[3];
    float uvs[3][2];
    for (int k = 0; k < 3; ++k) {
        const float* p = g.verts + 3 * ((size_t)vb + ix[k]);
        v[k] = {p[0], p[1], p[2]};
        const float* q = S.uvs + 2 * ((size_t)vb + ix[k]);
        uvs[k][0] = q[0];
        uvs[k][1] = q[1];
    }
    const V3 dx0 = v[1] - v[0];
    const V3 dx1 = v[2] - v[0];
    const float du0x = uvs[1][0] - uvs[0][0], du0y = uvs[1][1] - uvs[0][1];
    const float du1x = uvs[2][0] - uvs[0][0], du1y = uvs[2][1] - uvs[0][1];
    const float det = du0x * du1y - du0y * du1x;
    V3 tangentX;
    if (det >= 1e-5f || det <= -1e5f)
        tangentX = (dx0 * du1y - dx1 * du0y) / det;
    else
        tangentX = dx0;
    V3 geoN = cross(dx0, dx1);
    const float w2 = 1.f - h.w0 - h.w1;
    r.u = uvs[0][0] * h.w0 + uvs[1][0] * h.w1 + uvs[2][0] * w2;
    r.v = uvs[0][1] * h.w0 + uvs[1][1] * h.w1 + uvs[2][1] * w2;
    const M34 M = mat_of(g, h.inst);
    r.geoN = normalize(xform_normal(M, geoN));
    r.tangentX = normalize(xform_normal(M, tangentX));
    return r;
}

// any_intersection (intersection.cu:624-753)
bool any_intersection(const OrcPtScene& S, V3 a, V3 b, V3 geoNA, V3 geoNB, V3 dirAtoB) {
    add_epsilon(a, dirAtoB, geoNA);
    add_epsilon(b, -dirAtoB, geoNB);
    const V3 c = b - a;
    float tmax = len(c);
    const FastRay fr = make_fast_ray(a, c / tmax);
    Hit h;
    return traverse(S.geo, fr, tmax, h, true, nullptr, nullptr);
}

// tangent_space_geom_to_shader (tangent_space.hpp:10-60), triangles
TS tangent_space(const OrcPtScene& S, const HitInfo& hit) {
    const OrcScene& g = S.geo;
    const i32 lod = g.instLod[hit.inst];
    const i32 vb = g.lodVertOffset[lod], tb = g.lodTriOffset[lod];
    const u32* ix = g.tris + 3 * ((size_t)tb + hit.prim);
    V3 n[3];
    for (int k = 0; k < 3; ++k) {
        const float* p = S.normals + 3 * ((size_t)vb + ix[k]);
        n[k] = {p[0], p[1], p[2]};
    }
    V3 sn = hit.b0 * n[0] + hit.b1 * n[1] + (1.f - (hit.b0 + hit.b1)) * n[2];
    sn = normalize(xform_normal(mat_of(g, hit.inst), sn));
    const V3 tx = normalize(hit.tangentX - sn * dot(hit.tangentX, sn));
    const V3 ty = cross(sn, tx);
    return {sn, hit.geoN, tx, ty};
}

const OrcMaterial* material_of(const OrcPtScene& S, i32 inst, i32 prim) {
    const i32 lod = S.geo.instLod[inst];
    const i32 tb = S.geo.lodTriOffset[lod];
    return S.materials + S.matIdx[(size_t)tb + prim];
}

// ---------------------------------------------------------------- materials
struct PathSample {
    bool valid;
    V3 throughput, excident;
    float pdfF, pdfB;
};

// materials::sample (material_sampling.hpp:77-146) with Lambert (lambert.hpp:18-35)
PathSample mat_sample(const TS& ts, const OrcMaterial& m, V3 incident, float u0, float u1) {
    PathSample r{false, {0, 0, 0}, {0, 0, 0}, 0, 0};
    const float iDotG = -dot(incident, ts.geoN);
    const float iDotN = -dot(incident, ts.shadingN);
    if (iDotG * iDotN <= 0.0f) return r;
    const V3 inTS{-dot(incident, ts.shadingTX), -dot(incident, ts.shadingTY), iDotN};
    if (m.type != kLambert) return r;  // emissive: invalid sample
    // sample_dir_cosine (sampling.hpp:95-101)
    const float cosT = std::sqrt(u0);
    const float sinT = std::sqrt(1.0f - u0);
    const float phi = u1 * 2 * kPi;
    const float sphi = (float)std::sin((double)phi), cphi = (float)std::cos((double)phi);
    const V3 exTS0{sinT * sphi, sinT * cphi, cosT};
    const V3 exTS = exTS0 * sgnf(inTS.z);
    r.throughput = {m.albedo[0], m.albedo[1], m.albedo[2]};
    r.pdfF = exTS0.z / kPi;
    r.pdfB = std::fabs(inTS.z) / kPi;
    if (r.throughput.x == 0.0f && r.throughput.y == 0.0f && r.throughput.z == 0.0f) return r;
    V3 g = exTS.x * ts.shadingTX + exTS.y * ts.shadingTY + exTS.z * ts.shadingN;
    g = normalize(g);
    const float eDotG = dot(g, ts.geoN);
    const float eDotN = exTS.z;
    if (eDotG * eDotN <= 0.0f) return r;
    r.excident = g;
    r.pdfF = mn(r.pdfF, 1e9f);
    r.pdfB = mn(r.pdfB, 1e9f);
    r.valid = true;
    return r;
}

struct EvalValue {
    V3 value;
    float cosOut, pdfF, pdfB;
};

// materials::evaluate (material_sampling.hpp:160-237) with Lambert (lambert.hpp:37-49)
EvalValue mat_evaluate(const TS& ts, const OrcMaterial& m, V3 incident, V3 excident) {
    EvalValue z{{0, 0, 0}, 0, 0, 0};
    const float iDotN = -dot(incident, ts.shadingN);
    const float eDotN = dot(excident, ts.shadingN);
    // flags: LAMBERT is REFLECTIVE only, EMISSIVE neither
    const bool reflective = m.type == kLambert;
    if (!reflective && iDotN * eDotN > 0.0f) return z;
    if (iDotN * eDotN < 0.0f) return z;  // not refractive
    const float iDotG = -dot(incident, ts.geoN);
    const float eDotG = dot(excident, ts.geoN);
    if (eDotG * eDotN <= 0.0f || iDotG * iDotN <= 0.0f) return z;
    const V3 inTS{-dot(incident, ts.shadingTX), -dot(incident, ts.shadingTY), iDotN};
    const V3 exTS{dot(excident, ts.shadingTX), dot(excident, ts.shadingTY), eDotN};
    if (m.type != kLambert) return z;
    if (inTS.z * exTS.z < 0.0f) return z;
    EvalValue r;
    r.value = {m.albedo[0] / kPi, m.albedo[1] / kPi, m.albedo[2] / kPi};
    r.pdfF = std::fabs(exTS.z) / kPi;
    r.pdfB = std::fabs(inTS.z) / kPi;
    if (r.value.x == 0.0f && r.value.y == 0.0f && r.value.z == 0.0f) return z;
    r.pdfF = mn(r.pdfF, 1e9f);
    r.pdfB = mn(r.pdfB, 1e9f);
    r.cosOut = std::fabs(eDotN);
    return r;
}

// emission of a material (emissive.hpp:39-44); (0,0,0) for others
V3 mat_emission(const OrcMaterial& m, V3 geoN, V3 excident, float* pdf) {
    if (m.type != kEmissive) return {0, 0, 0};
    const float cosOut = dot(geoN, excident);
    if (cosOut <= 0.0f) return {0, 0, 0};
    if (pdf) *pdf = cosOut / kPi;
    return {m.radiance[0] * m.scale, m.radiance[1] * m.scale, m.radiance[2] * m.scale};
}

// ---------------------------------------------------------------- lights
struct Nee {
    V3 position;
    float creationPdf;
    V3 dir;
    float dirPdf;
    float cosOut;
    V3 diffIrradiance;
    float dist;
    V3 geoNormal;
    float distSq;
};

struct TriLight {  // AreaLightTriangle record
    float posV[3][3];
    uint16_t material;
    uint16_t pad;
    float uvV[3][2];
};
struct PointLightRec {
    float position[3];
    uint16_t medium, pad;
    float intensity[3];
    float pad1;
};
struct NodeRec {
    float leftFlux;
    uint32_t leftOffset;
    uint16_t leftType, rightType;
    uint32_t rightOffset;
    float rightFlux;
    float center[3];
};

Nee connect_light(const OrcPtScene& S, uint16_t type, const uint8_t* light, V3 pos, float u0, float u1) {
    Nee n;
    if (type == 0) {  // point (light_sampling.hpp:308-327)
        PointLightRec L;
        std::memcpy(&L, light, sizeof(L));
        const V3 lp{L.position[0], L.position[1], L.position[2]};
        V3 d = lp - pos;
        const float distSq = dot(d, d) + kDistSqEps;
        const float dist = std::sqrt(distSq);
        d = d / dist;
        n.position = lp;
        n.creationPdf = kInfPdf;
        n.dir = d;
        n.dirPdf = 1.0f / (4 * kPi);
        n.cosOut = 0.0f;
        n.diffIrradiance = {L.intensity[0] / distSq, L.intensity[1] / distSq, L.intensity[2] / distSq};
        n.dist = dist;
        n.geoNormal = {0, 0, 0};
        n.distSq = distSq;
        return n;
    }
    // area triangle (light_sampling.hpp:118-140, 350-375)
    TriLight L;
    std::memcpy(&L, light, sizeof(L));
    const V3 p0{L.posV[0][0], L.posV[0][1], L.posV[0][2]};
    const V3 e1{L.posV[1][0], L.posV[1][1], L.posV[1][2]};
    const V3 e2{L.posV[2][0], L.posV[2][1], L.posV[2][2]};
    V3 normal = cross(e1, e2);
    const float area2 = len(normal);
    normal = normal / area2;
    const float area = area2 * 0.5f;
    const float s = 1.f - std::sqrt(1.f - u0);
    const float t = (1.f - s) * u1;
    const V3 position = p0 + e1 * s + e2 * t;
    const OrcMaterial& m = S.materials[L.material];
    const V3 rad = mat_emission(m, {0, 0, 1}, {0, 0, 1}, nullptr);
    const float pdf = 1.0f / area;
    const V3 intensity = rad * area;
    V3 d = position - pos;
    const float distSq = dot(d, d) + kDistSqEps;
    const float dist = std::sqrt(distSq);
    d = d / dist;
    const float cosOut = dot(normal, -d);
    V3 value{0, 0, 0};
    float vc = 0.0f, vpdf = 0.0f;
    if (cosOut > 0.0f) {
        value = intensity;
        vc = cosOut;
        vpdf = cosOut / kPi;
    }
    n.position = position;
    n.creationPdf = pdf;
    n.dir = d;
    n.dirPdf = vpdf;
    n.cosOut = vc;
    n.diffIrradiance = {value.x / distSq, value.y / distSq, value.z / distSq};
    n.dist = dist;
    n.geoNormal = normal;
    n.distSq = distSq;
    return n;
}

// lights::connect (light_tree_sampling.hpp:195-329), positional lights only
Nee connect(const OrcPtScene& S, u64 index, u64 numIndices, u64 seed, V3 pos, float u0, float u1) {
    u64 rndChoice = numIndices > 0 ? seed + index * (~0ull / numIndices) : seed;
    if (rndChoice == ~0ull) --rndChoice;
    const float fluxSum = S.bgFlux + S.dirRootFlux + S.posRootFlux;
    const float envProb = S.bgFlux / fluxSum;
    const float dirProb = S.dirRootFlux / fluxSum;
    const float posProb = S.posRootFlux / fluxSum;
    const u64 rightEnv = percentage_of(~0ull, envProb);
    u64 right = percentage_of(~0ull, envProb + dirProb);
    u64 left = rightEnv;
    (void)rightEnv;
    // (env and directional subtrees are empty in the supported scenes)
    left = right;
    right = ~0ull;
    float lightProb = posProb;
    uint16_t type = (uint16_t)S.posRootType;
    uint32_t offset = 0;
    while (type == kInternalNode) {
        NodeRec node;
        std::memcpy(&node, S.lightMem + offset, sizeof(node));
        const float probLeft = node.leftFlux / (node.leftFlux + node.rightFlux);
        const u64 boundary = left + percentage_of(right - left, probLeft);
        if (rndChoice < boundary) {
            type = node.leftType;
            offset = node.leftOffset;
            right = boundary;
            lightProb *= probLeft;
        } else {
            type = node.rightType;
            offset = node.rightOffset;
            left = boundary;
            lightProb *= (1.0f - probLeft);
        }
    }
    Nee n = connect_light(S, type, S.lightMem + offset, pos, u0, u1);
    n.creationPdf *= lightProb;  // adjustPdf (light_tree_sampling.hpp:16-19)
    n.diffIrradiance = n.diffIrradiance / lightProb;
    return n;
}

// light_pdf (light_tree_sampling.hpp:340-390) -> connect pdf of a hit emitter
float light_pdf(const OrcPtScene& S, i32 inst, i32 prim) {
    const float treeFlux = S.dirRootFlux + S.posRootFlux + S.bgFlux;
    float p = S.posRootFlux / treeFlux;
    const u64 key = ((u64)(u32)inst << 32) | (u64)(u32)prim;
    i32 lo = 0, hi = S.mapCount - 1, found = -1;
    while (lo <= hi) {
        const i32 mid = (lo + hi) / 2;
        if (S.mapKeys[mid] == key) {
            found = mid;
            break;
        }
        if (S.mapKeys[mid] < key)
            lo = mid + 1;
        else
            hi = mid - 1;
    }
    if (found < 0) return 0.0f;
    u32 code = S.mapCodes[found];
    uint32_t offset = 0;
    uint16_t type = (uint16_t)S.posRootType;
    while (type == kInternalNode) {
        NodeRec node;
        std::memcpy(&node, S.lightMem + offset, sizeof(node));
        const float pLeft = node.leftFlux / (node.leftFlux + node.rightFlux);
        if (code & 0x80000000u) {
            p *= (1.0f - pLeft);
            type = node.rightType;
            offset = node.rightOffset;
        } else {
            p *= pLeft;
            type = node.leftType;
            offset = node.leftOffset;
        }
        code <<= 1;
    }
    TriLight L;
    std::memcpy(&L, S.lightMem + offset, sizeof(L));
    const V3 e1{L.posV[1][0], L.posV[1][1], L.posV[1][2]};
    const V3 e2{L.posV[2][0], L.posV[2][1], L.posV[2][2]};
    const float area = len(cross(e1, e2)) / 2.0f;
    return p / area;
}

// ---------------------------------------------------------------- accumulation
struct PixelAcc {
    float rad[3] = {0, 0, 0};
    float hit[2] = {0, 0};
    u32 nans = 0;
    void add(float* dst, const float* v, int c) {
        for (int k = 0; k < c; ++k) {
            float x = v[k];
            if (std::isnan(x)) {
                ++nans;
                x = 0.0f;
            }
            dst[k] += x;
        }
    }
};

// pt_sample (pt_common.hpp:75-156)
void pt_sample(const OrcPtScene& S, const int params[4], int px, int py, Pcg64& rng, PixelAcc& acc) {
    const int minPL = params[1], maxPL = params[2], neeCount = params[3];
    V3 throughput{1.0f, 1.0f, 1.0f};
    Vertex vtx;
    vtx.position = {S.camPos[0], S.camPos[1], S.camPos[2]};
    vtx.incidentPdf = 1.0f;  // mis_start_pdf(inf, inf, 1)
    rng.next();              // create_camera's RndSet2 (unused by the pinhole)
    const V3 viewDir{S.camDir[0], S.camDir[1], S.camDir[2]};
    const V3 up{S.camUp[0], S.camUp[1], S.camUp[2]};
    int pathLen = 0;
    do {
        if (pathLen > 0 && pathLen + 1 >= minPL && pathLen + 1 <= maxPL) {
            const u64 neeSeed = rng.next();
            for (int i = 0; i < neeCount; ++i) {
                const u64 x = rng.next();
                const float nu0 = sample_uniform((u32)x), nu1 = sample_uniform((u32)(x >> 32));
                const Nee nee = connect(S, (u64)i, (u64)neeCount, neeSeed, vtx.position, nu0, nu1);
                EvalValue value = mat_evaluate(vtx.ts, *vtx.mat, vtx.incident_dir(), nee.dir);
                if (nee.cosOut != 0) value.cosOut *= nee.cosOut;
                const V3 radiance = value.value * nee.diffIrradiance;
                if ((radiance.x > 0.0f || radiance.y > 0.0f || radiance.z > 0.0f) && value.cosOut > 0.0f) {
                    const bool occluded =
                        any_intersection(S, vtx.position, nee.position, vtx.ts.geoN, nee.geoNormal, nee.dir);
                    if (!occluded) {
                        const float hitPdf = value.pdfF * std::fabs(nee.cosOut) / nee.distSq;
                        const float mis = 1.0f / ((float)neeCount + hitPdf / nee.creationPdf);
                        const V3 c = throughput * value.cosOut * radiance * mis;
                        const float cv[3] = {c.x, c.y, c.z};
                        acc.add(acc.rad, cv, 3);
                    }
                }
            }
        }
        // walk (random_walk.hpp:38-112)
        const u64 x0 = rng.next();
        const u64 x1 = rng.next();
        (void)x1;
        const float u0 = sample_uniform((u32)x0), u1 = sample_uniform((u32)(x0 >> 32));
        const float rndRoulette = sample_uniform((u32)rng.next());
        V3 excident, origin = vtx.position;
        float pdfF;
        V3 sThroughput;
        if (vtx.isCamera) {  // pinholecam_sample_ray (pinhole.hpp:56-84)
            const float w = (float)S.width, h = (float)S.height;
            float cx = ((float)px + u0) / w * 2.0f - 1.0f;
            float cy = ((float)py + u1) / h * 2.0f - 1.0f;
            cx *= S.tanVFov;
            cy *= S.tanVFov;
            const float aspect = w / h;
            const V3 xAxis = cross(viewDir, up);
            excident = normalize(xAxis * (cx * aspect) + up * cy + viewDir * 1.0f);
            const float pixelArea = (2 * S.tanVFov) * (2 * S.tanVFov) * aspect;
            const float cosOut = dot(viewDir, excident);
            pdfF = 1.0f / (pixelArea * cosOut * cosOut * cosOut);
            sThroughput = {1.0f, 1.0f, 1.0f};
        } else {
            const PathSample ps = mat_sample(vtx.ts, *vtx.mat, vtx.incident_dir(), u0, u1);
            if (!ps.valid) break;  // CANCEL
            excident = ps.excident;
            pdfF = ps.pdfF;
            sThroughput = ps.throughput;
        }
        throughput = throughput * sThroughput;
        const float cont = mn(mx(mx(throughput.x, throughput.y), throughput.z) + 0.1f, 1.0f);
        if (rndRoulette >= cont) break;  // CANCEL
        throughput = throughput / cont;
        V3 rayOrigin = origin;
        const HitInfo hit = first_intersection(S, rayOrigin, excident, vtx.geo_normal(), kMaxSceneSize);
        Vertex next;
        next.isCamera = false;
        if (!hit.hit) {
            next.isVoid = true;
            next.incEnc = excident;
            next.incidentPdf = pdfF * 1.0f;
        } else {
            next.position = rayOrigin + excident * hit.t;
            next.ts = tangent_space(S, hit);
            next.incEnc = {excident.x, excident.y, hit.t * sgnf(excident.z)};
            next.inst = hit.inst;
            next.prim = hit.prim;
            next.sp0 = hit.b0;
            next.sp1 = hit.b1;
            next.mat = material_of(S, hit.inst, hit.prim);
            const float inCosAbs = std::fabs(dot(excident, next.ts.shadingN));
            next.incidentPdf = pdfF * std::fabs(inCosAbs) / (hit.t * hit.t);
        }
        vtx = next;
        if (pathLen == 0) {
            const float hv[2] = {(float)vtx.inst, (float)vtx.prim};
            acc.add(acc.hit, hv, 2);
        }
        ++pathLen;
        if (pathLen >= minPL) {
            V3 emission{0, 0, 0};
            if (!vtx.isVoid) {
                const V3 inc = vtx.incident_dir();
                emission = mat_emission(*vtx.mat, vtx.ts.geoN, -inc, nullptr);
                if (emission.x != 0.0f || emission.y != 0.0f || emission.z != 0.0f) {
                    const float connectPdf = light_pdf(S, vtx.inst, vtx.prim);
                    if (pathLen > 1) {
                        const float mis = 1.0f / (1.0f + (float)neeCount * (connectPdf / vtx.incidentPdf));
                        emission = emission * mis;
                    }
                }
            }
            const V3 c = throughput * emission;
            const float cv[3] = {c.x, c.y, c.z};
            acc.add(acc.rad, cv, 3);
        }
        if (vtx.isVoid) break;
    } while (pathLen < maxPL);
}

}  // namespace

extern "C" {

// Runs `iterations` iterations from a reset (fresh per-pixel RNGs) and returns
// the running means of the Radiance (3 floats / pixel) and Hit ID (2 floats /
// pixel) targets, row-major, plus the number of NaN contributions dropped.
int orc_pt_render(const OrcPtScene* S, int32_t seed, int32_t minPathLength, int32_t maxPathLength, int32_t neeCount,
                  int32_t iterations, float* outRadiance, float* outHitId, uint32_t* outNans) {
    const int W = S->width, H = S->height;
    const int num = W * H;
    std::vector<Pcg64> rngs;
    rngs.reserve(num);
    const int s = seed * (num + 1);  // cpu_pt.cpp:40-45
    for (int i = 0; i < num; ++i) rngs.emplace_back((u32)(i + s));
    std::memset(outRadiance, 0, sizeof(float) * 3 * (size_t)num);
    if (outHitId) std::memset(outHitId, 0, sizeof(float) * 2 * (size_t)num);
    const int params[4] = {seed, minPathLength, maxPathLength, neeCount};
    u32 nans = 0;
    for (int it = 0; it < iterations; ++it) {
#pragma omp parallel for schedule(dynamic, 256) reduction(+ : nans)
        for (int pixel = 0; pixel < num; ++pixel) {
            PixelAcc acc;
            pt_sample(*S, params, pixel % W, pixel / W, rngs[pixel], acc);
            nans += acc.nans;
            const float denom = mx(1.0f, (float)it + 1);
            for (int c = 0; c < 3; ++c) {
                float& cum = outRadiance[3 * (size_t)pixel + c];
                cum += (acc.rad[c] - cum) / denom;
            }
            if (outHitId)
                for (int c = 0; c < 2; ++c) {
                    float& cum = outHitId[2 * (size_t)pixel + c];
                    cum += (acc.hit[c] - cum) / denom;
                }
        }
    }
    if (outNans) *outNans = nans;
    return 0;
}

}  // extern "C"
