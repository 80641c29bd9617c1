// Oracle — test infrastructure only.
//
// CPU restatement of Mufflon's pure path tracer iteration for the supported
// scene content (triangles; EMISSIVE, LAMBERT, TORRANCE and MICROFACET
// materials with constant values or textures; homogeneous media; area-triangle
// and point lights through the light tree; COLORED or polar ENVMAP
// background; pinhole camera), following
//   pt_common.hpp:75-156       pt_sample (RNG consumption order, NEE, MIS)
//   random_walk.hpp:38-112     walk (sample, Russian roulette, first hit,
//                              medium transmission)
//   path_util.hpp:117-748      PathVertex (incident encoding, evaluate, sample,
//                              get_emission, create_surface / create_void)
//   pt_common.hpp:27-68        PtVertexExt (incidentPdf for MIS)
//   pinhole.hpp:43-84          pinholecam_sample_ray
//   material_sampling.hpp:33-237 fetch / sample / evaluate wrappers
//   lambert.hpp, emissive.hpp, microfacet_base.hpp, microfacet_specular.hpp
//   (TORRANCE), microfacet_full.hpp (MICROFACET), blend_fresnel.hpp:16-23
//   cputexture.cpp:372-411     CPU texture sampling (nearest / linear)
//   light_tree_sampling.hpp:195-329 (connect), :340-425 (light_pdf,
//                              background_pdf)
//   light_sampling.hpp:118-140,225-303,308-375,447-520 (area triangle, point,
//                              background connect_light, evaluate_background)
//   texture_sampling.hpp:113-155 (importance_sample_texture)
//   textures/interface.hpp:180-200 (direction_to_uv, sample_polar_map)
//   intersection.cu:366-621 (hit attributes), tangent_space.hpp:10-60
//   rng.hpp:12-93, sampling.hpp:23-127,175-200 (PCG64, samplers,
//                              percentage_of, rescale_sample)
//   render_target.hpp:62-76,189-195 (contribute, NaN check)
//   output_handler_variance.hpp:7-31 (running mean), cpu_pt.cpp:18-45 (seeds)
// Every float expression keeps the reference's operand order; the build uses
// -ffp-contract=off. Functions of the absent Epsilon library and libm calls
// have this project's definitions, shared with the device code:
//   normalize(v) = v / len(v); sgn(x) = x < 0 ? -1 : 1; min/max are the
//   compare-selects mn / mx; sq(x) = x * x; sdiv(x, d) = x / (d + 1e-20 sgn d);
//   floor(float) -> int; mod(a, m) is the non-negative remainder;
//   sin, cos, acos, atan2, exp, log and pow of a float are evaluated in double
//   and rounded once; float -> u64 conversion maps NaN and negatives to 0.
// Media are looked up by value: each material carries its outer medium and
// (MICROFACET) its inner medium, the camera and point lights sit in vacuum.
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
constexpr u64 kU64Max = ~0ull;

// Planted estimator mistakes, for the closed-form estimator tests only
// (orc_set_mutation): each test switches one on and checks that some
// closed-form comparison then fails. 0 = the restatement as written.
enum Mutation : int {
    kMutNone = 0,
    kMutDropLightCos = 1,      // NEE to an area light without the cosine at the light
    kMutNeeMisOne = 2,         // NEE MIS denominator with 1 instead of the NEE count
    kMutHitMisNoCount = 3,     // hit MIS without the NEE count
    kMutLightPdfNoTree = 4,    // light_pdf without the light-tree choice probability
    kMutNoRouletteDiv = 5,     // Russian roulette survivors not divided by the continuation probability
    kMutPointNoFalloff = 6,    // point light without the 1/d^2 falloff
    kMutBackgroundNoPdf = 7,   // background NEE without the division by the direction pdf
    kMutNeeMisNoLightCos = 8,  // NEE MIS hit pdf without the cosine at the light
    kMutIncidentSolidAngle = 9, // hit MIS with the incident pdf left in solid angle
    kMutLtNoChoiceProb = 10,    // light tracer: emission not divided by the light's choice probability
    kMutLtNoShadingCorr = 11,   // light tracer: adjoint BSDF without the shading normal correction
    kMutLtNoEtaSq = 12,         // light tracer: adjoint refraction without eta^2
    kMutLtImportanceNoCos3 = 13,// light tracer: camera importance without 1 / cos^3
    kMutLtAreaNoCos = 14        // light tracer: area light seen by the camera without its cosine
};
int g_mutation = kMutNone;

// Work counts of the paths a render traces, for the path tracer's
// algorithmic-bytes figure (orc_pt_path_stats): set per thread while a
// counted path runs, null otherwise.
struct PathStats {
    u64 extRays = 0, extNodes = 0, extTris = 0;           // closest-hit rays and their traversal work
    u64 shadowRays = 0, shadowNodes = 0, shadowTris = 0;  // any-hit rays
    u64 surfaceVertices = 0;                              // hit attributes + material fetches
    u64 texelTaps = 0;                                    // texture lookups (1x1 values included)
    u64 treeNodes = 0;                                    // light-tree node records read
    u64 lightRecords = 0;                                 // point / triangle light records read
    u64 paths = 0;
};
static_assert(sizeof(PathStats) == 11 * sizeof(u64), "PathStats is read as 11 counters");
thread_local PathStats* tl_stats = nullptr;

inline float sgnf(float x) { return x < 0.0f ? -1.0f : 1.0f; }
inline float clampf(float v, float lo, float hi) { return mn(mx(v, lo), hi); }
inline float sq(float x) { return x * x; }
inline V3 normalize(V3 v) { return v / len(v); }
inline float sdiv(float x, float d) { return x / (d + 1e-20f * sgnf(d)); }
inline bool nonzero(V3 v) { return v.x != 0.0f || v.y != 0.0f || v.z != 0.0f; }
inline float fsin(float x) { return (float)std::sin((double)x); }
inline float fcos(float x) { return (float)std::cos((double)x); }
inline float facos(float x) { return (float)std::acos((double)x); }
inline float fatan2(float y, float x) { return (float)std::atan2((double)y, (double)x); }
inline float fexp(float x) { return (float)std::exp((double)x); }
inline float flog(float x) { return (float)std::log((double)x); }
inline float fpow(float x, float y) { return (float)std::pow((double)x, (double)y); }
inline int ifloor(float x) { return (int)std::floor(x); }
inline int imod(int a, int m) {
    const int r = a % m;
    return r < 0 ? r + m : r;
}
inline u64 f2u64(float v) {
    if (!(v > 0.0f)) return 0;  // NaN, negative, zero
    if (v >= 18446744073709551616.0f) return kU64Max;
    return (u64)v;
}
inline u64 d2u64(double v) {
    if (!(v > 0.0)) return 0;
    if (v >= 18446744073709551616.0) return kU64Max;
    return (u64)v;
}

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
// sampling.hpp:39-46
inline u64 percentage_of(u64 num, float p) {
    const u64 pfix = f2u64(p * 2147483648.0f);
    return ((pfix * (num & 0x7fffffff)) >> 31) + (pfix * (num >> 31));
}
// sampling.hpp:72-85 (an empty interval maps to its left end)
inline u64 rescale_sample(u64 x, u64 pLeft, u64 pRight) {
    const u64 interval = pRight - pLeft;
    const u64 xrel = x - pLeft;
    if (interval == 0) return 0;
    return xrel * (kU64Max / interval) + d2u64((double)xrel * (double)(kU64Max % interval) / (double)interval);
}

}  // namespace

extern "C" {
struct OrcMaterial {  // mfh_material, 128 bytes
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
struct OrcTexture {
    int32_t width, height, channels, sampling;  // sampling: 0 nearest, 1 linear
    const float* texels;                        // 4 floats per texel (RGBA), row-major
};
struct OrcBackground {
    int32_t type;  // 0 colored, 1 envmap
    float color[3];
    float scale[3];
    int32_t envmapTexture;
    const float* sat;  // width * height (envmap)
    float flux[3];
};
struct OrcPtScene {
    OrcScene geo;
    const float* normals;     // 3 per vertex, concatenated like geo.verts
    const float* uvs;         // 2 per vertex
    const uint16_t* matIdx;   // 1 per primitive, concatenated like geo.primIds
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
    const OrcTexture* textures;
    int32_t numTextures;
    OrcBackground background;
    int32_t posGuide;  // "NEE position guide" (PNeePositionGuide, parameter.hpp:359-364)
};
}

namespace {

enum Mat : int32_t { kEmissive = 0, kLambert = 1, kTorrance = 3, kWalter = 4, kMicrofacet = 10 };
enum Ndf : int32_t { kGgx = 0, kBeckmann = 1 };
enum Shadowing : int32_t { kVCavity = 0, kSmith = 1 };

struct TS {  // TangentSpace
    V3 shadingN{0, 0, 0}, geoN{0, 0, 0}, shadingTX{0, 0, 0}, shadingTY{0, 0, 0};
};

struct Medium {
    float ior;
    V3 absorption;
};
const Medium kVacuum{1.0f, {0.0f, 0.0f, 0.0f}};
// Medium::get_transmission (medium.hpp:40-42)
inline V3 transmission(const Medium& m, float distance) {
    const V3 e = m.absorption * -distance;
    return {fexp(e.x), fexp(e.y), fexp(e.z)};
}

// ParameterPack of a fetched material (material_sampling.hpp:33-53)
struct MatParams {
    int32_t type = -1;
    V3 albedo{0, 0, 0};
    V3 radiance{0, 0, 0};
    float rx = 0, ry = 0;
    int32_t ndf = kGgx, shadowing = kVCavity;
    Medium inner = kVacuum, outer = kVacuum;
    // MaterialDescriptorBase::get_medium (material_definitions.hpp:53-56)
    const Medium& medium(float xDotN) const { return xDotN < 0.0f ? inner : outer; }
    bool reflective() const { return type == kLambert || type == kTorrance || type == kWalter || type == kMicrofacet; }
    bool refractive() const { return type == kWalter || type == kMicrofacet; }
    bool halfvector() const { return type == kTorrance || type == kWalter || type == kMicrofacet; }
};

// ---------------------------------------------------------------- textures
struct F4 {
    float x, y, z, w;
};
inline F4 texel(const OrcTexture& t, int x, int y) {
    const int ix = imod(x, t.width), iy = imod(y, t.height);
    const float* p = t.texels + 4 * ((size_t)iy * t.width + ix);
    return {p[0], p[1], p[2], p[3]};
}
// CpuTexture::sample_nearest / sample_linear / sample111_* (cputexture.cpp:372-411)
F4 tex_sample(const OrcTexture& t, float u, float v) {
    if (tl_stats) ++tl_stats->texelTaps;
    if (t.width == 1 && t.height == 1) return texel(t, 0, 0);
    if (t.sampling == 0) return texel(t, ifloor(u * (float)t.width), ifloor(v * (float)t.height));
    float fx = u * (float)t.width, fy = v * (float)t.height;
    const int ix = ifloor(fx), iy = ifloor(fy);
    fx -= (float)ix;
    fy -= (float)iy;
    const F4 a = texel(t, ix, iy), b = texel(t, ix + 1, iy), c = texel(t, ix, iy + 1), d = texel(t, ix + 1, iy + 1);
    F4 s;
    s.x = a.x * (1.0f - fx) * (1.0f - fy);
    s.y = a.y * (1.0f - fx) * (1.0f - fy);
    s.z = a.z * (1.0f - fx) * (1.0f - fy);
    s.w = a.w * (1.0f - fx) * (1.0f - fy);
    s.x += b.x * fx * (1.0f - fy);
    s.y += b.y * fx * (1.0f - fy);
    s.z += b.z * fx * (1.0f - fy);
    s.w += b.w * fx * (1.0f - fy);
    s.x += c.x * (1.0f - fx) * fy;
    s.y += c.y * (1.0f - fx) * fy;
    s.z += c.z * (1.0f - fx) * fy;
    s.w += c.w * (1.0f - fx) * fy;
    s.x += d.x * fx * fy;
    s.y += d.y * fx * fy;
    s.z += d.z * fx * fy;
    s.w += d.w * fx * fy;
    return s;
}

// fetch (material_sampling.hpp:33-53) with the per-type fetch of lambert.hpp:10-15,
// emissive.hpp:9-14, microfacet_specular.hpp:11-22, microfacet_full.hpp:12-25
MatParams fetch(const OrcPtScene& S, const OrcMaterial& m, float u, float v) {
    MatParams p;
    p.type = m.type;
    p.outer = {m.outerIor, {m.outerAbsorption[0], m.outerAbsorption[1], m.outerAbsorption[2]}};
    // Material::compute_medium: WALTER and MICROFACET carry their inner medium
    // (material_definitions.hpp:226-228, 269-271)
    p.inner = (m.type == kWalter || m.type == kMicrofacet)
                  ? Medium{m.ior, {m.absorption[0], m.absorption[1], m.absorption[2]}}
                  : p.outer;
    if (m.type == kLambert || m.type == kTorrance) {
        if (m.albedoTexture >= 0) {
            const F4 t = tex_sample(S.textures[m.albedoTexture], u, v);
            p.albedo = {t.x, t.y, t.z};
        } else {
            p.albedo = {m.albedo[0], m.albedo[1], m.albedo[2]};
        }
    }
    if (m.type == kTorrance || m.type == kWalter || m.type == kMicrofacet) {
        if (m.roughnessTexture >= 0) {
            const OrcTexture& tx = S.textures[m.roughnessTexture];
            const F4 t = tex_sample(tx, u, v);
            p.rx = t.x;
            p.ry = tx.channels > 1 ? t.y : t.x;
        } else {
            p.rx = m.roughness[0];
            p.ry = m.roughness[1];
        }
        p.ndf = m.ndf;
        p.shadowing = m.shadowing;
    }
    if (m.type == kEmissive)
        p.radiance = {m.radiance[0] * m.scale[0], m.radiance[1] * m.scale[1], m.radiance[2] * m.scale[2]};
    return p;
}

// ---------------------------------------------------------------- microfacets
// compute_half_vector (microfacet_base.hpp:15-27)
V3 half_vector(float nIn, float nEx, V3 in, V3 ex) {
    const V3 h = nIn * in + nEx * ex;
    const float l = len(h) * sgnf(h.z);
    return {sdiv(h.x, l), sdiv(h.y, l), sdiv(h.z, l)};
}
// Boundary::get_halfTS (microfacet_base.hpp:45-52)
V3 boundary_half(const Medium& incident, const Medium& other, V3 in, V3 ex) {
    const Medium& exM = in.z * ex.z > 0.0f ? incident : other;
    return half_vector(incident.ior, exM.ior, in, ex);
}

// microfacet_base.hpp:73-80
float g_vcavity(float wDotH, float wDotN, float hDotN) {
    if (wDotH * wDotN <= 0.0f) return 0.0f;
    return mn(1.0f, 2.0f * std::fabs(sdiv(hDotN * wDotN, wDotH)));
}
// microfacet_base.hpp:83-102 (GGX)
float g_smith(float wDotH, V3 w, float rx, float ry) {
    if (wDotH * w.z <= 0.0f) return 0.0f;
    const float oneOverASquared = sdiv(sq(w.x * rx) + sq(w.y * ry), w.z * w.z);
    return 2.0f / (1.0f + std::sqrt(1.f + oneOverASquared));
}
// microfacet_base.hpp:113-120
float modify_g(float g) {
    return g < 0.11f ? fpow(g + 0.74299714456847f, 8.0f) - 0.092874643071059f
                     : ((1.139400367544197f * g - 3.317658394008854f) * g + 3.317115685385115f) * g -
                           0.13885765892046f;
}
float g_reflection(int shadowing, float gi, float ge) {
    return shadowing == kSmith ? gi * ge : modify_g(mn(gi, ge));
}
float g_transmission(int shadowing, float gi, float ge) {
    return shadowing == kSmith ? gi * ge : mx(0.0f, modify_g(gi) + modify_g(ge) - 1.0f);
}
float g_term(const MatParams& p, float wDotH, V3 w, float hDotN) {
    return p.shadowing == kSmith ? g_smith(wDotH, w, p.rx, p.ry) : g_vcavity(wDotH, w.z, hDotN);
}

// eval_ndf (microfacet_base.hpp:281-310)
float eval_ndf(const MatParams& p, V3 h) {
    if (p.ndf == kGgx) {
        const float norm = mx(kPi * p.rx * p.ry, 1e-30f);
        const float cosThetaSq = h.z * h.z;
        const float rx = sdiv(h.x, p.rx);
        const float ry = sdiv(h.y, p.ry);
        const float tmp = rx * rx + ry * ry + cosThetaSq;
        return sdiv(1.0f, norm * tmp * tmp);
    }
    const float norm = kPi * p.rx * p.ry;
    const float rx = sdiv(h.x, p.rx);
    const float ry = sdiv(h.y, p.ry);
    const float nDotHSq = h.z * h.z;
    const float e = fexp(-(rx * rx + ry * ry) / nDotHSq);
    return sdiv(e, norm * nDotHSq * nDotHSq);
}

struct DirSample {
    V3 dir;
    float pdf;
};
// sample_ndf (microfacet_base.hpp:312-366)
DirSample sample_ndf(const MatParams& p, float u0, float u1) {
    if (p.ndf == kGgx) {
        const float phi = 2.0f * kPi * u0;
        const float e = std::sqrt(u1 / (1.0f - u1));
        float slopeX = e * fcos(phi);
        float slopeY = e * fsin(phi);
        const float norm = kPi * p.rx * p.ry;
        const float tmp = 1.0f + slopeX * slopeX + slopeY * slopeY;
        slopeX *= p.rx;
        slopeY *= p.ry;
        const V3 dir = normalize({-slopeX, -slopeY, 1.0f});
        return {dir, 1.0f / mx(norm * tmp * tmp * dir.z * dir.z * dir.z, 1e-20f)};
    }
    const float phi = 2 * kPi * u0;
    const float xi = u1 + 1e-20f;
    const float e = std::sqrt(-flog(xi));
    const float slopeX = e * fcos(phi);
    const float slopeY = e * fsin(phi);
    const V3 dir = normalize({-p.rx * slopeX, -p.ry * slopeY, 1.0f});
    return {dir, xi / mx(kPi * p.rx * p.ry * dir.z * dir.z * dir.z, 1e-20f)};
}

// sample_visible_normal_vcavity (microfacet_base.hpp:137-156)
V3 visible_normal_vcavity(V3 in, V3 cav, u64& rnd, float& cosI) {
    const V3 adj{-cav.x, -cav.y, cav.z};
    const float iDotHOut = dot(in, cav);
    const float iDotHadj = dot(in, adj);
    const float t = in.z > 0.0f ? mx(0.0f, iDotHadj) / (mx(0.0f, iDotHOut) + mx(0.0f, iDotHadj))
                                : mx(0.0f, -iDotHadj) / (mx(0.0f, -iDotHOut) + mx(0.0f, -iDotHadj));
    const u64 probAdj = percentage_of(kU64Max, t);
    if (rnd < probAdj) {
        rnd = rescale_sample(rnd, 0, probAdj - 1);
        cosI = iDotHadj;
        return adj;
    }
    rnd = rescale_sample(rnd, probAdj, kU64Max);
    cosI = iDotHOut;
    return cav;
}

// sample_visible_slopes_ggx (microfacet_base.hpp:213-244)
void visible_slopes_ggx(float tanTheta, float u0, float u1, u64& rnd2, float& sx, float& sy) {
    if (tanTheta < 0.0001f) {
        const float r = std::sqrt(sdiv(u0, (1.f - u0)));
        const float phi = 6.28318530718f * u1;
        sx = r * fcos(phi);
        sy = r * fsin(phi);
        return;
    }
    const float a1 = 1.0f / tanTheta;
    const float g1 = 2.f / (1.f + std::sqrt(1.0f + 1.0f / (a1 * a1)));
    const float a = 2.0f * u0 / g1 - 1.0f;
    const float tmp = sdiv(1.0f, a * a - 1.0f);
    const float d = tmp * std::sqrt(tanTheta * tanTheta - (a * a - tanTheta * tanTheta) / tmp);
    const float slopeX1 = tanTheta * tmp - d;
    const float slopeX2 = tanTheta * tmp + d;
    sx = (a < 0.f || slopeX2 > 1.0f / tanTheta) ? slopeX1 : slopeX2;
    float s;
    if (rnd2 & (1ull << 63)) {
        rnd2 *= 2;
        s = 1.0f;
    } else {
        rnd2 *= 2;
        s = -1.0f;
    }
    const float z = sdiv((u1 * (u1 * (u1 * 0.27385f - 0.73369f) + 0.46341f)),
                         (u1 * (u1 * (u1 * 0.093073f + 0.309420f) - 1.000000f) + 0.597999f));
    sy = s * z * std::sqrt(1.0f + sx * sx);
}

// sample_visible_normal_smith (microfacet_base.hpp:247-278), GGX
V3 visible_normal_smith(const MatParams& p, V3 in, float u0, float u1, u64& rnd2) {
    V3 st = in * V3{p.rx, p.ry, 1.0f};
    st = normalize(st);
    if (in.z < 0.0f) st = -st;
    const float sinTheta = std::sqrt((1.0f - st.z) * (1.0f + st.z));
    const float tanTheta = sinTheta / st.z;
    const float cosPhi = sinTheta > 1e-5f ? st.x / sinTheta : 0.0f;
    const float sinPhi = sinTheta > 1e-5f ? st.y / sinTheta : 0.0f;
    float sx, sy;
    visible_slopes_ggx(tanTheta, u0, u1, rnd2, sx, sy);
    const float tmp = cosPhi * sx - sinPhi * sy;
    sy = sinPhi * sx + cosPhi * sy;
    sx = tmp;
    sx *= p.rx;
    sy *= p.ry;
    return normalize({-sx, -sy, 1.0f});
}

struct Refraction {
    float f, cosTAbs;
};
// fresnel_dielectric (blend_fresnel.hpp:16-23)
Refraction fresnel_dielectric(float n_i, float n_t, float cosIAbs) {
    const float eta = sq(n_i / n_t);
    const float cosTAbs = std::sqrt(mx(0.0f, 1.0f - eta * (1.0f - cosIAbs * cosIAbs)));
    const float rParl = sdiv(n_t * cosIAbs - n_i * cosTAbs, n_t * cosIAbs + n_i * cosTAbs);
    const float rPerp = sdiv(n_t * cosTAbs - n_i * cosIAbs, n_t * cosTAbs + n_i * cosIAbs);
    return {0.5f * (rParl * rParl + rPerp * rPerp), cosTAbs};
}

// ---------------------------------------------------------------- BSDF in tangent space
struct TsSample {
    bool valid = false;
    V3 throughput{0, 0, 0};
    V3 excident{0, 0, 0};
    float pdfF = 0, pdfB = 0;
};

// Shared half-vector sampling of TORRANCE and MICROFACET.
V3 sample_half(const MatParams& p, V3 in, float u0, float u1, u64& rnd, float& cavityPdf, float& iDotH) {
    if (p.shadowing == kSmith) {
        const V3 h = visible_normal_smith(p, in, u0, u1, rnd);
        cavityPdf = eval_ndf(p, h);
        cavityPdf *= h.z;
        iDotH = dot(in, h);
        return h;
    }
    const DirSample cav = sample_ndf(p, u0, u1);
    float cosI;
    const V3 h = visible_normal_vcavity(in, cav.dir, rnd, cosI);
    iDotH = cosI;
    cavityPdf = cav.pdf;
    return h;
}

// TORRANCE sample (microfacet_specular.hpp:26-83)
TsSample torrance_sample(const MatParams& p, V3 in, float u0, float u1, u64 i0) {
    TsSample r;
    u64 rnd = i0;
    float cavityPdf, iDotH;
    const V3 h = sample_half(p, in, u0, u1, rnd, cavityPdf, iDotH);
    const V3 ex = (2.0f * iDotH) * h - in;
    float ge, gi, g;
    ge = g_term(p, iDotH, ex, h.z);
    gi = g_term(p, iDotH, in, h.z);
    g = g_reflection(p.shadowing, gi, ge);
    if (ge == 0.0f || gi == 0.0f) return r;
    r.throughput = p.albedo * sdiv(g, gi);
    r.excident = ex;
    r.pdfF = cavityPdf * sdiv(gi, std::fabs(4.0f * in.z * h.z));
    r.pdfB = cavityPdf * sdiv(ge, std::fabs(4.0f * ex.z * h.z));
    r.valid = true;
    return r;
}

// MICROFACET sample (microfacet_full.hpp:35-126)
TsSample microfacet_sample(const MatParams& p, V3 in, const Medium& incident, const Medium& other, float u0,
                           float u1, u64 i0, bool adjoint = false) {
    TsSample r;
    u64 rnd = i0;
    float cavityPdf, iDotH;
    const V3 h = sample_half(p, in, u0, u1, rnd, cavityPdf, iDotH);
    const float n_i = incident.ior;
    const float n_t = other.ior;
    const float eta = n_i / n_t;
    const float iDotHabs = std::fabs(iDotH);
    const Refraction f = fresnel_dielectric(n_i, n_t, iDotHabs);
    const bool reflect = ((float)rnd / (float)(kU64Max - 1)) < f.f;
    float eDotH, eDotHabs;
    V3 ex;
    if (reflect) {
        ex = (2.0f * iDotH) * h - in;
        eDotH = iDotH;
        eDotHabs = iDotHabs;
    } else {
        eDotHabs = f.cosTAbs;
        eDotH = eDotHabs * -sgnf(iDotH);
        ex = (sgnf(iDotH) * (eta * iDotHabs - eDotHabs)) * h - eta * in;
    }
    const float ge = g_term(p, eDotH, ex, h.z);
    const float gi = g_term(p, iDotH, in, h.z);
    if (ge == 0.0f || gi == 0.0f) return r;
    float throughput, pdfF, pdfB;
    if (reflect) {
        const float g = g_reflection(p.shadowing, gi, ge);
        throughput = sdiv(g, gi);
        const float common = cavityPdf * sdiv(f.f, 4.0f * h.z);
        pdfF = common * sdiv(gi, std::fabs(in.z));
        pdfB = common * sdiv(ge, std::fabs(ex.z));
    } else {
        const float g = g_transmission(p.shadowing, gi, ge);
        throughput = sdiv(g, gi);
        if (adjoint && g_mutation != kMutLtNoEtaSq) throughput *= eta * eta;  // light sub-paths (microfacet_full.hpp:114-115)
        const float common =
            cavityPdf * (1.0f - f.f) * sdiv(iDotHabs * eDotHabs, sq(n_i * iDotH + n_t * eDotH) * h.z);
        pdfF = common * sdiv(gi * n_t * n_t, std::fabs(in.z));
        pdfB = common * sdiv(ge * n_i * n_i, std::fabs(ex.z));
    }
    r.throughput = {throughput, throughput, throughput};
    r.excident = ex;
    r.pdfF = pdfF;
    r.pdfB = pdfB;
    r.valid = true;
    return r;
}

// WALTER sample (microfacet_refractive.hpp:26-121): the refracted direction
// about the sampled half vector, or its mirror under total internal
// reflection; no Fresnel choice (the view-path throughput has no eta^2).
TsSample walter_sample(const MatParams& p, V3 in, const Medium& incident, const Medium& other, float u0, float u1,
                       u64 i0, bool adjoint = false) {
    TsSample r;
    u64 rnd = i0;
    float cavityPdf, iDotH;
    const V3 h = sample_half(p, in, u0, u1, rnd, cavityPdf, iDotH);
    const float n_i = incident.ior;
    const float n_e = other.ior;
    const float eta = n_i / n_e;
    const float etaSq = eta * eta;
    const float iDotHabs = std::fabs(iDotH);
    const float t = 1.0f - etaSq * (1.0f - iDotHabs * iDotHabs);  // Snell: cos^2 of the refracted angle
    const bool total = t <= 0.0f;
    V3 ex;
    float eDotH, eDotHabs;
    if (total) {
        ex = (2.0f * iDotH) * h - in;
        eDotH = iDotH;
        eDotHabs = iDotHabs;
    } else {
        eDotHabs = std::sqrt(t);
        eDotH = eDotHabs * -sgnf(iDotH);
        ex = (sgnf(iDotH) * (eta * iDotHabs - eDotHabs)) * h - eta * in;
    }
    const float ge = g_term(p, eDotH, ex, h.z);
    const float gi = g_term(p, iDotH, in, h.z);
    if (ge == 0.0f || gi == 0.0f) return r;
    float throughput, pdfF, pdfB;
    if (total) {
        throughput = sdiv(g_reflection(p.shadowing, gi, ge), gi);
        pdfF = cavityPdf * sdiv(gi, std::fabs(4.0f * in.z * h.z));
        pdfB = cavityPdf * sdiv(ge, std::fabs(4.0f * ex.z * h.z));
    } else {
        throughput = sdiv(g_transmission(p.shadowing, gi, ge), gi);
        if (adjoint && g_mutation != kMutLtNoEtaSq) throughput *= eta * eta;  // light sub-paths (microfacet_refractive.hpp:109-110)
        const float common = cavityPdf * sdiv(iDotHabs * eDotHabs, sq(n_i * iDotH + n_e * eDotH) * h.z);
        pdfF = common * sdiv(gi * n_e * n_e, std::fabs(in.z));
        pdfB = common * sdiv(ge * n_i * n_i, std::fabs(ex.z));
    }
    r.throughput = {throughput, throughput, throughput};
    r.excident = ex;
    r.pdfF = pdfF;
    r.pdfB = pdfB;
    r.valid = true;
    return r;
}

// Lambert sample (lambert.hpp:18-35), sample_dir_cosine (sampling.hpp:95-101)
TsSample lambert_sample(const MatParams& p, V3 in, float u0, float u1) {
    TsSample r;
    const float cosT = std::sqrt(u0);
    const float sinT = std::sqrt(1.0f - u0);
    const float phi = u1 * 2 * kPi;
    const V3 ex0{sinT * fsin(phi), sinT * fcos(phi), cosT};
    r.excident = ex0 * sgnf(in.z);
    r.throughput = p.albedo;
    r.pdfF = ex0.z / kPi;
    r.pdfB = std::fabs(in.z) / kPi;
    r.valid = true;
    return r;
}

TsSample sample_ts(const MatParams& p, V3 inTS, float u0, float u1, u64 i0, bool adjoint = false) {
    const Medium& incident = p.medium(inTS.z);
    const Medium& other = p.medium(-inTS.z);
    switch (p.type) {
        case kLambert: return lambert_sample(p, inTS, u0, u1);
        case kTorrance: return torrance_sample(p, inTS, u0, u1, i0);
        case kWalter: return walter_sample(p, inTS, incident, other, u0, u1, i0, adjoint);
        case kMicrofacet: return microfacet_sample(p, inTS, incident, other, u0, u1, i0, adjoint);
        default: return TsSample{};  // EMISSIVE: not sampleable
    }
}

struct TsEval {
    V3 value{0, 0, 0};
    float pdfF = 0, pdfB = 0;
};

TsEval evaluate_ts(const MatParams& p, V3 in, V3 ex) {
    TsEval r;
    const Medium& incident = p.medium(in.z);
    const Medium& other = p.medium(-in.z);
    if (p.type == kLambert) {  // lambert.hpp:37-49
        if (in.z * ex.z < 0.0f) return r;
        r.value = p.albedo / kPi;
        r.pdfF = std::fabs(ex.z) / kPi;
        r.pdfB = std::fabs(in.z) / kPi;
        return r;
    }
    if (p.type == kTorrance) {  // microfacet_specular.hpp:86-122
        if (in.z * ex.z < 0.0f) return r;
        const V3 h = boundary_half(incident, other, in, ex);
        const float iDotH = dot(in, h);
        const float ge = g_term(p, iDotH, ex, h.z);
        const float gi = g_term(p, iDotH, in, h.z);
        const float g = g_reflection(p.shadowing, gi, ge);
        if (ge == 0.0f || gi == 0.0f) return r;
        const float d = eval_ndf(p, h);
        r.value = p.albedo * sdiv(g * d, std::fabs(4.0f * in.z * ex.z));
        r.pdfF = sdiv(gi * d, std::fabs(4.0f * in.z));
        r.pdfB = sdiv(ge * d, std::fabs(4.0f * ex.z));
        return r;
    }
    if (p.type == kWalter) {  // microfacet_refractive.hpp:124-187
        const bool isReflection = in.z * ex.z > 0.0f;
        const V3 h = boundary_half(incident, other, in, ex);
        const float iDotH = dot(in, h);
        const float eDotH = dot(ex, h);
        const float n_i = incident.ior;
        const float n_e = other.ior;
        // a reflection exists only under total internal reflection
        if (isReflection && sq(n_i / n_e) * (1.0f - iDotH * iDotH) < 1.0f) return r;
        const float ge = g_term(p, eDotH, ex, h.z);
        const float gi = g_term(p, iDotH, in, h.z);
        if (ge == 0.0f || gi == 0.0f) return r;
        const float d = eval_ndf(p, h);
        if (isReflection) {
            const float v = sdiv(g_reflection(p.shadowing, gi, ge) * d, 4.0f * in.z * ex.z);
            r.value = {v, v, v};
            r.pdfF = sdiv(gi * d, 4.0f * std::fabs(in.z));
            r.pdfB = sdiv(ge * d, 4.0f * std::fabs(ex.z));
            return r;
        }
        const float common = sdiv(std::fabs(d * iDotH * eDotH), sq(n_i * iDotH + n_e * eDotH));
        const float bsdf = g_transmission(p.shadowing, gi, ge) * common * sdiv(n_e * n_e, std::fabs(in.z * ex.z));
        r.value = {bsdf, bsdf, bsdf};
        r.pdfF = gi * common * sdiv(n_e * n_e, std::fabs(in.z));
        r.pdfB = ge * common * sdiv(n_i * n_i, std::fabs(ex.z));
        return r;
    }
    if (p.type == kMicrofacet) {  // microfacet_full.hpp:129-187
        const bool isReflection = in.z * ex.z > 0.0f;
        const V3 h = boundary_half(incident, other, in, ex);
        const float iDotH = dot(in, h);
        const float eDotH = dot(ex, h);
        const float n_i = incident.ior;
        const float n_t = other.ior;
        const float ge = g_term(p, eDotH, ex, h.z);
        const float gi = g_term(p, iDotH, in, h.z);
        if (ge == 0.0f || gi == 0.0f) return r;
        const float d = eval_ndf(p, h);
        const float f = fresnel_dielectric(n_i, n_t, std::fabs(iDotH)).f;
        if (isReflection) {
            const float g = g_reflection(p.shadowing, gi, ge);
            const float common = d * f / 4.0f;
            const float v = sdiv(g * common, in.z * ex.z);
            r.value = {v, v, v};
            r.pdfF = sdiv(gi * common, std::fabs(in.z));
            r.pdfB = sdiv(ge * common, std::fabs(ex.z));
            return r;
        }
        const float common = sdiv((1 - f) * std::fabs(d * iDotH * eDotH), sq(n_i * iDotH + n_t * eDotH));
        const float g = g_transmission(p.shadowing, gi, ge);
        const float bsdf = g * common * sdiv(n_t * n_t, std::fabs(in.z * ex.z));
        r.value = {bsdf, bsdf, bsdf};
        r.pdfF = gi * common * sdiv(n_t * n_t, std::fabs(in.z));
        r.pdfB = ge * common * sdiv(n_i * n_i, std::fabs(ex.z));
        return r;
    }
    return r;  // EMISSIVE
}

// ---------------------------------------------------------------- world-space wrappers
struct PathSample {
    bool valid;
    V3 throughput, excident;
    float pdfF, pdfB;
};

// materials::sample (material_sampling.hpp:68-146), view paths (adjoint = false)
constexpr float kShadingNormalEps = 1e-3f;  // SHADING_NORMAL_EPS (material_sampling.hpp:56)

PathSample mat_sample(const TS& ts, const MatParams& m, V3 incident, float u0, float u1, u64 i0,
                      bool adjoint = false) {
    PathSample r{false, {0, 0, 0}, {0, 0, 0}, 0, 0};
    const float iDotG = -dot(incident, ts.geoN);
    const float iDotN = -dot(incident, ts.shadingN);
    if (iDotG * iDotN <= 0.0f) return r;
    const V3 inTS{-dot(incident, ts.shadingTX), -dot(incident, ts.shadingTY), iDotN};
    const TsSample s = sample_ts(m, inTS, u0, u1, i0, adjoint);
    if (!s.valid || !nonzero(s.throughput)) return r;
    V3 g = s.excident.x * ts.shadingTX + s.excident.y * ts.shadingTY + s.excident.z * ts.shadingN;
    g = normalize(g);
    const float eDotG = dot(g, ts.geoN);
    const float eDotN = s.excident.z;
    if (eDotG * eDotN <= 0.0f) return r;
    r.throughput = s.throughput;
    r.excident = g;
    r.pdfF = mn(s.pdfF, 1e9f);
    r.pdfB = mn(s.pdfB, 1e9f);
    // shading normal correction of light sub-paths (material_sampling.hpp:131-135)
    if (adjoint && g_mutation != kMutLtNoShadingCorr)
        r.throughput = r.throughput * ((kShadingNormalEps + std::fabs(iDotN * eDotG)) /
                                       (kShadingNormalEps + std::fabs(iDotG * eDotN)));
    r.valid = true;
    return r;
}

struct EvalValue {
    V3 value;
    float cosOut, pdfF, pdfB;
};

// materials::evaluate (material_sampling.hpp:160-237), connections (no lightNormal)
EvalValue mat_evaluate(const TS& ts, const MatParams& m, V3 incident, V3 excident) {
    EvalValue z{{0, 0, 0}, 0, 0, 0};
    const float iDotN = -dot(incident, ts.shadingN);
    const float eDotN = dot(excident, ts.shadingN);
    if (!m.reflective() && iDotN * eDotN > 0.0f) return z;
    if (!m.refractive() && iDotN * eDotN < 0.0f) return z;
    const float iDotG = -dot(incident, ts.geoN);
    const float eDotG = dot(excident, ts.geoN);
    if (eDotG * eDotN <= 0.0f || iDotG * iDotN <= 0.0f) return z;
    const V3 inTS{-dot(incident, ts.shadingTX), -dot(incident, ts.shadingTY), iDotN};
    const V3 exTS{dot(excident, ts.shadingTX), dot(excident, ts.shadingTY), eDotN};
    const TsEval e = evaluate_ts(m, inTS, exTS);
    if (!nonzero(e.value)) return z;
    EvalValue r;
    r.value = e.value;
    r.cosOut = std::fabs(eDotN);
    r.pdfF = mn(e.pdfF, 1e9f);
    r.pdfB = mn(e.pdfB, 1e9f);
    return r;
}

// materials::evaluate with adjoint = true (material_sampling.hpp:160-237):
// the connections of light sub-paths evaluate the swapped pair and apply the
// shading normal correction; cosOut = |eDotN|.
EvalValue mat_evaluate_adjoint(const TS& ts, const MatParams& m, V3 incident, V3 excident) {
    EvalValue z{{0, 0, 0}, 0, 0, 0};
    const float iDotN = -dot(incident, ts.shadingN);
    const float eDotN = dot(excident, ts.shadingN);
    if (!m.reflective() && iDotN * eDotN > 0.0f) return z;
    if (!m.refractive() && iDotN * eDotN < 0.0f) return z;
    const float iDotG = -dot(incident, ts.geoN);
    const float eDotG = dot(excident, ts.geoN);
    if (eDotG * eDotN <= 0.0f || iDotG * iDotN <= 0.0f) return z;
    const V3 inTS{-dot(incident, ts.shadingTX), -dot(incident, ts.shadingTY), iDotN};
    const V3 exTS{dot(excident, ts.shadingTX), dot(excident, ts.shadingTY), eDotN};
    const TsEval e = evaluate_ts(m, exTS, inTS);
    if (!nonzero(e.value)) return z;
    EvalValue r;
    r.value = g_mutation == kMutLtNoShadingCorr
                  ? e.value
                  : e.value * ((kShadingNormalEps + std::fabs(iDotN * eDotG)) /
                               (kShadingNormalEps + std::fabs(iDotG * eDotN)));
    r.cosOut = std::fabs(eDotN);
    r.pdfF = mn(e.pdfB, 1e9f);
    r.pdfB = mn(e.pdfF, 1e9f);
    return r;
}

// emission (emissive.hpp:39-44); (0,0,0) for other types
V3 mat_emission(const MatParams& m, V3 geoN, V3 excident) {
    if (m.type != kEmissive) return {0, 0, 0};
    const float cosOut = dot(geoN, excident);
    if (cosOut <= 0.0f) return {0, 0, 0};
    return m.radiance;
}

// ---------------------------------------------------------------- roughness derivatives
// d ln value and d ln pdf of mat_evaluate with respect to the roughness
// (rx, ry), in double, for the adjoint. value = G * D * (roughness-free
// terms), pdf = G1(in) * D * (roughness-free terms) (evaluate_ts); the
// V-cavity G is roughness-free, Smith's is G1(in) * G1(ex) with
// G1 = 2 / (1 + sqrt(1 + A)), A = ((wx rx)^2 + (wy ry)^2) / wz^2; GGX
// D = 1 / (pi rx ry t^2), t = (hx/rx)^2 + (hy/ry)^2 + hz^2; Beckmann
// D = exp(-((hx/rx)^2 + (hy/ry)^2) / hz^2) / (pi rx ry hz^4).
void dln_ndf_d(const MatParams& p, V3 h, double& dx, double& dy) {
    const double rx = p.rx, ry = p.ry;
    const double x = h.x / rx, y = h.y / ry, cz2 = (double)h.z * h.z;
    if (p.ndf == kGgx) {
        const double t = x * x + y * y + cz2;
        dx = -1.0 / rx + 4.0 * x * x / (rx * t);
        dy = -1.0 / ry + 4.0 * y * y / (ry * t);
    } else {
        dx = 2.0 * x * x / (rx * cz2) - 1.0 / rx;
        dy = 2.0 * y * y / (ry * cz2) - 1.0 / ry;
    }
}
void dln_g1_d(const MatParams& p, V3 w, double& dx, double& dy) {
    dx = dy = 0.0;
    if (p.shadowing != kSmith) return;
    const double wz2 = (double)w.z * w.z;
    const double A = (((double)w.x * p.rx) * ((double)w.x * p.rx) + ((double)w.y * p.ry) * ((double)w.y * p.ry)) / wz2;
    const double sq1 = std::sqrt(1.0 + A);
    const double k = -1.0 / (2.0 * sq1 * (1.0 + sq1));
    dx = k * 2.0 * w.x * w.x * p.rx / wz2;
    dy = k * 2.0 * w.y * w.y * p.ry / wz2;
}
bool dln_mat_evaluate_d(const TS& ts, const MatParams& m, V3 incident, V3 excident, double dv[2], double dp[2]) {
    dv[0] = dv[1] = dp[0] = dp[1] = 0.0;
    if (!m.halfvector()) return false;
    if (!(m.rx > 0.0f && m.ry > 0.0f)) return false;
    const EvalValue e = mat_evaluate(ts, m, incident, excident);
    if (!nonzero(e.value)) return false;
    const V3 in{-dot(incident, ts.shadingTX), -dot(incident, ts.shadingTY), -dot(incident, ts.shadingN)};
    const V3 ex{dot(excident, ts.shadingTX), dot(excident, ts.shadingTY), dot(excident, ts.shadingN)};
    const V3 h = boundary_half(m.medium(in.z), m.medium(-in.z), in, ex);
    double ddx, ddy, gix, giy, gex, gey;
    dln_ndf_d(m, h, ddx, ddy);
    dln_g1_d(m, in, gix, giy);
    dln_g1_d(m, ex, gex, gey);
    dv[0] = ddx + gix + gex;
    dv[1] = ddy + giy + gey;
    if (e.pdfF < 1e9f) {
        dp[0] = ddx + gix;
        dp[1] = ddy + giy;
    }
    return true;
}

// ---------------------------------------------------------------- scene access
struct Vertex {
    bool isCamera = true, isVoid = false;
    V3 position{0, 0, 0};
    V3 incEnc{0, 0, 0};  // (x, y, dist * sgn(z)) for surfaces, raw dir for void
    TS ts;
    i32 inst = -1, prim = -1;
    float sp0 = 0.0f, sp1 = 0.0f;  // surface parameters of the hit (a quad's (u, v))
    MatParams mat;
    float incidentPdf = 1.0f;  // PtVertexExt
    V3 incident_dir() const {
        const float z = sgnf(incEnc.z) * std::sqrt(clampf(1.0f - incEnc.x * incEnc.x - incEnc.y * incEnc.y, 0.0f, 1.0f));
        return {incEnc.x, incEnc.y, z};
    }
    V3 geo_normal() const { return (isCamera || isVoid) ? V3{0, 0, 0} : ts.geoN; }
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
    int kind;        // 0 triangle, 1 quad, 2 sphere
    float b0, b1;    // barycentrics of v0 and v1, or the quad's (u, v)
    V3 geoN, tangentX;
    float u, v;
};

// Primitive counts and offsets of a LoD (triangles, quads, spheres in id order).
inline i32 lod_tris(const OrcScene& g, i32 l) { return g.lodTriOffset[l + 1] - g.lodTriOffset[l]; }
inline i32 lod_quads(const OrcScene& g, i32 l) {
    return g.lodQuadOffset ? g.lodQuadOffset[l + 1] - g.lodQuadOffset[l] : 0;
}
inline i32 lod_prim_offset(const OrcScene& g, i32 l) { return g.lodPrimOffset ? g.lodPrimOffset[l] : g.lodTriOffset[l]; }
inline int prim_kind(const OrcScene& g, i32 l, i32 prim) {
    const i32 nt = lod_tris(g, l);
    return prim < nt ? 0 : (prim < nt + lod_quads(g, l) ? 1 : 2);
}
inline const u32* quad_of(const OrcScene& g, i32 l, i32 prim) {
    return g.quads + 4 * ((size_t)g.lodQuadOffset[l] + (prim - lod_tris(g, l)));
}
inline const float* sphere_of(const OrcScene& g, i32 l, i32 prim) {
    return g.spheres + 4 * ((size_t)g.lodSphereOffset[l] + (prim - lod_tris(g, l) - lod_quads(g, l)));
}

// first_intersection (intersection.cu:366-621): the hit and its attributes
// for triangles (:543-571), quads (:572-610) and spheres (:524-541; the
// sphere of primitive id p is sphere p - (triangles + quads) of the LoD, as
// in accel_structs_commons.hpp:56-60).
HitInfo first_intersection(const OrcPtScene& S, V3& origin, V3 dir, V3 geoNormal, float tmax) {
    add_epsilon(origin, dir, geoNormal);
    const FastRay fr = make_fast_ray(origin, dir);
    float hitT = tmax;
    Hit h;
    if (tl_stats) {
        ++tl_stats->extRays;
        traverse(S.geo, fr, hitT, h, false, &tl_stats->extNodes, &tl_stats->extTris);
    } else {
        traverse(S.geo, fr, hitT, h, false, nullptr, nullptr);
    }
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
    r.kind = prim_kind(g, lod, h.prim);
    const M34 M = mat_of(g, h.inst);
    if (r.kind == 2) {
        const float* sp = sphere_of(g, lod, h.prim);
        const V3 hitPoint = origin + dir * hitT;
        const V3 local = xform_point(M, hitPoint);
        const V3 n = normalize(local - V3{sp[0], sp[1], sp[2]});
        V3 tangentX{1.0f, 0.0f, 0.0f};
        if (!(n.x == 0.0f && n.y == 0.0f)) {
            const float l2 = std::sqrt(n.y * n.y + n.x * n.x);
            tangentX = {n.y / l2, -n.x / l2, 0.0f};
        }
        r.u = fatan2(n.z, n.x) / (2.0f * kPi) + 0.5f;
        r.v = facos(-n.y) / kPi;
        r.geoN = normalize(xform_normal(M, n));
        r.tangentX = normalize(xform_normal(M, tangentX));
        return r;
    }
    if (r.kind == 1) {
        const u32* q = quad_of(g, lod, h.prim);
        V3 v[4];
        float uv[4][2];
        for (int k = 0; k < 4; ++k) {
            const float* p = g.verts + 3 * ((size_t)vb + q[k]);
            v[k] = {p[0], p[1], p[2]};
            const float* w = S.uvs + 2 * ((size_t)vb + q[k]);
            uv[k][0] = w[0];
            uv[k][1] = w[1];
        }
        const float su = h.w0, sv = h.w1;
        const V3 dxds = (1.f - su) * (v[3] - v[0]) + su * (v[2] - v[1]);
        const V3 dxdt = (1.f - sv) * (v[1] - v[0]) + sv * (v[2] - v[3]);
        const float dudsx = (1.f - su) * (uv[3][0] - uv[0][0]) + su * (uv[2][0] - uv[1][0]);
        const float dudsy = (1.f - su) * (uv[3][1] - uv[0][1]) + su * (uv[2][1] - uv[1][1]);
        const float dudtx = (1.f - sv) * (uv[1][0] - uv[0][0]) + sv * (uv[2][0] - uv[3][0]);
        const float dudty = (1.f - sv) * (uv[1][1] - uv[0][1]) + sv * (uv[2][1] - uv[3][1]);
        const float det = dudsx * dudty - dudtx * dudsy;
        V3 tangentX;
        if (det >= 1e-9f || det <= -1e-9f) {
            const float a = dudty / det, b = -dudsy / det;
            tangentX = {dxds.x * a + dxdt.x * b, dxds.y * a + dxdt.y * b, dxds.z * a + dxdt.z * b};
        } else {
            tangentX = dxds;
        }
        const V3 geoN = cross(dxdt, dxds);
        r.u = bilerp(uv[0][0], uv[1][0], uv[3][0], uv[2][0], su, sv);
        r.v = bilerp(uv[0][1], uv[1][1], uv[3][1], uv[2][1], su, sv);
        r.geoN = normalize(xform_normal(M, geoN));
        r.tangentX = normalize(xform_normal(M, tangentX));
        return r;
    }
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
    if (tl_stats) {
        ++tl_stats->shadowRays;
        return traverse(S.geo, fr, tmax, h, true, &tl_stats->shadowNodes, &tl_stats->shadowTris);
    }
    return traverse(S.geo, fr, tmax, h, true, nullptr, nullptr);
}

// tangent_space_geom_to_shader (tangent_space.hpp:10-60): spheres keep the
// geometric frame; triangles and quads interpolate the vertex normals
// (barycentric, bilinear) and orthogonalize the tangent.
TS tangent_space(const OrcPtScene& S, const HitInfo& hit) {
    if (hit.kind == 2) return {hit.geoN, hit.geoN, hit.tangentX, cross(hit.geoN, hit.tangentX)};
    const OrcScene& g = S.geo;
    const i32 lod = g.instLod[hit.inst];
    const i32 vb = g.lodVertOffset[lod], tb = g.lodTriOffset[lod];
    V3 sn;
    if (hit.kind == 1) {
        const u32* q = quad_of(g, lod, hit.prim);
        V3 n[4];
        for (int k = 0; k < 4; ++k) {
            const float* p = S.normals + 3 * ((size_t)vb + q[k]);
            n[k] = {p[0], p[1], p[2]};
        }
        sn = bilerp(n[0], n[1], n[3], n[2], hit.b0, hit.b1);
    } else {
        const u32* ix = g.tris + 3 * ((size_t)tb + hit.prim);
        V3 n[3];
        for (int k = 0; k < 3; ++k) {
            const float* p = S.normals + 3 * ((size_t)vb + ix[k]);
            n[k] = {p[0], p[1], p[2]};
        }
        sn = hit.b0 * n[0] + hit.b1 * n[1] + (1.f - (hit.b0 + hit.b1)) * n[2];
    }
    sn = normalize(xform_normal(mat_of(g, hit.inst), sn));
    const V3 tx = normalize(hit.tangentX - sn * dot(hit.tangentX, sn));
    const V3 ty = cross(sn, tx);
    return {sn, hit.geoN, tx, ty};
}

int material_index(const OrcPtScene& S, i32 inst, i32 prim) {
    return S.matIdx[(size_t)lod_prim_offset(S.geo, S.geo.instLod[inst]) + prim];
}
const OrcMaterial& material_of(const OrcPtScene& S, i32 inst, i32 prim) {
    return S.materials[material_index(S, inst, prim)];
}

// ---------------------------------------------------------------- background
inline float projected_area(V3 dir, const float box[6]) {  // sampling.hpp:160-167
    const V3 sides{box[3] - box[0], box[4] - box[1], box[5] - box[2]};
    const float projAx = sides.y * sides.z * std::fabs(dir.x);
    const float projAy = sides.x * sides.z * std::fabs(dir.y);
    const float projAz = sides.x * sides.y * std::fabs(dir.z);
    return projAx + projAy + projAz;
}
inline float box_diag_len(const float box[6]) {
    return len(V3{box[3] - box[0], box[4] - box[1], box[5] - box[2]});
}

const V3 kLumWeight{0.212671f, 0.715160f, 0.072169f};

// summed_details::lower_bound (texture_sampling.hpp:21-37)
int sat_lower_bound(const float* sat, int width, float val, int max, int constCoord, bool row) {
    int left = 0;
    int count = max + 1;
    while (count > 0) {
        const int step = count / 2;
        const int curr = left + step;
        const float x = row ? sat[(size_t)curr * width + constCoord] : sat[(size_t)constCoord * width + curr];
        if (x < val) {
            left = curr + 1;
            count -= step + 1;
        } else {
            count = step;
        }
    }
    return left;
}

struct EnvSample {
    int tx, ty;
    float u, v, pdf;
};
// importance_sample_texture (texture_sampling.hpp:119-155)
EnvSample importance_sample_sat(const float* sat, int W, int H, float u0, float u1) {
    const int bx = W - 1, by = H - 1;
    const float highestRowwise = sat[(size_t)by * W + bx];
    const float x = highestRowwise * u0;
    const int row = sat_lower_bound(sat, W, x, by, bx, true);
    const float vr0 = row == 0 ? 0.0f : sat[(size_t)(row - 1) * W + bx];
    const float vr1 = sat[(size_t)row * W + bx];
    const float rowVal = ((float)row + (x - vr0) / (vr1 - vr0)) / (float)H;
    const float highestColumnwise = vr1 - vr0;
    const float y = highestColumnwise * u1;
    const int column = sat_lower_bound(sat, W, y, bx, row, false);
    const float vc0 = column == 0 ? 0.f : sat[(size_t)row * W + column - 1];
    const float vc1 = column == bx ? highestRowwise : sat[(size_t)row * W + column];
    const float columnVal = ((float)column + (y - vc0) / (vc1 - vc0)) / (float)W;
    const float pdf = (vc1 - vc0) * (float)W * (float)H / highestRowwise;
    return {column, row, columnVal, rowVal, pdf};
}

struct EnvDir {
    V3 dir;
    float pdf;
    V3 radiance;
};
// sample_light_dir (light_sampling.hpp:225-290): COLORED and polar ENVMAP
EnvDir background_sample_dir(const OrcPtScene& S, float u0, float u1) {
    const OrcBackground& bg = S.background;
    EnvDir r;
    if (bg.type == 0) {  // sample_dir_sphere_uniform (sampling.hpp:101-109)
        const float cosTheta = u0 * 2.0f - 1.0f;
        const float sinTheta = std::sqrt((1.0f - cosTheta) * (1.0f + cosTheta));
        const float phi = u1 * 2.0f * kPi;
        r.dir = {sinTheta * fsin(phi), sinTheta * fcos(phi), cosTheta};
        r.pdf = 1.f / (4.f * kPi);
        r.radiance = V3{bg.color[0], bg.color[1], bg.color[2]} * V3{bg.scale[0], bg.scale[1], bg.scale[2]};
        return r;
    }
    const OrcTexture& env = S.textures[bg.envmapTexture];
    const EnvSample s = importance_sample_sat(bg.sat, env.width, env.height, u0, u1);
    const float phi = s.u * 2.f * kPi;
    const float theta = (1.f - s.v) * kPi;
    const float sinTheta = fsin(theta);
    const float cosTheta = fcos(theta);
    const float sinPhi = fsin(phi);
    const float cosPhi = fcos(phi);
    r.dir = {sinTheta * cosPhi, cosTheta, sinTheta * sinPhi};
    r.pdf = s.pdf / (2.f * kPi * kPi * sinTheta);
    const F4 t = texel(env, s.tx, s.ty);
    r.radiance = V3{t.x, t.y, t.z} * V3{bg.scale[0], bg.scale[1], bg.scale[2]};
    return r;
}

struct BgEval {
    V3 value;
    float pdfBack;
};
// evaluate_background (light_sampling.hpp:467-515): COLORED and polar ENVMAP
BgEval evaluate_background(const OrcPtScene& S, V3 direction) {
    const OrcBackground& bg = S.background;
    if (bg.type == 0)
        return {V3{bg.color[0], bg.color[1], bg.color[2]} * V3{bg.scale[0], bg.scale[1], bg.scale[2]},
                1.0f / (4.0f * kPi)};
    const OrcTexture& env = S.textures[bg.envmapTexture];
    // sample_polar_map / direction_to_uv (textures/interface.hpp:180-200)
    const float v0 = facos(direction.y) / kPi;
    const float u = fatan2(direction.z, direction.x) / (kPi * 2.0f);
    float v = 1.0f - v0;
    const float halfTexel = 0.5f / (float)env.height;
    v = clampf(v, halfTexel, 1.0f - halfTexel);
    const F4 t = tex_sample(env, u, v);
    V3 radiance{t.x, t.y, t.z};
    const int W = env.width, H = env.height;
    const float cdf = bg.sat[(size_t)(H - 1) * W + (W - 1)] / (float)(H * W);
    float pdfScale = dot(kLumWeight, radiance);
    const int pixelY = ifloor(v * (float)H);
    const float sinPixel = fsin(kPi * (float)((float)pixelY + 0.5f) / (float)H);
    const float sinJac = std::sqrt(1.0f - direction.y * direction.y);
    pdfScale *= sinPixel / (2.0f * kPi * kPi * sinJac);
    radiance = radiance * V3{bg.scale[0], bg.scale[1], bg.scale[2]};
    return {radiance, pdfScale / cdf};
}

// ---------------------------------------------------------------- lights
struct Nee {
    V3 position;
    float creationPdf;
    V3 dir;
    float cosOut;
    V3 diffIrradiance;
    float dist;
    V3 geoNormal;
    float distSq;
    int lightMaterial = -1;  // material of an area light (for the emission gradient)
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
struct QuadLight {  // AreaLightQuad record (lights.hpp:108-124), 96 bytes
    float posV[4][3];  // v0, v3 - v0, v1 - v0, v2 - v3 - (v1 - v0)
    float uvV[4][2];
    uint16_t material;
    uint16_t pad[7];
};
struct SphereLight {  // AreaLightSphere record (lights.hpp:125-137), 32 bytes
    float position[3];
    float radius;
    uint16_t material;
    uint16_t pad[7];
};
static_assert(sizeof(QuadLight) == 96 && sizeof(SphereLight) == 32, "light records");
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
    if (tl_stats) ++tl_stats->lightRecords;
    if (type == 0) {  // point (light_sampling.hpp:308-327); point lights sit in vacuum
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
        n.cosOut = 0.0f;
        n.diffIrradiance = {L.intensity[0] / distSq, L.intensity[1] / distSq, L.intensity[2] / distSq};
        if (g_mutation == kMutPointNoFalloff) n.diffIrradiance = {L.intensity[0], L.intensity[1], L.intensity[2]};
        n.dist = dist;
        n.geoNormal = {0, 0, 0};
        n.distSq = distSq;
        return n;
    }
    if (type == 3) {  // area quad (light_sampling.hpp:142-171, 376-399)
        QuadLight L;
        std::memcpy(&L, light, sizeof(L));
        V3 pv[4];
        for (int k = 0; k < 4; ++k) pv[k] = {L.posV[k][0], L.posV[k][1], L.posV[k][2]};
        const V3 tangentX = pv[1] + u0 * pv[3];
        const V3 tangentY = pv[2] + u1 * pv[3];
        V3 normal = cross(tangentY, tangentX);
        const float area = len(normal);
        normal = normal / area;
        const V3 position = pv[0] + tangentX * u1 + pv[2] * u0;
        const float lu = L.uvV[0][0] + L.uvV[1][0] * u1 + L.uvV[2][0] * u0 + L.uvV[3][0] * (u0 * u1);
        const float lv = L.uvV[0][1] + L.uvV[1][1] * u1 + L.uvV[2][1] * u0 + L.uvV[3][1] * (u0 * u1);
        const MatParams m = fetch(S, S.materials[L.material], lu, lv);
        const V3 rad = mat_emission(m, {0, 0, 1}, {0, 0, 1});
        const V3 intensity = rad * area;
        V3 d = position - pos;
        const float distSq = dot(d, d) + kDistSqEps;
        const float dist = std::sqrt(distSq);
        d = d / dist;
        const float cosOut = dot(normal, -d);
        V3 value{0, 0, 0};
        float vc = 0.0f;
        if (cosOut > 0.0f) {
            value = intensity;
            vc = cosOut;
        }
        n.position = position;
        n.creationPdf = 1.0f / area;
        n.dir = d;
        n.cosOut = vc;
        n.diffIrradiance = V3{value.x / distSq, value.y / distSq, value.z / distSq} * transmission(m.outer, dist);
        n.dist = dist;
        n.geoNormal = normal;
        n.distSq = distSq;
        n.lightMaterial = L.material;
        return n;
    }
    if (type == 4) {  // area sphere (light_sampling.hpp:400-432): the visible cap, uniform in area
        SphereLight L;
        std::memcpy(&L, light, sizeof(L));
        const V3 c{L.position[0], L.position[1], L.position[2]};
        V3 centerDir = pos - c;
        const float cd = len(centerDir);
        n.lightMaterial = L.material;
        if (cd <= L.radius) {  // inside: no estimate
            n.position = {0, 0, 0};
            n.creationPdf = 0.0f;
            n.dir = {0, 0, 0};
            n.cosOut = 0.0f;
            n.diffIrradiance = {0, 0, 0};
            n.dist = 0.0f;
            n.geoNormal = {0, 0, 0};
            n.distSq = 0.0f;
            return n;
        }
        centerDir = centerDir / cd;
        const float cosSphere = L.radius / cd;
        // sample_cone_uniform (sampling.hpp:112-119)
        const float cosTheta = (1.f - u0) + u0 * cosSphere;
        const float sinTheta = std::sqrt(1.f - cosTheta * cosTheta);
        const float phi = 2.f * u1 * kPi;
        const V3 local{fcos(phi) * sinTheta, fsin(phi) * sinTheta, cosTheta};
        const float conePdf = 1.f / (2.f * kPi * (1.f - cosSphere));
        const V3 tangentX = normalize(perpendicular(centerDir));
        const V3 tangentY = cross(centerDir, tangentX);
        const V3 globalDir = local.x * tangentX + local.y * tangentY + local.z * centerDir;
        const V3 surfPos = c + L.radius * globalDir;
        V3 cdir = surfPos - pos;
        const float cDistSq = dot(cdir, cdir) + kDistSqEps;
        const float cDist = std::sqrt(cDistSq);
        cdir = cdir / cDist;
        // direction_to_uv (textures/interface.hpp:169-174)
        const float tv = facos(globalDir.y) / kPi;
        const float tu = fatan2(globalDir.z, globalDir.x) / (kPi * 2.0f);
        const MatParams m = fetch(S, S.materials[L.material], tu, 1.0f - tv);
        V3 radiance = mat_emission(m, {0, 0, 1}, {0, 0, 1});
        const float sampleArea = L.radius * L.radius / conePdf;
        const float cosOut = mx(0.0f, -dot(globalDir, cdir));
        radiance = radiance * (sampleArea / cDistSq);
        radiance = radiance * transmission(m.outer, cDist);
        n.position = surfPos;
        n.creationPdf = 1.0f / sampleArea;
        n.dir = cdir;
        n.cosOut = cosOut;
        n.diffIrradiance = radiance;
        n.dist = cDist;
        n.geoNormal = globalDir;
        n.distSq = cDistSq;
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
    const float lu = L.uvV[0][0] + L.uvV[1][0] * s + L.uvV[2][0] * t;
    const float lv = L.uvV[0][1] + L.uvV[1][1] * s + L.uvV[2][1] * t;
    const MatParams m = fetch(S, S.materials[L.material], lu, lv);
    const V3 rad = mat_emission(m, {0, 0, 1}, {0, 0, 1});
    const float pdf = 1.0f / area;
    const V3 intensity = rad * area;
    V3 d = position - pos;
    const float distSq = dot(d, d) + kDistSqEps;
    const float dist = std::sqrt(distSq);
    d = d / dist;
    const float cosOut = dot(normal, -d);
    V3 value{0, 0, 0};
    float vc = 0.0f;
    if (cosOut > 0.0f) {
        value = intensity;
        vc = cosOut;
    }
    V3 diffIrradiance{value.x / distSq, value.y / distSq, value.z / distSq};
    diffIrradiance = diffIrradiance * transmission(m.outer, dist);
    n.position = position;
    n.creationPdf = pdf;
    n.dir = d;
    n.cosOut = vc;
    n.diffIrradiance = diffIrradiance;
    n.dist = dist;
    n.geoNormal = normal;
    n.distSq = distSq;
    n.lightMaterial = L.material;
    return n;
}

// connect_light(BackgroundDesc) (light_sampling.hpp:447-462)
Nee connect_background(const OrcPtScene& S, V3 pos, float u0, float u1) {
    const EnvDir s = background_sample_dir(S, u0, u1);
    Nee n;
    n.position = pos + s.dir * box_diag_len(S.sceneBox);
    n.creationPdf = s.pdf * std::fabs(1.0f) / (kMaxSceneSize * kMaxSceneSize);
    n.dir = s.dir;
    n.cosOut = 1.0f;
    n.diffIrradiance = s.radiance / s.pdf;
    if (g_mutation == kMutBackgroundNoPdf) n.diffIrradiance = s.radiance;
    n.dist = kMaxSceneSize;
    n.geoNormal = {0, 0, 0};
    n.distSq = kMaxSceneSize * kMaxSceneSize;
    return n;
}

inline Nee adjust_pdf(Nee n, float p) {  // light_tree_sampling.hpp:16-19
    n.creationPdf *= p;
    n.diffIrradiance = n.diffIrradiance / p;
    return n;
}

// get_center (light_tree.hpp:284-296; lights.hpp:210-228) of a light tree
// child: an internal node's center, a point light's position, a triangle's
// posV[0] + (posV[1] + posV[2]) / 3 (its record holds the edges).
V3 light_center(const OrcPtScene& S, uint32_t offset, uint16_t type) {
    const uint8_t* p = S.lightMem + offset;
    if (type == kInternalNode) {
        NodeRec n;
        std::memcpy(&n, p, sizeof(n));
        return {n.center[0], n.center[1], n.center[2]};
    }
    if (type == 0 || type == 4) {  // point light position, sphere center
        PointLightRec L;
        std::memcpy(&L, p, sizeof(L));
        return {L.position[0], L.position[1], L.position[2]};
    }
    if (type == 3) {  // quad: posV[0] + (posV[1] + posV[2]) / 2 + posV[3] / 4
        QuadLight q;
        std::memcpy(&q, p, sizeof(q));
        V3 pv[4];
        for (int k = 0; k < 4; ++k) pv[k] = {q.posV[k][0], q.posV[k][1], q.posV[k][2]};
        return pv[0] + (pv[1] + pv[2]) * 0.5f + pv[3] * 0.25f;
    }
    TriLight t;
    std::memcpy(&t, p, sizeof(t));
    const V3 p0{t.posV[0][0], t.posV[0][1], t.posV[0][2]};
    const V3 e1{t.posV[1][0], t.posV[1][1], t.posV[1][2]};
    const V3 e2{t.posV[2][0], t.posV[2][1], t.posV[2][2]};
    return p0 + (e1 + e2) / 3.0f;
}

// guide_flux / guide_flux_pos (light_tree_sampling.hpp:58-71): probability of
// the left child, by flux or by flux over the squared distance to ref.
float guide_left(const OrcPtScene& S, const NodeRec& node, V3 ref) {
    float leftFlux = node.leftFlux, rightFlux = node.rightFlux;
    if (S.posGuide) {
        const V3 dl = light_center(S, node.leftOffset, node.leftType) - ref;
        const V3 dr = light_center(S, node.rightOffset, node.rightType) - ref;
        leftFlux /= dot(dl, dl);
        rightFlux /= dot(dr, dr);
    }
    return leftFlux / (leftFlux + rightFlux);
}

// lights::connect (light_tree_sampling.hpp:195-329); the directional subtree is empty
Nee connect(const OrcPtScene& S, u64 index, u64 numIndices, u64 seed, V3 pos, float u0, float u1,
            int32_t* outLeaf = nullptr) {
    if (outLeaf) *outLeaf = -1;
    u64 rndChoice = numIndices > 0 ? seed + index * (~0ull / numIndices) : seed;
    if (rndChoice == ~0ull) --rndChoice;
    const float fluxSum = S.bgFlux + S.dirRootFlux + S.posRootFlux;
    const float envProb = S.bgFlux / fluxSum;
    const float dirProb = S.dirRootFlux / fluxSum;
    const float posProb = S.posRootFlux / fluxSum;
    const u64 rightEnv = percentage_of(~0ull, envProb);
    if (rndChoice < rightEnv) return adjust_pdf(connect_background(S, pos, u0, u1), envProb);
    u64 right = percentage_of(~0ull, envProb + dirProb);
    u64 left = rightEnv;
    (void)left;
    left = right;
    right = ~0ull;
    float lightProb = posProb;
    uint16_t type = (uint16_t)S.posRootType;
    uint32_t offset = 0;
    while (type == kInternalNode) {
        NodeRec node;
        std::memcpy(&node, S.lightMem + offset, sizeof(node));
        if (tl_stats) ++tl_stats->treeNodes;
        const float probLeft = guide_left(S, node, pos);
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
    if (outLeaf) *outLeaf = (int32_t)offset;
    return adjust_pdf(connect_light(S, type, S.lightMem + offset, pos, u0, u1), lightProb);
}

// light_pdf (light_tree_sampling.hpp:340-413) -> connect pdf of a hit emitter
// seen from refPos, the previous vertex of the path (path_util.hpp:443-444);
// (sp0, sp1): the hit's surface parameters (a quad's (u, v)).
float light_pdf(const OrcPtScene& S, i32 inst, i32 prim, V3 refPos, float sp0, float sp1) {
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
        if (tl_stats) ++tl_stats->treeNodes;
        const float pLeft = guide_left(S, node, refPos);
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
    if (type == 3) {  // the local area density at (sp0, sp1)
        QuadLight q;
        std::memcpy(&q, S.lightMem + offset, sizeof(q));
        V3 pv[4];
        for (int k = 0; k < 4; ++k) pv[k] = {q.posV[k][0], q.posV[k][1], q.posV[k][2]};
        const V3 tangentX = pv[1] + sp0 * pv[3];
        const V3 tangentY = pv[2] + sp1 * pv[3];
        return p / len(cross(tangentY, tangentX));
    }
    if (type == 4) {  // connect pdf: uniform over the cap visible from refPos
        SphereLight sl;
        std::memcpy(&sl, S.lightMem + offset, sizeof(sl));
        const float cosSphere = sl.radius / len(refPos - V3{sl.position[0], sl.position[1], sl.position[2]});
        const float solidAngle = 2 * kPi * (1.0f - cosSphere);
        const float rsq = sl.radius * sl.radius;
        return p / (solidAngle * rsq);
    }
    TriLight L;
    std::memcpy(&L, S.lightMem + offset, sizeof(L));
    const V3 e1{L.posV[1][0], L.posV[1][1], L.posV[1][2]};
    const V3 e2{L.posV[2][0], L.posV[2][1], L.posV[2][2]};
    const float area = len(cross(e1, e2)) / 2.0f;
    if (g_mutation == kMutLightPdfNoTree) return 1.0f / area;
    return p / area;
}

// background_pdf (light_tree_sampling.hpp:419-423)
float background_pdf(const OrcPtScene& S, float pdfBack) {
    const float p = S.bgFlux / (S.dirRootFlux + S.posRootFlux + S.bgFlux);
    return pdfBack * p;
}

// ---------------------------------------------------------------- accumulation
// Targets of PtTargets (pt_params.hpp:16-26) in this library's numbering:
// Radiance 0 (3), Hit ID 1 (2), Position 2 (3), Depth 3 (1), Normal 4 (3),
// UV 5 (2), Albedo 6 (3), Lightness 7 (1).
constexpr int kNumTargets = 8;
constexpr int kTargetChannels[kNumTargets] = {3, 2, 3, 1, 3, 2, 3, 1};
constexpr int kTargetOffset[kNumTargets] = {0, 3, 5, 8, 9, 12, 14, 17};
struct PixelAcc {
    u32 mask = 1u;
    float v[18] = {};
    u32 nans = 0;
    // RenderBuffer::contribute with check_nan (render_target.hpp:62-76,189-195)
    void add(int target, const float* x) {
        if (!(mask & (1u << target))) return;
        for (int k = 0; k < kTargetChannels[target]; ++k) {
            float y = x[k];
            if (std::isnan(y)) {
                ++nans;
                y = 0.0f;
            }
            v[kTargetOffset[target] + k] += y;
        }
    }
    void add3(int target, V3 a) {
        const float x[3] = {a.x, a.y, a.z};
        add(target, x);
    }
    void add1(int target, float a) { add(target, &a); }
};

// materials::albedo (lambert, microfacet_specular, microfacet_full, emissive)
V3 mat_albedo(const MatParams& m) {
    if (m.type == kLambert || m.type == kTorrance) return m.albedo;
    if (m.type == kWalter || m.type == kMicrofacet)  // microfacet_refractive.hpp:190-194
        return {1.0f / (1.0f + m.inner.absorption.x), 1.0f / (1.0f + m.inner.absorption.y),
                1.0f / (1.0f + m.inner.absorption.z)};
    return {0, 0, 0};
}

// Sampled BSDF directions of one path: recorded (mode 1) by a render of one
// scene and replayed (mode 2) by a render of another, whose walks then
// evaluate value * cos / pdf at the recorded directions instead of sampling.
// This is the estimator the adjoint differentiates with respect to the
// roughness (directions held fixed); finite differences of it check that
// derivative.
struct PathReplay {
    int mode = 0;
    std::vector<V3> dirs;
    size_t pos = 0;
};

// pt_sample (pt_common.hpp:75-156)
// Direction of pinholecam_sample_ray (pinhole.hpp:56-84) for pixel (px, py)
// jittered by (u0, u1).
V3 pinhole_dir(const OrcPtScene& S, int px, int py, float u0, float u1) {
    const V3 viewDir{S.camDir[0], S.camDir[1], S.camDir[2]};
    const V3 up{S.camUp[0], S.camUp[1], S.camUp[2]};
    const float w = (float)S.width, h = (float)S.height;
    float cx = ((float)px + u0) / w * 2.0f - 1.0f;
    float cy = ((float)py + u1) / h * 2.0f - 1.0f;
    cx *= S.tanVFov;
    cy *= S.tanVFov;
    const float aspect = w / h;
    const V3 xAxis = cross(viewDir, up);
    return normalize(xAxis * (cx * aspect) + up * cy + viewDir * 1.0f);
}

void pt_sample(const OrcPtScene& S, const int params[4], int px, int py, Pcg64& rng, PixelAcc& acc,
               PathReplay* replay = nullptr) {
    const int minPL = params[1], maxPL = params[2], neeCount = params[3];
    V3 throughput{1.0f, 1.0f, 1.0f};
    float guideWeight = 1.0f;
    float footA = 0.0f;  // FootprintV0::m_a (footprint.hpp:137-170); m_P stays 1
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
                EvalValue value = mat_evaluate(vtx.ts, vtx.mat, vtx.incident_dir(), nee.dir);
                if (nee.cosOut != 0 && g_mutation != kMutDropLightCos) value.cosOut *= nee.cosOut;
                const V3 radiance = value.value * nee.diffIrradiance;
                if ((radiance.x > 0.0f || radiance.y > 0.0f || radiance.z > 0.0f) && value.cosOut > 0.0f) {
                    const bool occluded =
                        any_intersection(S, vtx.position, nee.position, vtx.ts.geoN, nee.geoNormal, nee.dir);
                    if (!occluded) {
                        float hitPdf = value.pdfF * std::fabs(nee.cosOut) / nee.distSq;
                        if (g_mutation == kMutNeeMisNoLightCos) hitPdf = value.pdfF / nee.distSq;
                        float mis = 1.0f / ((float)neeCount + hitPdf / nee.creationPdf);
                        if (g_mutation == kMutNeeMisOne) mis = 1.0f / (1.0f + hitPdf / nee.creationPdf);
                        acc.add3(0, throughput * value.cosOut * radiance * mis);
                        acc.add1(7, guideWeight * value.cosOut);
                    }
                }
            }
        }
        // walk (random_walk.hpp:38-112)
        const u64 x0 = rng.next();
        const u64 x1 = rng.next();
        const float u0 = sample_uniform((u32)x0), u1 = sample_uniform((u32)(x0 >> 32));
        const float rndRoulette = sample_uniform((u32)rng.next());
        V3 excident, origin = vtx.position;
        float pdfF;
        V3 sThroughput;
        Medium medium = kVacuum;
        if (vtx.isCamera) {  // pinholecam_sample_ray (pinhole.hpp:56-84)
            const float w = (float)S.width, h = (float)S.height;
            const float aspect = w / h;
            excident = pinhole_dir(S, px, py, u0, u1);
            const float pixelArea = (2 * S.tanVFov) * (2 * S.tanVFov) * aspect;
            const float cosOut = dot(viewDir, excident);
            pdfF = 1.0f / (pixelArea * cosOut * cosOut * cosOut);
            sThroughput = {1.0f, 1.0f, 1.0f};
        } else if (replay && replay->mode == 2) {
            if (replay->pos >= replay->dirs.size()) break;
            excident = replay->dirs[replay->pos++];
            const EvalValue e = mat_evaluate(vtx.ts, vtx.mat, vtx.incident_dir(), excident);
            if (!nonzero(e.value) || !(e.pdfF > 0.0f)) break;
            sThroughput = e.value * (e.cosOut / e.pdfF);
            pdfF = e.pdfF;
            medium = vtx.mat.medium(dot(excident, vtx.ts.geoN));
        } else {
            const PathSample ps = mat_sample(vtx.ts, vtx.mat, vtx.incident_dir(), u0, u1, x1);
            if (!ps.valid) break;  // CANCEL
            excident = ps.excident;
            pdfF = ps.pdfF;
            sThroughput = ps.throughput;
            medium = vtx.mat.medium(dot(excident, vtx.ts.geoN));  // path_util.hpp:346-351
            if (replay && replay->mode == 1) replay->dirs.push_back(excident);
        }
        throughput = throughput * sThroughput;
        const float cont = mn(mx(mx(throughput.x, throughput.y), throughput.z) + 0.1f, 1.0f);
        if (rndRoulette >= cont) break;  // CANCEL
        if (g_mutation != kMutNoRouletteDiv) throughput = throughput / cont;
        V3 rayOrigin = origin;
        const HitInfo hit = first_intersection(S, rayOrigin, excident, vtx.geo_normal(), kMaxSceneSize);
        const float hitDist = hit.hit ? hit.t : kMaxSceneSize;
        throughput = throughput * transmission(medium, hitDist);
        // PtVertexExt::update (pt_common.hpp:45-61): FootprintV0::add_segment of a
        // perspective segment, guide weight 1 / (1 + solid angle)
        footA += 1.0f / std::sqrt(pdfF);
        guideWeight = 1.f / (1.f + footA * footA / 1.0f);
        float hu = 0.0f, hv = 0.0f;
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
            if (tl_stats) ++tl_stats->surfaceVertices;
            next.mat = fetch(S, material_of(S, hit.inst, hit.prim), hit.u, hit.v);
            hu = hit.u;
            hv = hit.v;
            const float inCosAbs = std::fabs(dot(excident, next.ts.shadingN));
            next.incidentPdf = pdfF * std::fabs(inCosAbs) / (hit.t * hit.t);
            if (g_mutation == kMutIncidentSolidAngle) next.incidentPdf = pdfF;
        }
        const V3 prevPos = vtx.position;
        vtx = next;
        if (pathLen == 0) {
            const float hid[2] = {(float)vtx.inst, (float)vtx.prim};
            acc.add(1, hid);
        }
        ++pathLen;
        if (pathLen >= minPL) {
            V3 emission{0, 0, 0};
            float connectPdf = 0.0f;
            if (vtx.isVoid) {  // get_emission, VOID (path_util.hpp:415-424)
                const BgEval b = evaluate_background(S, vtx.incEnc);
                emission = b.value;
                if (nonzero(emission)) connectPdf = background_pdf(S, b.pdfBack);
            } else {
                emission = mat_emission(vtx.mat, vtx.ts.geoN, -vtx.incident_dir());
                if (nonzero(emission)) connectPdf = light_pdf(S, vtx.inst, vtx.prim, prevPos, vtx.sp0, vtx.sp1);
            }
            if (nonzero(emission) && pathLen > 1) {
                float mis = 1.0f / (1.0f + (float)neeCount * (connectPdf / vtx.incidentPdf));
                if (g_mutation == kMutHitMisNoCount) mis = 1.0f / (1.0f + connectPdf / vtx.incidentPdf);
                emission = emission * mis;
            }
            acc.add3(0, throughput * emission);
            const V3 pos = vtx.isVoid ? V3{0, 0, 0} : vtx.position;
            acc.add3(2, guideWeight * pos);
            acc.add1(3, guideWeight * (vtx.isVoid ? 0.0f : std::fabs(vtx.incEnc.z)));
            acc.add3(4, guideWeight * (vtx.isVoid ? V3{0, 0, 0} : vtx.ts.shadingN));
            const float uvc[2] = {guideWeight * hu, guideWeight * hv};
            acc.add(5, uvc);
            acc.add3(6, guideWeight * (vtx.isVoid ? V3{0, 0, 0} : mat_albedo(vtx.mat)));
            acc.add1(7, guideWeight * ((emission.x + emission.y + emission.z) / 3.0f));
        }
        if (vtx.isVoid) break;
    } while (pathLen < maxPL);
}

MatParams params_of(const OrcMaterial& m, const OrcTexture* textures, float u, float v) {
    OrcPtScene S;
    std::memset(&S, 0, sizeof(S));
    S.textures = textures;
    return fetch(S, m, u, v);
}

// ---------------------------------------------------------------- adjoint
// Path replay of pt_sample recording what the derivative needs; the float
// path is the primal's, the gradient arithmetic is in double.
struct AdjRec {
    int tex = -1;
    int rtex = -1;  // roughness texture
    float hu = 0, hv = 0;
    V3 rho{0, 0, 0};
    double neeW[3] = {0, 0, 0};
    double emW[3] = {0, 0, 0};
    double misC = 0.0;  // 1 - MIS weight of this vertex's emission
    double G[3] = {0, 0, 0};
    float cont = 1.0f;
    float tp = 0.0f;
    int astar = 0;
    double rw[2] = {0, 0};  // d ln(walk weight) / d(rx, ry), sampled direction held fixed
    double rp[2] = {0, 0};  // d ln(pdf of the sampled direction) / d(rx, ry)
    double rn[2] = {0, 0};  // next-event estimates: weighted contribution * d ln / d(rx, ry)
};

struct AdjOut {
    double* const* texGrad;  // per texture id: 4 doubles per texel or null
    double* emGrad;          // 3 per material
    int fixedDecisions;      // 0: derivative of the expected image; 1: of this realization
};

void texel_weights(const OrcTexture& t, float u, float v, int* idx, double* w, int& n) {
    if (t.width == 1 && t.height == 1) {
        idx[0] = 0;
        w[0] = 1.0;
        n = 1;
        return;
    }
    if (t.sampling == 0) {
        idx[0] = imod(ifloor(v * (float)t.height), t.height) * t.width + imod(ifloor(u * (float)t.width), t.width);
        w[0] = 1.0;
        n = 1;
        return;
    }
    float fx = u * (float)t.width, fy = v * (float)t.height;
    const int ix = ifloor(fx), iy = ifloor(fy);
    fx -= (float)ix;
    fy -= (float)iy;
    const int x0 = imod(ix, t.width), x1 = imod(ix + 1, t.width), y0 = imod(iy, t.height), y1 = imod(iy + 1, t.height);
    idx[0] = y0 * t.width + x0;
    idx[1] = y0 * t.width + x1;
    idx[2] = y1 * t.width + x0;
    idx[3] = y1 * t.width + x1;
    w[0] = (1.0 - fx) * (1.0 - fy);
    w[1] = fx * (1.0 - fy);
    w[2] = (1.0 - fx) * fy;
    w[3] = (double)fx * fy;
    n = 4;
}

void add_em_grad(const OrcPtScene& S, const AdjOut& out, int mat, const double* w) {
    const OrcMaterial& m = S.materials[mat];
    for (int c = 0; c < 3; ++c)
        if (m.scale[c] != 0.0f) out.emGrad[3 * mat + c] += w[c] / m.scale[c];
}

void pt_adjoint_path(const OrcPtScene& S, const int params[4], int px, int py, Pcg64& rng, const double delta[3],
                     const AdjOut& out) {
    const int minPL = params[1], maxPL = params[2], neeCount = params[3];
    std::vector<AdjRec> rec(maxPL + 2);
    V3 throughput{1.0f, 1.0f, 1.0f};
    float footA = 0.0f;  // footprint (feeds only the auxiliary targets)
    Vertex vtx;
    vtx.position = {S.camPos[0], S.camPos[1], S.camPos[2]};
    rng.next();
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
                EvalValue value = mat_evaluate(vtx.ts, vtx.mat, vtx.incident_dir(), nee.dir);
                if (nee.cosOut != 0) value.cosOut *= nee.cosOut;
                const V3 radiance = value.value * nee.diffIrradiance;
                if ((radiance.x > 0.0f || radiance.y > 0.0f || radiance.z > 0.0f) && value.cosOut > 0.0f) {
                    if (!any_intersection(S, vtx.position, nee.position, vtx.ts.geoN, nee.geoNormal, nee.dir)) {
                        const float hitPdf = value.pdfF * std::fabs(nee.cosOut) / nee.distSq;
                        const float mis = 1.0f / ((float)neeCount + hitPdf / nee.creationPdf);
                        const V3 c = throughput * value.cosOut * radiance * mis;
                        const double wgt[3] = {delta[0] * c.x, delta[1] * c.y, delta[2] * c.z};
                        AdjRec& r = rec[pathLen];
                        for (int k = 0; k < 3; ++k) r.neeW[k] += wgt[k];
                        if (nee.lightMaterial >= 0) add_em_grad(S, out, nee.lightMaterial, wgt);
                        double dv[2], dp[2];
                        if (r.rtex >= 0 && dln_mat_evaluate_d(vtx.ts, vtx.mat, vtx.incident_dir(), nee.dir, dv, dp)) {
                            // d ln mis / d ln pdf = -mis * hitPdf / creationPdf
                            const double misF = (double)mis * ((double)hitPdf / nee.creationPdf);
                            for (int q = 0; q < 2; ++q) r.rn[q] += (wgt[0] + wgt[1] + wgt[2]) * (dv[q] - misF * dp[q]);
                        }
                    }
                }
            }
        }
        const u64 x0 = rng.next();
        const u64 x1 = rng.next();
        const float u0 = sample_uniform((u32)x0), u1 = sample_uniform((u32)(x0 >> 32));
        const float rndRoulette = sample_uniform((u32)rng.next());
        V3 excident, origin = vtx.position;
        float pdfF;
        V3 sThroughput;
        Medium medium = kVacuum;
        if (vtx.isCamera) {
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
            const PathSample ps = mat_sample(vtx.ts, vtx.mat, vtx.incident_dir(), u0, u1, x1);
            if (!ps.valid) break;
            excident = ps.excident;
            pdfF = ps.pdfF;
            sThroughput = ps.throughput;
            medium = vtx.mat.medium(dot(excident, vtx.ts.geoN));
            AdjRec& r = rec[pathLen];
            double dv[2], dp[2];
            if (r.rtex >= 0 && dln_mat_evaluate_d(vtx.ts, vtx.mat, vtx.incident_dir(), excident, dv, dp)) {
                for (int q = 0; q < 2; ++q) {
                    // d ln f (expected image), or d ln (f / pdf) with the direction held (this realization)
                    r.rw[q] = out.fixedDecisions ? dv[q] - dp[q] : dv[q];
                    r.rp[q] = pdfF < 1e9f ? dp[q] : 0.0;
                }
            }
        }
        throughput = throughput * sThroughput;
        const float cont = mn(mx(mx(throughput.x, throughput.y), throughput.z) + 0.1f, 1.0f);
        if (pathLen > 0) {
            AdjRec& r = rec[pathLen];
            r.cont = cont;
            const V3 t = throughput;
            r.astar = (t.x < t.y) ? ((t.y < t.z) ? 2 : 1) : ((t.x < t.z) ? 2 : 0);
            r.tp = r.astar == 0 ? t.x : (r.astar == 1 ? t.y : t.z);
        }
        if (rndRoulette >= cont) break;
        throughput = throughput / cont;
        V3 rayOrigin = origin;
        const HitInfo hit = first_intersection(S, rayOrigin, excident, vtx.geo_normal(), kMaxSceneSize);
        const float hitDist = hit.hit ? hit.t : kMaxSceneSize;
        throughput = throughput * transmission(medium, hitDist);
        footA += 1.0f / std::sqrt(pdfF);
        Vertex next;
        next.isCamera = false;
        int matIdx = -1;
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
            matIdx = material_index(S, hit.inst, hit.prim);
            next.mat = fetch(S, S.materials[matIdx], hit.u, hit.v);
            const float inCosAbs = std::fabs(dot(excident, next.ts.shadingN));
            next.incidentPdf = pdfF * std::fabs(inCosAbs) / (hit.t * hit.t);
        }
        const V3 prevPos = vtx.position;
        vtx = next;
        ++pathLen;
        AdjRec& r = rec[pathLen];
        if (!vtx.isVoid) {
            r.hu = hit.u;
            r.hv = hit.v;
            r.rho = vtx.mat.albedo;
            if (vtx.mat.type == kLambert || vtx.mat.type == kTorrance) r.tex = S.materials[matIdx].albedoTexture;
            if (vtx.mat.halfvector()) r.rtex = S.materials[matIdx].roughnessTexture;
        }
        if (pathLen >= minPL) {
            V3 emission{0, 0, 0};
            float connectPdf = 0.0f;
            if (vtx.isVoid) {
                const BgEval b = evaluate_background(S, vtx.incEnc);
                emission = b.value;
                if (nonzero(emission)) connectPdf = background_pdf(S, b.pdfBack);
            } else {
                emission = mat_emission(vtx.mat, vtx.ts.geoN, -vtx.incident_dir());
                if (nonzero(emission)) connectPdf = light_pdf(S, vtx.inst, vtx.prim, prevPos, vtx.sp0, vtx.sp1);
            }
            if (nonzero(emission) && pathLen > 1) {
                const float mis = 1.0f / (1.0f + (float)neeCount * (connectPdf / vtx.incidentPdf));
                emission = emission * mis;
                r.misC = 1.0 - (double)mis;  // d ln mis / d ln(incident pdf)
            }
            const V3 c = throughput * emission;
            r.emW[0] = delta[0] * c.x;
            r.emW[1] = delta[1] * c.y;
            r.emW[2] = delta[2] * c.z;
            if (!vtx.isVoid && vtx.mat.type == kEmissive) add_em_grad(S, out, matIdx, r.emW);
        }
        if (vtx.isVoid) break;
    } while (pathLen < maxPL);
    const int last = pathLen;
    for (int v = last - 1; v >= 0; --v)
        for (int k = 0; k < 3; ++k) rec[v].G[k] = rec[v + 1].G[k] + rec[v + 1].emW[k] + rec[v + 1].neeW[k];
    for (int j = 1; j <= last; ++j) {
        const AdjRec& r = rec[j];
        if (r.rtex >= 0 && out.texGrad[r.rtex]) {
            // roughness of vertex j: the walk's BSDF value scales everything
            // after walk j, the next vertex's emission MIS sees the sampled
            // pdf, the next-event estimates at j carry their own terms; with
            // the decisions held fixed, the roulette probabilities of walks
            // m >= j follow the throughput
            const double Gj = r.G[0] + r.G[1] + r.G[2];
            const double em1 = j < last ? (rec[j + 1].emW[0] + rec[j + 1].emW[1] + rec[j + 1].emW[2]) * rec[j + 1].misC
                                        : 0.0;
            double g[2];
            for (int q = 0; q < 2; ++q) {
                g[q] = r.rw[q] * Gj + r.rn[q] + em1 * r.rp[q];
                double sumD = 0.0;
                for (int m = j; out.fixedDecisions && m < last; ++m) {
                    if (!(rec[m].cont < 1.0f)) continue;
                    const double D = (double)rec[m].tp / rec[m].cont * (r.rw[q] - sumD);
                    g[q] -= D * (rec[m].G[0] + rec[m].G[1] + rec[m].G[2]);
                    sumD += D;
                }
            }
            const OrcTexture& tx = S.textures[r.rtex];
            int idx[4];
            double w[4];
            int nt = 0;
            texel_weights(tx, r.hu, r.hv, idx, w, nt);
            const double gc[2] = {tx.channels > 1 ? g[0] : g[0] + g[1], g[1]};
            for (int ch = 0; ch < (tx.channels > 1 ? 2 : 1); ++ch)
                for (int q = 0; q < nt; ++q) out.texGrad[r.rtex][4 * (size_t)idx[q] + ch] += w[q] * gc[ch];
        }
        if (r.tex < 0 || !out.texGrad[r.tex]) continue;
        int idx[4];
        double w[4];
        int nt = 0;
        texel_weights(S.textures[r.tex], r.hu, r.hv, idx, w, nt);
        const double rho[3] = {r.rho.x, r.rho.y, r.rho.z};
        for (int c = 0; c < 3; ++c) {
            if (rho[c] == 0.0) continue;
            double g = (r.G[c] + r.neeW[c]) / rho[c];
            double sumD = 0.0;
            for (int m = j; out.fixedDecisions && m < last; ++m) {
                if (!(rec[m].cont < 1.0f)) continue;
                const double D = (double)rec[m].tp / rec[m].cont * ((rec[m].astar == c ? 1.0 / rho[c] : 0.0) - sumD);
                g -= D * (rec[m].G[0] + rec[m].G[1] + rec[m].G[2]);
                sumD += D;
            }
            for (int q = 0; q < nt; ++q) out.texGrad[r.tex][4 * (size_t)idx[q] + c] += w[q] * g;
        }
    }
}


// ---------------------------------------------------------------- light tracer
// lt_sample (lt_common.hpp:17-78) for trees of positional lights.
struct LtEmitter {  // Emitter (light_sampling.hpp:21-52)
    V3 position, intensity, normal, absorption;
    bool area;
};

// sample_light_pos (light_sampling.hpp:80-199): point, triangle, quad, sphere.
LtEmitter light_position(const OrcPtScene& S, uint16_t type, const uint8_t* light, float u0, float u1) {
    LtEmitter e;
    e.absorption = {0, 0, 0};
    if (type == 0) {  // point
        PointLightRec L;
        std::memcpy(&L, light, sizeof(L));
        e.position = {L.position[0], L.position[1], L.position[2]};
        e.intensity = {L.intensity[0], L.intensity[1], L.intensity[2]};
        e.normal = {0, 0, 0};
        e.area = false;
        return e;
    }
    e.area = true;
    if (type == 3) {  // area quad
        QuadLight L;
        std::memcpy(&L, light, sizeof(L));
        const V3 pv[4] = {{L.posV[0][0], L.posV[0][1], L.posV[0][2]}, {L.posV[1][0], L.posV[1][1], L.posV[1][2]},
                          {L.posV[2][0], L.posV[2][1], L.posV[2][2]}, {L.posV[3][0], L.posV[3][1], L.posV[3][2]}};
        const V3 tangentX = pv[1] + u0 * pv[3];
        const V3 tangentY = pv[2] + u1 * pv[3];
        const V3 normal = cross(tangentY, tangentX);
        const float area = len(normal);
        e.normal = normal / area;
        e.position = pv[0] + tangentX * u1 + pv[2] * u0;
        const float lu = L.uvV[0][0] + L.uvV[1][0] * u1 + L.uvV[2][0] * u0 + L.uvV[3][0] * (u0 * u1);
        const float lv = L.uvV[0][1] + L.uvV[1][1] * u1 + L.uvV[2][1] * u0 + L.uvV[3][1] * (u0 * u1);
        const MatParams m = fetch(S, S.materials[L.material], lu, lv);
        e.intensity = mat_emission(m, {0, 0, 1}, {0, 0, 1}) * area;
        e.absorption = m.outer.absorption;
        return e;
    }
    if (type == 4) {  // area sphere
        SphereLight L;
        std::memcpy(&L, light, sizeof(L));
        // sample_dir_sphere_uniform (sampling.hpp:101-109)
        const float cosTheta = u0 * 2.0f - 1.0f;
        const float sinTheta = std::sqrt((1.0f - cosTheta) * (1.0f + cosTheta));
        const float phi = u1 * 2.0f * kPi;
        const V3 n{sinTheta * fsin(phi), sinTheta * fcos(phi), cosTheta};
        const MatParams m = fetch(S, S.materials[L.material], u1, facos(n.z));  // uv (u1, acos(n.z))
        const V3 radiance = mat_emission(m, {0, 0, 1}, {0, 0, 1});
        e.absorption = m.outer.absorption;
        const float area = 4 * kPi * L.radius * L.radius;
        e.position = V3{L.position[0], L.position[1], L.position[2]} + n * L.radius;
        e.normal = n;
        e.intensity = radiance * area;
        return e;
    }
    TriLight L;
    std::memcpy(&L, light, sizeof(L));
    const V3 p0{L.posV[0][0], L.posV[0][1], L.posV[0][2]};
    const V3 e1{L.posV[1][0], L.posV[1][1], L.posV[1][2]}, e2{L.posV[2][0], L.posV[2][1], L.posV[2][2]};
    const V3 normal = cross(e1, e2);
    const float area2 = len(normal);
    e.normal = normal / area2;
    const float area = area2 * 0.5f;
    const float bs = 1.f - std::sqrt(1.f - u0);  // sample_barycentric (sampling.hpp:123-128)
    const float bt = (1.f - bs) * u1;
    e.position = p0 + e1 * bs + e2 * bt;
    const float lu = L.uvV[0][0] + L.uvV[1][0] * bs + L.uvV[2][0] * bt;
    const float lv = L.uvV[0][1] + L.uvV[1][1] * bs + L.uvV[2][1] * bt;
    const MatParams m = fetch(S, S.materials[L.material], lu, lv);
    e.intensity = mat_emission(m, {0, 0, 1}, {0, 0, 1}) * area;
    e.absorption = m.outer.absorption;
    return e;
}

// emit (light_tree_sampling.hpp:81-187) without emitting background or
// directional lights: the tree is descended on flux alone and the intensity
// divided by the chosen light's probability (adjustPdf).
LtEmitter emit(const OrcPtScene& S, u64 index, u64 numIndices, u64 seed, float u0, float u1) {
    u64 rndChoice = numIndices > 0 ? seed + index * (~0ull / numIndices) : seed;
    if (rndChoice == ~0ull) --rndChoice;
    const float fluxSum = S.bgFlux + S.dirRootFlux + S.posRootFlux;
    const float envProb = S.bgFlux / fluxSum;
    const float dirProb = S.dirRootFlux / fluxSum;
    const float posProb = S.posRootFlux / fluxSum;
    u64 left = percentage_of(~0ull, envProb + dirProb);
    u64 right = ~0ull;
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
    LtEmitter e = light_position(S, type, S.lightMem + offset, u0, u1);
    if (g_mutation != kMutLtNoChoiceProb) e.intensity = e.intensity / lightProb;
    return e;
}

// pinholecam_project (pinhole.hpp:86-116)
bool project(const OrcPtScene& S, V3 dir, int& px, int& py, float& w) {
    const V3 viewDir{S.camDir[0], S.camDir[1], S.camDir[2]};
    const V3 up{S.camUp[0], S.camUp[1], S.camUp[2]};
    const float cosOut = dot(viewDir, dir);
    if (cosOut < 0.0f) return false;
    const V3 xAxis = cross(viewDir, up);
    const float den = cosOut * S.tanVFov;
    float ux = dot(xAxis, dir) / den;
    const float uy = dot(up, dir) / den;
    const float aspect = (float)S.width / (float)S.height;
    ux /= aspect;
    if (!(ux > -1 && ux <= 1 && uy > -1 && uy <= 1)) return false;
    px = (int)std::floor((ux * 0.5f + 0.5f) * (float)S.width);
    py = (int)std::floor((uy * 0.5f + 0.5f) * (float)S.height);
    if (px >= S.width) px = S.width - 1;
    if (py >= S.height) py = S.height - 1;
    const float pixelArea = (2.0f * S.tanVFov) * (2.0f * S.tanVFov) * aspect;
    w = g_mutation == kMutLtImportanceNoCos3 ? 1.0f / pixelArea : 1.0f / (pixelArea * cosOut * cosOut * cosOut);
    return true;
}

// One light path: splats (Radiance, Lightness, Density) into img, 5 per pixel.
void lt_path(const OrcPtScene& S, int minPL, int maxPL, bool density, int index, Pcg64& rng, float* img) {
    const int num = S.width * S.height;
    const V3 camPos{S.camPos[0], S.camPos[1], S.camPos[2]};
    V3 throughput{1.0f, 1.0f, 1.0f};
    const u64 xStart = rng.next();
    const u64 treeSeed = rng.next();
    const LtEmitter light = emit(S, (u64)index, (u64)num, treeSeed, sample_uniform((u32)xStart),
                                 sample_uniform((u32)(xStart >> 32)));
    rng.next();  // create_camera's RndSet2
    // the current surface vertex
    V3 pos{0, 0, 0}, incident{0, 0, 0};
    TS ts;
    MatParams mat;
    int pathLen = 0;
    do {
        if (pathLen + 1 >= minPL && pathLen + 1 <= maxPL) {
            const V3 vpos = pathLen == 0 ? light.position : pos;
            const V3 conn = vpos - camPos;
            const float distSq = dot(conn, conn);
            const float dist = std::sqrt(distSq);
            const V3 dir = conn / (dist + 1e-20f * sgnf(dist));  // sdiv (get_connection, path_util.hpp:384-399)
            int px, py;
            float w;
            if (project(S, dir, px, py, w)) {
                V3 lval, geoN;
                float lcos, geomFactor;
                if (pathLen == 0) {
                    if (light.area) {  // evaluate_area (lights.hpp:286-295)
                        const float c = dot(light.normal, -1.0f * dir);
                        lval = c <= 0.0f ? V3{0, 0, 0} : light.intensity;
                        lcos = c <= 0.0f ? 0.0f : (g_mutation == kMutLtAreaNoCos ? 1.0f : c);
                        geoN = light.normal;
                        geomFactor = dot(dir, light.normal);
                    } else {  // evaluate_point (lights.hpp:262-265)
                        lval = light.intensity;
                        lcos = 1.0f;
                        geoN = {0, 0, 0};
                        geomFactor = 0.0f;  // get_geometric_factor has no point-light case
                    }
                } else {
                    const EvalValue ev = mat_evaluate_adjoint(ts, mat, incident, -1.0f * dir);
                    lval = ev.value;
                    lcos = ev.cosOut;
                    geoN = ts.geoN;
                    geomFactor = dot(dir, ts.shadingN);
                }
                V3 bxdfProd{w * lval.x, w * lval.y, w * lval.z};
                const float cosProd = 1.0f * lcos;
                if (density || ((bxdfProd.x > 0.0f || bxdfProd.y > 0.0f || bxdfProd.z > 0.0f) && cosProd > 0.0f)) {
                    if (!any_intersection(S, camPos, vpos, {0, 0, 0}, geoN, dir)) {
                        bxdfProd = bxdfProd / distSq;
                        const V3 rad = throughput * cosProd * bxdfProd;
                        float* o = img + 5 * ((size_t)px + (size_t)py * S.width);
                        o[0] += rad.x;
                        o[1] += rad.y;
                        o[2] += rad.z;
                        o[3] += (throughput.x + throughput.y + throughput.z) / 3.0f * cosProd;
                        if (density) o[4] += w * 1.0f / distSq * std::fabs(geomFactor);
                    }
                }
            }
        }
        // walk (random_walk.hpp:38-112), adjoint
        const u64 x0 = rng.next();
        const u64 x1 = rng.next();
        const float rndRoulette = sample_uniform((u32)rng.next());
        const float u0 = sample_uniform((u32)x0), u1 = sample_uniform((u32)(x0 >> 32));
        V3 dirOut, sThroughput, absorption, geoOut, origin;
        if (pathLen == 0) {
            if (light.area) {  // sample_light_dir_area (light_sampling.hpp:201-211)
                const float cosT = std::sqrt(u0);
                const float sinT = std::sqrt(1.0f - u0);
                const float phi = u1 * 2 * kPi;
                const float lx = sinT * fsin(phi), ly = sinT * fcos(phi);
                const V3 tx = normalize(perpendicular(light.normal));
                const V3 ty = cross(light.normal, tx);
                dirOut = lx * tx + ly * ty + cosT * light.normal;
                sThroughput = light.intensity * kPi;
                geoOut = light.normal;
            } else {  // sample_light_dir_point (light_sampling.hpp:85-89)
                const float cosT = u0 * 2.0f - 1.0f;
                const float sinT = std::sqrt((1.0f - cosT) * (1.0f + cosT));
                const float phi = u1 * 2.0f * kPi;
                dirOut = {sinT * fsin(phi), sinT * fcos(phi), cosT};
                sThroughput = light.intensity * 4 * kPi;
                geoOut = {0, 0, 0};
            }
            origin = light.position;
            absorption = light.absorption;
        } else {
            const PathSample ps = mat_sample(ts, mat, incident, u0, u1, x1, true);
            if (!ps.valid) break;
            dirOut = ps.excident;
            sThroughput = ps.throughput;
            absorption = mat.medium(dot(dirOut, ts.geoN)).absorption;
            origin = pos;
            geoOut = ts.geoN;
        }
        throughput = throughput * sThroughput;
        const float cont = mn(mx(mx(throughput.x, throughput.y), throughput.z) + 0.1f, 1.0f);
        if (rndRoulette >= cont) break;
        throughput = throughput / cont;
        V3 rayOrigin = origin;
        const HitInfo hit = first_intersection(S, rayOrigin, dirOut, geoOut, kMaxSceneSize);
        Medium med;
        med.ior = 1.0f;
        med.absorption = absorption;
        throughput = throughput * transmission(med, hit.hit ? hit.t : kMaxSceneSize);
        if (!hit.hit) break;
        pos = rayOrigin + dirOut * hit.t;
        incident = dirOut;
        ts = tangent_space(S, hit);
        mat = fetch(S, S.materials[material_index(S, hit.inst, hit.prim)], hit.u, hit.v);
        ++pathLen;
    } while (pathLen < maxPL);
}

}  // namespace

extern "C" {

// Adjoint of the Radiance image of `iterations` iterations from a reset:
// delta = dLoss/dRadiance (3 floats per pixel). texGrad[t] (4 doubles per
// texel, or null) and emGrad (3 per material) receive the gradients; the sum
// runs serially in pixel order. fixedDecisions selects the derivative of this
// realization (1) instead of the derivative of the expected image (0).
int orc_pt_adjoint(const OrcPtScene* S, int32_t seed, int32_t minPathLength, int32_t maxPathLength, int32_t neeCount,
                   int32_t iterations, const float* delta, double* const* texGrad, double* emGrad,
                   int32_t fixedDecisions) {
    const int W = S->width, H = S->height;
    const int num = W * H;
    const int s = seed * (num + 1);
    const int params[4] = {seed, minPathLength, maxPathLength, neeCount};
    const AdjOut out{texGrad, emGrad, fixedDecisions ? 1 : 0};
    for (int pixel = 0; pixel < num; ++pixel) {
        const double d[3] = {(double)(delta[3 * (size_t)pixel] * (1.0f / (float)iterations)),
                             (double)(delta[3 * (size_t)pixel + 1] * (1.0f / (float)iterations)),
                             (double)(delta[3 * (size_t)pixel + 2] * (1.0f / (float)iterations))};
        if (d[0] == 0.0 && d[1] == 0.0 && d[2] == 0.0) continue;
        Pcg64 rng((u32)(pixel + s));
        for (int it = 0; it < iterations; ++it) pt_adjoint_path(*S, params, pixel % W, pixel / W, rng, d, out);
    }
    return 0;
}

// Runs `iterations` iterations from a reset (fresh per-pixel RNGs) and returns
// the running means of the enabled targets (targetMask bit t -> outTargets[t],
// kTargetChannels[t] floats per pixel, row-major), the running variances of
// the targets in varianceMask (-> outVariance[t]; UpdateIter,
// output_handler_variance.hpp:7-31, called with the iteration index of
// output_handler.hpp:138-139,223-225) plus the number of NaN contributions
// dropped.
int orc_pt_render_targets_var(const OrcPtScene* S, int32_t seed, int32_t minPathLength, int32_t maxPathLength,
                              int32_t neeCount, int32_t iterations, uint32_t targetMask, float** outTargets,
                              uint32_t varianceMask, float** outVariance, uint32_t* outNans) {
    varianceMask &= targetMask;
    const int W = S->width, H = S->height;
    const int num = W * H;
    std::vector<Pcg64> rngs;
    rngs.reserve(num);
    const int s = seed * (num + 1);  // cpu_pt.cpp:40-45
    for (int i = 0; i < num; ++i) rngs.emplace_back((u32)(i + s));
    for (int t = 0; t < kNumTargets; ++t) {
        if (targetMask & (1u << t)) std::memset(outTargets[t], 0, sizeof(float) * kTargetChannels[t] * (size_t)num);
        if (varianceMask & (1u << t))
            std::memset(outVariance[t], 0, sizeof(float) * kTargetChannels[t] * (size_t)num);
    }
    const int params[4] = {seed, minPathLength, maxPathLength, neeCount};
    u32 nans = 0;
    for (int it = 0; it < iterations; ++it) {
#pragma omp parallel for schedule(dynamic, 16) reduction(+ : nans)
        for (int pixel = 0; pixel < num; ++pixel) {
            PixelAcc acc;
            acc.mask = targetMask;
            pt_sample(*S, params, pixel % W, pixel / W, rngs[pixel], acc);
            nans += acc.nans;
            const float denom = mx(1.0f, (float)it + 1);
            const float varDenom = mx(1.0f, (float)it);
            for (int t = 0; t < kNumTargets; ++t) {
                if (!(targetMask & (1u << t))) continue;
                for (int c = 0; c < kTargetChannels[t]; ++c) {
                    const size_t idx = (size_t)kTargetChannels[t] * pixel + c;
                    const float iter = acc.v[kTargetOffset[t] + c];
                    float& cum = outTargets[t][idx];
                    const float diff = iter - cum;
                    cum += diff / denom;
                    if (varianceMask & (1u << t)) {
                        float& var = outVariance[t][idx];
                        const float varIter = diff * (iter - cum);
                        var += (varIter - var) / varDenom;
                    }
                }
            }
        }
    }
    if (outNans) *outNans = nans;
    return 0;
}

int orc_pt_render_targets(const OrcPtScene* S, int32_t seed, int32_t minPathLength, int32_t maxPathLength,
                          int32_t neeCount, int32_t iterations, uint32_t targetMask, float** outTargets,
                          uint32_t* outNans) {
    return orc_pt_render_targets_var(S, seed, minPathLength, maxPathLength, neeCount, iterations, targetMask,
                                     outTargets, 0u, nullptr, outNans);
}

// Work counts of one iteration of the paths of a pixel window (PathStats, in
// its field order: extension rays, their node visits and triangle tests,
// shadow rays, their node visits and triangle tests, surface vertices, texel
// taps, light-tree nodes, light records, paths), summed over the window.
int orc_pt_path_stats(const OrcPtScene* S, int32_t seed, int32_t minPathLength, int32_t maxPathLength,
                      int32_t neeCount, int32_t x0, int32_t y0, int32_t w, int32_t h, uint64_t* out11) {
    const int W = S->width, H = S->height;
    if (x0 < 0 || y0 < 0 || w < 0 || h < 0 || x0 + w > W || y0 + h > H) return 1;
    const int s = seed * (W * H + 1);
    const int params[4] = {seed, minPathLength, maxPathLength, neeCount};
    PathStats total;
#pragma omp parallel
    {
        PathStats mine;
        tl_stats = &mine;
#pragma omp for schedule(dynamic, 4)
        for (int k = 0; k < w * h; ++k) {
            const int px = x0 + k % w, py = y0 + k / w;
            Pcg64 rng((u32)(px + py * W + s));
            PixelAcc acc;
            pt_sample(*S, params, px, py, rng, acc);
            ++mine.paths;
        }
        tl_stats = nullptr;
#pragma omp critical
        {
            const u64* a = &mine.extRays;
            u64* b = &total.extRays;
            for (int i = 0; i < 11; ++i) b[i] += a[i];
        }
    }
    std::memcpy(out11, &total, sizeof(total));
    return 0;
}

// Radiance (3 floats per pixel, row-major over the window) of `iterations`
// iterations from a reset of the pixels x0 <= x < x0 + w, y0 <= y < y0 + h
// only. Every pixel runs from its own RNG, seeded as the reset seeds it, so
// the window equals the same pixels of a full render.
int orc_pt_render_window(const OrcPtScene* S, int32_t seed, int32_t minPathLength, int32_t maxPathLength,
                         int32_t neeCount, int32_t iterations, int32_t x0, int32_t y0, int32_t w, int32_t h,
                         float* outRadiance) {
    const int W = S->width, H = S->height;
    if (x0 < 0 || y0 < 0 || w < 0 || h < 0 || x0 + w > W || y0 + h > H) return 1;
    const int s = seed * (W * H + 1);  // cpu_pt.cpp:40-45
    const int params[4] = {seed, minPathLength, maxPathLength, neeCount};
    std::memset(outRadiance, 0, sizeof(float) * 3 * (size_t)w * h);
#pragma omp parallel for schedule(dynamic, 4)
    for (int k = 0; k < w * h; ++k) {
        const int px = x0 + k % w, py = y0 + k / w;
        const int pixel = px + py * W;
        Pcg64 rng((u32)(pixel + s));
        float* cum = outRadiance + 3 * (size_t)k;
        for (int it = 0; it < iterations; ++it) {
            PixelAcc acc;
            acc.mask = 1u;
            pt_sample(*S, params, px, py, rng, acc);
            const float denom = mx(1.0f, (float)it + 1);
            for (int c = 0; c < 3; ++c) cum[c] += (acc.v[c] - cum[c]) / denom;
        }
    }
    return 0;
}

// Radiance of `iterations` iterations of `eval` along the BSDF directions
// that the same paths (same seeds) sampled in `base` (PathReplay).
int orc_pt_render_replay(const OrcPtScene* base, const OrcPtScene* eval, int32_t seed, int32_t minPathLength,
                         int32_t maxPathLength, int32_t neeCount, int32_t iterations, float* outRadiance) {
    const int W = base->width, H = base->height;
    const int num = W * H;
    const int s = seed * (num + 1);
    const int params[4] = {seed, minPathLength, maxPathLength, neeCount};
#pragma omp parallel for schedule(dynamic, 64)
    for (int pixel = 0; pixel < num; ++pixel) {
        Pcg64 rng((u32)(pixel + s));
        float cum[3] = {0.0f, 0.0f, 0.0f};
        for (int it = 0; it < iterations; ++it) {
            const Pcg64 start = rng;
            PathReplay rp;
            rp.mode = 1;
            PixelAcc ignored;
            pt_sample(*base, params, pixel % W, pixel / W, rng, ignored, &rp);
            rp.mode = 2;
            Pcg64 r2 = start;
            PixelAcc acc;
            pt_sample(*eval, params, pixel % W, pixel / W, r2, acc, &rp);
            const float denom = mx(1.0f, (float)it + 1);
            for (int c = 0; c < 3; ++c) cum[c] += (acc.v[c] - cum[c]) / denom;
        }
        for (int c = 0; c < 3; ++c) outRadiance[3 * (size_t)pixel + c] = cum[c];
    }
    return 0;
}

// Radiance and Hit ID only.
int orc_pt_render(const OrcPtScene* S, int32_t seed, int32_t minPathLength, int32_t maxPathLength, int32_t neeCount,
                  int32_t iterations, float* outRadiance, float* outHitId, uint32_t* outNans) {
    float* t[kNumTargets] = {outRadiance, outHitId, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr};
    return orc_pt_render_targets(S, seed, minPathLength, maxPathLength, neeCount, iterations, outHitId ? 3u : 1u, t,
                                 outNans);
}

// BSDF of one material in tangent space (shading frame = identity), for the
// closed-form checks: sample -> (valid, throughput[3], excident[3], pdfF, pdfB)
// and evaluate -> (value[3], pdfF, pdfB). uv selects texels.
int orc_bsdf_sample(const OrcMaterial* m, const OrcTexture* textures, float u, float v, const float* inTS, float u0,
                    float u1, uint64_t i0, float* out9) {
    const MatParams p = params_of(*m, textures, u, v);
    const TsSample s = sample_ts(p, {inTS[0], inTS[1], inTS[2]}, u0, u1, i0);
    out9[0] = s.valid ? 1.0f : 0.0f;
    out9[1] = s.throughput.x;
    out9[2] = s.throughput.y;
    out9[3] = s.throughput.z;
    out9[4] = s.excident.x;
    out9[5] = s.excident.y;
    out9[6] = s.excident.z;
    out9[7] = s.pdfF;
    out9[8] = s.pdfB;
    return 0;
}
int orc_bsdf_evaluate(const OrcMaterial* m, const OrcTexture* textures, float u, float v, const float* inTS,
                      const float* exTS, float* out5) {
    const MatParams p = params_of(*m, textures, u, v);
    const TsEval e = evaluate_ts(p, {inTS[0], inTS[1], inTS[2]}, {exTS[0], exTS[1], exTS[2]});
    out5[0] = e.value.x;
    out5[1] = e.value.y;
    out5[2] = e.value.z;
    out5[3] = e.pdfF;
    out5[4] = e.pdfB;
    return 0;
}
// orc_bsdf_evaluate over n (in, ex) pairs (3 floats each) -> 5 floats each,
// for quadratures over the hemisphere.
int orc_bsdf_evaluate_batch(const OrcMaterial* m, const OrcTexture* textures, float u, float v, int32_t n,
                            const float* inTS, const float* exTS, float* out5) {
    const MatParams p = params_of(*m, textures, u, v);
#pragma omp parallel for schedule(static)
    for (int32_t i = 0; i < n; ++i) {
        const float* a = inTS + 3 * (size_t)i;
        const float* b = exTS + 3 * (size_t)i;
        const TsEval e = evaluate_ts(p, {a[0], a[1], a[2]}, {b[0], b[1], b[2]});
        float* o = out5 + 5 * (size_t)i;
        o[0] = e.value.x;
        o[1] = e.value.y;
        o[2] = e.value.z;
        o[3] = e.pdfF;
        o[4] = e.pdfB;
    }
    return 0;
}

// eval_ndf and fresnel_dielectric for the closed-form checks.
float orc_eval_ndf(int32_t ndf, float rx, float ry, const float* h) {
    MatParams p;
    p.ndf = ndf;
    p.rx = rx;
    p.ry = ry;
    return eval_ndf(p, {h[0], h[1], h[2]});
}
float orc_fresnel_dielectric(float n_i, float n_t, float cosIAbs, float* cosTAbs) {
    const Refraction r = fresnel_dielectric(n_i, n_t, cosIAbs);
    if (cosTAbs) *cosTAbs = r.cosTAbs;
    return r.f;
}
// Texture sampling (CPU semantics).
int orc_texture_sample(const OrcTexture* t, float u, float v, float* out4) {
    const F4 s = tex_sample(*t, u, v);
    out4[0] = s.x;
    out4[1] = s.y;
    out4[2] = s.z;
    out4[3] = s.w;
    return 0;
}
// Background: sample a direction (dir[3], pdf, radiance[3]) and evaluate one
// (radiance[3], pdf.back). Needs only S->background, S->textures, S->sceneBox.
int orc_background_sample(const OrcPtScene* S, float u0, float u1, float* out7) {
    const EnvDir d = background_sample_dir(*S, u0, u1);
    out7[0] = d.dir.x;
    out7[1] = d.dir.y;
    out7[2] = d.dir.z;
    out7[3] = d.pdf;
    out7[4] = d.radiance.x;
    out7[5] = d.radiance.y;
    out7[6] = d.radiance.z;
    return 0;
}
// RNG hooks: scramble_seed (SplitMix64 finalizer) and the first n outputs of
// PCG64(seed) (rng.hpp:12-17, 70-93).
// The wireframe renderer (GpuWireframe::iterate, gpu_wireframe.cpp:33-48;
// sample_wireframe, wireframe_common.hpp:94-181) for triangle meshes:
// `iterations` iterations, each drawing one u32 seed per pixel (pixel order)
// from the renderer's PCG64 state *rngState (updated), into the running mean
// (and, if outVariance, variance) of the 3-channel Border target.
int orc_wireframe_render(const OrcPtScene* S, int32_t iterations, float lineWidth, int32_t maxTraceDepth,
                         uint64_t* rngState, float* outBorder, float* outVariance) {
    const int W = S->width, H = S->height;
    const int num = W * H;
    std::memset(outBorder, 0, sizeof(float) * 3 * (size_t)num);
    if (outVariance) std::memset(outVariance, 0, sizeof(float) * 3 * (size_t)num);
    Pcg64 gen(0u);
    gen.s = *rngState;
    std::vector<u32> seeds(num);
    const V3 camPos{S->camPos[0], S->camPos[1], S->camPos[2]};
    for (int it = 0; it < iterations; ++it) {
        for (int i = 0; i < num; ++i) seeds[i] = (u32)gen.next();
#pragma omp parallel for schedule(dynamic, 256)
        for (int pixel = 0; pixel < num; ++pixel) {
            const int px = pixel % W, py = pixel / W;
            Pcg64 rng(seeds[pixel]);
            rng.next();  // create_camera's RndSet2 (unused by the pinhole)
            const u64 x0 = rng.next();
            rng.next();
            const V3 dir = pinhole_dir(*S, px, py, sample_uniform((u32)x0), sample_uniform((u32)(x0 >> 32)));
            V3 origin = camPos;
            float value[3] = {0.0f, 0.0f, 0.0f};
            float totalDistance = 0.0f;
            for (int i = 0; i < maxTraceDepth; ++i) {
                V3 rayOrigin = origin;
                const HitInfo hit = first_intersection(*S, rayOrigin, dir, V3{0, 0, 0}, kMaxSceneSize);
                if (!hit.hit) {
                    const BgEval bg = evaluate_background(*S, dir);
                    if (bg.value.x > 0.0f || bg.value.y > 0.0f || bg.value.z > 0.0f) {
                        value[0] = bg.value.x;
                        value[1] = bg.value.y;
                        value[2] = bg.value.z;
                    }
                    break;
                }
                const float hitDist = hit.t < 0.001f ? 0.001f : hit.t;
                const V3 p = rayOrigin + dir * hitDist;
                totalDistance += hitDist;
                const float pixelSize = totalDistance * S->tanVFov / (float)H;
                const OrcScene& g = S->geo;
                const M34 M = instance_to_world(g.worldToInst + 12 * (size_t)hit.inst);
                const i32 lod = g.instLod[hit.inst];
                auto world_vert = [&](u32 i) {
                    const float* q = g.verts + 3 * ((size_t)g.lodVertOffset[lod] + i);
                    return xform_point(M, V3{q[0], q[1], q[2]});
                };
                float projDistToRim;
                if (hit.kind == 2) {
                    // computeDistToRim (wireframe_common.hpp:79-89), scaled by 0.05 (:164-169)
                    const float* sp = sphere_of(g, lod, hit.prim);
                    const V3 center = xform_point(M, V3{sp[0], sp[1], sp[2]});
                    projDistToRim = 0.05f * (sp[3] * dot(normalize(center - p), dir));
                } else {
                    V3 closest;
                    if (hit.kind == 1) {
                        // computeClosestLinePoint, quads (wireframe_common.hpp:42-77):
                        // A = v0, B = v1, C = v3, D = v2; edges AB, AC, BD, CD
                        const u32* q = quad_of(g, lod, hit.prim);
                        const V3 A = world_vert(q[0]), B = world_vert(q[1]), Cc = world_vert(q[3]),
                                 D = world_vert(q[2]);
                        const V3 AB = B - A, AC = Cc - A, BD = D - B, CD = D - Cc;
                        const V3 AP = p - A, BP = p - B, CP = p - Cc;
                        const V3 onAB = A + AB * (dot(AP, AB) / dot(AB, AB));
                        const V3 onAC = A + AC * (dot(AP, AC) / dot(AC, AC));
                        const V3 onBD = B + BD * (dot(BP, BD) / dot(BD, BD));
                        const V3 onCD = Cc + CD * (dot(CP, CD) / dot(CD, CD));
                        const float dAB = dot(onAB - p, onAB - p), dAC = dot(onAC - p, onAC - p),
                                    dBD = dot(onBD - p, onBD - p), dCD = dot(onCD - p, onCD - p);
                        if (dAB <= dAC && dAB <= dBD && dAB <= dCD)
                            closest = onAB;
                        else if (dAC <= dAB && dAC <= dBD && dAC <= dCD)
                            closest = onAC;
                        else if (dBD <= dAB && dBD <= dAC && dBD <= dCD)
                            closest = onBD;
                        else
                            closest = onCD;
                    } else {
                        const u32* ix = g.tris + 3 * ((size_t)g.lodTriOffset[lod] + hit.prim);
                        V3 v[3];
                        for (int k = 0; k < 3; ++k) v[k] = world_vert(ix[k]);
                        // computeClosestLinePoint (wireframe_common.hpp:16-40)
                        const V3 AB = v[1] - v[0], AC = v[2] - v[0], BC = v[2] - v[1];
                        const V3 AP = p - v[0], BP = p - v[1];
                        const V3 onAB = v[0] + AB * (dot(AP, AB) / dot(AB, AB));
                        const V3 onAC = v[0] + AC * (dot(AP, AC) / dot(AC, AC));
                        const V3 onBC = v[1] + BC * (dot(BP, BC) / dot(BC, BC));
                        const float dAB = dot(onAB - p, onAB - p), dAC = dot(onAC - p, onAC - p),
                                    dBC = dot(onBC - p, onBC - p);
                        closest = (dAB <= dAC && dAB <= dBC) ? onAB : ((dAC <= dAB && dAC <= dBC) ? onAC : onBC);
                    }
                    const V3 lineSegment = closest - camPos;
                    const V3 projected = lineSegment - dir * dot(lineSegment, dir);
                    projDistToRim = len(projected);
                }
                if (projDistToRim > lineWidth * pixelSize) {
                    origin = rayOrigin + dir * hitDist;
                } else {
                    value[0] = value[1] = value[2] = 1.0f;
                    break;
                }
            }
            const float denom = mx(1.0f, (float)it + 1);
            const float varDenom = mx(1.0f, (float)it);
            for (int c = 0; c < 3; ++c) {
                const float x = std::isnan(value[c]) ? 0.0f : value[c];
                float& cum = outBorder[3 * (size_t)pixel + c];
                const float diff = x - cum;
                cum += diff / denom;
                if (outVariance) {
                    float& var = outVariance[3 * (size_t)pixel + c];
                    var += (diff * (x - cum) - var) / varDenom;
                }
            }
        }
    }
    *rngState = gen.s;
    return 0;
}

uint64_t orc_scramble_seed(uint32_t x) { return Pcg64::scramble(x); }

// Switches on one planted estimator mistake (Mutation; 0 = none) for the
// closed-form estimator tests; returns the previous setting.
int orc_set_mutation(int32_t m) {
    const int old = g_mutation;
    g_mutation = m;
    return old;
}
int orc_pcg_sequence(uint32_t seed, int32_t n, uint64_t* out) {
    Pcg64 r(seed);
    for (int32_t i = 0; i < n; ++i) out[i] = r.next();
    return 0;
}
// Light tree sampling hooks: one NEE connection from pos (light position
// xyz, creation pdf, diffIrradiance xyz, byte offset of the chosen
// positional light or -1 for the background) and the connect pdf of a hit
// emitter triangle seen from refPos.
int orc_light_connect(const OrcPtScene* S, uint64_t index, uint64_t numIndices, uint64_t seed, const float* pos,
                      float u0, float u1, float* out8) {
    int32_t leaf = -1;
    const Nee n = connect(*S, index, numIndices, seed, {pos[0], pos[1], pos[2]}, u0, u1, &leaf);
    out8[0] = n.position.x;
    out8[1] = n.position.y;
    out8[2] = n.position.z;
    out8[3] = n.creationPdf;
    out8[4] = n.diffIrradiance.x;
    out8[5] = n.diffIrradiance.y;
    out8[6] = n.diffIrradiance.z;
    out8[7] = (float)leaf;
    return 0;
}
float orc_light_pdf(const OrcPtScene* S, int32_t inst, int32_t prim, const float* refPos, float sp0, float sp1) {
    return light_pdf(*S, inst, prim, {refPos[0], refPos[1], refPos[2]}, sp0, sp1);
}
int orc_background_evaluate(const OrcPtScene* S, const float* dir, float* out4) {
    const BgEval b = evaluate_background(*S, {dir[0], dir[1], dir[2]});
    out4[0] = b.value.x;
    out4[1] = b.value.y;
    out4[2] = b.value.z;
    out4[3] = b.pdfBack;
    return 0;
}

// The light tracer (GpuLightTracer, gpu_lt.cpp:25-37; lt_sample,
// lt_common.hpp:17-78): `iterations` iterations of width * height light
// paths from generators seeded i + seed * (pixels + 1); path i's splats are
// added in path order. out5: running means of Radiance (3), Lightness and
// Density per pixel; outVar5 (optional) their variances.
int orc_lt_render(const OrcPtScene* S, int32_t seed, int32_t minPathLength, int32_t maxPathLength,
                  int32_t iterations, int32_t density, float* out5, float* outVar5) {
    const int num = S->width * S->height;
    std::vector<Pcg64> rngs;
    rngs.reserve(num);
    const int s = seed * (num + 1);
    for (int i = 0; i < num; ++i) rngs.emplace_back((u32)(i + s));
    std::memset(out5, 0, sizeof(float) * 5 * (size_t)num);
    if (outVar5) std::memset(outVar5, 0, sizeof(float) * 5 * (size_t)num);
    std::vector<float> img(5 * (size_t)num);
    for (int it = 0; it < iterations; ++it) {
        std::fill(img.begin(), img.end(), 0.0f);
        for (int i = 0; i < num; ++i) lt_path(*S, minPathLength, maxPathLength, density != 0, i, rngs[i], img.data());
        const float denom = mx(1.0f, (float)it + 1);
        const float varDenom = mx(1.0f, (float)it);
        for (size_t k = 0; k < img.size(); ++k) {
            const float iter = img[k];
            const float diff = iter - out5[k];
            const float cum = out5[k] + diff / denom;
            out5[k] = cum;
            if (outVar5) outVar5[k] += (diff * (iter - cum) - outVar5[k]) / varDenom;
        }
    }
    return 0;
}

}  // extern "C"
