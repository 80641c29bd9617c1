// Host light-tree builder: LightTreeBuilder::build (light_tree.cpp:340-447),
// create_light_tree (:119-266), fill_map (:268-302), get_flux (lights.cpp:11-81)
// and the Hilbert ordering (sfcurves.hpp), producing Mufflon's byte layout.
#include "light_tree.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <numeric>

namespace mfh {

namespace {

using u16 = uint16_t;
using u32 = uint32_t;
using u64 = uint64_t;

constexpr float kPi = 3.14159265358979323846f;
constexpr size_t kNodeBytes = 32;
constexpr size_t kTriBytes = 64;
constexpr size_t kPointBytes = 32;

struct V3 {
    float x, y, z;
};
inline V3 add(V3 a, V3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
inline V3 sub(V3 a, V3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline V3 mul(V3 a, float s) { return {a.x * s, a.y * s, a.z * s}; }
inline V3 divs(V3 a, float s) { return {a.x / s, a.y / s, a.z / s}; }
inline V3 cross(V3 a, V3 b) { return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
inline float len(V3 a) { return std::sqrt(a.x * a.x + a.y * a.y + a.z * a.z); }

// sfcurves.hpp:16-24, 40-71
u64 part_by_two16(u16 x) {
    u64 r = x;
    r = (r | (r << 16)) & 0x0000ff0000ffull;
    r = (r | (r << 8)) & 0x00f00f00f00full;
    r = (r | (r << 4)) & 0x0c30c30c30c3ull;
    r = (r | (r << 2)) & 0x249249249249ull;
    return r;
}
u64 gray_to_binary(u64 n) {
    n ^= n >> 32;
    n ^= n >> 16;
    n ^= n >> 8;
    n ^= n >> 4;
    n ^= n >> 2;
    n ^= n >> 1;
    return n;
}
u64 binary_to_gray(u64 n) { return n ^ (n >> 1); }
u64 hilbert_code(u16 a, u16 b, u16 c) {
    return gray_to_binary(part_by_two16(a) | (part_by_two16(b) << 1u) | (part_by_two16(c) << 2u));
}

// rng.hpp:12-17 and the GoldenRatio2D sampler (rng.hpp:114-140)
u64 scramble_seed(u32 x) {
    u64 z = (u64)x + 0x9e3779b97f4a7c15ull;
    z = (z ^ (z >> 30u)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27u)) * 0x94d049bb133111ebull;
    return z ^ (z >> 31u);
}
struct GoldenRatio2D {
    u64 state;
    explicit GoldenRatio2D(u32 seed) : state(scramble_seed(seed)) {}
    u64 next() {
        u64 s = state;
        state += 11400714819323198485ull;
        s = binary_to_gray(s);
        s = ((s >> 1) & 0x2222222222222222ull) | ((s << 1) & 0x4444444444444444ull) | (s & 0x9999999999999999ull);
        s = ((s >> 2) & 0x0c0c0c0c0c0c0c0cull) | ((s << 2) & 0x3030303030303030ull) | (s & 0xc3c3c3c3c3c3c3c3ull);
        s = ((s >> 4) & 0x00f000f000f000f0ull) | ((s << 4) & 0x0f000f000f000f00ull) | (s & 0xf00ff00ff00ff00full);
        s = ((s >> 8) & 0x0000ff000000ff00ull) | ((s << 8) & 0x00ff000000ff0000ull) | (s & 0xff0000ffff0000ffull);
        s = ((s >> 16) & 0x00000000ffff0000ull) | ((s << 16) & 0x0000ffff00000000ull) | (s & 0xffff00000000ffffull);
        return s;
    }
};
inline float sample_uniform(u32 u) { return (float)u / 4294967810.0f; }

// float -> u32 as the reference's x86-64 build converts it (through a signed
// 64-bit integer, then truncation).
inline u32 f2u(float v) { return (u32)(int64_t)v; }

struct PosLight {
    bool isArea;
    HostAreaLight area;
    HostPointLight point;
};

inline V3 point_of(const float p[3]) { return {p[0], p[1], p[2]}; }

// The sort key reads the light description as the stored record
// (light_tree.cpp:366-379 passes the variant's storage to get_center): for a
// triangle points[0] + (points[1] + points[2]) / 3, for a quad points[0] +
// (points[1] + points[2]) / 2 + points[3] / 4 (lights.hpp:216-224 over the
// description's points), a sphere's center, a point light's position.
V3 sort_center(const PosLight& l) {
    if (!l.isArea) return point_of(l.point.position);
    const HostAreaLight& a = l.area;
    if (a.kind == kAreaSphere) return point_of(a.points[0]);
    const V3 p0 = point_of(a.points[0]), p1 = point_of(a.points[1]), p2 = point_of(a.points[2]);
    if (a.kind == kAreaQuad) return add(add(p0, mul(add(p1, p2), 0.5f)), mul(point_of(a.points[3]), 0.25f));
    return add(p0, divs(add(p1, p2), 3.0f));
}

struct TriRecord {  // AreaLightTriangle (lights.hpp:91-107), 64 bytes
    float posV[3][3];
    u16 material;
    u16 pad0;
    float uvV[3][2];
};
static_assert(sizeof(TriRecord) == 64, "AreaLightTriangle record");
struct QuadRecord {  // AreaLightQuad (lights.hpp:108-124), 96 bytes
    float posV[4][3];  // v0, v3 - v0, v1 - v0, v2 - v3 - (v1 - v0)
    float uvV[4][2];
    u16 material;
    u16 pad[7];
};
static_assert(sizeof(QuadRecord) == 96, "AreaLightQuad record");
struct SphereRecord {  // AreaLightSphere (lights.hpp:125-137), 32 bytes
    float position[3];
    float radius;
    u16 material;
    u16 pad[7];
};
static_assert(sizeof(SphereRecord) == 32, "AreaLightSphere record");
constexpr size_t kQuadBytes = 96;
constexpr size_t kSphereBytes = 32;
struct PointRecord {  // PointLight (lights.hpp:46-50), 32 bytes
    float position[3];
    u16 mediumIndex;
    u16 pad0;
    float intensity[3];
    float pad1;
};
static_assert(sizeof(PointRecord) == 32, "PointLight record");
struct NodeRecord {  // LightSubTree::Node (light_tree.hpp:57-90), 32 bytes packed
    float leftFlux;
    u32 leftOffset;
    u16 leftType;
    u16 rightType;
    u32 rightOffset;
    float rightFlux;
    float center[3];
};
static_assert(sizeof(NodeRecord) == 32, "LightSubTree::Node record");

struct Builder {
    std::vector<uint8_t>& mem;
    const EmissionLookup& emission;

    float flux_of(u32 offset, u16 type) const {
        const uint8_t* p = mem.data() + offset;
        if (type == kInternalNode) {
            NodeRecord n;
            std::memcpy(&n, p, sizeof(n));
            return n.leftFlux + n.rightFlux;
        }
        if (type == kPointLight) {
            PointRecord r;
            std::memcpy(&r, p, sizeof(r));
            const float f0 = r.intensity[0] * 4.f * kPi, f1 = r.intensity[1] * 4.f * kPi,
                        f2 = r.intensity[2] * 4.f * kPi;
            return f0 + f1 + f2;
        }
        if (type == kAreaLightQuad) return quad_flux(p);
        if (type == kAreaLightSphere) return sphere_flux(p);
        // area triangle: 128 golden-ratio samples of the emission (lights.cpp:24-41)
        TriRecord r;
        std::memcpy(&r, p, sizeof(r));
        const V3 e1{r.posV[1][0], r.posV[1][1], r.posV[1][2]}, e2{r.posV[2][0], r.posV[2][1], r.posV[2][2]};
        const float area = len(cross(e1, e2)) / 2.0f;
        u32 seed;
        std::memcpy(&seed, &area, 4);
        GoldenRatio2D gen(seed);
        float sum[3] = {0.0f, 0.0f, 0.0f};
        for (int i = 0; i < 128; ++i) {
            const u64 x = gen.next();
            const float ux = sample_uniform((u32)(x & 0xffffffffull)), uy = sample_uniform((u32)(x >> 32));
            const float s = 1.f - std::sqrt(1.f - ux);
            const float t = (1.f - s) * uy;
            const float u = r.uvV[0][0] + r.uvV[1][0] * s + r.uvV[2][0] * t;
            const float v = r.uvV[0][1] + r.uvV[1][1] * s + r.uvV[2][1] * t;
            float rad[3];
            emission.radiance(r.material, u, v, rad);
            for (int k = 0; k < 3; ++k) sum[k] += rad[k];
        }
        float f[3];
        for (int k = 0; k < 3; ++k) f[k] = sum[k] / 128.0f * area * 2 * kPi;
        return f[0] + f[1] + f[2];
    }
    // lights.cpp:43-60: 128 golden-ratio samples of the local area (the
    // parametric density |dP/du1 x dP/du0|) times the emission. The reference
    // seeds the sampler with the record's memory address, which no rerun
    // reproduces; this build seeds it with the bits of the parallelogram area
    // |posV[1] x posV[2]| (the triangle's rule, lights.cpp:28).
    float quad_flux(const uint8_t* p) const {
        QuadRecord r;
        std::memcpy(&r, p, sizeof(r));
        V3 pv[4];
        for (int k = 0; k < 4; ++k) pv[k] = point_of(r.posV[k]);
        const float seedArea = len(cross(pv[1], pv[2]));
        u32 seed;
        std::memcpy(&seed, &seedArea, 4);
        GoldenRatio2D gen(seed);
        float sum[3] = {0.0f, 0.0f, 0.0f};
        for (int i = 0; i < 128; ++i) {
            const u64 x = gen.next();
            const float ux = sample_uniform((u32)(x & 0xffffffffull)), uy = sample_uniform((u32)(x >> 32));
            const float u = r.uvV[0][0] + r.uvV[1][0] * uy + r.uvV[2][0] * ux + r.uvV[3][0] * (ux * uy);
            const float v = r.uvV[0][1] + r.uvV[1][1] * uy + r.uvV[2][1] * ux + r.uvV[3][1] * (ux * uy);
            const V3 tangentX = add(pv[1], mul(pv[3], ux));
            const V3 tangentY = add(pv[2], mul(pv[3], uy));
            const float area = len(cross(tangentY, tangentX));
            float rad[3];
            emission.radiance(r.material, u, v, rad);
            for (int k = 0; k < 3; ++k) sum[k] += area * rad[k];
        }
        float f[3];
        for (int k = 0; k < 3; ++k) f[k] = sum[k] / 128.0f * 2 * kPi;
        return f[0] + f[1] + f[2];
    }
    // lights.cpp:63-80: the emission sampled at 128 golden-ratio lon-lat
    // points, times the surface 4 pi r^2 (the role of ei::surface(Sphere)).
    float sphere_flux(const uint8_t* p) const {
        SphereRecord r;
        std::memcpy(&r, p, sizeof(r));
        const float area = 4 * kPi * r.radius * r.radius;
        u32 seed;
        std::memcpy(&seed, &area, 4);
        GoldenRatio2D gen(seed);
        float sum[3] = {0.0f, 0.0f, 0.0f};
        for (int i = 0; i < 128; ++i) {
            const u64 x = gen.next();
            const float ux = sample_uniform((u32)(x & 0xffffffffull)), uy = sample_uniform((u32)(x >> 32));
            const float theta = std::acos(ux * 2.0f - 1.0f) / kPi;
            const float phi = uy;
            float rad[3];
            emission.radiance(r.material, phi, theta, rad);
            for (int k = 0; k < 3; ++k) sum[k] += rad[k];
        }
        float f[3];
        for (int k = 0; k < 3; ++k) f[k] = sum[k] / 128.0f * area * 2 * kPi;
        return f[0] + f[1] + f[2];
    }
    V3 center_of(u32 offset, u16 type) const {
        const uint8_t* p = mem.data() + offset;
        if (type == kInternalNode) {
            NodeRecord n;
            std::memcpy(&n, p, sizeof(n));
            return {n.center[0], n.center[1], n.center[2]};
        }
        if (type == kPointLight || type == kAreaLightSphere) {  // position first in both records
            PointRecord r;
            std::memcpy(&r, p, sizeof(r));
            return {r.position[0], r.position[1], r.position[2]};
        }
        if (type == kAreaLightQuad) {  // lights.hpp:221-224
            QuadRecord r;
            std::memcpy(&r, p, sizeof(r));
            return add(add(point_of(r.posV[0]), mul(add(point_of(r.posV[1]), point_of(r.posV[2])), 0.5f)),
                       mul(point_of(r.posV[3]), 0.25f));
        }
        TriRecord r;
        std::memcpy(&r, p, sizeof(r));
        const V3 v0{r.posV[0][0], r.posV[0][1], r.posV[0][2]}, v1{r.posV[1][0], r.posV[1][1], r.posV[1][2]},
            v2{r.posV[2][0], r.posV[2][1], r.posV[2][2]};
        return add(v0, divs(add(v1, v2), 3.0f));
    }
    void make_node(size_t index, u32 lOff, u16 lType, u32 rOff, u16 rType) {
        NodeRecord n;
        n.leftFlux = flux_of(lOff, lType);
        n.leftOffset = lOff;
        n.leftType = lType;
        n.rightType = rType;
        n.rightOffset = rOff;
        n.rightFlux = flux_of(rOff, rType);
        const V3 c = divs(add(center_of(lOff, lType), center_of(rOff, rType)), 2.0f);
        n.center[0] = c.x;
        n.center[1] = c.y;
        n.center[2] = c.z;
        std::memcpy(mem.data() + index * kNodeBytes, &n, sizeof(n));
    }
};

size_t num_internal_nodes(size_t n) {
    if (n <= 1) return 0;
    const size_t height = (size_t)std::log2((double)n);
    return ((size_t(1) << height) - 1) + (n - (size_t(1) << height));
}

}  // namespace

void build_light_tree(const std::vector<HostAreaLight>& area, const std::vector<HostPointLight>& points,
                      const float sceneBox[6], const EmissionLookup& emission, LightTreeHost& out) {
    out = LightTreeHost();
    std::vector<PosLight> lights;
    for (const auto& t : area) {
        lights.push_back(PosLight{true, t, {}});
        if (t.kind == kAreaSphere) out.sphereLights = true;
    }
    for (const auto& p : points) lights.push_back(PosLight{false, {}, p});
    const size_t n = lights.size();
    out.posLightCount = (u32)n;
    if (n == 0) return;

    const V3 scale{sceneBox[3] - sceneBox[0], sceneBox[4] - sceneBox[1], sceneBox[5] - sceneBox[2]};
    std::sort(lights.begin(), lights.end(), [&scale](const PosLight& a, const PosLight& b) {
        const V3 ca = sort_center(a), cb = sort_center(b);
        const u32 x[3] = {f2u(ca.x * scale.x), f2u(ca.y * scale.y), f2u(ca.z * scale.z)};
        const u32 y[3] = {f2u(cb.x * scale.x), f2u(cb.y * scale.y), f2u(cb.z * scale.z)};
        u64 codeA = hilbert_code((u16)(x[0] >> 16u), (u16)(x[1] >> 16u), (u16)(x[2] >> 16u));
        u64 codeB = hilbert_code((u16)(y[0] >> 16u), (u16)(y[1] >> 16u), (u16)(y[2] >> 16u));
        if (codeA == codeB) {
            codeA = hilbert_code((u16)(x[0] & 0xffff), (u16)(x[1] & 0xffff), (u16)(x[2] & 0xffff));
            codeB = hilbert_code((u16)(y[0] & 0xffff), (u16)(y[1] & 0xffff), (u16)(y[2] & 0xffff));
        }
        return codeA < codeB;
    });

    const size_t nInt = num_internal_nodes(n);
    out.posInternalCount = (u32)nInt;
    std::vector<u32> offsets(n);
    std::vector<u16> types(n);
    size_t memSize = nInt * kNodeBytes;
    for (size_t i = 0; i < n; ++i) {
        offsets[i] = (u32)memSize;
        const int kind = lights[i].isArea ? lights[i].area.kind : -1;
        types[i] = kind == kAreaTriangle ? kAreaLightTriangle
                                         : (kind == kAreaQuad ? kAreaLightQuad
                                                              : (kind == kAreaSphere ? kAreaLightSphere : kPointLight));
        memSize += kind == kAreaTriangle ? kTriBytes
                                         : (kind == kAreaQuad ? kQuadBytes : (kind == kAreaSphere ? kSphereBytes : kPointBytes));
    }
    out.mem.assign(memSize, 0);
    for (size_t i = 0; i < n; ++i) {
        uint8_t* dst = out.mem.data() + offsets[i];
        if (lights[i].isArea && lights[i].area.kind == kAreaQuad) {
            // AreaLightQuad::operator= (lights.hpp:112-123)
            const HostAreaLight& d = lights[i].area;
            QuadRecord r;
            std::memset(&r, 0, sizeof(r));
            for (int k = 0; k < 3; ++k) {
                r.posV[0][k] = d.points[0][k];
                r.posV[1][k] = d.points[3][k] - d.points[0][k];
                r.posV[2][k] = d.points[1][k] - d.points[0][k];
                r.posV[3][k] = d.points[2][k] - d.points[3][k] - r.posV[2][k];
            }
            for (int k = 0; k < 2; ++k) {
                r.uvV[0][k] = d.uv[0][k];
                r.uvV[1][k] = d.uv[3][k] - d.uv[0][k];
                r.uvV[2][k] = d.uv[1][k] - d.uv[0][k];
                r.uvV[3][k] = d.uv[2][k] - d.uv[3][k] - r.uvV[2][k];
            }
            r.material = d.material;
            std::memcpy(dst, &r, sizeof(r));
        } else if (lights[i].isArea && lights[i].area.kind == kAreaSphere) {
            const HostAreaLight& d = lights[i].area;
            SphereRecord r;
            std::memset(&r, 0, sizeof(r));
            std::memcpy(r.position, d.points[0], 12);
            r.radius = d.radius;
            r.material = d.material;
            std::memcpy(dst, &r, sizeof(r));
        } else if (lights[i].isArea) {
            const HostAreaLight& d = lights[i].area;
            TriRecord r;
            std::memset(&r, 0, sizeof(r));
            for (int k = 0; k < 3; ++k) {
                r.posV[0][k] = d.points[0][k];
                r.posV[1][k] = d.points[1][k] - d.points[0][k];
                r.posV[2][k] = d.points[2][k] - d.points[0][k];
            }
            for (int k = 0; k < 2; ++k) {
                r.uvV[0][k] = d.uv[0][k];
                r.uvV[1][k] = d.uv[1][k] - d.uv[0][k];
                r.uvV[2][k] = d.uv[2][k] - d.uv[0][k];
            }
            r.material = d.material;
            std::memcpy(dst, &r, sizeof(r));
        } else {
            const HostPointLight& d = lights[i].point;
            PointRecord r;
            std::memset(&r, 0, sizeof(r));
            std::memcpy(r.position, d.position, 12);
            std::memcpy(r.intensity, d.intensity, 12);
            std::memcpy(dst, &r, sizeof(r));
        }
    }

    Builder b{out.mem, emission};
    if (n == 1) {
        out.posRootType = types[0];
        out.posRootFlux = b.flux_of(offsets[0], types[0]);
    } else {
        const size_t height = (size_t)std::log2((double)n);
        const size_t extraNodes = n - (size_t(1) << height);
        if (extraNodes > 0) {
            const size_t startNode = (size_t(1) << height) - 1;
            const size_t startLight = n - 2 * extraNodes;
            for (size_t i = 0; i < extraNodes; ++i) {
                const size_t l = startLight + 2 * i, r = l + 1;
                b.make_node(startNode + i, offsets[l], types[l], offsets[r], types[r]);
            }
            const size_t startInnerNode = (size_t(1) << (height - 1)) - 1;
            for (size_t i = 0; i < extraNodes / 2; ++i) {
                const size_t l = startNode + 2 * i, r = l + 1;
                b.make_node(startInnerNode + i, (u32)(l * kNodeBytes), kInternalNode, (u32)(r * kNodeBytes),
                            kInternalNode);
            }
            if (extraNodes % 2 != 0) {
                const size_t l = startNode + extraNodes - 1;
                b.make_node(startInnerNode + extraNodes / 2, (u32)(l * kNodeBytes), kInternalNode, offsets[0],
                            types[0]);
            }
        }
        const size_t startLight = (extraNodes % 2 == 0) ? 0 : 1;
        const size_t levelNodeCount = n - 2 * extraNodes - startLight;
        const size_t nodeCount = levelNodeCount / 2;
        const size_t startNodeIndex = (size_t(1) << height) - nodeCount - 1;
        for (size_t i = 0; i < nodeCount; ++i) {
            const size_t l = startLight + 2 * i, r = l + 1;
            b.make_node(startNodeIndex + i, offsets[l], types[l], offsets[r], types[r]);
        }
        if (height > 1) {
            for (size_t level = height - 1; level >= 1; --level) {
                const size_t nodes = size_t(1) << level;
                const size_t innerNode = nodes / 2 - 1;
                for (size_t i = 0; i < nodes / 2; ++i) {
                    const size_t l = nodes - 1 + 2 * i, r = l + 1;
                    b.make_node(innerNode + i, (u32)(l * kNodeBytes), kInternalNode, (u32)(r * kNodeBytes),
                                kInternalNode);
                }
            }
        }
        out.posRootType = kInternalNode;
        NodeRecord root;
        std::memcpy(&root, out.mem.data(), sizeof(root));
        out.posRootFlux = root.leftFlux + root.rightFlux;
    }

    // Primitive -> tree path (fill_map, light_tree.cpp:268-302); u32 wrap-around as in the reference.
    std::vector<std::pair<u64, u32>> map;
    {
        int height = 0;
        while ((size_t(1) << (height + 1)) <= n) ++height;  // ilog2
        const u32 extraNodes = (u32)n - (1u << height);
        const u32 lowerLightCount = (1u << height) - extraNodes;
        u32 i = 0;
        auto key = [](const HostAreaLight& t) { return ((u64)(u32)t.instance << 32) | (u64)(u32)t.primitive; };
        if (extraNodes > 0) {
            for (; i < lowerLightCount; ++i) {
                if (!lights[i].isArea) continue;
                u32 code = ((1u << height) - 1) - (lowerLightCount - i - 1);
                code <<= (32 - height) & 31;
                map.push_back({key(lights[i].area), code});
            }
            height += 1;
        }
        for (; i < (u32)n; ++i) {
            if (!lights[i].isArea) continue;
            const u32 code = (u32)(i - lowerLightCount) << ((32 - height) & 31);
            map.push_back({key(lights[i].area), code});
        }
    }
    std::sort(map.begin(), map.end());
    for (const auto& kv : map) {
        out.mapKeys.push_back(kv.first);
        out.mapCodes.push_back(kv.second);
    }
}

void build_background(int32_t type, const float color[3], const float scale[3], const float* envRgba, int32_t width,
                      int32_t height, const float sceneBox[6], BackgroundHost& out) {
    out = BackgroundHost{};
    out.type = type;
    for (int k = 0; k < 3; ++k) {
        out.color[k] = color[k];
        out.scale[k] = scale[k];
    }
    const V3 diag{sceneBox[3] - sceneBox[0], sceneBox[4] - sceneBox[1], sceneBox[5] - sceneBox[2]};
    if (type == 0) {
        // compute_constant_flux (background.cpp:125-135); the COLORED descriptor
        // multiplies by the color only (:36-40)
        const float solidAngle = 4 * kPi;
        const float area = 2 * kPi * (diag.x * diag.y + diag.x * diag.z + diag.y * diag.z);
        const float f = solidAngle * area;
        for (int k = 0; k < 3; ++k) out.flux[k] = f * color[k];
    } else {
        out.width = width;
        out.height = height;
        out.sat.assign((size_t)width * height, 0.0f);
        auto read = [&](int x, int y) { return envRgba + 4 * ((size_t)y * width + x); };
        // create_summed_area_table, polar map (texture_sampling.hpp:84-101)
        const float lum[3] = {0.212671f, 0.715160f, 0.072169f};
        float accumLumY = 0.f;
        for (int y = 0; y < height; ++y) {
            const float sinTheta = std::sin(kPi * static_cast<float>(y + 0.5f) / static_cast<float>(height));
            float accumLumX = 0.f;
            for (int x = 0; x < width; ++x) {
                const float* t = read(x, y);
                const float luminance = lum[0] * t[0] + lum[1] * t[1] + lum[2] * t[2] + 0.0f * t[3];
                accumLumX += luminance * sinTheta;
                out.sat[(size_t)y * width + x] = accumLumX;
            }
            accumLumY += accumLumX;
            out.sat[(size_t)y * width + width - 1] = accumLumY;
        }
        // compute_envmap_flux, polar map (background.cpp:78-122); the ENVMAP
        // descriptor scales the flux once more (:46-54)
        const float avgPixelSolidAngle = 4 * kPi / (float)(width * height);
        float flux[3] = {0.0f, 0.0f, 0.0f};
        // cos / sin of a column's phi do not depend on the row: evaluated once
        // per column (the same floats the per-texel evaluation gives)
        std::vector<float> cosPhi(width), sinPhi(width);
        for (int x = 0; x < width; ++x) {
            const float phi = 2 * kPi * (x + 0.5f) / float(width);
            cosPhi[x] = (float)std::cos((double)phi);
            sinPhi[x] = (float)std::sin((double)phi);
        }
        for (int y = 0; y < height; ++y) {
            const float sinTheta = std::sin(kPi * (y + 0.5f) / float(height));
            const float dirY = std::sqrt(1.0f - sinTheta * sinTheta);
            const float w = avgPixelSolidAngle / (2.0f * kPi * kPi * sinTheta);  // per row as well
            for (int x = 0; x < width; ++x) {
                const float* t = read(x, y);
                const V3 dir{sinTheta * cosPhi[x], dirY, sinTheta * sinPhi[x]};
                const float surface = diag.y * diag.z * std::fabs(dir.x) + diag.x * diag.z * std::fabs(dir.y) +
                                      diag.x * diag.y * std::fabs(dir.z);
                for (int k = 0; k < 3; ++k) flux[k] += (t[k] * scale[k]) * w * surface;
            }
        }
        for (int k = 0; k < 3; ++k) out.flux[k] = flux[k] * scale[k];
    }
    out.fluxSum = out.flux[0] + out.flux[1] + out.flux[2];
}

}  // namespace mfh
