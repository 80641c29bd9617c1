// Oracle — test infrastructure only.
//
// An independent FP64 Monte Carlo estimator of the same pixel integral the
// path tracer estimates, for converged reference images and for checking the
// float restatement (pt_oracle.cpp) in expectation. It shares no code with
// that restatement: its own BVH, ray/triangle test, random numbers, light
// choice and estimator. Only the integrand is common, and it is the one the
// reference defines:
//   * pinhole camera, box pixel filter: a pixel is the mean radiance over
//     directions through uniformly jittered points of the pixel
//     (pinhole.hpp:56-84);
//   * paths of at most maxPathLength segments, the camera segment counted
//     (pt_common.hpp:75-156: NEE at vertex k needs k + 1 <= max, the walk
//     stops at max);
//   * Lambert: f = albedo / pi, cosine taken against the shading normal, zero
//     unless the incident and excident directions lie on the same side of the
//     shading normal and each agrees in side with the geometric normal
//     (lambert.hpp:37-49, material_sampling.hpp:68-237);
//   * emissive triangles emit radiance * scale on the side their geometric
//     normal faces (emissive.hpp:39-44) and do not scatter;
//   * point lights: intensity / d^2 in vacuum (light_sampling.hpp:308-327);
//   * a uniform background radiance seen by rays that leave the scene.
// Strategy (unbiased, no MIS, no Russian roulette): emitters are reached by
// next-event estimation only (an emissive triangle chosen in proportion to
// its area, a uniform point on it; every point light), except when the camera
// sees them directly; the background is reached by BSDF sampling only
// (cosine-weighted about the shading normal).
// Everything is in double; random numbers are SplitMix64 (Steele, Lea and
// Flood 2014) per (pixel, sample) stream.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

namespace {

struct D3 {
    double x, y, z;
};
inline D3 operator+(D3 a, D3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
inline D3 operator-(D3 a, D3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline D3 operator*(D3 a, double s) { return {a.x * s, a.y * s, a.z * s}; }
inline D3 mul(D3 a, D3 b) { return {a.x * b.x, a.y * b.y, a.z * b.z}; }
inline double dot(D3 a, D3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline D3 cross(D3 a, D3 b) { return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
inline D3 unit(D3 a) { return a * (1.0 / std::sqrt(dot(a, a))); }
inline double comp(D3 a, int k) { return k == 0 ? a.x : (k == 1 ? a.y : a.z); }

constexpr double kPi = 3.14159265358979323846;

inline uint64_t mix64(uint64_t z) {  // SplitMix64 output function
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

struct Rng {  // SplitMix64
    uint64_t s;
    uint64_t next() { return mix64(s += 0x9E3779B97F4A7C15ull); }
    double uniform() { return (double)(next() >> 11) * (1.0 / 9007199254740992.0); }
};

struct Bvh {
    struct Node {
        double lo[3], hi[3];
        int32_t left, right;  // children, or (-1 - first, count) for a leaf
    };
    std::vector<Node> nodes;
    std::vector<int32_t> order;
};

struct Scene {
    std::vector<D3> v;       // vertices
    std::vector<D3> n;       // vertex normals
    std::vector<int32_t> t;  // 3 per triangle
    std::vector<D3> geoN;    // unit face normal, cross(v1 - v0, v2 - v0)
    std::vector<int32_t> mat;
    std::vector<int32_t> matType;  // 0 emissive, 1 lambert
    std::vector<D3> albedo, emission;
    std::vector<D3> plPos, plI;
    std::vector<int32_t> emitters;  // emissive triangles
    std::vector<double> emitterCdf;
    double emitterArea = 0.0;
    D3 background{0, 0, 0};
    D3 camPos, camDir, camUp;
    double tanVFov;
    int width, height;
    double eps;
    Bvh bvh;
};

void tri_box(const Scene& S, int tri, double lo[3], double hi[3]) {
    for (int k = 0; k < 3; ++k) {
        lo[k] = 1e300;
        hi[k] = -1e300;
    }
    for (int c = 0; c < 3; ++c) {
        const D3 p = S.v[S.t[3 * tri + c]];
        for (int k = 0; k < 3; ++k) {
            lo[k] = std::min(lo[k], comp(p, k));
            hi[k] = std::max(hi[k], comp(p, k));
        }
    }
}

int build_node(Scene& S, std::vector<D3>& cent, int first, int count) {
    Bvh::Node node;
    for (int k = 0; k < 3; ++k) {
        node.lo[k] = 1e300;
        node.hi[k] = -1e300;
    }
    double clo[3] = {1e300, 1e300, 1e300}, chi[3] = {-1e300, -1e300, -1e300};
    for (int i = first; i < first + count; ++i) {
        double lo[3], hi[3];
        const int tri = S.bvh.order[i];
        tri_box(S, tri, lo, hi);
        for (int k = 0; k < 3; ++k) {
            node.lo[k] = std::min(node.lo[k], lo[k]);
            node.hi[k] = std::max(node.hi[k], hi[k]);
            clo[k] = std::min(clo[k], comp(cent[tri], k));
            chi[k] = std::max(chi[k], comp(cent[tri], k));
        }
    }
    const int id = (int)S.bvh.nodes.size();
    S.bvh.nodes.push_back(node);
    int axis = 0;
    for (int k = 1; k < 3; ++k)
        if (chi[k] - clo[k] > chi[axis] - clo[axis]) axis = k;
    if (count <= 4 || chi[axis] - clo[axis] <= 0.0) {
        S.bvh.nodes[id].left = -1 - first;
        S.bvh.nodes[id].right = count;
        return id;
    }
    const int mid = first + count / 2;
    std::nth_element(S.bvh.order.begin() + first, S.bvh.order.begin() + mid, S.bvh.order.begin() + first + count,
                     [&](int a, int b) { return comp(cent[a], axis) < comp(cent[b], axis); });
    const int l = build_node(S, cent, first, mid - first);
    const int r = build_node(S, cent, mid, first + count - mid);
    S.bvh.nodes[id].left = l;
    S.bvh.nodes[id].right = r;
    return id;
}

inline bool box_hit(const Bvh::Node& nd, D3 o, D3 inv, double tmax) {
    double t0 = 0.0, t1 = tmax;
    for (int k = 0; k < 3; ++k) {
        double a = (nd.lo[k] - comp(o, k)) * comp(inv, k);
        double b = (nd.hi[k] - comp(o, k)) * comp(inv, k);
        if (a > b) std::swap(a, b);
        t0 = std::max(t0, a);
        t1 = std::min(t1, b);
        if (t0 > t1) return false;
    }
    return true;
}

// Moeller-Trumbore in double: t and the weights of v1 (u) and v2 (w).
inline bool tri_hit(const Scene& S, int tri, D3 o, D3 d, double& t, double& u, double& w) {
    const D3 v0 = S.v[S.t[3 * tri]];
    const D3 e1 = S.v[S.t[3 * tri + 1]] - v0;
    const D3 e2 = S.v[S.t[3 * tri + 2]] - v0;
    const D3 p = cross(d, e2);
    const double det = dot(e1, p);
    if (det == 0.0) return false;
    const double inv = 1.0 / det;
    const D3 s = o - v0;
    u = dot(s, p) * inv;
    if (u < 0.0 || u > 1.0) return false;
    const D3 q = cross(s, e1);
    w = dot(d, q) * inv;
    if (w < 0.0 || u + w > 1.0) return false;
    t = dot(e2, q) * inv;
    return t > 0.0;
}

struct Hit {
    int tri = -1;
    double t, u, w;
};

// Closest hit (anyHit = false) or any hit before tmax (anyHit = true).
Hit trace(const Scene& S, D3 o, D3 d, double tmax, bool anyHit) {
    Hit h;
    h.t = tmax;
    const D3 inv{1.0 / d.x, 1.0 / d.y, 1.0 / d.z};
    int stack[128];
    int sp = 0;
    stack[sp++] = 0;
    while (sp > 0) {
        const Bvh::Node& nd = S.bvh.nodes[stack[--sp]];
        if (!box_hit(nd, o, inv, h.t)) continue;
        if (nd.left < 0) {
            const int first = -1 - nd.left;
            for (int i = first; i < first + nd.right; ++i) {
                const int tri = S.bvh.order[i];
                double t, u, w;
                if (tri_hit(S, tri, o, d, t, u, w) && t < h.t) {
                    h.tri = tri;
                    h.t = t;
                    h.u = u;
                    h.w = w;
                    if (anyHit) return h;
                }
            }
        } else {
            stack[sp++] = nd.left;
            stack[sp++] = nd.right;
        }
    }
    return h;
}

// Ray origin moved off a surface to the side of direction d.
inline D3 offset(const Scene& S, D3 p, D3 geoN, D3 d) { return p + geoN * (dot(geoN, d) >= 0.0 ? S.eps : -S.eps); }

inline bool visible(const Scene& S, D3 p, D3 geoN, D3 q, D3 qN, bool qIsSurface) {
    D3 a = offset(S, p, geoN, q - p);
    D3 b = qIsSurface ? offset(S, q, qN, p - q) : q;
    D3 d = b - a;
    const double dist = std::sqrt(dot(d, d));
    d = d * (1.0 / dist);
    return trace(S, a, d, dist, true).tri < 0;
}

// Lambert value f * |cos| at a surface point (shading normal sn, geometric
// normal gn) for a ray arriving along inDir and leaving along ex; 0 where the
// reference's side rules reject the pair.
inline double lambert_cos(D3 sn, D3 gn, D3 inDir, D3 ex) {
    const double iN = -dot(inDir, sn), eN = dot(ex, sn);
    const double iG = -dot(inDir, gn), eG = dot(ex, gn);
    if (iN * eN <= 0.0 || eG * eN <= 0.0 || iG * iN <= 0.0) return 0.0;
    return std::fabs(eN) / kPi;
}

D3 sample_path(const Scene& S, int px, int py, int maxLen, Rng& rng) {
    D3 L{0, 0, 0};
    const double u0 = rng.uniform(), u1 = rng.uniform();
    const double w = S.width, h = S.height;
    const double cx = ((px + u0) / w * 2.0 - 1.0) * S.tanVFov * (w / h);
    const double cy = ((py + u1) / h * 2.0 - 1.0) * S.tanVFov;
    D3 d = unit(cross(S.camDir, S.camUp) * cx + S.camUp * cy + S.camDir);
    D3 o = S.camPos;
    D3 beta{1, 1, 1};
    D3 prevGeoN{0, 0, 0};
    bool fromSurface = false;
    for (int k = 1; k <= maxLen; ++k) {
        const D3 ro = fromSurface ? offset(S, o, prevGeoN, d) : o;
        const Hit hit = trace(S, ro, d, 1e30, false);
        if (hit.tri < 0) {
            L = L + mul(beta, S.background);
            break;
        }
        const int tri = hit.tri;
        const D3 x = ro + d * hit.t;
        const D3 gn = S.geoN[tri];
        const int m = S.mat[tri];
        if (S.matType[m] == 0) {  // emitter: counted here only when the camera sees it
            if (k == 1 && dot(gn, d) < 0.0) L = L + mul(beta, S.emission[m]);
            break;
        }
        const int* ix = &S.t[3 * tri];
        const D3 sn = unit(S.n[ix[0]] * (1.0 - hit.u - hit.w) + S.n[ix[1]] * hit.u + S.n[ix[2]] * hit.w);
        if (-dot(d, gn) * -dot(d, sn) <= 0.0) break;  // incident side mismatch: nothing leaves
        const D3 rho = S.albedo[m];
        if (k + 1 <= maxLen) {
            // area lights: one emitter point, chosen by area
            if (!S.emitters.empty()) {
                const double pick = rng.uniform() * S.emitterArea;
                const int j = (int)(std::upper_bound(S.emitterCdf.begin(), S.emitterCdf.end(), pick) -
                                    S.emitterCdf.begin());
                const int lt = S.emitters[std::min(j, (int)S.emitters.size() - 1)];
                double a = rng.uniform(), b = rng.uniform();
                if (a + b > 1.0) {
                    a = 1.0 - a;
                    b = 1.0 - b;
                }
                const D3 v0 = S.v[S.t[3 * lt]];
                const D3 y = v0 + (S.v[S.t[3 * lt + 1]] - v0) * a + (S.v[S.t[3 * lt + 2]] - v0) * b;
                D3 wl = y - x;
                const double dist2 = dot(wl, wl);
                wl = wl * (1.0 / std::sqrt(dist2));
                const double cosL = -dot(S.geoN[lt], wl);
                const double fc = lambert_cos(sn, gn, d, wl);
                if (cosL > 0.0 && fc > 0.0 && visible(S, x, gn, y, S.geoN[lt], true)) {
                    const double g = fc * cosL / dist2 * S.emitterArea;
                    L = L + mul(mul(beta, rho), S.emission[S.mat[lt]]) * g;
                }
            }
            for (size_t p = 0; p < S.plPos.size(); ++p) {
                D3 wl = S.plPos[p] - x;
                const double dist2 = dot(wl, wl);
                wl = wl * (1.0 / std::sqrt(dist2));
                const double fc = lambert_cos(sn, gn, d, wl);
                if (fc > 0.0 && visible(S, x, gn, S.plPos[p], {0, 0, 0}, false))
                    L = L + mul(mul(beta, rho), S.plI[p]) * (fc / dist2);
            }
        }
        if (k == maxLen) break;
        // cosine-weighted continuation about the shading normal, on the incident side
        const double r1 = rng.uniform(), r2 = rng.uniform();
        const double ct = std::sqrt(r1), st = std::sqrt(1.0 - r1), ph = 2.0 * kPi * r2;
        const D3 nz = dot(d, sn) < 0.0 ? sn : sn * -1.0;
        const D3 tx = unit(std::fabs(nz.x) > 0.5 ? cross(nz, D3{0, 1, 0}) : cross(nz, D3{1, 0, 0}));
        const D3 ty = cross(nz, tx);
        const D3 ex = unit(tx * (st * std::cos(ph)) + ty * (st * std::sin(ph)) + nz * ct);
        if (dot(ex, gn) * dot(ex, sn) <= 0.0) break;  // leaves below the geometric surface
        beta = mul(beta, rho);
        o = x;
        d = ex;
        prevGeoN = gn;
        fromSurface = true;
    }
    return L;
}

}  // namespace

extern "C" {

// Geometry (one object in world space), materials as type / albedo /
// emission (radiance * scale) triples, point lights as (position,
// intensity), a uniform background radiance and a pinhole camera.
struct Fp64Scene {
    const float* verts;
    int32_t numVerts;
    const uint32_t* tris;
    int32_t numTris;
    const float* normals;
    const uint16_t* matIdx;
    const int32_t* matType;  // 0 emissive, 1 lambert
    const float* albedo;     // 3 per material
    const float* emission;   // 3 per material
    int32_t numMaterials;
    const float* pointLights;  // 6 per light
    int32_t numPointLights;
    float background[3];
    float camPos[3], camDir[3], camUp[3];
    float tanVFov;
    int32_t width, height;
};

// spp samples per pixel; mean (3 per pixel) and the sample variance of one
// sample (3 per pixel), row-major; pixels run in parallel, each from its own
// streams, so the result depends only on the seed.
int orc64_render(const Fp64Scene* in, int32_t maxPathLength, int32_t spp, uint64_t seed, double* mean,
                 double* var) {
    Scene S;
    S.v.resize(in->numVerts);
    S.n.resize(in->numVerts);
    double lo = 1e300, hi = -1e300;
    for (int i = 0; i < in->numVerts; ++i) {
        S.v[i] = {in->verts[3 * i], in->verts[3 * i + 1], in->verts[3 * i + 2]};
        S.n[i] = {in->normals[3 * i], in->normals[3 * i + 1], in->normals[3 * i + 2]};
        lo = std::min({lo, S.v[i].x, S.v[i].y, S.v[i].z});
        hi = std::max({hi, S.v[i].x, S.v[i].y, S.v[i].z});
    }
    S.eps = 1e-7 * std::max(1.0, hi - lo);
    S.t.assign(in->tris, in->tris + 3 * (size_t)in->numTris);
    S.mat.assign(in->matIdx, in->matIdx + in->numTris);
    for (int m = 0; m < in->numMaterials; ++m) {
        if (in->matType[m] != 0 && in->matType[m] != 1) return 1;
        S.matType.push_back(in->matType[m]);
        S.albedo.push_back({in->albedo[3 * m], in->albedo[3 * m + 1], in->albedo[3 * m + 2]});
        S.emission.push_back({in->emission[3 * m], in->emission[3 * m + 1], in->emission[3 * m + 2]});
    }
    std::vector<D3> cent(in->numTris);
    for (int i = 0; i < in->numTris; ++i) {
        const D3 a = S.v[S.t[3 * i]], b = S.v[S.t[3 * i + 1]], c = S.v[S.t[3 * i + 2]];
        const D3 cr = cross(b - a, c - a);
        const double len2 = dot(cr, cr);
        S.geoN.push_back(len2 > 0.0 ? cr * (1.0 / std::sqrt(len2)) : D3{0, 0, 0});
        cent[i] = (a + b + c) * (1.0 / 3.0);
        if (S.mat[i] >= in->numMaterials) return 2;
        if (S.matType[S.mat[i]] == 0 && len2 > 0.0) {
            S.emitters.push_back(i);
            S.emitterArea += 0.5 * std::sqrt(len2);
            S.emitterCdf.push_back(S.emitterArea);
        }
    }
    for (int p = 0; p < in->numPointLights; ++p) {
        const float* q = in->pointLights + 6 * p;
        S.plPos.push_back({q[0], q[1], q[2]});
        S.plI.push_back({q[3], q[4], q[5]});
    }
    S.background = {in->background[0], in->background[1], in->background[2]};
    S.camPos = {in->camPos[0], in->camPos[1], in->camPos[2]};
    S.camDir = {in->camDir[0], in->camDir[1], in->camDir[2]};
    S.camUp = {in->camUp[0], in->camUp[1], in->camUp[2]};
    S.tanVFov = in->tanVFov;
    S.width = in->width;
    S.height = in->height;
    S.bvh.order.resize(in->numTris);
    for (int i = 0; i < in->numTris; ++i) S.bvh.order[i] = i;
    if (in->numTris == 0) return 3;
    build_node(S, cent, 0, in->numTris);
    const int numPixels = S.width * S.height;
#pragma omp parallel for schedule(dynamic, 16)
    for (int pixel = 0; pixel < numPixels; ++pixel) {
        double sum[3] = {0, 0, 0}, sum2[3] = {0, 0, 0};
        for (int s = 0; s < spp; ++s) {
            // one stream per (pixel, sample): the state is a hash of both
            Rng rng{mix64(seed + mix64(((uint64_t)pixel << 32) | (uint32_t)s))};
            const D3 L = sample_path(S, pixel % S.width, pixel / S.width, maxPathLength, rng);
            for (int k = 0; k < 3; ++k) {
                const double x = comp(L, k);
                sum[k] += x;
                sum2[k] += x * x;
            }
        }
        for (int k = 0; k < 3; ++k) {
            const double m = sum[k] / spp;
            mean[3 * (size_t)pixel + k] = m;
            var[3 * (size_t)pixel + k] = spp > 1 ? std::max(0.0, (sum2[k] - spp * m * m) / (spp - 1)) : 0.0;
        }
    }
    return 0;
}

}  // extern "C"
