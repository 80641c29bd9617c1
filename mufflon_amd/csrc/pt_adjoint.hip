// Adjoint (path replay) of the path tracer for gfx950: gradients of a scalar
// loss of the Radiance target with respect to albedo textures (LAMBERT and
// TORRANCE albedo), roughness textures (TORRANCE, WALTER, MICROFACET) and the
// emission scale of every emissive material.
//
// The primal image of `iterations` iterations from a reset is
// I = mean_s L_s over one path per pixel and iteration. Given
// delta = dLoss/dI per pixel and channel, the kernel re-seeds the per-pixel
// RNGs exactly as the reset did and replays the same paths with the same
// code as the megakernel (pt_path.hpp), recording per vertex the albedo
// parameter, the delta-weighted next-event and emission contributions and
// the Russian-roulette data. A backward pass over the recorded vertices then
// forms one of two derivatives (AdjointState::fixedDecisions).
//
// Derivative of the expected image (fixedDecisions = 0, the default). Each
// contribution C of the estimator is w f / p: a path's measurement f, the
// pdf p of the random decisions that produced it (BSDF and light sampling,
// the light-tree choice, the Russian-roulette survivals) and its MIS weight
// w. With the decisions' distribution depending on the parameter theta,
//   d/dtheta E[C] = E[C (d ln w + d ln f)]
// (the score d ln p of the sampling and the derivative -d ln p of 1/p
// cancel), so every contribution is multiplied by the d ln of the BSDF
// values it contains, plus its MIS weight's d ln:
//   * albedo: one factor albedo_j[c] per vertex j that scatters toward C
//     (the walk from j or the next-event estimate at j): d ln f = 1/albedo;
//   * roughness: d ln value of the walk's BSDF at its sampled direction,
//     d ln value of a connection's BSDF, the hit MIS weight
//     (1 - mis) d ln pdf and the next-event MIS weight
//     -mis pdf_hit / pdf_light d ln pdf (dln_mat_evaluate);
//   * emission scale: d ln f = 1/scale for a contribution of that emitter
//     (the light-tree choice probabilities need no term).
// Its expectation is the derivative of the expected image.
//
// Derivative of this realization (fixedDecisions = 1): the estimator with
// every random number, light and lobe choice, roulette decision and sampled
// direction held fixed, which is what central differences with common random
// numbers see (where no decision flips) of renders that replay the same
// directions. On top of the factors above, the walk weight value * cos / pdf
// of a roughness-textured vertex contributes d ln value - d ln pdf, and the
// continuation probability of each later walk m, cont_m =
// min(max_a T'_m[a] + 0.1, 1), follows the throughput: with
//   D_m = T'_m[a*] / cont_m * (d ln walk_j[a*] - sum_{j<=l<m} D_l)
// (zero when cont_m = 1) every contribution after walk m picks up -D_m.
// These terms have a non-zero mean: the roulette's boundary (a decision that
// flips) and the moving sampling pdf carry the rest of the derivative of the
// expectation, which the default form accounts for in expectation.
//
// Texel gradients follow the texture's own sampling weights (bilinear or
// nearest, 1x1 constant) and are added with float atomics.
#include "pt_adjoint.hpp"
#include "pt_path.hpp"

namespace mfh {

namespace {

using namespace shade;
using namespace path;

constexpr int kAdjThreads = 128;

struct AdjVertex {
    int tex;        // albedo texture of the vertex's material, -1 when it has no albedo parameter
    int rtex;       // roughness texture, -1 when the roughness is not textured
    float hu, hv;   // uv of the hit
    V3 rho;         // fetched albedo
    V3 neeW;        // delta-weighted next-event contributions at this vertex
    V3 emW;         // delta-weighted emission contribution of this vertex
    float misC;     // 1 - MIS weight of that emission (0 without MIS)
    float cont;     // continuation probability of the walk from this vertex
    float tp;       // throughput before the roulette in its largest channel
    int astar;      // that channel
    V3 G;           // weighted contributions after the walk from this vertex
    float rw[2];    // d ln(walk weight) / d(rx, ry) at the sampled direction
    float rp[2];    // d ln(pdf of the sampled direction) / d(rx, ry)
    float rn[2];    // sum over next-event estimates: weighted contribution * d ln / d(rx, ry)
};

__device__ __forceinline__ void add_emission_grad(const PtScene& S, const AdjointState& A, int mat, V3 weighted) {
    const PtMaterialRecord& m = S.materials[mat];
    const float w[3] = {weighted.x, weighted.y, weighted.z};
#pragma unroll
    for (int c = 0; c < 3; ++c)
        if (m.scale[c] != 0.0f && w[c] != 0.0f) atomicAdd(A.emissionGrad + 3 * mat + c, w[c] / m.scale[c]);
}

// Texel weights of a texture lookup, as tex_sample forms them.
__device__ __forceinline__ int texel_weights(const DevTexture& t, float u, float v, int* idx, float* w) {
    if (t.width == 1 && t.height == 1) {
        idx[0] = 0;
        w[0] = 1.0f;
        return 1;
    }
    if (t.sampling == 0) {
        idx[0] = imod(ifloor(v * (float)t.height), t.height) * t.width + imod(ifloor(u * (float)t.width), t.width);
        w[0] = 1.0f;
        return 1;
    }
    float fx = u * (float)t.width, fy = v * (float)t.height;
    const int ix = ifloor(fx), iy = ifloor(fy);
    fx -= (float)ix;
    fy -= (float)iy;
    const int x0 = imod(ix, t.width), x1 = imod(ix + 1, t.width);
    const int y0 = imod(iy, t.height), y1 = imod(iy + 1, t.height);
    idx[0] = y0 * t.width + x0;
    idx[1] = y0 * t.width + x1;
    idx[2] = y1 * t.width + x0;
    idx[3] = y1 * t.width + x1;
    w[0] = (1.0f - fx) * (1.0f - fy);
    w[1] = fx * (1.0f - fy);
    w[2] = (1.0f - fx) * fy;
    w[3] = fx * fy;
    return 4;
}

template <int kMode>
__global__ __launch_bounds__(kAdjThreads) void k_pt_adjoint(PtScene S, AdjointState A) {
    __shared__ int sStack[kLdsDepth * kAdjThreads];
    const int pixel = blockIdx.x * kAdjThreads + threadIdx.x;
    if (pixel >= S.width * S.height) return;
    const int px = pixel % S.width, py = pixel / S.width;
    int* const lstack = sStack + threadIdx.x;
    const float* dl = A.delta + 3 * (size_t)pixel;
    const float inv = 1.0f / (float)A.iterations;
    const V3 delta = v3(dl[0] * inv, dl[1] * inv, dl[2] * inv);
    if (!nonzero(delta)) return;
    u64 rng = A.rngs[pixel];
    AdjVertex rec[kAdjMaxVertices + 1];
    for (int it = 0; it < A.iterations; ++it) {
        // ---- forward replay (the megakernel's pt_sample)
        float guideWeight = 1.0f, footA = 0.0f;
        V3 throughput = v3(1.0f, 1.0f, 1.0f);
        Vertex vx;
        pcg_next(rng);
        int pathLen = 0;
        rec[0].neeW = rec[0].emW = v3(0, 0, 0);
        rec[0].tex = rec[0].rtex = -1;
        rec[0].cont = 1.0f;
        rec[0].misC = 0.0f;
        do {
            if (pathLen > 0 && pathLen + 1 >= S.minPathLength && pathLen + 1 <= S.maxPathLength) {
                const u64 neeSeed = pcg_next(rng);
                for (int i = 0; i < S.neeCount; ++i) {
                    const u64 x = pcg_next(rng);
                    NeeCandidate c;
                    if (nee_candidate(S, vx, throughput, guideWeight, i, neeSeed, x, c)) {
                        const TraceHit sh = trace_mode<kMode, true, kAdjThreads, kLdsDepth, kPtLeafBatch, false>(
                            S.trace, c.a, c.dir, c.tmax, lstack);
                        if (sh.inst == -1) {
                            AdjVertex& r = rec[pathLen];
                            const V3 wgt = delta * c.radiance;
                            r.neeW = r.neeW + wgt;
                            if (c.lightMaterial >= 0) add_emission_grad(S, A, c.lightMaterial, wgt);
                            float dv[2], dp[2];
                            if (r.rtex >= 0 && dln_mat_evaluate(vx.ts, vx.mat, vx.incident(), c.lightDir, dv, dp)) {
                                const float w = wgt.x + wgt.y + wgt.z;
                                r.rn[0] += w * (dv[0] - c.misFactor * dp[0]);
                                r.rn[1] += w * (dv[1] - c.misFactor * dp[1]);
                            }
                        }
                    }
                }
            }
            const u64 x0 = pcg_next(rng);
            const u64 x1 = pcg_next(rng);
            const float rndRoulette = sample_uniform((u32)pcg_next(rng));
            V3 rayOrigin, dir, absorption, tPrime;
            float pdfF, cont = 1.0f;
            if (pathLen == 0) {
                camera_ray(S, px, py, x0, dir, pdfF);
                rayOrigin = v3(S.camPos[0], S.camPos[1], S.camPos[2]);
                absorption = v3(0.0f, 0.0f, 0.0f);
            } else {
                const bool ok =
                    walk_surface(vx, x0, x1, rndRoulette, throughput, rayOrigin, dir, pdfF, absorption, tPrime, cont);
                AdjVertex& r = rec[pathLen];
                r.cont = cont;
                if (ok && r.rtex >= 0) {
                    float dv[2], dp[2];
                    dln_mat_evaluate(vx.ts, vx.mat, vx.incident(), dir, dv, dp);
                    r.rw[0] = A.fixedDecisions ? dv[0] - dp[0] : dv[0];
                    r.rw[1] = A.fixedDecisions ? dv[1] - dp[1] : dv[1];
                    r.rp[0] = pdfF < 1e9f ? dp[0] : 0.0f;
                    r.rp[1] = pdfF < 1e9f ? dp[1] : 0.0f;
                }
                r.astar = (tPrime.x < tPrime.y) ? ((tPrime.y < tPrime.z) ? 2 : 1) : ((tPrime.x < tPrime.z) ? 2 : 0);
                r.tp = r.astar == 0 ? tPrime.x : (r.astar == 1 ? tPrime.y : tPrime.z);
                if (!ok) break;
            }
            const TraceHit hit = trace_mode<kMode, false, kAdjThreads, kLdsDepth, kPtLeafBatch, false>(
                S.trace, rayOrigin, dir, kMaxSceneSize, lstack);
            throughput = throughput * transmission(Medium{1.0f, absorption}, hit.inst == -1 ? kMaxSceneSize : hit.t);
            footA += 1.0f / fsqrt(pdfF);
            guideWeight = 1.f / (1.f + footA * footA / 1.0f);
            const V3 prevPos = pathLen == 0 ? rayOrigin : vx.position;
            make_vertex(S, rayOrigin, dir, pdfF, hit.t, hit.inst, hit.prim, hit.w0, hit.w1, vx);
            ++pathLen;
            AdjVertex& r = rec[pathLen];
            r.neeW = r.emW = v3(0, 0, 0);
            r.tex = r.rtex = -1;
            r.cont = 1.0f;
            r.misC = 0.0f;
            r.rw[0] = r.rw[1] = r.rp[0] = r.rp[1] = r.rn[0] = r.rn[1] = 0.0f;
            r.hu = vx.hu;
            r.hv = vx.hv;
            r.rho = vx.mat.albedo;
            int matIdx = -1;
            if (!vx.isVoid) {
                matIdx = material_index(S, vx.inst, vx.prim);
                if (vx.mat.type == kMatLambert || vx.mat.type == kMatTorrance) r.tex = S.materials[matIdx].albedoTexture;
                if (mat_has_roughness(vx.mat.type))
                    r.rtex = S.materials[matIdx].roughnessTexture;
            }
            if (pathLen >= S.minPathLength) {
                V3 em;
                float mis;
                const V3 cc = emission_contribution(S, vx, pathLen, throughput, em, prevPos, mis);
                r.emW = delta * cc;
                r.misC = 1.0f - mis;
                if (matIdx >= 0 && vx.mat.type == kMatEmissive) add_emission_grad(S, A, matIdx, r.emW);
            }
            if (vx.isVoid) break;
        } while (pathLen < S.maxPathLength && pathLen < kAdjMaxVertices);
        // ---- backward: G[v] = sum of the weighted contributions after the walk from v
        const int last = pathLen;
        rec[last].G = v3(0, 0, 0);
        for (int v = last - 1; v >= 0; --v) rec[v].G = rec[v + 1].G + rec[v + 1].emW + rec[v + 1].neeW;
        for (int j = 1; j <= last; ++j) {
            if (rec[j].rtex >= 0 && A.texGrad[rec[j].rtex]) {
                // dL / d(rx, ry) of vertex j (see the header)
                const float Gj = rec[j].G.x + rec[j].G.y + rec[j].G.z;
                const float em1 = j < last ? (rec[j + 1].emW.x + rec[j + 1].emW.y + rec[j + 1].emW.z) * rec[j + 1].misC
                                           : 0.0f;
                float g[2];
                for (int q = 0; q < 2; ++q) {
                    g[q] = rec[j].rw[q] * Gj + rec[j].rn[q] + em1 * rec[j].rp[q];
                    float sumD = 0.0f;
                    for (int m = j; A.fixedDecisions && m < last; ++m) {
                        if (!(rec[m].cont < 1.0f)) continue;
                        const float D = rec[m].tp / rec[m].cont * (rec[j].rw[q] - sumD);
                        g[q] -= D * (rec[m].G.x + rec[m].G.y + rec[m].G.z);
                        sumD += D;
                    }
                }
                const DevTexture& tx = S.textures[rec[j].rtex];
                float* grad = A.texGrad[rec[j].rtex];
                int idx[4];
                float w[4];
                const int nt = texel_weights(tx, rec[j].hu, rec[j].hv, idx, w);
                // one channel drives both roughnesses; otherwise x and y
                const float gc[2] = {tx.channels > 1 ? g[0] : g[0] + g[1], g[1]};
                for (int ch = 0; ch < (tx.channels > 1 ? 2 : 1); ++ch)
                    if (gc[ch] != 0.0f && gc[ch] == gc[ch])
                        for (int q = 0; q < nt; ++q) atomicAdd(grad + 4 * (size_t)idx[q] + ch, w[q] * gc[ch]);
            }
            if (rec[j].tex < 0) continue;
            float* grad = A.texGrad[rec[j].tex];
            if (!grad) continue;
            int idx[4];
            float w[4];
            const int nt = texel_weights(S.textures[rec[j].tex], rec[j].hu, rec[j].hv, idx, w);
            const float rho[3] = {rec[j].rho.x, rec[j].rho.y, rec[j].rho.z};
            const float gdir[3] = {rec[j].G.x + rec[j].neeW.x, rec[j].G.y + rec[j].neeW.y, rec[j].G.z + rec[j].neeW.z};
            for (int c = 0; c < 3; ++c) {
                if (rho[c] == 0.0f) continue;
                float g = gdir[c] / rho[c];
                float sumD = 0.0f;
                for (int m = j; A.fixedDecisions && m < last; ++m) {  // walks from m that continued to vertex m + 1
                    if (!(rec[m].cont < 1.0f)) continue;
                    const float D = rec[m].tp / rec[m].cont * ((rec[m].astar == c ? 1.0f / rho[c] : 0.0f) - sumD);
                    g -= D * (rec[m].G.x + rec[m].G.y + rec[m].G.z);
                    sumD += D;
                }
                if (g != 0.0f && g == g)
                    for (int q = 0; q < nt; ++q) atomicAdd(grad + 4 * (size_t)idx[q] + c, w[q] * g);
            }
        }
    }
}

__global__ void k_adjoint_rng(u64* rngs, int num, int seed) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= num) return;
    const u32 s = (u32)(i + seed);
    u64 st = scramble_seed(s);
    pcg_next(st);
    st += s;
    pcg_next(st);
    rngs[i] = st;
}

}  // namespace

void launch_pt_adjoint(const PtScene& S, const AdjointState& A, int seed, hipStream_t stream) {
    const int num = S.width * S.height;
    hipLaunchKernelGGL(k_adjoint_rng, dim3((num + 255) / 256), dim3(256), 0, stream, A.rngs, num, seed);
    const dim3 grid((num + kAdjThreads - 1) / kAdjThreads);
    const int mode = trace_mode_of(S.trace);
    if (mode == kTraceSingle)
        hipLaunchKernelGGL(HIP_KERNEL_NAME(k_pt_adjoint<kTraceSingle>), grid, dim3(kAdjThreads), 0, stream, S, A);
    else if (mode == kTraceMulti)
        hipLaunchKernelGGL(HIP_KERNEL_NAME(k_pt_adjoint<kTraceMulti>), grid, dim3(kAdjThreads), 0, stream, S, A);
    else
        hipLaunchKernelGGL(HIP_KERNEL_NAME(k_pt_adjoint<kTraceMixed>), grid, dim3(kAdjThreads), 0, stream, S, A);
    MFH_HIP(hipGetLastError());
}

}  // namespace mfh
