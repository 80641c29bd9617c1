// Device state of the wavefront path tracer: per-path state, double-buffered
// ray / hit queues, the shadow queue and the counters the persistent kernels
// drain, all in HBM.
#pragma once
#include "pt_device.hpp"

namespace mfh {

struct WfState {
    int numPaths;
    uint32_t slotMask;  // accumulator slots of the enabled targets
    int chunk;          // queue entries grabbed per wave (multiple of 64)
    int numCounters;
    int qBase;          // counters[qBase + k]: length of queue k (k = 0 .. maxPathLength)
    int shBase;         // counters[shBase + k]: shadow slots of bounce k
    int workBase;       // counters[workBase + j]: work counter of launch j
    int* counters;
    int* iteration;     // device iteration counter (index of the iteration being run)
    float4* tp;         // per path: throughput xyz, footprint
    float* acc;         // per path: 18 accumulator slots, slot-major
    float4* rayO[2];    // per queue slot: origin xyz, forward pdf
    float4* rayD[2];    // direction xyz, path index (int bits)
    float4* rayM[2];    // medium absorption xyz, hit barycentric w1
    float4* rayP[2];    // position of the vertex the ray leaves (camera or surface); only
                        // with the NEE position guide or sphere lights, null otherwise
    float4* hit[2];     // t, instance, primitive (int bits), w0
    float4* shO;        // per shadow entry (slot * neeCount + i): origin xyz, tmax
    float4* shD;        // direction xyz
    float4* shC;        // radiance xyz, lightness
    int2* shPath;       // per shadow slot: path index, candidate bits
    // ray sorting: keys / ids of the radix sort (numPaths each), result in sortIds[0]
    uint64_t* sortKeys[2];
    int32_t* sortIds[2];
    int neeStride;      // shadow entries per shadow slot (the largest NEE count)
    // batches of scenes: path p belongs to scene s with sceneOffset[s] <= p < sceneOffset[s + 1]
    const PtScene* scenes;
    const int* sceneOffset;
    int numScenes;
};

struct WfTraceShape {
    int block, lds, leafBatch;
};
constexpr int kWfNumTraceVariants = 14;
extern const WfTraceShape kWfTraceShapes[kWfNumTraceVariants];

struct WfTuning {
    int traceVariant = 5;  // 256 threads, 16 LDS stack entries, leaf batch 1
    int chunk = 256;
    int sortBounces = 0;   // extension / shadow queues of the first bounces traced in sorted order
    int refill = 0;        // > 0: trace stages refill idle lanes every `refill` traversal steps
    int tailAfter = 0;     // > 0: after the extension rays of this bounce, one launch finishes every path
    int traceGrid = 0;
    int shadeGrid = 0;
    int tailGrid = 0;
};

// Grid of the persistent trace kernels for a variant (occupancy x CUs) and of the shade kernel.
int wf_grid(const PtScene& S, const WfTuning& tune, int numCUs);
int wf_shade_grid(int numCUs);
int wf_tail_grid(int numCUs);
// One iteration: begin, primary, maxPathLength x (shade, shadow, extend), finish.
void launch_pt_wavefront(const PtScene& S, const WfState& W, const WfTuning& tune, BuildScratch& scratch,
                         hipStream_t stream);
// One iteration of a batch of scenes (W.scenes, W.sceneOffset): every stage is
// a single launch over the paths of all scenes; maxLen = the longest maximum
// path length of the batch.
void launch_pt_wavefront_batch(const WfState& W, const WfTuning& tune, int maxLen, hipStream_t stream);
int wf_grid_batch(const WfTuning& tune, int numCUs);

}  // namespace mfh
