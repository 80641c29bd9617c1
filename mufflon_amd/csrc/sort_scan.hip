// Device-wide prefix sums and the stable LSD radix sort of (u64 key, i32 id)
// pairs used by the LBVH builder (the role of CUB's DeviceRadixSort::SortPairs
// and DeviceScan::InclusiveSum in Mufflon, cu_lib_wrapper.cu:13-58,364-405).
//
// Radix sort: 8-bit digits, per pass
//   1. rs_hist    : per-tile digit counts (LDS atomics), digit-major layout;
//   2. exclusive scan of the digit-major counts -> global destination of
//      every (digit, tile) run;
//   3. rs_scatter : stable rank of each key inside its tile (wave64 ballot
//      match over the 8 digit bits + per-wave counts in LDS), the tile is
//      reordered in LDS by digit, then written out as contiguous runs.
// Stability holds because ranks follow (round, wave, lane) order, which is
// the element order inside a tile, and tiles are laid out in order.
#include <algorithm>

#include "mfh_device.hpp"
#include "mfh_internal.hpp"

namespace mfh {

namespace {

constexpr int kScanThreads = 256;
constexpr int kScanIpt = 8;
constexpr int kScanChunk = kScanThreads * kScanIpt;

__device__ __forceinline__ int wave_incl_scan(int v) {
    const int lane = threadIdx.x & 63;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const int t = __shfl_up(v, o, 64);
        if (lane >= o) v += t;
    }
    return v;
}

// Exclusive scan of one value per thread across a 256-thread block; returns
// the block total through `total`.
__device__ __forceinline__ int block_excl_scan(int v, int* sWave, int& total) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int incl = wave_incl_scan(v);
    if (lane == 63) sWave[wave] = incl;
    __syncthreads();
    int waveOff = 0, tot = 0;
#pragma unroll
    for (int w = 0; w < kScanThreads / 64; ++w) {
        const int s = sWave[w];
        if (w < wave) waveOff += s;
        tot += s;
    }
    __syncthreads();
    total = tot;
    return waveOff + incl - v;
}

// The scan kernels take up to two arrays of n values; blockIdx.y picks the
// array (its partial sums follow the first array's in `partials`).
struct ScanArrays {
    const int* in[2];
    int* out[2];
};

__global__ __launch_bounds__(kScanThreads) void k_scan_partials(ScanArrays a, int n, int* partials) {
    __shared__ int sWave[kScanThreads / 64];
    const int* in = a.in[blockIdx.y];
    partials += blockIdx.y * gridDim.x;
    const int base = blockIdx.x * kScanChunk + threadIdx.x * kScanIpt;
    int s = 0;
#pragma unroll
    for (int i = 0; i < kScanIpt; ++i) {
        const int idx = base + i;
        if (idx < n) s += in[idx];
    }
    int total;
    block_excl_scan(s, sWave, total);
    if (threadIdx.x == 0) partials[blockIdx.x] = total;
}

// Single block per array: exclusive scan of the partial sums in place.
__global__ __launch_bounds__(kScanThreads) void k_scan_top(int* partials, int n) {
    __shared__ int sWave[kScanThreads / 64];
    partials += blockIdx.y * n;
    int carry = 0;
    for (int base = 0; base < n; base += kScanChunk) {
        int v[kScanIpt];
        int s = 0;
        const int b = base + threadIdx.x * kScanIpt;
#pragma unroll
        for (int i = 0; i < kScanIpt; ++i) {
            v[i] = (b + i < n) ? partials[b + i] : 0;
            s += v[i];
        }
        int total;
        int off = block_excl_scan(s, sWave, total) + carry;
#pragma unroll
        for (int i = 0; i < kScanIpt; ++i) {
            if (b + i < n) partials[b + i] = off;
            off += v[i];
        }
        carry += total;
        __syncthreads();
    }
}

template <bool kInclusive>
__global__ __launch_bounds__(kScanThreads) void k_scan_apply(ScanArrays a, int n, const int* partials) {
    __shared__ int sWave[kScanThreads / 64];
    const int* in = a.in[blockIdx.y];
    int* out = a.out[blockIdx.y];
    partials += blockIdx.y * gridDim.x;
    const int base = blockIdx.x * kScanChunk + threadIdx.x * kScanIpt;
    int v[kScanIpt];
    int s = 0;
#pragma unroll
    for (int i = 0; i < kScanIpt; ++i) {
        v[i] = (base + i < n) ? in[base + i] : 0;
        s += v[i];
    }
    int total;
    int off = block_excl_scan(s, sWave, total) + partials[blockIdx.x];
#pragma unroll
    for (int i = 0; i < kScanIpt; ++i) {
        if (kInclusive) off += v[i];
        if (base + i < n) out[base + i] = off;
        if (!kInclusive) off += v[i];
    }
}

template <bool kInclusive>
void scan_i32(const ScanArrays& a, int arrays, int n, BuildScratch& s, hipStream_t stream) {
    if (n <= 0) return;
    const int blocks = (n + kScanChunk - 1) / kScanChunk;
    s.partials.reserve(sizeof(int) * (size_t)blocks * arrays);
    int* partials = s.partials.as<int>();
    hipLaunchKernelGGL(k_scan_partials, dim3(blocks, arrays), dim3(kScanThreads), 0, stream, a, n, partials);
    hipLaunchKernelGGL(k_scan_top, dim3(1, arrays), dim3(kScanThreads), 0, stream, partials, blocks);
    hipLaunchKernelGGL(k_scan_apply<kInclusive>, dim3(blocks, arrays), dim3(kScanThreads), 0, stream, a, n,
                       partials);
    MFH_HIP(hipGetLastError());
}

__global__ void k_zero_u32(u32* p, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0u;
}

// ---------------------------------------------------------------- radix sort
constexpr int kRsThreads = 256;
constexpr int kRsTile = kRadixTileKeys;
constexpr int kRsIpt = kRsTile / kRsThreads;
static_assert(kRsIpt * kRsThreads == kRsTile, "radix tile");
constexpr int kRsWaves = kRsThreads / 64;

// Sorts of at most kFusedTiles tiles skip the separate scan: the histogram
// pass writes tile-major counts and adds them into per-group digit sums
// (kGroupTiles tiles a group), and each scatter block derives its runs'
// destinations from the group sums and the counts of the earlier tiles of
// its group.
constexpr int kGroupTiles = 32;
constexpr int kFusedTiles = 1024;

__global__ __launch_bounds__(kRsThreads) void k_rs_hist(const u64* __restrict__ keys, int n, int shift,
                                                       u32* __restrict__ counts, int numTiles,
                                                       u32* __restrict__ groupSums) {
    __shared__ u32 sHist[256];
    sHist[threadIdx.x] = 0;
    __syncthreads();
    const int base = blockIdx.x * kRsTile;
#pragma unroll 4
    for (int r = 0; r < kRsIpt; ++r) {
        const int idx = base + r * kRsThreads + threadIdx.x;
        if (idx < n) {
            const u32 d = (u32)(keys[idx] >> shift) & 255u;
            atomicAdd(&sHist[d], 1u);
        }
    }
    __syncthreads();
    const u32 c = sHist[threadIdx.x];
    if (groupSums) {
        counts[blockIdx.x * 256 + threadIdx.x] = c;
        if (c) atomicAdd(&groupSums[(blockIdx.x / kGroupTiles) * 256 + threadIdx.x], c);
    } else {
        counts[threadIdx.x * numTiles + blockIdx.x] = c;
    }
}

__global__ __launch_bounds__(kRsThreads) void k_rs_scatter(const u64* __restrict__ kin,
                                                          const i32* __restrict__ vin,
                                                          u64* __restrict__ kout, i32* __restrict__ vout,
                                                          int n, int shift, const u32* __restrict__ offs,
                                                          int numTiles, const u32* __restrict__ groupSums) {
    __shared__ u64 sKeys[kRsTile];
    __shared__ i32 sVals[kRsTile];
    __shared__ u32 sRun[256];
    __shared__ u32 sWave[kRsWaves][256];
    __shared__ u32 sStart[256];
    __shared__ u32 sGlob[256];
    __shared__ int sScanWave[kRsWaves];

    const int tid = threadIdx.x;
    const int lane = tid & 63, wave = tid >> 6;
    const int base = blockIdx.x * kRsTile;
    const u64 ltMask = (lane == 0) ? 0ull : ((~0ull) >> (64 - lane));

    sRun[tid] = 0;
    if (groupSums) {
        // destination of this tile's run of digit tid: all keys of smaller
        // digits, then this digit's keys in earlier groups and earlier tiles
        const int g = blockIdx.x / kGroupTiles, groups = (numTiles + kGroupTiles - 1) / kGroupTiles;
        u32 total = 0, before = 0;
#pragma unroll 8
        for (int gg = 0; gg < groups; ++gg) {
            const u32 v = groupSums[gg * 256 + tid];
            total += v;
            before += gg < g ? v : 0u;
        }
#pragma unroll 8
        for (int t = g * kGroupTiles; t < (int)blockIdx.x; ++t) before += offs[t * 256 + tid];
        int all;
        const int digitStart = block_excl_scan((int)total, sScanWave, all);
        sGlob[tid] = (u32)digitStart + before;
    } else {
        sGlob[tid] = offs[tid * numTiles + blockIdx.x];
    }

    u64 key[kRsIpt];
    i32 val[kRsIpt];
    u32 rank[kRsIpt];
#pragma unroll
    for (int r = 0; r < kRsIpt; ++r) {
        const int idx = base + r * kRsThreads + tid;
        key[r] = (idx < n) ? kin[idx] : ~0ull;
        val[r] = (idx < n) ? vin[idx] : 0;
    }
    __syncthreads();

#pragma unroll
    for (int r = 0; r < kRsIpt; ++r) {
        const bool valid = base + r * kRsThreads + tid < n;
        const u32 d = (u32)(key[r] >> shift) & 255u;
        u64 peers = __ballot(valid);
#pragma unroll
        for (int b = 0; b < 8; ++b) {
            const bool bit = (d >> b) & 1u;
            const u64 m = __ballot(bit);
            peers &= bit ? m : ~m;
        }
        const u64 before = peers & ltMask;
#pragma unroll
        for (int w = 0; w < kRsWaves; ++w) sWave[w][tid] = 0;
        __syncthreads();
        if (valid && before == 0) sWave[wave][d] = (u32)__popcll(peers);
        __syncthreads();
        if (valid) {
            u32 rk = sRun[d] + (u32)__popcll(before);
            for (int w = 0; w < wave; ++w) rk += sWave[w][d];
            rank[r] = rk;
        }
        __syncthreads();
        u32 add = 0;
#pragma unroll
        for (int w = 0; w < kRsWaves; ++w) add += sWave[w][tid];
        sRun[tid] += add;
        __syncthreads();
    }

    // Tile-local start of each digit run.
    {
        int total;
        const int start = block_excl_scan((int)sRun[tid], sScanWave, total);
        sStart[tid] = (u32)start;
    }
    __syncthreads();
#pragma unroll
    for (int r = 0; r < kRsIpt; ++r) {
        const bool valid = base + r * kRsThreads + tid < n;
        if (valid) {
            const u32 d = (u32)(key[r] >> shift) & 255u;
            const u32 pos = sStart[d] + rank[r];
            sKeys[pos] = key[r];
            sVals[pos] = val[r];
        }
    }
    __syncthreads();
    const int tileCount = min(kRsTile, n - base);
#pragma unroll
    for (int r = 0; r < kRsIpt; ++r) {
        const int i = r * kRsThreads + tid;
        if (i < tileCount) {
            const u64 k = sKeys[i];
            const u32 d = (u32)(k >> shift) & 255u;
            const u32 dst = sGlob[d] + ((u32)i - sStart[d]);
            kout[dst] = k;
            vout[dst] = sVals[i];
        }
    }
}

}  // namespace

void scan_inclusive_i32(const int32_t* in, int32_t* out, int n, BuildScratch& s, hipStream_t stream) {
    scan_i32<true>(ScanArrays{{in, in}, {out, out}}, 1, n, s, stream);
}
void scan_inclusive_i32_pair(int32_t* a, int32_t* b, int n, BuildScratch& s, hipStream_t stream) {
    scan_i32<true>(ScanArrays{{a, b}, {a, b}}, 2, n, s, stream);
}
void scan_exclusive_u32(const uint32_t* in, uint32_t* out, int n, BuildScratch& s, hipStream_t stream) {
    const int* i = reinterpret_cast<const int*>(in);
    int* o = reinterpret_cast<int*>(out);
    scan_i32<false>(ScanArrays{{i, i}, {o, o}}, 1, n, s, stream);
}

size_t radix_sort_hist_bytes(int n, int beginBit, int endBit) {
    const size_t numTiles = ((size_t)std::max(n, 1) + kRsTile - 1) / kRsTile;
    const size_t groups = (numTiles + kGroupTiles - 1) / kGroupTiles;
    const size_t passes8 = (size_t)(endBit - beginBit + 7) / 8;
    return sizeof(u32) * 256 * (numTiles + (numTiles <= (size_t)kFusedTiles ? groups * passes8 : 0));
}

size_t radix_sort_partials_bytes(int n) {
    // the longest scan of a sort is its digit counts (unfused sorts)
    const size_t scanned = 256 * (((size_t)std::max(n, 1) + kRsTile - 1) / kRsTile);
    return sizeof(int) * ((scanned + kScanChunk - 1) / kScanChunk + 1);
}

void radix_sort_pairs(uint64_t* keysA, uint64_t* keysB, int32_t* idsA, int32_t* idsB, int n, int beginBit,
                      int endBit, BuildScratch& s, hipStream_t stream) {
    if (n <= 1) return;
    const int numTiles = (n + kRsTile - 1) / kRsTile;
    const bool fused = numTiles <= kFusedTiles;
    const int groups = (numTiles + kGroupTiles - 1) / kGroupTiles;
    const int passes8 = (endBit - beginBit + 7) / 8;
    // digit counts, then (fused) one set of group sums per pass
    s.hist.reserve(radix_sort_hist_bytes(n, beginBit, endBit));
    u32* counts = s.hist.as<u32>();
    u32* groupSums = fused ? counts + 256 * (size_t)numTiles : nullptr;
    if (fused) {
        const int words = 256 * groups * passes8;
        hipLaunchKernelGGL(k_zero_u32, dim3((words + kScanThreads - 1) / kScanThreads), dim3(kScanThreads), 0,
                           stream, groupSums, words);
    }
    u64* kin = keysA;
    u64* kout = keysB;
    i32* vin = idsA;
    i32* vout = idsB;
    int passes = 0;
    for (int shift = beginBit; shift < endBit; shift += 8) {
        u32* gs = fused ? groupSums + 256 * (size_t)groups * passes : nullptr;
        hipLaunchKernelGGL(k_rs_hist, dim3(numTiles), dim3(kRsThreads), 0, stream, kin, n, shift, counts,
                           numTiles, gs);
        if (!fused) scan_exclusive_u32(counts, counts, 256 * numTiles, s, stream);
        hipLaunchKernelGGL(k_rs_scatter, dim3(numTiles), dim3(kRsThreads), 0, stream, kin, vin, kout, vout, n,
                           shift, counts, numTiles, gs);
        MFH_HIP(hipGetLastError());
        std::swap(kin, kout);
        std::swap(vin, vout);
        ++passes;
    }
    if (passes & 1) {
        MFH_HIP(hipMemcpyAsync(keysA, kin, sizeof(u64) * (size_t)n, hipMemcpyDeviceToDevice, stream));
        MFH_HIP(hipMemcpyAsync(idsA, vin, sizeof(i32) * (size_t)n, hipMemcpyDeviceToDevice, stream));
    }
}

}  // namespace mfh
