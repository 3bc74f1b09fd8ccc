// Stable counting sort of a hit queue by the material of each hit, misses
// last: sortLightingQueue's key (Integrators/PathTracingIntegrator.cu:50-92,
// MaterialType per task, thrust::stable_sort_by_key) without moving the
// tasks. The output is a permutation perm[j] = queue slot; shading kernels
// read their task through it, so the lanes of a wave share one BSDF.
//
// The live count stays on the device: the launches cover the queue's
// capacity and blocks past the count contribute nothing.
#include "../../include/cavalry.h"
#include "render.h"

namespace cav {
namespace {

constexpr int kSortBlock = 1024;
constexpr int kKeys = 8;  // material ids 0..5 (MaterialType), 7 = miss
constexpr int kWaves = kSortBlock / 64;
constexpr int kScanTiles = kSortBlock;  // tiles of the histogram scan (one block scans their sums)

__device__ inline int hit_key(const SceneDev& sc, const int32_t* hit_prim, int64_t i) {
    const int prim = hit_prim[i];
    if (prim < 0) return kKeys - 1;
    const int m = sc.meshes[sc.tri_mesh[prim]].material;
    return (m >= 0 && m < kKeys - 1) ? m : kKeys - 1;
}

__device__ inline uint64_t lanes_below() {
    const int lane = threadIdx.x & 63;
    return lane == 0 ? 0ull : (~0ull >> (64 - lane));
}

// per-block key counts: hist[key * nblk + block]
__global__ __launch_bounds__(kSortBlock) void k_sort_hist(SceneDev sc, const int32_t* __restrict__ hit_prim,
                                                          const uint32_t* count, uint32_t* __restrict__ hist,
                                                          int nblk) {
    __shared__ uint32_t cnt[kKeys];
    if (threadIdx.x < kKeys) cnt[threadIdx.x] = 0;
    __syncthreads();
    const int64_t i = (int64_t)blockIdx.x * kSortBlock + threadIdx.x;
    const int key = i < (int64_t)*count ? hit_key(sc, hit_prim, i) : -1;
#pragma unroll
    for (int k = 0; k < kKeys; ++k) {
        const uint64_t m = __ballot(key == k);
        if ((threadIdx.x & 63) == 0 && m) atomicAdd(&cnt[k], (uint32_t)__popcll(m));
    }
    __syncthreads();
    if (threadIdx.x < kKeys) hist[threadIdx.x * nblk + blockIdx.x] = cnt[threadIdx.x];
}

// Exclusive scan of v over the block (kSortBlock threads); `total` gets the
// block's sum. Every thread of the block must call it.
__device__ inline uint32_t block_excl_scan(uint32_t v, uint32_t& total) {
    __shared__ uint32_t wsum[kWaves];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    uint32_t x = v;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const uint32_t y = __shfl_up(x, o, 64);
        if (lane >= o) x += y;
    }
    if (lane == 63) wsum[wv] = x;
    __syncthreads();
    if (wv == 0) {
        uint32_t w = lane < kWaves ? wsum[lane] : 0u;
#pragma unroll
        for (int o = 1; o < kWaves; o <<= 1) {
            const uint32_t y = __shfl_up(w, o, 64);
            if (lane >= o) w += y;
        }
        if (lane < kWaves) wsum[lane] = w;  // inclusive over the waves
    }
    __syncthreads();
    const uint32_t before = wv == 0 ? 0u : wsum[wv - 1];
    total = wsum[kWaves - 1];
    __syncthreads();  // wsum is reused by the next call
    return before + x - v;
}

// The key-major exclusive scan of hist in three launches (reduce, scan of
// the tile sums, rescan with the tile's offset) over at most kScanTiles tiles
// of `tile` consecutive words: one block walking the whole histogram waited
// on its dependent loads for ~300 us per bounce at 2^24 queue slots.
__global__ __launch_bounds__(kSortBlock) void k_scan_reduce(const uint32_t* __restrict__ hist, int total, int tile,
                                                            uint32_t* __restrict__ tsum) {
    const int lo = blockIdx.x * tile, hi = min(lo + tile, total);
    uint32_t v = 0;
    for (int k = lo + (int)threadIdx.x; k < hi; k += kSortBlock) v += hist[k];
    uint32_t t;
    (void)block_excl_scan(v, t);
    if (threadIdx.x == 0) tsum[blockIdx.x] = t;
}

__global__ __launch_bounds__(kSortBlock) void k_scan_tiles(uint32_t* __restrict__ tsum, int ntiles) {
    const uint32_t v = (int)threadIdx.x < ntiles ? tsum[threadIdx.x] : 0u;
    uint32_t t;
    const uint32_t e = block_excl_scan(v, t);
    if ((int)threadIdx.x < ntiles) tsum[threadIdx.x] = e;
}

__global__ __launch_bounds__(kSortBlock) void k_scan_apply(uint32_t* __restrict__ hist, int total, int tile,
                                                           const uint32_t* __restrict__ tsum) {
    const int lo = blockIdx.x * tile, hi = min(lo + tile, total);
    uint32_t carry = tsum[blockIdx.x];
    for (int base = lo; base < hi; base += kSortBlock) {
        const int k = base + (int)threadIdx.x;
        const uint32_t v = k < hi ? hist[k] : 0u;
        uint32_t t;
        const uint32_t e = block_excl_scan(v, t);
        if (k < hi) hist[k] = carry + e;
        carry += t;
    }
}

// perm[offset(key, block) + rank of the thread among its block's equal keys] = slot
__global__ __launch_bounds__(kSortBlock) void k_sort_scatter(SceneDev sc, const int32_t* __restrict__ hit_prim,
                                                             const uint32_t* count, const uint32_t* __restrict__ hist,
                                                             int nblk, uint32_t* __restrict__ perm) {
    __shared__ uint32_t wcnt[kWaves][kKeys];
    const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int64_t i = (int64_t)blockIdx.x * kSortBlock + threadIdx.x;
    const int key = i < (int64_t)*count ? hit_key(sc, hit_prim, i) : -1;
    uint32_t rank = 0;
#pragma unroll
    for (int k = 0; k < kKeys; ++k) {
        const uint64_t m = __ballot(key == k);
        if (key == k) rank = (uint32_t)__popcll(m & lanes_below());
        if (lane == 0) wcnt[wv][k] = (uint32_t)__popcll(m);
    }
    __syncthreads();
    if (threadIdx.x < kKeys) {  // exclusive prefix over the waves, per key
        uint32_t run = 0;
        for (int w = 0; w < kWaves; ++w) {
            const uint32_t v = wcnt[w][threadIdx.x];
            wcnt[w][threadIdx.x] = run;
            run += v;
        }
    }
    __syncthreads();
    if (key >= 0) perm[hist[key * nblk + blockIdx.x] + wcnt[wv][key] + rank] = (uint32_t)i;
}

}  // namespace

hipError_t sort_hits_by_material(const SceneDev& sc, const int32_t* hit_prim, const uint32_t* count, int64_t cap,
                                 uint32_t* hist, uint32_t* perm, hipStream_t s) {
    const int nblk = (int)((cap + kSortBlock - 1) / kSortBlock);
    if (nblk == 0) return hipSuccess;
    hipLaunchKernelGGL(k_sort_hist, dim3(nblk), dim3(kSortBlock), 0, s, sc, hit_prim, count, hist, nblk);
    // histogram words, then kScanTiles words of tile sums (sort_hist_words)
    const int total = nblk * kKeys;
    int tile = (total + kScanTiles - 1) / kScanTiles;
    tile = (tile + kSortBlock - 1) / kSortBlock * kSortBlock;
    const int ntiles = (total + tile - 1) / tile;
    uint32_t* tsum = hist + total;
    hipLaunchKernelGGL(k_scan_reduce, dim3(ntiles), dim3(kSortBlock), 0, s, hist, total, tile, tsum);
    hipLaunchKernelGGL(k_scan_tiles, dim3(1), dim3(kSortBlock), 0, s, tsum, ntiles);
    hipLaunchKernelGGL(k_scan_apply, dim3(ntiles), dim3(kSortBlock), 0, s, hist, total, tile, tsum);
    hipLaunchKernelGGL(k_sort_scatter, dim3(nblk), dim3(kSortBlock), 0, s, sc, hit_prim, count, hist, nblk, perm);
    return hipGetLastError();
}

int64_t sort_hist_words(int64_t cap) { return ((cap + kSortBlock - 1) / kSortBlock) * kKeys + kScanTiles; }

}  // namespace cav
