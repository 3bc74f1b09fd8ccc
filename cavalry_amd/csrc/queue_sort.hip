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

// exclusive scan of hist in key-major order, one block
__global__ __launch_bounds__(kSortBlock) void k_sort_scan(uint32_t* __restrict__ hist, int total) {
    __shared__ uint32_t part[kSortBlock];
    const int chunk = (total + kSortBlock - 1) / kSortBlock;
    const int lo = threadIdx.x * chunk, hi = min(lo + chunk, total);
    uint32_t s = 0;
    for (int k = lo; k < hi; ++k) s += hist[k];
    part[threadIdx.x] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t run = 0;
        for (int t = 0; t < kSortBlock; ++t) {
            const uint32_t v = part[t];
            part[t] = run;
            run += v;
        }
    }
    __syncthreads();
    uint32_t run = part[threadIdx.x];
    for (int k = lo; k < hi; ++k) {
        const uint32_t v = hist[k];
        hist[k] = run;
        run += v;
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
    hipLaunchKernelGGL(k_sort_hist, dim3(nblk), dim3(kSortBlock), 0, s, sc, hit_prim, count, hist, nblk);
    hipLaunchKernelGGL(k_sort_scan, dim3(1), dim3(kSortBlock), 0, s, hist, nblk * kKeys);
    hipLaunchKernelGGL(k_sort_scatter, dim3(nblk), dim3(kSortBlock), 0, s, sc, hit_prim, count, hist, nblk, perm);
    return hipGetLastError();
}

int64_t sort_hist_words(int64_t cap) { return ((cap + kSortBlock - 1) / kSortBlock) * kKeys; }

}  // namespace cav
