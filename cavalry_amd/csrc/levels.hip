// Bottom-up scheduling by node height, used by the treelet rounds.
//
// The reference orders its treelet rounds by letting each leaf thread climb
// and the second visitor of a node continue (BVH/Optimization.cu:467-507).
// A treelet reads and rewrites up to 13 nodes, far more than the box that
// the build's climb passes between arrivals, so here the climb only computes
// heights, and every height level is then processed by its own launch;
// kernel boundaries order the levels and no data fences are needed:
//
//   1. climb: every leaf walks up; each visitor of a node adds
//      (1 << 16 | its height + 1) to the node's word with one atomic, so the
//      second visitor finds the first one's height in the returned value
//      (the build's climb leaves the same words for the first round);
//   2. bucket the internal nodes by height (block-aggregated counting sort);
//   3. level h = 3..H in order: a launch over the nodes of height h, the few
//      nodes of the top levels in a single workgroup with barriers.
//
// Children always have a lower height than their parent, so this is a valid
// bottom-up order; nodes of one level have disjoint subtrees.
#include "dev_common.h"
#include "lbvh.h"

namespace cav {
namespace {

constexpr int kBlock = 256;
constexpr int kLdsBins = 256;

// Climb from one leaf. Each visitor adds (1 << 16 | its height + 1) to the
// parent's word; the second visitor reads the first one's height from the
// returned value, stores the parent's final (2 << 16 | height) and goes on.
__device__ __forceinline__ void climb_heights(const int4* __restrict__ topo, uint32_t* hw, int leaf,
                                              int max_steps) {
    int cur = leaf;
    uint32_t h = 0;
    for (int step = 0; cur != 0 && step < max_steps; ++step) {
        const int p = topo[cur].z;
        const uint32_t old = atomicAdd(&hw[p], (1u << 16) | (h + 1));
        if ((old >> 16) != 1u) return;  // first visitor stops here
        h = max(old & 0xffffu, h + 1);
        hw[p] = (2u << 16) | h;
        cur = p;
    }
}

__global__ __launch_bounds__(kBlock) void k_heights(const int4* __restrict__ topo,
                                                    const uint32_t* __restrict__ leaf_ids,
                                                    int nleaves, uint32_t* hw) {
    const int k = blockIdx.x * kBlock + threadIdx.x;
    if (k >= nleaves) return;
    climb_heights(topo, hw, (int)leaf_ids[k], nleaves);
}

__global__ void k_fill(uint32_t* p, int64_t n, uint32_t v) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

constexpr int kPer = 16;  // nodes per thread in the bucketing kernels

// cnt[h] = number of internal nodes of height h (h < kMaxLevels). Counts are
// aggregated per workgroup of 4096 nodes before one global add per height.
__global__ __launch_bounds__(kBlock) void k_level_hist(const int4* __restrict__ topo,
                                                       const uint32_t* __restrict__ hw, int64_t m,
                                                       uint32_t* __restrict__ cnt,
                                                       uint32_t* __restrict__ err) {
    __shared__ uint32_t h[kLdsBins];
    h[threadIdx.x] = 0;
    __syncthreads();
    const int64_t base = (int64_t)blockIdx.x * kBlock * kPer + threadIdx.x;
    for (int j = 0; j < kPer; ++j) {
        const int64_t id = base + (int64_t)j * kBlock;
        if (id < m && topo[id].w < 0) {
            const uint32_t lv = hw[id] & 0xffffu;
            if (lv < (uint32_t)kLdsBins) atomicAdd(&h[lv], 1u);
            else if (lv < (uint32_t)kMaxLevels) atomicAdd(&cnt[lv], 1u);
            else atomicAdd(err, 1u);
        }
    }
    __syncthreads();
    if (h[threadIdx.x]) atomicAdd(&cnt[threadIdx.x], h[threadIdx.x]);
}

// list[cursor[h] ...] = internal node ids of height h; each workgroup
// reserves one range per height it holds.
__global__ __launch_bounds__(kBlock) void k_level_scatter(const int4* __restrict__ topo,
                                                          const uint32_t* __restrict__ hw,
                                                          int64_t m, uint32_t* cursor,
                                                          uint32_t* __restrict__ list) {
    __shared__ uint32_t lc[kLdsBins];
    __shared__ uint32_t base[kLdsBins];
    lc[threadIdx.x] = 0;
    __syncthreads();
    const int64_t first = (int64_t)blockIdx.x * kBlock * kPer + threadIdx.x;
    uint32_t lv[kPer], r[kPer];
    for (int j = 0; j < kPer; ++j) {
        const int64_t id = first + (int64_t)j * kBlock;
        lv[j] = 0xffffffffu;
        r[j] = 0;
        if (id < m && topo[id].w < 0) {
            lv[j] = hw[id] & 0xffffu;
            if (lv[j] < (uint32_t)kLdsBins) r[j] = atomicAdd(&lc[lv[j]], 1u);
            else if (lv[j] < (uint32_t)kMaxLevels) list[atomicAdd(&cursor[lv[j]], 1u)] = (uint32_t)id;
        }
    }
    __syncthreads();
    if (lc[threadIdx.x]) base[threadIdx.x] = atomicAdd(&cursor[threadIdx.x], lc[threadIdx.x]);
    __syncthreads();
    for (int j = 0; j < kPer; ++j)
        if (lv[j] < (uint32_t)kLdsBins) list[base[lv[j]] + r[j]] = (uint32_t)(first + (int64_t)j * kBlock);
}

// Depth of every leaf (parent hops to the root, at most n) and their maximum.
__global__ __launch_bounds__(kBlock) void k_max_depth(const int4* __restrict__ topo, int64_t m, int n,
                                                      uint32_t* __restrict__ out) {
    const int64_t id = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    uint32_t d = 0;
    if (id < m && topo[id].w >= 0) {
        int cur = (int)id;
        for (int step = 0; cur != 0 && step < n; ++step) {
            cur = topo[cur].z;
            ++d;
        }
    }
#pragma unroll
    for (int dd = 32; dd >= 1; dd >>= 1) d = max(d, (uint32_t)__shfl_xor((int)d, dd, 64));
    if ((threadIdx.x & 63) == 0 && d > 0) atomicMax(out, d);
}

}  // namespace

hipError_t levels_max_depth(const LbvhBuffers& b, int n, uint32_t* out, hipStream_t s) {
    const int64_t m = 2 * (int64_t)n - 1;
    hipLaunchKernelGGL(k_fill, dim3(1), dim3(64), 0, s, out, (int64_t)1, 0u);
    hipLaunchKernelGGL(k_max_depth, dim3((unsigned)((m + kBlock - 1) / kBlock)), dim3(kBlock), 0, s, b.topo, m, n,
                       out);
    return hipGetLastError();
}

hipError_t levels_from_list(LbvhBuffers& b, int n, hipStream_t s, Marker* mk) {
    const int64_t m = 2 * (int64_t)n - 1;
    hipLaunchKernelGGL(k_fill, dim3((unsigned)((m + 255) / 256)), dim3(256), 0, s, b.visited, m, 0u);
    const int nb = (n + kBlock - 1) / kBlock;
    hipLaunchKernelGGL(k_heights, dim3(nb), dim3(kBlock), 0, s, b.topo, b.leaf_ids, n, b.visited);
    mark(mk, "treelet_heights");
    return hipGetLastError();
}

hipError_t levels_bucket(LbvhBuffers& b, int n, hipStream_t s, Marker* mk) {
    const int64_t m = 2 * (int64_t)n - 1;
    const unsigned nb = (unsigned)((m + kBlock * kPer - 1) / (kBlock * kPer));
    hipLaunchKernelGGL(k_fill, dim3((kMaxLevels + 2 + 255) / 256), dim3(256), 0, s, b.lvl_cnt,
                       (int64_t)kMaxLevels + 2, 0u);
    hipLaunchKernelGGL(k_level_hist, dim3(nb), dim3(kBlock), 0, s, b.topo, b.visited, m, b.lvl_cnt,
                       b.lvl_cnt + kMaxLevels + 1);
    hipError_t e = hipMemcpyAsync(b.h_lvl, b.lvl_cnt, sizeof(uint32_t) * (kMaxLevels + 2),
                                  hipMemcpyDeviceToHost, s);
    if (e != hipSuccess) return e;
    if ((e = hipStreamSynchronize(s)) != hipSuccess) return e;
    if (b.h_lvl[kMaxLevels + 1] != 0) return hipErrorInvalidValue;  // height beyond kMaxLevels
    int top = 0;
    for (int h = 0; h < kMaxLevels; ++h)
        if (b.h_lvl[h]) top = h;
    b.height = top;
    uint32_t run = 0;
    for (int h = 0; h <= top + 1; ++h) {
        b.h_off[h] = run;
        if (h <= top) run += b.h_lvl[h];
    }
    e = hipMemcpyAsync(b.lvl_off, b.h_off, sizeof(uint32_t) * (top + 2), hipMemcpyHostToDevice, s);
    if (e != hipSuccess) return e;
    e = hipMemcpyAsync(b.lvl_cursor, b.h_off, sizeof(uint32_t) * (top + 2), hipMemcpyHostToDevice, s);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_level_scatter, dim3(nb), dim3(kBlock), 0, s, b.topo, b.visited, m,
                       b.lvl_cursor, b.lvl_list);
    mark(mk, "levels");
    return hipGetLastError();
}

}  // namespace cav
