// Bottom-up scheduling by node height, used by the refit and the treelet
// rounds.
//
// The reference orders its bottom-up passes by letting each leaf thread climb
// and the second visitor of a node continue (BVH/BVH.cu:217-259,
// BVH/Optimization.cu:467-507). Publishing a node's data to a visitor on
// another XCD needs an agent-scope release/acquire, which on gfx950 writes
// back and invalidates L2 -- far too costly once per node. Here the climb only
// computes heights, and every height level is then processed by its own
// launch, so kernel boundaries order the levels and no data fences are needed:
//
//   1. climb: every leaf walks up; each visitor of a node adds
//      (1 << 16 | its height + 1) to the node's word with one atomic, so the
//      second visitor finds the first one's height in the returned value;
//   2. bucket the internal nodes by height (block-aggregated counting sort);
//   3. level h = 1..H in order: a launch over the nodes of height h, the few
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

__device__ __forceinline__ void load_v(const float* __restrict__ t, int64_t tri, v3& a, v3& b, v3& c) {
    const float* p = t + tri * 9;
    a = mk3(p[0], p[1], p[2]);
    b = mk3(p[3], p[4], p[5]);
    c = mk3(p[6], p[7], p[8]);
}

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

// computeBounds (BVH/BVH.cu:217-259): the box, area and cost of a node from
// its left and right children's.
__device__ __forceinline__ NodeBox combine(const NodeBox& bl, const NodeBox& br) {
    const v3 lo = mk3(fmn(bl.a.x, br.a.x), fmn(bl.a.y, br.a.y), fmn(bl.a.z, br.a.z));
    const v3 hi = mk3(fmx(bl.b.x, br.b.x), fmx(bl.b.y, br.b.y), fmx(bl.b.z, br.b.z));
    const float sa = surface_area(lo, hi);
    return NodeBox{make_float4(lo.x, lo.y, lo.z, sa), make_float4(hi.x, hi.y, hi.z, internal_cost3(sa, bl.b.w, br.b.w))};
}

// fillLeafBoundingBoxes (BVH/BVH.cu:46-56) for sorted position k, written to
// leaf node N-1+k, followed by the height climb. In a Karras tree the leaves
// under internal node p are a contiguous range rng[p] with p at one end, so a
// node whose range lies within this workgroup's leaves has both climbers here
// and is gated by an LDS word; only nodes whose range leaves the workgroup use
// the global word, as in climb_heights. Either way the second visitor stores
// the node's final (2 << 16 | height). At a workgroup-local node each climber
// also leaves its child's box in LDS before the gate, and the second one
// refits the node from the two (the reference's computeBounds as a climb,
// BVH/BVH.cu:217-259, with LDS in place of a device-wide fence). Nodes above
// a boundary-crossing node all cross too; levels_refit does those.
__global__ __launch_bounds__(kBlock) void k_leaves(const float* __restrict__ tris, int n,
                                                   const uint32_t* __restrict__ order,
                                                   NodeBox* __restrict__ nbox, int4* topo,
                                                   uint32_t* hw, const int2* __restrict__ rng) {
    __shared__ uint32_t gate[kBlock];
    __shared__ NodeBox sbox[2 * kBlock];  // per local node: its left, right child's box
    const int k0 = blockIdx.x * kBlock;
    const int k = k0 + threadIdx.x;
    gate[threadIdx.x] = 0u;
    __syncthreads();
    if (k >= n) return;
    const int t = (int)order[k];
    v3 a, b, c;
    load_v(tris, t, a, b, c);
    const v3 lo = mk3(fmn(a.x, fmn(b.x, c.x)), fmn(a.y, fmn(b.y, c.y)), fmn(a.z, fmn(b.z, c.z)));
    const v3 hi = mk3(fmx(a.x, fmx(b.x, c.x)), fmx(a.y, fmx(b.y, c.y)), fmx(a.z, fmx(b.z, c.z)));
    const float sa = surface_area(lo, hi);
    const int node = n - 1 + k;
    const int parent = topo[node].z;
    NodeBox cb{make_float4(lo.x, lo.y, lo.z, sa), make_float4(hi.x, hi.y, hi.z, leaf_cost(sa))};
    nbox[node] = cb;
    topo[node] = make_int4(-1, -1, parent, t);
    int cur = node, p = parent, first = k;  // first = the first leaf under cur
    uint32_t h = 0;
    for (int step = 0; cur != 0 && step < n; ++step) {
        const int2 r = rng[p];
        const uint32_t add = (1u << 16) | (h + 1);
        const bool local = r.x >= k0 && r.y < k0 + kBlock;
        const int side = first == r.x ? 0 : 1;  // cur is p's left child iff their ranges start together
        uint32_t old;
        if (local) {
            sbox[2 * (p - k0) + side] = cb;
            __threadfence_block();
            old = atomicAdd(&gate[p - k0], add);
        } else {
            old = atomicAdd(&hw[p], add);
        }
        if ((old >> 16) != 1u) return;  // first visitor stops here
        h = max(old & 0xffffu, h + 1);
        hw[p] = (2u << 16) | h;
        if (local) {
            __threadfence_block();
            cb = side == 0 ? combine(cb, sbox[2 * (p - k0) + 1]) : combine(sbox[2 * (p - k0)], cb);
            nbox[p] = cb;
        }
        cur = p;
        first = r.x;
        if (cur != 0) p = topo[cur].z;
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

// computeBounds (BVH/BVH.cu:217-259) for one internal node whose children are final.
__device__ __forceinline__ void refit_node(int id, NodeBox* nbox, const int4* __restrict__ topo) {
    const int4 tp = topo[id];
    nbox[id] = combine(nbox[tp.x], nbox[tp.y]);
}

// Nodes whose leaves lie within one k_leaves workgroup were refitted there.
constexpr int kClimbLeaves = kBlock;
__device__ __forceinline__ bool climb_refitted(const int2* rng, int id) {
    if (!rng) return false;
    const int2 r = rng[id];
    return r.x / kClimbLeaves == r.y / kClimbLeaves;
}

__global__ __launch_bounds__(kBlock) void k_refit_level(const uint32_t* __restrict__ list, int count,
                                                        NodeBox* nbox, const int4* __restrict__ topo,
                                                        const int2* __restrict__ rng) {
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i < count) {
        const int id = (int)list[i];
        if (!climb_refitted(rng, id)) refit_node(id, nbox, topo);
    }
}

// Levels h0..h1 in one workgroup; a barrier separates the levels.
__global__ __launch_bounds__(1024) void k_refit_tail(const uint32_t* __restrict__ list,
                                                     const uint32_t* __restrict__ off, int h0, int h1,
                                                     NodeBox* nbox, const int4* __restrict__ topo,
                                                     const int2* __restrict__ rng) {
    for (int h = h0; h <= h1; ++h) {
        const uint32_t b = off[h], e = off[h + 1];
        for (uint32_t i = b + threadIdx.x; i < e; i += 1024)
            if (!climb_refitted(rng, (int)list[i])) refit_node((int)list[i], nbox, topo);
        __syncthreads();
    }
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

hipError_t levels_from_leaves(LbvhBuffers& b, const float* tris, int n, hipStream_t s, Marker* mk) {
    const int nb = (n + kBlock - 1) / kBlock;
    hipLaunchKernelGGL(k_leaves, dim3(nb), dim3(kBlock), 0, s, tris, n, b.order, b.nbox, b.topo, b.visited,
                       reinterpret_cast<const int2*>(b.ref));
    mark(mk, "refit_leaves");
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

hipError_t levels_refit(LbvhBuffers& b, hipStream_t s, Marker* mk) {
    const int top = b.height;
    for (int h = 1; h <= top; ++h) {
        const uint32_t rest = b.h_off[top + 1] - b.h_off[h];
        if (rest <= (uint32_t)kTailNodes) {
            hipLaunchKernelGGL(k_refit_tail, dim3(1), dim3(1024), 0, s, b.lvl_list, b.lvl_off, h, top,
                               b.nbox, b.topo, reinterpret_cast<const int2*>(b.ref));
            break;
        }
        const int cnt = (int)b.h_lvl[h];
        if (cnt == 0) continue;
        hipLaunchKernelGGL(k_refit_level, dim3((cnt + kBlock - 1) / kBlock), dim3(kBlock), 0, s,
                           b.lvl_list + b.h_off[h], cnt, b.nbox, b.topo, reinterpret_cast<const int2*>(b.ref));
    }
    mark(mk, "refit");
    return hipGetLastError();
}

}  // namespace cav
