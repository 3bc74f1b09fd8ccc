// LBVH build kernels: scene bounds, Morton codes, Karras radix tree, refit,
// and the flatten into the traversal layout.
//
// Reference: BVH/BVH.cu:46-370 (fillLeafBoundingBoxes, fillLeafMortonCodes,
// buildRadixTree, computeBounds, mergeNodesArray, copyToBVH) and the scene
// bounds loop of Core/Scene.cu:131-136.
//
// Data layout in HBM (N triangles, M = 2N-1 nodes, unified id space: internal
// node i at i, the leaf of sorted position k at N-1+k before restructuring):
//   tris   : float[N][9]  v0, v1, v2 (caller-owned, original order)
//   keys   : u32[N] or u64[N] Morton codes (sorted in place by the radix sort)
//   order  : u32[N] sorted position -> original triangle index
//   nbox   : NodeBox[M]   (lo, sa | hi, cost), 32 B
//   ntopo  : int4[M]      (left, right, parent, prim), 16 B
//   visited: u32[M]       refit / treelet climb gates
#include "dev_common.h"
#include "lbvh.h"

namespace cav {
namespace {

constexpr int kBlock = 256;

__device__ __forceinline__ v3 load_v(const float* __restrict__ t, int64_t tri, int j) {
    const float* p = t + tri * 9 + j * 3;
    return mk3(p[0], p[1], p[2]);
}

// Core/Triangle.h:242-268
__device__ __forceinline__ void tri_box(v3 a, v3 b, v3 c, v3& lo, v3& hi) {
    lo = mk3(fmn(a.x, fmn(b.x, c.x)), fmn(a.y, fmn(b.y, c.y)), fmn(a.z, fmn(b.z, c.z)));
    hi = mk3(fmx(a.x, fmx(b.x, c.x)), fmx(a.y, fmx(b.y, c.y)), fmx(a.z, fmx(b.z, c.z)));
}

constexpr int kBoundsBlocks = 1024;

__device__ __forceinline__ uint32_t wave_min(uint32_t v) {
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) v = min(v, (uint32_t)__shfl_xor((int)v, d, 64));
    return v;
}
__device__ __forceinline__ uint32_t wave_max(uint32_t v) {
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) v = max(v, (uint32_t)__shfl_xor((int)v, d, 64));
    return v;
}

__device__ __forceinline__ void acc(uint32_t* mn, uint32_t* mx, int axis, float f) {
    if (f == f) {
        const uint32_t o = f2ord(f);
        mn[axis] = min(mn[axis], o);
        mx[axis] = max(mx[axis], o);
    }
}

// Block-reduces 6 per-thread values (3 minima, 3 maxima) and writes them.
__device__ __forceinline__ void block_store6(uint32_t* mn, uint32_t* mx, uint32_t* out) {
    __shared__ uint32_t red[kBlock / 64][6];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
    for (int a = 0; a < 3; ++a) { mn[a] = wave_min(mn[a]); mx[a] = wave_max(mx[a]); }
    if (lane == 0)
        for (int a = 0; a < 3; ++a) { red[wave][a] = mn[a]; red[wave][3 + a] = mx[a]; }
    __syncthreads();
    if (threadIdx.x < 6) {
        uint32_t v = red[0][threadIdx.x];
        for (int w = 1; w < kBlock / 64; ++w)
            v = threadIdx.x < 3 ? min(v, red[w][threadIdx.x]) : max(v, red[w][threadIdx.x]);
        out[threadIdx.x] = v;
    }
}

// Union of all triangle boxes as order-preserving integers (NaN skipped), so
// the result does not depend on reduction order. Stage 1: each workgroup
// reduces a grid-strided share of the soup, four triangles (nine 16-byte
// loads) per thread and step, into one partial.
__global__ __launch_bounds__(kBlock) void k_bounds(const float* __restrict__ tris, int64_t n,
                                                   uint32_t* __restrict__ part) {
    uint32_t mn[3] = {0xffffffffu, 0xffffffffu, 0xffffffffu}, mx[3] = {0u, 0u, 0u};
    const int64_t quads = n / 4;
    const float4* q4 = reinterpret_cast<const float4*>(tris);
    for (int64_t q = (int64_t)blockIdx.x * kBlock + threadIdx.x; q < quads;
         q += (int64_t)gridDim.x * kBlock) {
        float4 v[9];
#pragma unroll
        for (int k = 0; k < 9; ++k) v[k] = q4[q * 9 + k];
#pragma unroll
        for (int k = 0; k < 9; ++k) {
            acc(mn, mx, (4 * k + 0) % 3, v[k].x);
            acc(mn, mx, (4 * k + 1) % 3, v[k].y);
            acc(mn, mx, (4 * k + 2) % 3, v[k].z);
            acc(mn, mx, (4 * k + 3) % 3, v[k].w);
        }
    }
    if (blockIdx.x == 0)
        for (int64_t t = quads * 4 + threadIdx.x; t < n; t += kBlock)
            for (int k = 0; k < 9; ++k) acc(mn, mx, k % 3, tris[t * 9 + k]);
    block_store6(mn, mx, part + (int64_t)blockIdx.x * 6);
}

// Stage 2: one workgroup reduces the partials.
__global__ __launch_bounds__(kBlock) void k_bounds_final(const uint32_t* __restrict__ part,
                                                         int nparts, uint32_t* __restrict__ ob) {
    uint32_t mn[3] = {0xffffffffu, 0xffffffffu, 0xffffffffu}, mx[3] = {0u, 0u, 0u};
    for (int i = threadIdx.x; i < nparts; i += kBlock)
        for (int a = 0; a < 3; ++a) {
            mn[a] = min(mn[a], part[i * 6 + a]);
            mx[a] = max(mx[a], part[i * 6 + 3 + a]);
        }
    block_store6(mn, mx, ob);
}

__device__ __forceinline__ uint64_t sat_u64(float v) {
    if (!(v > 0.f)) return 0;
    if (v >= 18446744073709551616.0f) return ~0ull;
    return (uint64_t)v;
}
__device__ __forceinline__ uint32_t sat_u32(float v) {
    if (!(v > 0.f)) return 0;
    if (v >= 4294967296.0f) return ~0u;
    return (uint32_t)v;
}
// BVH/BVH.cu:59-76
__device__ __forceinline__ uint64_t spread21(uint64_t x) {
    if (x >= (1ull << 21)) x = 1ull << 20;
    x = (x | x << 32) & 0x1f00000000ffffull;
    x = (x | x << 16) & 0x1f0000ff0000ffull;
    x = (x | x << 8) & 0x100f00f00f00f00full;
    x = (x | x << 4) & 0x10c30c30c30c30c3ull;
    x = (x | x << 2) & 0x1249249249249249ull;
    return x;
}
__device__ __forceinline__ uint32_t spread10(uint32_t x) {
    if (x >= (1u << 10)) x = 1u << 9;
    x = (x | (x << 16)) & 0x030000ffu;
    x = (x | (x << 8)) & 0x0300f00fu;
    x = (x | (x << 4)) & 0x030c30c3u;
    x = (x | (x << 2)) & 0x09249249u;
    return x;
}

template <typename K>
__device__ __forceinline__ K morton_code(v3 c);
template <>
__device__ __forceinline__ uint64_t morton_code<uint64_t>(v3 c) {
    const float f = 2097152.0f;
    return (spread21(sat_u64(c.x * f)) << 2) | (spread21(sat_u64(c.y * f)) << 1) |
           spread21(sat_u64(c.z * f));
}
template <>
__device__ __forceinline__ uint32_t morton_code<uint32_t>(v3 c) {
    const float f = 1024.0f;
    return (spread10(sat_u32(c.x * f)) << 2) | (spread10(sat_u32(c.y * f)) << 1) |
           spread10(sat_u32(c.z * f));
}

// BVH/BVH.cu:79-92: code of the triangle-box centroid normalised by the scene
// box; values[i] = i so the stable sort yields the original index. Also
// counts the codes' 8-bit digits for every radix-sort pass (one LDS add per
// digit and code, one global add per digit and workgroup, over a bounded
// grid) and clears the sort's look-back words.
constexpr int kMortonBlocks = 1024;
constexpr int kMortonUnroll = 4;
template <typename K>
__global__ __launch_bounds__(kBlock) void k_morton(const float* __restrict__ tris, int64_t n,
                                                   const uint32_t* __restrict__ ob,
                                                   K* __restrict__ keys,
                                                   uint32_t* __restrict__ vals,
                                                   uint32_t* __restrict__ hist, int passes,
                                                   uint32_t* __restrict__ zero, int64_t nzero) {
    __shared__ uint32_t h[8 * 256];
    for (int j = threadIdx.x; j < passes * 256; j += kBlock) h[j] = 0u;
    const int64_t stride = (int64_t)gridDim.x * kBlock;
    for (int64_t j = (int64_t)blockIdx.x * kBlock + threadIdx.x; j < nzero; j += stride) zero[j] = 0u;
    __syncthreads();
    const v3 slo = mk3(ord2f(ob[0]), ord2f(ob[1]), ord2f(ob[2]));
    const v3 shi = mk3(ord2f(ob[3]), ord2f(ob[4]), ord2f(ob[5]));
    const v3 sc = scale3(add3(slo, shi), 0.5f);
    const v3 se = sub3(shi, slo);
    // the triangles of kMortonUnroll trips are loaded before any is used
    // (past the end: the last triangle again), so their loads overlap
    const int64_t trips = (n + stride - 1) / stride;
    const int64_t first = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    for (int64_t it = 0; it < trips; it += kMortonUnroll) {
        v3 va[kMortonUnroll], vb[kMortonUnroll], vc[kMortonUnroll];
#pragma unroll
        for (int u = 0; u < kMortonUnroll; ++u) {
            const int64_t i = (it + u) * stride + first;
            const int64_t ic = i < n ? i : n - 1;
            va[u] = load_v(tris, ic, 0);
            vb[u] = load_v(tris, ic, 1);
            vc[u] = load_v(tris, ic, 2);
        }
#pragma unroll
        for (int u = 0; u < kMortonUnroll; ++u) {
            const int64_t i = (it + u) * stride + first;
            const bool valid = i < n;
            K key = 0;
            if (valid) {
                v3 lo, hi;
                tri_box(va[u], vb[u], vc[u], lo, hi);
                const v3 c = div3(sub3(scale3(add3(lo, hi), 0.5f), sc), se);
                key = morton_code<K>(c);
                keys[i] = key;
                vals[i] = (uint32_t)i;
            }
            if (valid)
                for (int p = 0; p < passes; ++p) atomicAdd(&h[p * 256 + ((uint32_t)(key >> (8 * p)) & 255u)], 1u);
        }
    }
    __syncthreads();
    for (int j = threadIdx.x; j < passes * 256; j += kBlock)
        if (h[j]) atomicAdd(&hist[j], h[j]);
}

__global__ void k_zero_words(uint32_t* p, int n) {
    for (int i = threadIdx.x; i < n; i += blockDim.x) p[i] = 0u;
}

// ---- device-wide exclusive scan of leaf flags (restructured trees) ----
constexpr int kScanItems = 16;
constexpr int kScanTile = kBlock * kScanItems;

__device__ __forceinline__ uint32_t wave_incl(uint32_t v) {
    const int lane = threadIdx.x & 63;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        uint32_t o = __shfl_up(v, d, 64);
        if (lane >= d) v += o;
    }
    return v;
}
__device__ __forceinline__ uint32_t block_excl(uint32_t v, uint32_t* ws, uint32_t& total) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const uint32_t inc = wave_incl(v);
    if (lane == 63) ws[wave] = inc;
    __syncthreads();
    uint32_t off = 0;
    total = 0;
    for (int w = 0; w < kBlock / 64; ++w) { if (w < wave) off += ws[w]; total += ws[w]; }
    __syncthreads();
    return off + inc - v;
}

// exclusive count of set flags before this thread's in the block, and the
// block's total
__device__ __forceinline__ uint32_t block_flag_excl(bool f, uint32_t* ws, uint32_t& total) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const uint64_t b = __ballot(f);
    if (lane == 0) ws[wave] = (uint32_t)__popcll(b);
    __syncthreads();
    uint32_t off = 0;
    total = 0;
    for (int w = 0; w < kBlock / 64; ++w) { if (w < wave) off += ws[w]; total += ws[w]; }
    __syncthreads();
    return off + (uint32_t)__popcll(b & ((1ull << lane) - 1ull));
}

// Tile t covers ids [t * kScanTile, (t + 1) * kScanTile); pass j of a tile
// reads ids tile + j * kBlock + thread, so every load is coalesced.
__global__ __launch_bounds__(kBlock) void k_leafcount(const int4* __restrict__ topo, int64_t m,
                                                      uint32_t* __restrict__ partial) {
    __shared__ uint32_t ws[kBlock / 64];
    const int64_t tile = (int64_t)blockIdx.x * kScanTile;
    uint32_t c = 0;
    for (int j = 0; j < kScanItems; ++j) {
        const int64_t id = tile + (int64_t)j * kBlock + threadIdx.x;
        if (id < m && topo[id].w >= 0) ++c;
    }
    uint32_t tot;
    block_excl(c, ws, tot);
    if (threadIdx.x == 0) partial[blockIdx.x] = tot;
}
__global__ __launch_bounds__(kBlock) void k_partials(uint32_t* partial, int np) {
    __shared__ uint32_t ws[kBlock / 64];
    const int per = (np + kBlock - 1) / kBlock;
    const int t0 = threadIdx.x * per;
    uint32_t s = 0;
    for (int t = t0; t < t0 + per && t < np; ++t) s += partial[t];
    uint32_t tot;
    uint32_t run = block_excl(s, ws, tot);
    for (int t = t0; t < t0 + per && t < np; ++t) { uint32_t c = partial[t]; partial[t] = run; run += c; }
}
// ref[id] = ~(leaf rank) for leaves, id - (leaf rank) for internal nodes.
__global__ __launch_bounds__(kBlock) void k_refs(const int4* __restrict__ topo, int64_t m,
                                                 const uint32_t* __restrict__ partial,
                                                 int32_t* __restrict__ ref) {
    __shared__ uint32_t ws[kBlock / 64];
    const int64_t tile = (int64_t)blockIdx.x * kScanTile;
    uint32_t run = partial[blockIdx.x];
    for (int j = 0; j < kScanItems; ++j) {
        const int64_t id = tile + (int64_t)j * kBlock + threadIdx.x;
        const bool leaf = id < m && topo[id].w >= 0;
        uint32_t tot;
        const uint32_t ex = run + block_flag_excl(leaf, ws, tot);
        if (id < m) ref[id] = leaf ? ~(int32_t)ex : (int32_t)(id - ex);
        run += tot;
    }
}

__device__ __forceinline__ int32_t ref_of(const int32_t* ref, int n, int id) {
    if (ref) return ref[id];
    return id < n - 1 ? id : ~(id - (n - 1));
}

// copyToBVH (BVH/BVH.cu:304-315) into the traversal layout: one 64-byte
// record per node in the sibling-pair layout (TRec), an internal node holding
// both child boxes, a leaf its triangle. A node's slot follows from its
// parent's rank (ref_of) and which child it is; the root is at slot 0.
__global__ __launch_bounds__(kBlock) void k_flatten(const NodeBox* __restrict__ nbox,
                                                    const int4* __restrict__ topo, int n,
                                                    const int32_t* __restrict__ ref,
                                                    const float* __restrict__ tris,
                                                    const int32_t* __restrict__ mesh,
                                                    TRec* __restrict__ rec,
                                                    float* __restrict__ root) {
    const int64_t id = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    const int64_t m = 2 * (int64_t)n - 1;
    if (id >= m) return;
    const int4 tp = topo[id];
    int64_t slot = 0;
    if (id == 0) {
        const NodeBox b = nbox[0];
        root[0] = b.a.x; root[1] = b.a.y; root[2] = b.a.z;
        root[3] = b.b.x; root[4] = b.b.y; root[5] = b.b.z;
    } else {
        const int par = tp.z;
        slot = 2 * (int64_t)ref_of(ref, n, par) + 2 + (topo[par].x == (int)id ? 0 : 1);
    }
    TRec o;
    if (tp.w < 0) {
        const NodeBox bl = nbox[tp.x], br = nbox[tp.y];
        const int32_t base = 2 * ref_of(ref, n, (int)id) + 2;
        const bool ll = topo[tp.x].w >= 0, rl = topo[tp.y].w >= 0;
        o.q0 = make_float4(bl.a.x, bl.a.y, bl.a.z, bl.b.x);
        o.q1 = make_float4(bl.b.y, bl.b.z, br.a.x, br.a.y);
        o.q2 = make_float4(br.a.z, br.b.x, br.b.y, br.b.z);
        o.q3 = make_float4(__int_as_float(ll ? ~base : base), __int_as_float(rl ? ~(base + 1) : base + 1), 0.f, 0.f);
    } else {
        const int p = tp.w;
        const v3 v0 = load_v(tris, p, 0), v1 = load_v(tris, p, 1), v2 = load_v(tris, p, 2);
        const v3 e1 = sub3(v1, v0), e2 = sub3(v2, v0);
        o.q0 = make_float4(v0.x, v0.y, v0.z, __int_as_float(p));
        o.q1 = make_float4(e1.x, e1.y, e1.z, __int_as_float(mesh ? mesh[p] : 0));
        o.q2 = make_float4(e2.x, e2.y, e2.z, 0.f);
        o.q3 = make_float4(0.f, 0.f, 0.f, 0.f);
    }
    rec[slot] = o;
}

template <typename K>
hipError_t build_impl(LbvhBuffers& b, const float* tris, int n, int keybits, hipStream_t s,
                      Marker* mk) {
    const int passes = (keybits + 7) / 8;
    const int64_t tiles = onesweep_tiles(n);
    OnesweepBuffers<K> ob;
    ob.keys[0] = reinterpret_cast<K*>(b.keys[0]);
    ob.keys[1] = reinterpret_cast<K*>(b.keys[1]);
    ob.vals[0] = b.vals[0];
    ob.vals[1] = b.vals[1];
    ob.hist = b.hist;
    ob.status = b.counts;
    ob.tile_ctr = b.hist + kSortHistWords;
    ob.err = b.hist + kSortHistWords + 8;
    int nb = (n + kBlock - 1) / kBlock;
    if (nb > kMortonBlocks) nb = kMortonBlocks;
    hipLaunchKernelGGL(k_zero_words, dim3(1), dim3(kBlock), 0, s, b.hist, kSortHistWords + 16);
    hipLaunchKernelGGL(k_morton<K>, dim3(nb), dim3(kBlock), 0, s, tris, (int64_t)n, b.scene_ord, ob.keys[0],
                       ob.vals[0], b.hist, passes, b.counts, (int64_t)passes * 256 * tiles);
    mark(mk, "morton");
    hipError_t e = onesweep_sort_pairs(ob, n, keybits, s, mk);
    if (e != hipSuccess) return e;
    b.sorted = ob.result;
    const K* codes = ob.keys[ob.result];
    b.order = ob.vals[ob.result];
    // radix tree + refit (BVH/BVH.cu:122-259) in one climb from the leaves
    return lbvh_bottom_up(b, codes, keybits, tris, n, s, mk);
}

}  // namespace

hipError_t lbvh_scene_bounds(const float* tris, int n, uint32_t* scene_ord, uint32_t* part,
                             hipStream_t s, Marker* mk) {
    if ((reinterpret_cast<uintptr_t>(tris) & 15u) != 0) return hipErrorInvalidValue;
    int blocks = (int)(((int64_t)n / 4 + kBlock * 4 - 1) / (kBlock * 4));
    if (blocks > kBoundsBlocks) blocks = kBoundsBlocks;
    if (blocks < 1) blocks = 1;
    hipLaunchKernelGGL(k_bounds, dim3(blocks), dim3(kBlock), 0, s, tris, (int64_t)n, part);
    hipLaunchKernelGGL(k_bounds_final, dim3(1), dim3(kBlock), 0, s, part, blocks, scene_ord);
    mark(mk, "bounds");
    return hipGetLastError();
}

hipError_t lbvh_build(LbvhBuffers& b, const float* tris, int n, int morton_bits, hipStream_t s,
                      Marker* mk) {
    if (morton_bits == 63) return build_impl<uint64_t>(b, tris, n, 63, s, mk);
    return build_impl<uint32_t>(b, tris, n, 30, s, mk);
}

hipError_t lbvh_refs(LbvhBuffers& b, int n, hipStream_t s, Marker* mk, const char* label) {
    const int64_t m = 2 * (int64_t)n - 1;
    const int np = (int)((m + kScanTile - 1) / kScanTile);
    hipLaunchKernelGGL(k_leafcount, dim3(np), dim3(kBlock), 0, s, b.topo, m, b.partial);
    hipLaunchKernelGGL(k_partials, dim3(1), dim3(kBlock), 0, s, b.partial, np);
    hipLaunchKernelGGL(k_refs, dim3(np), dim3(kBlock), 0, s, b.topo, m, b.partial, b.ref);
    mark(mk, label);
    return hipGetLastError();
}

hipError_t lbvh_flatten(LbvhBuffers& b, const float* tris, const int32_t* mesh, int n,
                        bool identity_refs, TRec* rec, float* root, hipStream_t s,
                        Marker* mk) {
    const int64_t m = 2 * (int64_t)n - 1;
    const int nb = (int)((m + kBlock - 1) / kBlock);
    hipLaunchKernelGGL(k_flatten, dim3(nb), dim3(kBlock), 0, s, b.nbox, b.topo, n,
                       identity_refs ? nullptr : b.ref, tris, mesh, rec, root);
    mark(mk, "flatten");
    return hipGetLastError();
}

int64_t lbvh_partials(int64_t m) { return (m + kScanTile - 1) / kScanTile; }

}  // namespace cav
