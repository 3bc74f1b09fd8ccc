// CPU oracle: LBVH build (Morton codes, stable sort, Karras radix tree, refit).
//
// TEST INFRASTRUCTURE ONLY: a restatement of the reference algorithm used to
// check the HIP kernels and as the timed CPU baseline. Not part of the product.
//
// Node arrays use the reference's unified index space of 2N-1 nodes
// (BVH/BVH.cu:273-302): internal node i lives at i (0..N-2), the leaf of sorted
// position k lives at N-1+k. prim[] holds the original triangle index of a
// leaf and -1 for internal nodes; left/right are -1 for leaves.
#include "ref_api.h"
#include "ref_math.h"

#include <algorithm>
#include <cstdio>
#include <numeric>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#include <parallel/algorithm>
#endif

using namespace ref;

static inline f3 vtx(const float* t, int64_t i, int j) {
    const float* p = t + i * 9 + j * 3;
    return mk(p[0], p[1], p[2]);
}

// Core/Scene.cu:131-136: scene bounds = union of triangle boxes.
extern "C" void ref_scene_bounds(const float* tris, int n, float* out6) {
    Box b = tri_box(vtx(tris, 0, 0), vtx(tris, 0, 1), vtx(tris, 0, 2));
    for (int i = 1; i < n; ++i)
        b = union_box(b, tri_box(vtx(tris, i, 0), vtx(tris, i, 1), vtx(tris, i, 2)));
    out6[0] = b.lo.x; out6[1] = b.lo.y; out6[2] = b.lo.z;
    out6[3] = b.hi.x; out6[4] = b.hi.y; out6[5] = b.hi.z;
}

// float -> unsigned conversion with the saturating semantics of the PTX
// cvt.rzi.u{32,64}.f32 the reference compiles to: NaN and values <= 0 give 0,
// values beyond the type's range give its maximum, everything else truncates.
static inline uint64_t sat_u64(float v) {
    if (std::isnan(v) || v <= 0.f) return 0;
    if (v >= 18446744073709551616.0f) return ~0ull;
    return (uint64_t)v;
}
static inline uint32_t sat_u32(float v) {
    if (std::isnan(v) || v <= 0.f) return 0;
    if (v >= 4294967296.0f) return ~0u;
    return (uint32_t)v;
}

// BVH/BVH.cu:59-76 (63-bit branch). Inputs >= 2^21 are replaced by 2^20
// (the reference writes `1 << BITS - 1`, which parses as 1 << (BITS-1)).
static inline uint64_t spread21(uint64_t x) {
    if (x >= (1ull << 21)) x = 1ull << 20;
    x = (x | x << 32) & 0x1f00000000ffffull;
    x = (x | x << 16) & 0x1f0000ff0000ffull;
    x = (x | x << 8) & 0x100f00f00f00f00full;
    x = (x | x << 4) & 0x10c30c30c30c30c3ull;
    x = (x | x << 2) & 0x1249249249249249ull;
    return x;
}
// BVH/BVH.cu:59-76 (30-bit branch), clamp 2^10 -> 2^9.
static inline uint32_t spread10(uint32_t x) {
    if (x >= (1u << 10)) x = 1u << 9;
    x = (x | (x << 16)) & 0x030000ffu;
    x = (x | (x << 8)) & 0x0300f00fu;
    x = (x | (x << 4)) & 0x030c30c3u;
    x = (x | (x << 2)) & 0x09249249u;
    return x;
}

// BVH/BVH.cu:79-92: code of the leaf box centroid normalised by the scene box.
static inline uint64_t morton_of(const Box& leaf, const Box& scene, int bits) {
    f3 c = divv(centroid(leaf) - centroid(scene), extent(scene));
    if (bits == 63) {
        const float factor = 2097152.0f;  // 2^21
        uint64_t x = sat_u64(c.x * factor), y = sat_u64(c.y * factor), z = sat_u64(c.z * factor);
        return (spread21(x) << 2) | (spread21(y) << 1) | spread21(z);
    }
    const float factor = 1024.0f;  // 2^10
    uint32_t x = sat_u32(c.x * factor), y = sat_u32(c.y * factor), z = sat_u32(c.z * factor);
    return (uint64_t)((spread10(x) << 2) | (spread10(y) << 1) | spread10(z));
}

extern "C" void ref_morton(const float* tris, int n, const float* scene6, int bits,
                           uint64_t* codes) {
    Box scene{mk(scene6[0], scene6[1], scene6[2]), mk(scene6[3], scene6[4], scene6[5])};
#pragma omp parallel for schedule(static)
    for (int i = 0; i < n; ++i) {
        Box b = tri_box(vtx(tris, i, 0), vtx(tris, i, 1), vtx(tris, i, 2));
        codes[i] = morton_of(b, scene, bits);
    }
}

// BVH/BVH.cu:101-120. For distinct keys: number of equal leading bits of the
// (3*B)-bit code; for equal keys: 3*B + clz32(i ^ j); i == j gives 3*B + 32.
static inline int prefix_len(const uint64_t* codes, int keybits, uint32_t i, uint32_t j) {
    uint64_t a = codes[i], b = codes[j];
    if (a != b) {
        uint64_t x = a ^ b;
        int hi = 63 - __builtin_clzll(x);
        return keybits - 1 - hi;
    }
    uint32_t y = i ^ j;
    if (y == 0) return keybits + 32;
    return keybits + __builtin_clz(y);
}

// BVH/BVH.cu:122-215 for one internal node i. Writes children in the unified
// index space and the parent of each child.
static void karras_node(int n, const uint64_t* codes, int keybits, int i, int32_t* left,
                        int32_t* right, int32_t* parent) {
    int d = 1;
    if (i > 0) {
        int a = prefix_len(codes, keybits, i, i + 1) - prefix_len(codes, keybits, i, i - 1);
        d = a >= 0 ? 1 : -1;
    }
    int dmin = 0;
    if (i > 0) dmin = prefix_len(codes, keybits, i, i - d);

    // exponential search; the reference forms i + maxLength*d in unsigned
    // arithmetic, so a negative position wraps and fails the `< n` test.
    uint32_t maxlen = 2;
    while (true) {
        uint32_t j = (uint32_t)i + maxlen * (uint32_t)d;
        if (!(j < (uint32_t)n)) break;
        if (!(prefix_len(codes, keybits, i, j) > dmin)) break;
        maxlen *= 2;
    }
    uint32_t len = 0;
    for (int p = 1;; ++p) {
        uint32_t t = maxlen >> p;
        if (t == 0) break;
        int j = (int)((uint32_t)i + (len + t) * (uint32_t)d);
        if (j >= 0 && j < n && prefix_len(codes, keybits, i, j) > dmin) len += t;
    }
    int j = (int)((uint32_t)i + len * (uint32_t)d);
    if (i == 0) j = n - 1;
    int plen = prefix_len(codes, keybits, i, j);

    uint32_t split = 0;
    for (int p = 1;; ++p) {
        uint32_t t = (uint32_t)std::ceil((float)len / (float)(1 << p));
        int k = (int)((uint32_t)i + (split + t) * (uint32_t)d);
        if (k >= 0 && k < n && prefix_len(codes, keybits, i, k) > plen) split += t;
        if (t <= 1) break;
    }
    int spos = (int)((uint32_t)i + split * (uint32_t)d) + std::min(0, d);
    int lc = spos, rc = spos + 1;
    bool lleaf = std::min(i, j) == lc;
    bool rleaf = (std::max(i, j) == rc) || (rc == n - 1);
    int lid = lleaf ? (n - 1 + lc) : lc;
    int rid = rleaf ? (n - 1 + rc) : rc;
    left[i] = lid;
    right[i] = rid;
    parent[lid] = i;
    parent[rid] = i;
}

extern "C" int ref_lbvh_build(const float* tris, int n, int bits, const float* scene6_or_null,
                              uint64_t* sorted_codes, int32_t* order, float* box, float* sa,
                              float* cost, int32_t* left, int32_t* right, int32_t* parent,
                              int32_t* prim) {
    if (n < 2 || (bits != 30 && bits != 63)) return -1;
    float scene6[6];
    if (scene6_or_null) std::memcpy(scene6, scene6_or_null, sizeof scene6);
    else ref_scene_bounds(tris, n, scene6);
    const int keybits = bits == 63 ? 63 : 30;
    const int64_t m = 2 * (int64_t)n - 1;

    std::vector<uint64_t> codes(n);
    ref_morton(tris, n, scene6, bits, codes.data());

    // thrust::stable_sort_by_key (BVH/BVH.cu:335): ascending, equal keys keep
    // input order -- identical to sorting (key, index) pairs.
    std::vector<std::pair<uint64_t, int32_t>> kv(n);
    for (int i = 0; i < n; ++i) kv[i] = {codes[i], i};
#ifdef _OPENMP
    __gnu_parallel::sort(kv.begin(), kv.end());
#else
    std::sort(kv.begin(), kv.end());
#endif
    std::vector<uint64_t> sc(n);
    for (int k = 0; k < n; ++k) { sc[k] = kv[k].first; order[k] = kv[k].second; }
    if (sorted_codes) std::memcpy(sorted_codes, sc.data(), sizeof(uint64_t) * n);

    for (int64_t id = 0; id < m; ++id) { left[id] = -1; right[id] = -1; prim[id] = -1; }
    parent[0] = 0;  // root parent is 0 (zero-initialised array in the reference)

#pragma omp parallel for schedule(static)
    for (int i = 0; i < n - 1; ++i) karras_node(n, sc.data(), keybits, i, left, right, parent);

    // leaves: fillLeafBoundingBoxes (BVH/BVH.cu:46-56) after the sort
#pragma omp parallel for schedule(static)
    for (int k = 0; k < n; ++k) {
        int t = order[k];
        Box b = tri_box(vtx(tris, t, 0), vtx(tris, t, 1), vtx(tris, t, 2));
        int64_t id = n - 1 + k;
        float* bx = box + id * 6;
        bx[0] = b.lo.x; bx[1] = b.lo.y; bx[2] = b.lo.z; bx[3] = b.hi.x; bx[4] = b.hi.y; bx[5] = b.hi.z;
        sa[id] = surface_area(b);
        cost[id] = leaf_cost(sa[id], 1);
        prim[id] = t;
    }

    // computeBounds (BVH/BVH.cu:217-259): every leaf climbs; the second visitor
    // of an internal node unions the children. Each node is computed once from
    // its final children, so the sequential climb gives the parallel result.
    std::vector<uint32_t> visited(n - 1, 0);
    for (int k = 0; k < n; ++k) {
        int cur = parent[n - 1 + k];
        while (visited[cur]++ > 0) {
            int l = left[cur], r = right[cur];
            Box bl{mk(box[l * 6 + 0], box[l * 6 + 1], box[l * 6 + 2]),
                   mk(box[l * 6 + 3], box[l * 6 + 4], box[l * 6 + 5])};
            Box br{mk(box[r * 6 + 0], box[r * 6 + 1], box[r * 6 + 2]),
                   mk(box[r * 6 + 3], box[r * 6 + 4], box[r * 6 + 5])};
            Box u = union_box(bl, br);
            float* bx = box + (int64_t)cur * 6;
            bx[0] = u.lo.x; bx[1] = u.lo.y; bx[2] = u.lo.z; bx[3] = u.hi.x; bx[4] = u.hi.y; bx[5] = u.hi.z;
            sa[cur] = surface_area(u);
            cost[cur] = internal_cost3(sa[cur], cost[l], cost[r]);
            if (cur == 0) break;
            cur = parent[cur];
        }
    }
    return 0;
}
