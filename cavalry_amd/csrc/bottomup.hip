// LBVH topology and refit in one bottom-up pass over the sorted leaves.
//
// Reference: buildRadixTree + computeBounds (BVH/BVH.cu:122-259), with the
// leaf records of fillLeafBoundingBoxes (BVH/BVH.cu:46-56).
//
// The reference emits the Karras radix tree top-down (one thread per internal
// node, binary searches over the codes) and then refits it bottom-up. Both
// describe one tree: the binary radix tree over the sorted (code, index)
// pairs. With delta(j) the common-prefix length of sorted positions j and j+1
// (the reference's commonPrefixLength, index bits breaking equal codes),
//   * the node whose leaves are [l, r] has its split at the boundary of
//     smallest delta inside the range, and its parent's split at whichever of
//     l-1 and r has the larger delta (they are never equal; l = 0 or r = N-1
//     leaves only the other side): it is the parent's left child when that is
//     r, the right child when it is l-1;
//   * the reference numbers internal nodes by one end of their range: a left
//     child is node r, a right child node l, the root node 0; a child whose
//     range is one leaf k is leaf N-1+k.
// So one climb from every leaf, the second arrival at each split continuing
// (the reference's computeBounds gate), produces the reference's node ids,
// children, parents and boxes without the top-down searches.
//
// Two passes. Pass 1 gives each workgroup 256 consecutive sorted leaves. A
// split whose node's leaves all lie in the workgroup (nearer boundaries of
// smaller delta on both sides inside the workgroup, found by a min-scan over
// the workgroup's deltas) is gated in LDS: the first arrival leaves its
// child's box there, the second completes the node and climbs on. A climb
// ends at the first split that crosses workgroups, where the child's box,
// topology record and an arrival word (far leaf, height, the delta of its
// outer range boundary) are left in HBM; no wave waits on another
// workgroup. About 2% of the splits cross workgroups. Pass 2 climbs them:
// a split both of whose children pass 1 completed sits at a workgroup edge,
// and one thread per edge starts there. Its arrival at a split higher up
// stores the child's box and record with agent-scope stores, waits for them
// to complete and adds its arrival word to the split's 64-bit gate; the
// second arrival reads the first's word back from the add, clears the gate
// for the next build and loads the box and record with agent-scope loads.
// Every access that crosses workgroups within pass 2 is an agent-scope
// atomic, so no cache write-back or invalidate is needed. Results do not
// depend on which arrival is second: both children are final when their
// parent is computed, as in the reference.
#include "dev_common.h"
#include "lbvh.h"

namespace cav {
namespace {

constexpr int kBlock = 256;

__device__ __forceinline__ v3 load_tri_v(const float* __restrict__ t, int64_t tri, int j) {
    const float* p = t + tri * 9 + j * 3;
    return mk3(p[0], p[1], p[2]);
}

// BVH/BVH.cu:101-120 for adjacent positions j, j+1 (codes a, b).
template <typename K>
__device__ __forceinline__ int delta_of(K a, K b, int keybits, uint32_t j) {
    if (a != b) {
        const int w = (int)sizeof(K) * 8;
        const int clz = sizeof(K) == 8 ? __clzll((long long)(a ^ b)) : __clz((int)(a ^ b));
        return keybits - w + clz;
    }
    return keybits + __clz((int)(j ^ (j + 1u)));
}

// computeBounds (BVH/BVH.cu:217-259)
__device__ __forceinline__ NodeBox combine(const NodeBox& bl, const NodeBox& br) {
    const v3 lo = mk3(fmn(bl.a.x, br.a.x), fmn(bl.a.y, br.a.y), fmn(bl.a.z, br.a.z));
    const v3 hi = mk3(fmx(bl.b.x, br.b.x), fmx(bl.b.y, br.b.y), fmx(bl.b.z, br.b.z));
    const float sa = surface_area(lo, hi);
    return NodeBox{make_float4(lo.x, lo.y, lo.z, sa),
                   make_float4(hi.x, hi.y, hi.z, internal_cost3(sa, bl.b.w, br.b.w))};
}

// A node box as four 64-bit agent-scope atomic stores / loads: coherent for
// a reader on another XCD without flushing any cache.
__device__ __forceinline__ void store_box_agent(NodeBox* p, const NodeBox& b) {
    uint64_t w[4];
    __builtin_memcpy(w, &b, 32);
    uint64_t* q = reinterpret_cast<uint64_t*>(p);
#pragma unroll
    for (int i = 0; i < 4; ++i) __hip_atomic_store(q + i, w[i], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ NodeBox load_box_agent(const NodeBox* p) {
    uint64_t w[4];
    const uint64_t* q = reinterpret_cast<const uint64_t*>(p);
#pragma unroll
    for (int i = 0; i < 4; ++i) w[i] = __hip_atomic_load(q + i, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    NodeBox b;
    __builtin_memcpy(&b, w, 32);
    return b;
}

__device__ __forceinline__ void store_rec_agent(int4* p, int4 v) {
    uint64_t w[2];
    __builtin_memcpy(w, &v, 16);
    uint64_t* q = reinterpret_cast<uint64_t*>(p);
    __hip_atomic_store(q, w[0], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __hip_atomic_store(q + 1, w[1], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ int4 load_rec_agent(const int4* p) {
    uint64_t w[2];
    const uint64_t* q = reinterpret_cast<const uint64_t*>(p);
    w[0] = __hip_atomic_load(q, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    w[1] = __hip_atomic_load(q + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    int4 v;
    __builtin_memcpy(&v, w, 16);
    return v;
}

// The topology record but its parent of the node with leaves [l, r] and
// split g (internal; children as buildRadixTree numbers them,
// BVH/BVH.cu:186-190), or of a leaf (g = ~prim).
__device__ __forceinline__ int4 node_rec(int g, int l, int r, int n) {
    if (g < 0) return make_int4(-1, -1, 0, ~g);
    return make_int4(l == g ? n - 1 + g : g, g + 1 == r ? n - 1 + r : g + 1, 0, -1);
}

// A workgroup barrier for LDS data only: this wave's LDS operations are
// done (lgkmcnt(0)), while its global loads (the leaf's gather) stay in
// flight across the barrier.
__device__ __forceinline__ void lds_barrier() {
    __builtin_amdgcn_s_waitcnt(0xc07f);  // vmcnt(63) expcnt(7) lgkmcnt(0)
    __builtin_amdgcn_s_barrier();
}

// Inclusive min-scan of one value per thread over the workgroup.
__device__ __forceinline__ int block_min_scan(int v, int* wmin) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        const int o = __shfl_up(v, d, 64);
        if (lane >= d) v = min(v, o);
    }
    if (lane == 63) wmin[wave] = v;
    lds_barrier();
    for (int w = 0; w < wave; ++w) v = min(v, wmin[w]);
    lds_barrier();
    return v;
}

// A child's arrival at a split that crosses workgroups, in one word: an
// arrival bit, the delta of the child's outer range boundary + 1 (7 bits:
// deltas are at most 94, -1 for none), the child's height + 1 (16 bits) and
// the child's far leaf (its first leaf for a left child, last for a right).
constexpr uint64_t kArrive = 1ull << 63;
__device__ __forceinline__ uint64_t pack_arrival(int far, uint32_t hp, int dout) {
    return kArrive | ((uint64_t)(uint32_t)(dout + 1) << 56) | ((uint64_t)hp << 32) | (uint32_t)far;
}
__device__ __forceinline__ int arrival_far(uint64_t w) { return (int)(uint32_t)w; }
__device__ __forceinline__ uint32_t arrival_hp(uint64_t w) { return (uint32_t)(w >> 32) & 0xffffu; }
__device__ __forceinline__ int arrival_dout(uint64_t w) { return (int)((w >> 56) & 0x7fu) - 1; }

// Split g is the last boundary of a workgroup: both its children can lie
// inside workgroups (one each), so both can arrive in the local pass. Such a
// split's arrivals go to two words of their own (left, right child) behind
// the gates, and its gate counts arrivals.
__device__ __forceinline__ bool edge_split(int g) { return g % kBlock == kBlock - 1; }

// Pass 1: one workgroup per 256 consecutive sorted leaves. Every leaf climbs
// through the splits whose node lies inside the workgroup (gated in LDS, the
// second arrival continuing); a climb ends at the first split that crosses
// workgroups, where it leaves its node's box, topology record and arrival
// word in HBM for pass 2. No wave waits on another workgroup.
template <typename K>
__global__ __launch_bounds__(kBlock) void k_bottom_up(const K* __restrict__ codes, int n, int keybits,
                                                      const uint32_t* __restrict__ order,
                                                      const float* __restrict__ tris,
                                                      NodeBox* nbox, int4* topo, uint32_t* hw,
                                                      unsigned long long* gate64) {
    __shared__ K skey[kBlock + 2];        // codes of positions k0-1 .. k0+256
    __shared__ int sdel[kBlock + 1];      // delta of boundaries k0-1 .. k0+255 (-1: none)
    __shared__ int wmin[kBlock / 64];
    __shared__ uint8_t slocal[kBlock];    // split k0+i has all its leaves in this workgroup
    __shared__ uint32_t gate[kBlock];     // arrivals << 16 | height, per local split
    __shared__ NodeBox sbox[2 * kBlock];  // per local split: left, right child's box
    __shared__ int sfar[2 * kBlock];      // per local split: left child's first, right child's last leaf
    __shared__ int ssplit[2 * kBlock];    // per local split: left, right child's split, or ~prim for a leaf

    const int k0 = blockIdx.x * kBlock;
    const int t = threadIdx.x;
    const int k = k0 + t;
    // codes of positions k0-1+t and k0+255+t (the latter for t < 2)
    const int64_t j0 = (int64_t)k0 - 1 + t, j1 = j0 + kBlock;
    const K key0 = (j0 >= 0 && j0 < n) ? codes[j0] : (K)0;
    const K key1 = (t < 2 && j1 < n) ? codes[j1] : (K)0;
    // the leaf's triangle (BVH/BVH.cu:46-56), a random gather, issued by every
    // thread (past the end: the last leaf's) so that no branch forces a wait
    // for it before the setup below, which syncs on LDS only
    const int tri = (int)order[k < n ? k : n - 1];
    const v3 a = load_tri_v(tris, tri, 0), b = load_tri_v(tris, tri, 1), c = load_tri_v(tris, tri, 2);
    skey[t] = key0;
    if (t < 2) skey[kBlock + t] = key1;
    gate[t] = 0u;
    lds_barrier();
    for (int i = t; i <= kBlock; i += kBlock) {
        const int64_t j = (int64_t)k0 - 1 + i;
        sdel[i] = (j < 0 || j >= n - 1) ? -1 : delta_of<K>(skey[i], skey[i + 1], keybits, (uint32_t)j);
    }
    lds_barrier();
    // split g = k0 + t: a smaller delta among boundaries k0-1 .. g-1 (prefix
    // min of sdel[0..t]) and among g+1 .. k0+255 (suffix min of sdel[t+2..256])
    const int pre = block_min_scan(sdel[t], wmin);
    const int suf_rev = block_min_scan(sdel[kBlock - t], wmin);  // min of sdel[256-t .. 256]
    __shared__ int ssuf[kBlock];
    ssuf[kBlock - 1 - t] = suf_rev;  // ssuf[i] = min of sdel[i+1 .. 256]
    lds_barrier();
    {
        const int g = k0 + t;
        const int dg = sdel[t + 1];
        const int suf = t + 1 < kBlock ? ssuf[t + 1] : 0x7fffffff;  // min of sdel[t+2 .. 256]
        slocal[t] = (t < kBlock - 1 && g <= n - 2 && pre < dg && suf < dg) ? 1 : 0;
    }
    lds_barrier();

    if (k >= n) return;
    // delta of boundary j for j in [k0-1, k0+255]: every node this pass
    // completes lies inside the workgroup, so its range boundaries are there
    auto delta = [&](int j) -> int { return sdel[j - k0 + 1]; };

    const v3 lo = mk3(fmn(a.x, fmn(b.x, c.x)), fmn(a.y, fmn(b.y, c.y)), fmn(a.z, fmn(b.z, c.z)));
    const v3 hi = mk3(fmx(a.x, fmx(b.x, c.x)), fmx(a.y, fmx(b.y, c.y)), fmx(a.z, fmx(b.z, c.z)));
    const float sa0 = surface_area(lo, hi);
    NodeBox cb{make_float4(lo.x, lo.y, lo.z, sa0), make_float4(hi.x, hi.y, hi.z, leaf_cost(sa0))};
    int cur = n - 1 + k;
    // cur's split, or ~prim for a leaf: with its range it gives cur's whole
    // topology record but the parent. Each node's record is written once, by
    // the arrival that completes its parent (the only one that knows the
    // parent's id).
    int csplit = ~tri;
    int l = k, r = k;
    uint32_t h = 0;
    int dl = delta(k - 1), dr = delta(k);  // deltas of cur's range boundaries l-1 and r
    for (int step = 0; step < kBlock; ++step) {
        // side of cur in its parent: the left child iff delta(r) > delta(l-1)
        const bool left = dr > dl;
        const int g = left ? r : l - 1;  // the parent's split
        const int far = left ? l : r;
        const uint32_t hp = min(h + 1u, 0xffffu);
        const int li = g - k0;
        if (!(li >= 0 && li < kBlock && slocal[li])) {
            // a split that crosses workgroups: pass 2 completes it
            nbox[cur] = cb;
            topo[cur] = node_rec(csplit, l, r, n);
            const uint64_t w = pack_arrival(far, hp, left ? dl : dr);
            if (edge_split(g)) {
                gate64[n + 2 * ((g + 1) / kBlock) + (left ? 0 : 1)] = w;
                (void)__hip_atomic_fetch_add(gate64 + g, 1ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            } else {
                gate64[g] = w;  // the only arrival in this pass: the sibling crosses workgroups
            }
            return;
        }
        nbox[cur] = cb;
        sbox[2 * li + (left ? 0 : 1)] = cb;
        sfar[2 * li + (left ? 0 : 1)] = far;
        ssplit[2 * li + (left ? 0 : 1)] = csplit;
        __threadfence_block();
        const uint32_t old = atomicAdd(&gate[li], (1u << 16) | hp);
        if ((old >> 16) == 0u) return;  // first arrival: the sibling continues
        __threadfence_block();
        const uint32_t hs = old & 0xffffu;
        const NodeBox sib = sbox[2 * li + (left ? 1 : 0)];
        const int ofar = sfar[2 * li + (left ? 1 : 0)];
        const int os = ssplit[2 * li + (left ? 1 : 0)];
        // the parent: leaves [lp, rp], split g; the sibling's leaves are
        // [g + 1, ofar] (right) or [ofar, g] (left)
        const int lp = left ? l : ofar, rp = left ? ofar : r;
        const int lid = lp == g ? n - 1 + g : g;
        const int rid = g + 1 == rp ? n - 1 + rp : g + 1;
        const int4 srec_o = left ? node_rec(os, g + 1, ofar, n) : node_rec(os, ofar, g, n);
        const NodeBox pb = left ? combine(cb, sib) : combine(sib, cb);
        h = max(h + 1u, hs);
        const bool root = lp == 0 && rp == n - 1;
        const int dlp = delta(lp - 1), drp = delta(rp);
        const int pid = root ? 0 : (drp > dlp ? rp : lp);
        const int4 crec = node_rec(csplit, l, r, n);
        const int4 lrec = left ? crec : srec_o, rrec = left ? srec_o : crec;
        topo[lid] = make_int4(lrec.x, lrec.y, pid, lrec.w);
        topo[rid] = make_int4(rrec.x, rrec.y, pid, rrec.w);
        hw[pid] = (2u << 16) | min(h, 0xffffu);
        if (root) {  // all n leaves in this workgroup
            topo[0] = make_int4(lid, rid, 0, -1);
            nbox[0] = pb;
            return;
        }
        csplit = g;
        cur = pid;
        l = lp;
        r = rp;
        dl = dlp;
        dr = drp;
        cb = pb;
    }
}

// Pass 2: the splits that cross workgroups, climbed from the workgroup
// edges whose split had both children completed by pass 1 (one thread per
// edge). A climb stops at a split whose other child has not arrived yet;
// its arrival (box and record with agent-scope stores, waited for, then the
// arrival word added to the gate) lets that child's climb continue. The
// node ids, parents and boxes are those of pass 1's rule.
__global__ __launch_bounds__(kBlock) void k_climb_crossing(int n, int nedges, NodeBox* nbox, int4* topo,
                                                           uint32_t* hw, unsigned long long* gate64) {
    const int e = blockIdx.x * kBlock + threadIdx.x + 1;  // the edge between workgroups e-1 and e
    if (e > nedges) return;
    unsigned long long* edge = gate64 + n;  // two arrival words per edge split
    int g = e * kBlock - 1;
    // (pass 1's words are read with agent-scope loads like pass 2's own)
    if (__hip_atomic_load(gate64 + g, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 2ull)
        return;  // a child crosses workgroups: its climb continues here
    __hip_atomic_store(gate64 + g, 0ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    const uint64_t wl = __hip_atomic_load(edge + 2 * e, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    const uint64_t wr = __hip_atomic_load(edge + 2 * e + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    int lp = arrival_far(wl), rp = arrival_far(wr);
    int dlp = arrival_dout(wl), drp = arrival_dout(wr);  // deltas of boundaries lp-1 and rp
    uint32_t h = max(arrival_hp(wl), arrival_hp(wr));
    int lid = lp == g ? n - 1 + g : g;
    int rid = g + 1 == rp ? n - 1 + rp : g + 1;
    NodeBox bl = load_box_agent(nbox + lid), br = load_box_agent(nbox + rid);
    int4 lrec = load_rec_agent(topo + lid), rrec = load_rec_agent(topo + rid);
    for (int step = 0; step < n; ++step) {
        const NodeBox pb = combine(bl, br);
        const bool root = lp == 0 && rp == n - 1;
        const bool pleft = !root && drp > dlp;
        const int pid = root ? 0 : (pleft ? rp : lp);
        topo[lid] = make_int4(lrec.x, lrec.y, pid, lrec.w);
        topo[rid] = make_int4(rrec.x, rrec.y, pid, rrec.w);
        hw[pid] = (2u << 16) | min(h, 0xffffu);
        if (root) {
            topo[0] = make_int4(lid, rid, 0, -1);
            nbox[0] = pb;
            return;
        }
        // arrive at the split above node pid (leaves [lp, rp])
        const int g2 = pleft ? rp : lp - 1;
        const uint32_t hp = min(h + 1u, 0xffffu);
        const uint64_t w = pack_arrival(pleft ? lp : rp, hp, pleft ? dlp : drp);
        store_box_agent(nbox + pid, pb);
        store_rec_agent(topo + pid, make_int4(lid, rid, 0, -1));
        uint64_t other;
        if (edge_split(g2)) {
            const int e2 = (g2 + 1) / kBlock;
            __hip_atomic_store(edge + 2 * e2 + (pleft ? 0 : 1), w, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            __atomic_signal_fence(__ATOMIC_SEQ_CST);
            __builtin_amdgcn_s_waitcnt(0);  // the stores have completed
            __atomic_signal_fence(__ATOMIC_SEQ_CST);
            const uint64_t old = __hip_atomic_fetch_add(gate64 + g2, 1ull, __ATOMIC_RELAXED,
                                                        __HIP_MEMORY_SCOPE_AGENT);
            if (old == 0u) return;  // first arrival
            __atomic_signal_fence(__ATOMIC_SEQ_CST);
            __hip_atomic_store(gate64 + g2, 0ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            other = __hip_atomic_load(edge + 2 * e2 + (pleft ? 1 : 0), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        } else {
            __atomic_signal_fence(__ATOMIC_SEQ_CST);
            __builtin_amdgcn_s_waitcnt(0);  // the stores have completed
            __atomic_signal_fence(__ATOMIC_SEQ_CST);
            const uint64_t old = __hip_atomic_fetch_add(gate64 + g2, (unsigned long long)w, __ATOMIC_RELAXED,
                                                        __HIP_MEMORY_SCOPE_AGENT);
            if (old == 0u) return;  // first arrival
            __atomic_signal_fence(__ATOMIC_SEQ_CST);
            __hip_atomic_store(gate64 + g2, 0ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            other = old;
        }
        // the parent of pid and its sibling: leaves [lp', rp'], split g2
        const int ofar = arrival_far(other);
        const uint32_t hs = arrival_hp(other);
        const int ds = arrival_dout(other);
        if (pleft) {
            rp = ofar;
            drp = ds;
        } else {
            lp = ofar;
            dlp = ds;
        }
        const int4 prec = make_int4(lid, rid, 0, -1);  // pid's record but its parent
        g = g2;
        lid = lp == g ? n - 1 + g : g;
        rid = g + 1 == rp ? n - 1 + rp : g + 1;
        const int sid = pleft ? rid : lid;
        const NodeBox sb = load_box_agent(nbox + sid);
        const int4 sr = load_rec_agent(topo + sid);
        bl = pleft ? pb : sb;
        br = pleft ? sb : pb;
        lrec = pleft ? prec : sr;
        rrec = pleft ? sr : prec;
        h = max(hp, hs);
    }
}

}  // namespace

// Words of the split gates for n leaves: n-1 gates, then two arrival words
// per workgroup edge.
int64_t bottom_up_gate_words(int n) { return (int64_t)n + 2 * ((int64_t)(n + kBlock - 1) / kBlock) + 2; }

hipError_t lbvh_bottom_up(LbvhBuffers& b, const void* codes, int keybits, const float* tris, int n,
                          hipStream_t s, Marker* mk) {
    const int nb = (n + kBlock - 1) / kBlock;
    if (keybits > 32)
        hipLaunchKernelGGL(k_bottom_up<uint64_t>, dim3(nb), dim3(kBlock), 0, s,
                           static_cast<const uint64_t*>(codes), n, keybits, b.order, tris, b.nbox, b.topo,
                           b.visited, b.gate64);
    else
        hipLaunchKernelGGL(k_bottom_up<uint32_t>, dim3(nb), dim3(kBlock), 0, s,
                           static_cast<const uint32_t*>(codes), n, keybits, b.order, tris, b.nbox, b.topo,
                           b.visited, b.gate64);
    mark(mk, "refit_bottom_up");
    const int nedges = nb - 1;
    if (nedges > 0) {
        hipLaunchKernelGGL(k_climb_crossing, dim3((unsigned)((nedges + kBlock - 1) / kBlock)), dim3(kBlock), 0, s,
                           n, nedges, b.nbox, b.topo, b.visited, b.gate64);
        mark(mk, "refit_crossing");
    }
    return hipGetLastError();
}

}  // namespace cav
