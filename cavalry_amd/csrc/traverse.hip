// Closest-hit and any-hit BVH traversal.
//
// Reference: BVH::intersect and BVH::testVisibility (BVH/BVH.h:23-189), the
// slab test of Core/BoundingBox.h:55-74 and Moller-Trumbore of
// Core/Triangle.h:7-52. The closest-hit visiting order is the reference's
// exactly (both children tested at their parent, the far child stacked, the
// near one taken next, left first on ties), so hit ids and distances match
// bit for bit. Any-hit visits the reference's nodes up to the ray's limit
// (the limit cull below); its answer does not depend on the order.
//
// Layout read: TRec[2N] (64-byte records in the sibling-pair layout: an
// internal node's two child boxes and child refs, or a leaf's v0/e1/e2 +
// original index + mesh id), rays float[R][6].
#include <algorithm>
#include <cfloat>

#include "dev_common.h"
#include "launch.h"
#include "lbvh.h"

namespace cav {
namespace {

constexpr int kStack = 64;  // BVH/BVH.h:41,122

// Traversal stack of kStack entries per ray: the first D in LDS (one column
// per thread, so a wave's accesses to one depth hit 64 distinct banks), the
// rest in private memory, which only unusually deep traversals touch. The
// depth sp is a separate local so that it stays in a register: a depth held
// beside the dynamically indexed spill array would live in scratch memory
// with it, a scratch load and store on every push and pop.
template <int D, int B>
__device__ __forceinline__ void stack_push(int (*lds)[B], volatile int* spill, int& sp, int v) {
    if (sp < D) lds[sp][threadIdx.x] = v;
    else spill[sp - D] = v;
    ++sp;
}
// (the LDS read goes through an LDS-qualified pointer so that the compiler
// cannot merge it with the spill read into one flat load through a selected
// pointer)
typedef __attribute__((address_space(3))) int lds_int;
typedef __attribute__((address_space(3))) uint32_t lds_uint;
template <int D, int B>
__device__ __forceinline__ int stack_pop(int (*lds)[B], const volatile int* spill, int& sp) {
    --sp;
    if (sp < D) return *(const lds_int*)(&lds[sp][threadIdx.x]);
    return spill[sp - D];
}

// Core/BoundingBox.h:55-74; returns the entry distance or -1.
__device__ __forceinline__ float slab(float lx, float ly, float lz, float hx, float hy, float hz,
                                      v3 o, v3 inv) {
    const float t0x = (lx - o.x) * inv.x, t0y = (ly - o.y) * inv.y, t0z = (lz - o.z) * inv.z;
    const float t1x = (hx - o.x) * inv.x, t1y = (hy - o.y) * inv.y, t1z = (hz - o.z) * inv.z;
    const float maxmin = fmin_fmax(t0x, t1x, fmin_fmax(t0y, t1y, fmn(t0z, t1z)));
    const float minmax = fmax_fmin(t0x, t1x, fmax_fmin(t0y, t1y, fmx(t0z, t1z)));
    if (maxmin <= minmax && minmax > 0) return fmx(0.f, maxmin);
    return -1.f;
}

// Core/Triangle.h:7-52
__device__ __forceinline__ bool tri_hit(const TTri& tr, v3 o, v3 d, float& t, float& u, float& v) {
    const v3 v0 = mk3(tr.a.x, tr.a.y, tr.a.z);
    const v3 e1 = mk3(tr.b.x, tr.b.y, tr.b.z);
    const v3 e2 = mk3(tr.c.x, tr.c.y, tr.c.z);
    const v3 tvec = sub3(o, v0);
    const v3 pvec = cross3(d, e2);
    float det = dot3(e1, pvec);
    if (det == 0) return false;
    det = 1.f / det;
    u = dot3(tvec, pvec) * det;
    if (u < 0.0f || u > 1.0f) return false;
    const v3 qvec = cross3(tvec, e1);
    v = dot3(d, qvec) * det;
    if (v < 0.0f || (u + v) > 1.0f) return false;
    t = dot3(e2, qvec) * det;
    return t > 0.f;
}

__device__ __forceinline__ void load_ray(const float* __restrict__ rays, int64_t i, v3& o, v3& d) {
    const float2* r = reinterpret_cast<const float2*>(rays + i * 6);
    const float2 a = r[0], b = r[1], c = r[2];
    o = mk3(a.x, a.y, b.x);
    d = mk3(b.y, c.x, c.y);
}

// Persistent closest-hit traversal: a fixed grid whose lanes each keep one
// ray in flight and pull a new one as soon as theirs finishes, so a wave is
// not held by its slowest ray (nodes visited per ray vary by 1000x on scenes
// with many coincident Morton codes). Rays are handed out per wave from
// chunks taken with one atomic. Each ray's node order is the reference's, so
// results do not depend on which lane or wave traced it.
//
// Phase 1 of three: once every ray has been handed out, a ray that has taken
// `lr.after` steps (0: never) is handed, while the list has room, to the
// cooperative pass (k_closest_coop), which traces it again from the root
// with a whole wave; the few steps taken here are dropped. ids (optional)
// maps the dispensed index to the ray index.
struct LongRays {
    uint32_t* ids = nullptr;  // cap ray indices
    uint32_t* n = nullptr;    // rays handed over (past cap: the rest stayed on their lanes)
    uint32_t cap = 0;
    int after = 0;            // steps before a ray may be handed over (0: never)
};

template <int kLdsDepth, int kB>
__global__ __launch_bounds__(kB) void k_trace(const TRec* __restrict__ recs,
                                                  const float* __restrict__ root,
                                                  const float* __restrict__ rays, int64_t nrays_arg,
                                                  const uint32_t* __restrict__ nrays_dev,
                                                  const uint32_t* __restrict__ ids,
                                                  int32_t* __restrict__ oprim,
                                                  float* __restrict__ ot, float* __restrict__ ou,
                                                  float* __restrict__ ov,
                                                  uint32_t* __restrict__ overflow,
                                                  unsigned long long* __restrict__ next_ray,
                                                  unsigned long long* __restrict__ fetched,
                                                  LongRays lr, int max_steps, int chunk) {
    __shared__ int lstack[kLdsDepth][kB];
    const int lane = threadIdx.x & 63;
    const uint64_t lt = (1ull << lane) - 1ull;
    const int64_t nrays = nrays_dev ? (int64_t)*nrays_dev : nrays_arg;

    int64_t ri = -1;       // ray held by this lane; -1 = needs one, -2 = pool drained
    v3 o = mk3(0, 0, 0), d = o, inv = o;
    float best = FLT_MAX, bu = 0.f, bv = 0.f;
    int bp = -1, cur = 0, steps = 0;
    bool hit = false;
    bool may_hand = false;  // this lane's ray may still go to the cooperative pass
    bool drained = false;   // every ray has been handed out
    uint32_t ovf = 0;
    int maxsp = 0;  // deepest stack this lane reached
    uint64_t nfetch = 0;  // 64-byte records this lane fetched (one per step)
    volatile int spill[kStack - kLdsDepth];  // in scratch memory, not registers
    int sp = 0;
    int64_t chunk_base = 0;
    int chunk_avail = 0;

    while (true) {
        // ---- hand new rays to idle lanes (wave-uniform control flow)
        const uint64_t need = __ballot(ri == -1);
        if (need) {
            const int cnt = __popcll(need);
            const int rank = __popcll(need & lt);
            int64_t mine;
            if (chunk_avail < cnt) {
                unsigned long long nb = 0;
                const int first = __ffsll((long long)need) - 1;
                if (lane == first) nb = atomicAdd(next_ray, (unsigned long long)chunk);
                nb = __shfl(nb, first, 64);
                mine = rank < chunk_avail ? chunk_base + rank : (int64_t)nb + (rank - chunk_avail);
                chunk_base = (int64_t)nb + (cnt - chunk_avail);
                chunk_avail = chunk - (cnt - chunk_avail);
            } else {
                mine = chunk_base + rank;
                chunk_base += cnt;
                chunk_avail -= cnt;
            }
            if (ri == -1) {
                if (mine >= nrays) {
                    ri = -2;
                } else {
                    ri = ids ? (int64_t)ids[mine] : mine;
                    load_ray(rays, ri, o, d);
                    inv = div3(mk3(1.f, 1.f, 1.f), d);
                    best = FLT_MAX; hit = false; bp = -1; bu = 0.f; bv = 0.f;
                    sp = 0; cur = 0; steps = 0;
                    may_hand = lr.after > 0;
                    // BVH/BVH.h:34-38: a ray missing the root box has no hit
                    if (slab(root[0], root[1], root[2], root[3], root[4], root[5], o, inv) < 0.f) {
                        oprim[ri] = -1; ot[ri] = 0.f;
                        if (ou) ou[ri] = 0.f;
                        if (ov) ov[ri] = 0.f;
                        ri = -1;
                    }
                }
            }
        }
        if (!__ballot(ri >= 0)) {
            if (!__ballot(ri == -1)) break;  // every lane drained
            continue;                        // lanes whose ray missed the root refill
        }
        if (ri < 0) continue;

        if (may_hand && steps >= lr.after && !drained && ((steps - lr.after) & 255) == 0)
            drained = (int64_t)__hip_atomic_load(next_ray, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) >= nrays;
        if (may_hand && drained && steps >= lr.after) {
            // a long ray at the end: the cooperative pass traces it
            const uint32_t slot = atomicAdd(lr.n, 1u);
            if (slot < lr.cap) {
                lr.ids[slot] = (uint32_t)ri;
                ri = -1;
                continue;
            }
            may_hand = false;  // the list is full: finish it here
        }

        // ---- one traversal step for this lane's ray
        bool done = false;
        if (++steps > max_steps) {
            ++ovf;
            done = true;
        } else {
            // One 64-byte fetch for either kind of step, so that a wave with
            // lanes at nodes and lanes at triangles waits for memory once.
            const bool at_node = cur >= 0;
            ++nfetch;
            const float4* rec = reinterpret_cast<const float4*>(recs + (at_node ? cur : ~cur));
            float4 q0 = rec[0], q1 = rec[1], q2 = rec[2];
            int4 refs = reinterpret_cast<const int4*>(rec)[3];
            // (an empty asm that takes the node-only words keeps the compiler
            // from moving their loads into the node branch, a second wait)
            asm volatile("" : "+v"(q1.w), "+v"(q2.w), "+v"(refs.x), "+v"(refs.y));
            if (at_node) {  // BVH/BVH.h:72-95
                const float dl = slab(q0.x, q0.y, q0.z, q0.w, q1.x, q1.y, o, inv);
                const float dr = slab(q1.z, q1.w, q2.x, q2.y, q2.z, q2.w, o, inv);
                const bool hl = dl >= 0 && dl < best;
                const bool hr = dr >= 0 && dr < best;
                bool next = false;
                if (hl && hr) {
                    const bool lfar = dl > dr;
                    const int far = lfar ? refs.x : refs.y;
                    const int near = lfar ? refs.y : refs.x;
                    if (sp < kStack) stack_push<kLdsDepth, kB>(lstack, spill, sp, far); else ++ovf;
                    maxsp = max(maxsp, sp);
                    if (sp < kStack) { cur = near; next = true; } else ++ovf;
                } else if (hl || hr) {
                    if (sp < kStack) { cur = hl ? refs.x : refs.y; next = true; } else ++ovf;
                }
                if (!next) {
                    if (sp == 0) done = true;
                    else cur = stack_pop<kLdsDepth, kB>(lstack, spill, sp);
                }
            } else {  // BVH/BVH.h:58-70
                const TTri tr{q0, q1, q2};
                float t, u, v;
                if (tri_hit(tr, o, d, t, u, v) && (!hit || t < best)) {
                    hit = true; best = t; bu = u; bv = v; bp = __float_as_int(tr.a.w);
                }
                if (sp == 0) done = true;
                else cur = stack_pop<kLdsDepth, kB>(lstack, spill, sp);
            }
        }
        if (done) {
            oprim[ri] = hit ? bp : -1;
            ot[ri] = hit ? best : 0.f;
            if (ou) ou[ri] = hit ? bu : 0.f;
            if (ov) ov[ri] = hit ? bv : 0.f;
            ri = -1;
        }
    }
    if (ovf) atomicAdd(overflow, ovf);
    // deepest stack over the launch: overflow[6] (6 words past the overflow count)
#pragma unroll
    for (int dd = 32; dd >= 1; dd >>= 1) maxsp = max(maxsp, __shfl_xor(maxsp, dd, 64));
    if (lane == 0 && maxsp > 0) atomicMax(overflow + 6, (uint32_t)maxsp);
    // records fetched over the launch (summed over every launch until reset)
#pragma unroll
    for (int dd = 32; dd >= 1; dd >>= 1) nfetch += __shfl_xor(nfetch, dd, 64);
    if (lane == 0 && nfetch > 0) atomicAdd(fetched, (unsigned long long)nfetch);
}

// Any-hit limit cull. BVH::testVisibility (BVH/BVH.h:158-184) pushes every
// child box the ray's line crosses, also those that start beyond the ray's
// limit, whose triangles can never block it (t < limit). Here a child box
// that starts more than kLimitSlack times the limit along the ray is not
// visited. That changes no answer: a box's computed entry distance is
// within a few ulps of the exact one, which is at most the exact distance
// of any point of a triangle inside it, so a skipped triangle could still
// block only if Moller-Trumbore put its t more than 2^-10 (relative) short
// of the exact one, an error of order 1e-7 / (sin(smallest angle) *
// |cos(incidence)|): a ray within about 1e-4 rad of the plane of a sliver
// triangle that lies just past the limit. The visited set shrinks, so the
// reference's stack-overflow count can only fall. TraceLaunch::any_full
// restores the reference's whole-line visiting order.
constexpr float kLimitSlack = 1.0f + 1.0f / 1024.0f;
__device__ __forceinline__ float any_cut(int cull, bool use_lim, float lim) {
    return cull && use_lim ? lim * kLimitSlack : INFINITY;
}

// Any-hit traversal over sibling pairs.
//
// BVH::testVisibility (BVH/BVH.h:109-189) answers whether any triangle of a
// leaf reached through hit boxes blocks the ray before its limit; which one
// it meets first does not change the answer. So each step here takes a pair
// of siblings -- the two 64-byte records sharing one 128-byte line in the
// sibling-pair layout -- tests both nodes, and goes on with the pairs of
// their children: half the dependent steps of one node per step, and no
// half-used lines. The reference's 64-entry stack is kept exactly: each
// node carries the stack depth the reference would have at it (s = far
// siblings pending below it; a far child has its parent's s, a near child
// s + 1, and a near child that would need s = 64 is dropped and counted, as
// the reference's failed push is), so the same nodes are visited and the
// same overflows counted (all of them when no occluder ends the ray first;
// with the limit cull, which drops boxes beyond the ray's limit, the subset
// of them on the visited nodes).
//
// A pending pair is (base slot >> 1) << 2 | active mask (bit 0: the left
// sibling at the base slot, bit 1: the right one) and one byte: the s of its
// far or only member, | 128 when the near member is the right one. The root
// is the pair at slot 0 with only bit 0 active.
struct PairOut {
    uint32_t item[2];
    uint32_t byte[2];
    int n;
};

__device__ __forceinline__ void pair_step(const float4 (&r)[8], uint32_t item, uint32_t byte, v3 o, v3 d,
                                          v3 inv, int sm, int dm, bool use_lim, float lim, float cut, bool& hit,
                                          uint32_t& ovf, int& maxsp, PairOut& out) {
    const uint32_t mask = item & 3u;
    const int sb = (int)(byte & 63u);
    const bool near_right = (byte & 128u) != 0;
    out.n = 0;
#pragma unroll
    for (int m = 0; m < 2; ++m) {
        if (!((mask >> m) & 1u) || hit) continue;
        const float4 q0 = r[4 * m], q1 = r[4 * m + 1], q2 = r[4 * m + 2], q3 = r[4 * m + 3];
        const int s = (mask == 3u && (m == 1) == near_right) ? sb + 1 : sb;
        const int refl = __float_as_int(q3.x);
        if (refl == 0) {  // a leaf's record: its triangle
            const TTri tr{q0, q1, q2};
            const int mesh = __float_as_int(tr.b.w);
            float t, u, v;
            if (mesh != sm && mesh != dm && tri_hit(tr, o, d, t, u, v) && (!use_lim || t < lim)) hit = true;
        } else {
            const float dl = slab(q0.x, q0.y, q0.z, q0.w, q1.x, q1.y, o, inv);
            const float dr = slab(q1.z, q1.w, q2.x, q2.y, q2.z, q2.w, o, inv);
            const bool hl = dl >= 0 && dl <= cut, hr = dr >= 0 && dr <= cut;
            if (hl || hr) {
                const int cb = refl >= 0 ? refl : ~refl;  // the children's slots: cb, cb + 1
                uint32_t cm = (hl ? 1u : 0u) | (hr ? 2u : 0u);
                bool nr = false;
                if (cm == 3u) {
                    maxsp = max(maxsp, s + 1);  // the far child's push
                    nr = dl > dr;               // BVH/BVH.h:171-178: the nearer child is taken first
                    if (s + 1 >= kStack) {      // the near child would need a 65th entry
                        ++ovf;
                        cm = nr ? 1u : 2u;      // the far child alone
                    }
                }
                out.item[out.n] = ((uint32_t)(cb >> 1) << 2) | cm;
                out.byte[out.n] = (uint32_t)s | (nr && cm == 3u ? 128u : 0u);
                ++out.n;
            }
        }
    }
}

// Phase 1: one lane per ray. Once every ray has been handed out, a ray that
// has taken `spill_after` steps (0: never) is written to the spill pool,
// while it has room, with its pending pairs, and finished by phase 2: the
// long rays still in flight at the end are spread over whole waves instead
// of keeping one lane each busy while the rest of the device idles. ids
// (optional) maps the dispensed index to the ray index.
template <int kLdsDepth, int kB>
__global__ __launch_bounds__(kB) void k_trace_any_pairs(const TRec* __restrict__ recs,
                                                            const float* __restrict__ rays, int64_t nrays_arg,
                                                            const uint32_t* __restrict__ nrays_dev,
                                                            const uint32_t* __restrict__ ids,
                                                            const float* __restrict__ limit,
                                                            const int32_t* __restrict__ src,
                                                            const int32_t* __restrict__ dst,
                                                            uint8_t* __restrict__ ovis,
                                                            uint32_t* __restrict__ overflow,
                                                            unsigned long long* __restrict__ next_ray,
                                                            unsigned long long* __restrict__ fetched,
                                                            AnySpill* __restrict__ spill, uint32_t* spill_n,
                                                            uint32_t spill_cap, int spill_after, int cull,
                                                            int max_steps, int chunk) {
    __shared__ uint32_t litem[kLdsDepth][kB];
    __shared__ uint8_t lsb[kLdsDepth][kB];
    const int lane = threadIdx.x & 63;
    const uint64_t lt = (1ull << lane) - 1ull;
    const int64_t nrays = nrays_dev ? (int64_t)*nrays_dev : nrays_arg;

    int64_t ri = -1;  // ray held by this lane; -1 = needs one, -2 = pool drained
    v3 o = mk3(0, 0, 0), d = o, inv = o;
    float lim = 0.f, cut = INFINITY;
    int sm = -1, dm = -1, steps = 0;
    bool hit = false, use_lim = false, may_spill = false;
    bool drained = false;  // every ray has been handed out
    uint32_t ovf = 0;
    int maxsp = 0;  // deepest stack the reference's traversal reaches on this lane's rays
    uint64_t nfetch = 0;
    volatile uint32_t spill_i[kStack - kLdsDepth];
    volatile uint8_t spill_b[kStack - kLdsDepth];
    int sp = 0;
    uint32_t citem = 0;  // current pair
    uint32_t cbyte = 0;
    int64_t chunk_base = 0;
    int chunk_avail = 0;

    while (true) {
        // ---- hand new rays to idle lanes (wave-uniform control flow)
        const uint64_t need = __ballot(ri == -1);
        if (need) {
            const int cnt = __popcll(need);
            const int rank = __popcll(need & lt);
            int64_t mine;
            if (chunk_avail < cnt) {
                unsigned long long nb = 0;
                const int first = __ffsll((long long)need) - 1;
                if (lane == first) nb = atomicAdd(next_ray, (unsigned long long)chunk);
                nb = __shfl(nb, first, 64);
                mine = rank < chunk_avail ? chunk_base + rank : (int64_t)nb + (rank - chunk_avail);
                chunk_base = (int64_t)nb + (cnt - chunk_avail);
                chunk_avail = chunk - (cnt - chunk_avail);
            } else {
                mine = chunk_base + rank;
                chunk_base += cnt;
                chunk_avail -= cnt;
            }
            if (ri == -1) {
                if (mine >= nrays) {
                    ri = -2;
                } else {
                    ri = ids ? (int64_t)ids[mine] : mine;
                    load_ray(rays, ri, o, d);
                    inv = div3(mk3(1.f, 1.f, 1.f), d);
                    hit = false;
                    sp = 0;
                    steps = 0;
                    citem = 1u;  // the root alone: slot 0, s = 0
                    cbyte = 0u;
                    may_spill = spill_after > 0;
                    lim = limit[ri];
                    use_lim = !(lim == INFINITY);
                    cut = any_cut(cull, use_lim, lim);
                    sm = src ? src[ri] : -1;
                    dm = dst ? dst[ri] : -1;
                }
            }
        }
        if (!__ballot(ri >= 0)) break;  // every lane drained
        if (ri < 0) continue;

        if (may_spill && steps >= spill_after && !drained && ((steps - spill_after) & 255) == 0)
            drained = (int64_t)__hip_atomic_load(next_ray, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) >= nrays;
        if (may_spill && drained && steps >= spill_after) {
            // a long ray at the end: hand its pending pairs to the cooperative pass
            const uint32_t slot = atomicAdd(spill_n, 1u);
            if (slot < spill_cap) {
                AnySpill* rec = spill + slot;
                rec->ray = ri;
                rec->n = sp + 1;
                rec->item[0] = citem;
                rec->byte[0] = (uint8_t)cbyte;
                for (int i = 0; i < sp; ++i) {
                    rec->item[i + 1] = i < kLdsDepth ? *(const volatile lds_uint*)(&litem[i][threadIdx.x])
                                                     : spill_i[i - kLdsDepth];
                    rec->byte[i + 1] = i < kLdsDepth ? lsb[i][threadIdx.x] : spill_b[i - kLdsDepth];
                }
                ri = -1;
                continue;
            }
            may_spill = false;  // the pool is full: finish it here
        }

        // ---- one step: both members of the current pair
        bool done = false;
        if (++steps > max_steps) {
            ++ovf;
            done = true;
        } else {
            const float4* rec = reinterpret_cast<const float4*>(recs + ((int64_t)(citem >> 2) << 1));
            float4 r[8];
#pragma unroll
            for (int k = 0; k < 8; ++k) r[k] = rec[k];
            nfetch += (uint64_t)__popc(citem & 3u);
            PairOut po;
            pair_step(r, citem, cbyte, o, d, inv, sm, dm, use_lim, lim, cut, hit, ovf, maxsp, po);
            if (hit) {
                done = true;
            } else if (po.n > 0) {
                citem = po.item[0];
                cbyte = po.byte[0];
                if (po.n == 2) {
                    if (sp < kStack) {
                        if (sp < kLdsDepth) {
                            litem[sp][threadIdx.x] = po.item[1];
                            lsb[sp][threadIdx.x] = (uint8_t)po.byte[1];
                        } else {
                            spill_i[sp - kLdsDepth] = po.item[1];
                            spill_b[sp - kLdsDepth] = (uint8_t)po.byte[1];
                        }
                        ++sp;
                    } else {
                        ++ovf;  // only under trees deeper than the stack
                    }
                }
            } else if (sp == 0) {
                done = true;
            } else {
                --sp;
                if (sp < kLdsDepth) {
                    citem = *(const volatile lds_uint*)(&litem[sp][threadIdx.x]);
                    cbyte = lsb[sp][threadIdx.x];
                } else {
                    citem = spill_i[sp - kLdsDepth];
                    cbyte = spill_b[sp - kLdsDepth];
                }
            }
        }
        if (done) {
            ovis[ri] = hit ? 0 : 1;
            ri = -1;
        }
    }
    if (ovf) atomicAdd(overflow, ovf);
#pragma unroll
    for (int dd = 32; dd >= 1; dd >>= 1) {
        maxsp = max(maxsp, __shfl_xor(maxsp, dd, 64));
        nfetch += __shfl_xor(nfetch, dd, 64);
    }
    if (lane == 0 && maxsp > 0) atomicMax(overflow + 6, (uint32_t)maxsp);
    if (lane == 0 && nfetch > 0) atomicAdd(fetched, (unsigned long long)nfetch);
}

// Any-hit traversal one node per step, in the reference's order
// (BVH/BVH.h:109-189: both children tested at their parent, the far one
// stacked, the near one next). Same arguments and results as
// k_trace_any_pairs; it fetches one 64-byte record per dependent step
// instead of two, which is the cheaper trade on trees that stay in the
// caches (half the box tests per fetch and no second member to mask). A
// long ray at the end is spilled in the pair format: its current node and
// every stacked node as a one-member pair carrying its stack depth.
__device__ __forceinline__ uint32_t node_item(int ref) {
    const uint32_t slot = (uint32_t)(ref >= 0 ? ref : ~ref);
    return ((slot >> 1) << 2) | (1u << (slot & 1u));
}

template <int kLdsDepth, int kB>
__global__ __launch_bounds__(kB) void k_trace_any(const TRec* __restrict__ recs,
                                                      const float* __restrict__ rays, int64_t nrays_arg,
                                                      const uint32_t* __restrict__ nrays_dev,
                                                      const uint32_t* __restrict__ ids,
                                                      const float* __restrict__ limit,
                                                      const int32_t* __restrict__ src,
                                                      const int32_t* __restrict__ dst,
                                                      uint8_t* __restrict__ ovis,
                                                      uint32_t* __restrict__ overflow,
                                                      unsigned long long* __restrict__ next_ray,
                                                      unsigned long long* __restrict__ fetched,
                                                      AnySpill* __restrict__ spill, uint32_t* spill_n,
                                                      uint32_t spill_cap, int spill_after, int cull,
                                                      int max_steps, int chunk) {
    __shared__ int lstack[kLdsDepth][kB];
    const int lane = threadIdx.x & 63;
    const uint64_t lt = (1ull << lane) - 1ull;
    const int64_t nrays = nrays_dev ? (int64_t)*nrays_dev : nrays_arg;

    int64_t ri = -1;  // ray held by this lane; -1 = needs one, -2 = pool drained
    v3 o = mk3(0, 0, 0), d = o, inv = o;
    float lim = 0.f, cut = INFINITY;
    int sm = -1, dm = -1, cur = 0, steps = 0;
    bool use_lim = false, may_spill = false;
    bool drained = false;  // every ray has been handed out
    uint32_t ovf = 0;
    int maxsp = 0;  // deepest stack this lane reached
    uint64_t nfetch = 0;
    volatile int deep[kStack - kLdsDepth];  // in scratch memory, not registers
    int sp = 0;
    int64_t chunk_base = 0;
    int chunk_avail = 0;

    while (true) {
        // ---- hand new rays to idle lanes (wave-uniform control flow)
        const uint64_t need = __ballot(ri == -1);
        if (need) {
            const int cnt = __popcll(need);
            const int rank = __popcll(need & lt);
            int64_t mine;
            if (chunk_avail < cnt) {
                unsigned long long nb = 0;
                const int first = __ffsll((long long)need) - 1;
                if (lane == first) nb = atomicAdd(next_ray, (unsigned long long)chunk);
                nb = __shfl(nb, first, 64);
                mine = rank < chunk_avail ? chunk_base + rank : (int64_t)nb + (rank - chunk_avail);
                chunk_base = (int64_t)nb + (cnt - chunk_avail);
                chunk_avail = chunk - (cnt - chunk_avail);
            } else {
                mine = chunk_base + rank;
                chunk_base += cnt;
                chunk_avail -= cnt;
            }
            if (ri == -1) {
                if (mine >= nrays) {
                    ri = -2;
                } else {
                    ri = ids ? (int64_t)ids[mine] : mine;
                    load_ray(rays, ri, o, d);
                    inv = div3(mk3(1.f, 1.f, 1.f), d);
                    sp = 0;
                    cur = 0;
                    steps = 0;
                    may_spill = spill_after > 0;
                    lim = limit[ri];
                    use_lim = !(lim == INFINITY);
                    cut = any_cut(cull, use_lim, lim);
                    sm = src ? src[ri] : -1;
                    dm = dst ? dst[ri] : -1;
                }
            }
        }
        if (!__ballot(ri >= 0)) break;  // every lane drained
        if (ri < 0) continue;

        if (may_spill && steps >= spill_after && !drained && ((steps - spill_after) & 255) == 0)
            drained = (int64_t)__hip_atomic_load(next_ray, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) >= nrays;
        if (may_spill && drained && steps >= spill_after) {
            // a long ray at the end: hand its pending nodes to the cooperative pass
            const uint32_t slot = atomicAdd(spill_n, 1u);
            if (slot < spill_cap) {
                AnySpill* rec = spill + slot;
                rec->ray = ri;
                rec->n = sp + 1;
                rec->item[0] = node_item(cur);
                rec->byte[0] = (uint8_t)sp;
                for (int i = 0; i < sp; ++i) {
                    const int ref = i < kLdsDepth ? *(const lds_int*)(&lstack[i][threadIdx.x]) : deep[i - kLdsDepth];
                    rec->item[i + 1] = node_item(ref);
                    rec->byte[i + 1] = (uint8_t)i;
                }
                ri = -1;
                continue;
            }
            may_spill = false;  // the pool is full: finish it here
        }

        // ---- one traversal step for this lane's ray
        bool done = false, hit = false;
        if (++steps > max_steps) {
            ++ovf;
            done = true;
        } else {
            // one 64-byte fetch for either kind of step
            const bool at_node = cur >= 0;
            ++nfetch;
            const float4* rec = reinterpret_cast<const float4*>(recs + (at_node ? cur : ~cur));
            float4 q0 = rec[0], q1 = rec[1], q2 = rec[2];
            int4 refs = reinterpret_cast<const int4*>(rec)[3];
            asm volatile("" : "+v"(q1.w), "+v"(q2.w), "+v"(refs.x), "+v"(refs.y));
            if (at_node) {  // BVH/BVH.h:158-184
                const float dl = slab(q0.x, q0.y, q0.z, q0.w, q1.x, q1.y, o, inv);
                const float dr = slab(q1.z, q1.w, q2.x, q2.y, q2.z, q2.w, o, inv);
                const bool hl = dl >= 0 && dl <= cut, hr = dr >= 0 && dr <= cut;
                bool next = false;
                if (hl && hr) {
                    const bool lfar = dl > dr;
                    const int far = lfar ? refs.x : refs.y;
                    const int near = lfar ? refs.y : refs.x;
                    if (sp < kStack) stack_push<kLdsDepth, kB>(lstack, deep, sp, far); else ++ovf;
                    maxsp = max(maxsp, sp);
                    if (sp < kStack) { cur = near; next = true; } else ++ovf;
                } else if (hl || hr) {
                    if (sp < kStack) { cur = hl ? refs.x : refs.y; next = true; } else ++ovf;
                }
                if (!next) {
                    if (sp == 0) done = true;
                    else cur = stack_pop<kLdsDepth, kB>(lstack, deep, sp);
                }
            } else {  // BVH/BVH.h:136-156
                const TTri tr{q0, q1, q2};
                const int mesh = __float_as_int(tr.b.w);
                float t, u, v;
                hit = mesh != sm && mesh != dm && tri_hit(tr, o, d, t, u, v) && (!use_lim || t < lim);
                if (hit || sp == 0) done = true;
                else cur = stack_pop<kLdsDepth, kB>(lstack, deep, sp);
            }
        }
        if (done) {
            ovis[ri] = hit ? 0 : 1;
            ri = -1;
        }
    }
    if (ovf) atomicAdd(overflow, ovf);
#pragma unroll
    for (int dd = 32; dd >= 1; dd >>= 1) {
        maxsp = max(maxsp, __shfl_xor(maxsp, dd, 64));
        nfetch += __shfl_xor(nfetch, dd, 64);
    }
    if (lane == 0 && maxsp > 0) atomicMax(overflow + 6, (uint32_t)maxsp);
    if (lane == 0 && nfetch > 0) atomicAdd(fetched, (unsigned long long)nfetch);
}

// Phase 2: one wave per spilled ray. The wave keeps the ray's pending pairs
// in LDS and takes up to 64 of them at a time, newest first, one per lane;
// their child pairs go back on top. It stops at the first occluder or when
// nothing is pending. The visited set and overflow count are the reference's
// (the answer does not depend on the order). Should the list outgrow LDS --
// it holds at most 64 pairs per tree level, 3200 at the 50 levels of a
// 2^24-triangle tree -- the ray goes to the fail list and phase 3 traces it
// again from its root on one lane.
constexpr int kCoopCap = 4096;
__global__ __launch_bounds__(64) void k_any_coop(const TRec* __restrict__ recs, const float* __restrict__ rays,
                                                 const float* __restrict__ limit, const int32_t* __restrict__ src,
                                                 const int32_t* __restrict__ dst, uint8_t* __restrict__ ovis,
                                                 uint32_t* __restrict__ overflow,
                                                 unsigned long long* __restrict__ fetched,
                                                 const AnySpill* __restrict__ spill, const uint32_t* spill_n,
                                                 uint32_t spill_cap, uint32_t* next_spill,
                                                 uint32_t* __restrict__ fail_ids, uint32_t* fail_n, int cull) {
    __shared__ uint32_t witem[kCoopCap];
    __shared__ uint8_t wbyte[kCoopCap];
    const int lane = threadIdx.x;
    const uint64_t lt = (1ull << lane) - 1ull;
    const uint32_t total = min(*spill_n, spill_cap);
    uint32_t ovf = 0;
    int maxsp = 0;
    uint64_t nfetch = 0;
    while (true) {
        uint32_t idx = 0;
        if (lane == 0) idx = atomicAdd(next_spill, 1u);
        idx = __shfl(idx, 0, 64);
        if (idx >= total) break;
        const AnySpill* sr = spill + idx;
        const int64_t ri = sr->ray;
        const int n0 = sr->n;
        v3 o, d;
        load_ray(rays, ri, o, d);
        const v3 inv = div3(mk3(1.f, 1.f, 1.f), d);
        const float lim = limit[ri];
        const bool use_lim = !(lim == INFINITY);
        const float cut = any_cut(cull, use_lim, lim);
        const int sm = src ? src[ri] : -1, dm = dst ? dst[ri] : -1;
        for (int i = lane; i < n0; i += 64) {
            witem[i] = sr->item[i];
            wbyte[i] = sr->byte[i];
        }
        __builtin_amdgcn_wave_barrier();
        int count = n0;
        bool occluded = false, failed = false;
        while (count > 0) {
            const int k = count < 64 ? count : 64;
            const bool act = lane < k;
            uint32_t item = 0, byte = 0;
            if (act) {
                item = witem[count - 1 - lane];
                byte = wbyte[count - 1 - lane];
            }
            count -= k;
            __builtin_amdgcn_wave_barrier();
            bool hit = false;
            PairOut po;
            po.n = 0;
            if (act) {
                const float4* rec = reinterpret_cast<const float4*>(recs + ((int64_t)(item >> 2) << 1));
                float4 r[8];
#pragma unroll
                for (int q = 0; q < 8; ++q) r[q] = rec[q];
                nfetch += (uint64_t)__popc(item & 3u);
                pair_step(r, item, byte, o, d, inv, sm, dm, use_lim, lim, cut, hit, ovf, maxsp, po);
            }
            if (__ballot(hit)) {
                occluded = true;
                break;
            }
            // append the child pairs: a wave prefix over the lanes' counts
            int pos = po.n;
#pragma unroll
            for (int dd = 1; dd < 64; dd <<= 1) {
                const int v = __shfl_up(pos, dd, 64);
                if (lane >= dd) pos += v;
            }
            const int added = __shfl(pos, 63, 64);
            pos -= po.n;  // exclusive
            if (count + added > kCoopCap) {
                failed = true;
                break;
            }
            for (int j = 0; j < po.n; ++j) {
                witem[count + pos + j] = po.item[j];
                wbyte[count + pos + j] = (uint8_t)po.byte[j];
            }
            count += added;
            __builtin_amdgcn_wave_barrier();
        }
        if (lane == 0) {
            if (failed) fail_ids[atomicAdd(fail_n, 1u)] = (uint32_t)ri;
            else ovis[ri] = occluded ? 0 : 1;
        }
        (void)lt;
    }
    if (ovf) atomicAdd(overflow, ovf);
#pragma unroll
    for (int dd = 32; dd >= 1; dd >>= 1) {
        maxsp = max(maxsp, __shfl_xor(maxsp, dd, 64));
        nfetch += __shfl_xor(nfetch, dd, 64);
    }
    if (lane == 0 && maxsp > 0) atomicMax(overflow + 6, (uint32_t)maxsp);
    if (lane == 0 && nfetch > 0) atomicAdd(fetched, (unsigned long long)nfetch);
}

// Phase 2 of the closest-hit traversal: one wave per long ray, from the
// root. The wave keeps the pending nodes in LDS (with where their boxes
// start) and takes up to 64 of them at a time, newest first, one per lane;
// the hit children go back on top (the near one last, so it is taken first).
// The nearest hit so far lives in LDS (float bits, atomic min). A node whose
// box starts more than kTieMargin beyond it is not fetched, a child box that
// starts so far is not pushed.
//
// The reference's answer (BVH/BVH.h:58-95) is the nearest hit its
// depth-first order keeps: a later hit must be strictly nearer, and a child
// whose box starts at or beyond the nearest hit so far is skipped. So the
// nearest hit M is its answer unless a box on M's path starts at or beyond
// an earlier hit X; such a box starts at or before M's own box (e_M), so X
// lies at or before max(t_M, e_M). Every hit within kTieMargin of the
// running nearest is recorded (a full list drops those no longer within
// it), and M is the answer when no other recorded hit lies at or before
// max(t_M, e_M) and e_M is within half the margin of t_M (rounding puts a
// triangle's hit a hair before the start of its own box at most; a hit
// that far before it -- an incidence within about 1e-4 rad of its plane --
// would have to be a rival this pass prunes). Otherwise the ray goes to
// phase 3, the reference's order on one lane, as it does when the pending
// list outgrows LDS, when more than kCoopCands hits lie within the margin,
// or when a node could sit kCoopDeep or more entries deep in the
// reference's 64-entry stack (the depth counted here assumes every pair of
// intersected children is pushed, a bound on the reference's).
constexpr int kCoopNodes = 2048;
constexpr int kCoopWidth = 64;  // pending nodes a wave takes per step (at most 64)
constexpr int kCoopCands = 64;
constexpr int kCoopDeep = 56;
constexpr float kTieMargin = 1.0f + 1.0f / 1024.0f;
constexpr float kHalfMargin = 1.0f + 1.0f / 2048.0f;
__global__ __launch_bounds__(64) void k_closest_coop(const TRec* __restrict__ recs, const float* __restrict__ rays,
                                                     int32_t* __restrict__ oprim, float* __restrict__ ot,
                                                     float* __restrict__ ou, float* __restrict__ ov,
                                                     unsigned long long* __restrict__ fetched,
                                                     const uint32_t* __restrict__ long_ids, const uint32_t* long_n,
                                                     uint32_t long_cap, uint32_t* next_long,
                                                     uint32_t* __restrict__ fail_ids, uint32_t* fail_n) {
    __shared__ int32_t wref[kCoopNodes];
    __shared__ float wdist[kCoopNodes];  // where the node's box starts along the ray
    __shared__ uint8_t wdep[kCoopNodes];
    __shared__ float cand_t[kCoopCands];
    __shared__ int32_t cand_p[kCoopCands];
    __shared__ float cand_u[kCoopCands];
    __shared__ float cand_v[kCoopCands];
    __shared__ float cand_e[kCoopCands];  // where the hit triangle's own box starts
    __shared__ uint32_t sbest;  // float bits of the nearest hit so far (hits are at t > 0)
    __shared__ uint32_t ncand;
    const int lane = threadIdx.x;
    const uint64_t lt = (1ull << lane) - 1ull;
    const uint32_t total = min(*long_n, long_cap);
    uint64_t nfetch = 0;
    while (true) {
        uint32_t idx = 0;
        if (lane == 0) idx = atomicAdd(next_long, 1u);
        idx = __shfl(idx, 0, 64);
        if (idx >= total) break;
        const int64_t ri = long_ids[idx];
        v3 o, d;
        load_ray(rays, ri, o, d);
        const v3 inv = div3(mk3(1.f, 1.f, 1.f), d);
        if (lane == 0) {
            wref[0] = 0;  // the root (phase 1 found its box hit)
            wdist[0] = 0.f;
            wdep[0] = 0;
            sbest = __float_as_uint(FLT_MAX);
            ncand = 0u;
        }
        __builtin_amdgcn_wave_barrier();
        int count = 1;
        bool failed = false;
        while (count > 0) {
            const int k = count < kCoopWidth ? count : kCoopWidth;
            const float reach = __uint_as_float(*(volatile uint32_t*)&sbest) * kTieMargin;
            bool act = lane < k;
            int ref = 0, dep = 0;
            float edist = 0.f;
            if (act) {
                ref = wref[count - 1 - lane];
                dep = wdep[count - 1 - lane];
                edist = wdist[count - 1 - lane];
                // a node whose box starts beyond the nearest hit found since it
                // was pushed is not fetched
                act = edist <= reach;
            }
            count -= k;
            __builtin_amdgcn_wave_barrier();
            int nout = 0, out0 = 0, out1 = 0, dep0 = 0, dep1 = 0;
            float dist0 = 0.f, dist1 = 0.f;
            bool bad = false;
            bool pend = false;  // a recorded hit that did not fit in the full list
            float pt = 0.f, pu = 0.f, pv = 0.f, pe = 0.f;
            int pp = 0;
            if (act) {
                const float4* rec = reinterpret_cast<const float4*>(recs + (ref >= 0 ? ref : ~ref));
                const float4 q0 = rec[0], q1 = rec[1], q2 = rec[2];
                const int4 refs = reinterpret_cast<const int4*>(rec)[3];
                ++nfetch;
                if (ref >= 0) {  // BVH/BVH.h:72-95
                    const float dl = slab(q0.x, q0.y, q0.z, q0.w, q1.x, q1.y, o, inv);
                    const float dr = slab(q1.z, q1.w, q2.x, q2.y, q2.z, q2.w, o, inv);
                    const bool hl = dl >= 0 && dl <= reach;
                    const bool hr = dr >= 0 && dr <= reach;
                    // the reference's stack depth, bounded: with both children
                    // intersected, the near one is one entry deeper than the far one
                    const bool both = dl >= 0 && dr >= 0;
                    const bool lfar = dl > dr;
                    const int dnear = both ? dep + 1 : dep;
                    if (hl && hr) {  // far first, near last (taken first)
                        out0 = lfar ? refs.x : refs.y;
                        dist0 = lfar ? dl : dr;
                        dep0 = dep;
                        out1 = lfar ? refs.y : refs.x;
                        dist1 = lfar ? dr : dl;
                        dep1 = dnear;
                        nout = 2;
                    } else if (hl || hr) {
                        out0 = hl ? refs.x : refs.y;
                        dist0 = hl ? dl : dr;
                        dep0 = (hl != lfar) ? dnear : dep;  // hl != lfar: the kept child is the near one
                        nout = 1;
                    }
                    bad = dnear >= kCoopDeep;
                } else {  // BVH/BVH.h:58-70
                    const TTri tr{q0, q1, q2};
                    float t, u, v;
                    if (tri_hit(tr, o, d, t, u, v)) {
                        const uint32_t old = atomicMin(&sbest, __float_as_uint(t));
                        if (t <= __uint_as_float(old) * kTieMargin) {
                            const uint32_t c = atomicAdd(&ncand, 1u);
                            pt = t;
                            pp = __float_as_int(tr.a.w);
                            pu = u;
                            pv = v;
                            pe = edist;
                            if (c < (uint32_t)kCoopCands) {
                                cand_t[c] = pt;
                                cand_p[c] = pp;
                                cand_u[c] = pu;
                                cand_v[c] = pv;
                                cand_e[c] = pe;
                            } else {
                                pend = true;  // the list is full: compacted below
                            }
                        }
                    }
                }
            }
            if (__ballot(bad)) {
                failed = true;
                break;
            }
            if (__ballot(pend)) {
                // keep the recorded hits still within kTieMargin of the nearest
                // (the list holds kCoopCands = 64 of them, one per lane), then
                // add the ones that did not fit
                __builtin_amdgcn_wave_barrier();
                const float lim = __uint_as_float(sbest) * kTieMargin;
                const float ct = cand_t[lane], cu = cand_u[lane], cv = cand_v[lane], ce = cand_e[lane];
                const int cp = cand_p[lane];
                const bool keep = ct <= lim;
                const uint64_t km = __ballot(keep);
                const bool add = pend && pt <= lim;
                const uint64_t am = __ballot(add);
                const int nk = __popcll(km), na = __popcll(am);
                __builtin_amdgcn_wave_barrier();
                if (nk + na > kCoopCands) {
                    failed = true;  // more than 64 hits within a hair of the nearest
                    break;
                }
                if (keep) {
                    const int i = __popcll(km & lt);
                    cand_t[i] = ct;
                    cand_p[i] = cp;
                    cand_u[i] = cu;
                    cand_v[i] = cv;
                    cand_e[i] = ce;
                }
                if (add) {
                    const int i = nk + __popcll(am & lt);
                    cand_t[i] = pt;
                    cand_p[i] = pp;
                    cand_u[i] = pu;
                    cand_v[i] = pv;
                    cand_e[i] = pe;
                }
                if (lane == 0) ncand = (uint32_t)(nk + na);
                __builtin_amdgcn_wave_barrier();
            }
            // append the children: a wave prefix over the lanes' counts
            int pos = nout;
#pragma unroll
            for (int dd = 1; dd < 64; dd <<= 1) {
                const int v = __shfl_up(pos, dd, 64);
                if (lane >= dd) pos += v;
            }
            const int added = __shfl(pos, 63, 64);
            pos -= nout;  // exclusive
            if (count + added > kCoopNodes) {
                failed = true;
                break;
            }
            if (nout > 0) {
                wref[count + pos] = out0;
                wdist[count + pos] = dist0;
                wdep[count + pos] = (uint8_t)dep0;
            }
            if (nout > 1) {
                wref[count + pos + 1] = out1;
                wdist[count + pos + 1] = dist1;
                wdep[count + pos + 1] = (uint8_t)dep1;
            }
            count += added;
            __builtin_amdgcn_wave_barrier();
        }
        __builtin_amdgcn_wave_barrier();
        if (!failed) {
            // The nearest hit M is the reference's answer unless the
            // reference's order skips it: a box on M's path is skipped only
            // at or beyond a hit found before it, and every such box starts
            // at or before M's own box (e_M), so a competitor is a recorded
            // hit other than M at or before max(t_M, e_M). Rounding puts e_M
            // beyond t_M by a hair at most; past half the margin, phase 3.
            const float bt = __uint_as_float(sbest);
            const int nc = (int)min(ncand, (uint32_t)kCoopCands);
            const bool at_min = lane < nc && cand_t[lane] == bt;
            const uint64_t am = __ballot(at_min);
            if (am == 0ull) {  // no hit
                if (lane == 0) {
                    oprim[ri] = -1;
                    ot[ri] = 0.f;
                    if (ou) ou[ri] = 0.f;
                    if (ov) ov[ri] = 0.f;
                }
            } else if (__popcll(am) > 1) {
                failed = true;  // two triangles at the nearest distance: the order decides
            } else {
                const int m = __ffsll((long long)am) - 1;
                const float em = cand_e[m];
                const float bound = fmaxf(bt, em);
                const bool rival = lane < nc && lane != m && cand_t[lane] <= bound;
                if (__ballot(rival) != 0ull || em > bt * kHalfMargin) {
                    failed = true;
                } else if (lane == m) {
                    oprim[ri] = cand_p[m];
                    ot[ri] = bt;
                    if (ou) ou[ri] = cand_u[m];
                    if (ov) ov[ri] = cand_v[m];
                }
            }
        }
        if (failed && lane == 0) fail_ids[atomicAdd(fail_n, 1u)] = (uint32_t)ri;
        __builtin_amdgcn_wave_barrier();
    }
#pragma unroll
    for (int dd = 32; dd >= 1; dd >>= 1) nfetch += __shfl_xor(nfetch, dd, 64);
    if (lane == 0 && nfetch > 0) atomicAdd(fetched, (unsigned long long)nfetch);
}

// 64-bit count of records fetched by closest-hit (words 16-17) or any-hit
// (words 18-19) launches
__host__ __device__ inline unsigned long long* fetch_word(uint32_t* counters, bool closest) {
    return reinterpret_cast<unsigned long long*>(counters + (closest ? 16 : 18));
}

// zeroes n words; a kernel rather than hipMemsetAsync so the launch sequence
// can be captured into a hipGraph
__global__ void k_clear_words(uint32_t* p, int n) {
    if ((int)threadIdx.x < n) p[threadIdx.x] = 0u;
}

// Workgroups of a persistent launch of `threads`-thread workgroups:
// max_per_cu per CU (0: as many as fit).
int persistent_blocks(const void* kernel, int threads, int max_per_cu) {
    // (initialised once, thread-safely: contexts on several threads launch here)
    static const int cus = [] {
        int dev = 0, n = 0;
        if (hipGetDevice(&dev) != hipSuccess || hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || n <= 0)
            n = 256;
        return n;
    }();
    int per = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, kernel, threads, 0) != hipSuccess || per <= 0)
        per = 4 * 256 / threads;
    if (max_per_cu > 0 && per > max_per_cu) per = max_per_cu;
    return cus * per;
}

// One persistent kernel instance (stack depth D, B-thread workgroups) on
// its grid, for `work` rays at most.
template <int D, int B, typename... P, typename... A>
void launch_persistent(const char* name, void (*kernel)(P...), const TraceLaunch& L, int64_t work, LaunchSink& ls,
                       A... args) {
    int64_t blocks = persistent_blocks(reinterpret_cast<const void*>(kernel), B, L.per_cu);
    const int64_t need = (work + B - 1) / B;
    if (blocks > need) blocks = need;
    if (blocks < 1) blocks = 1;
    ls.launch(name, kernel, dim3((unsigned)blocks), dim3(B), args..., L.chunk);
}

// The closest-hit kernel of the launch's stack depth and workgroup size.
template <typename... A>
void launch_trace(const TraceLaunch& L, int64_t work, LaunchSink& ls, A... args) {
    const char* nm = "k_trace<closest>";
    const bool deep = L.lds_depth >= 48;
    if (L.block == 64) {
        if (deep) launch_persistent<48, 64>(nm, &k_trace<48, 64>, L, work, ls, args...);
        else launch_persistent<32, 64>(nm, &k_trace<32, 64>, L, work, ls, args...);
    } else if (L.block == 128) {
        if (deep) launch_persistent<48, 128>(nm, &k_trace<48, 128>, L, work, ls, args...);
        else launch_persistent<32, 128>(nm, &k_trace<32, 128>, L, work, ls, args...);
    } else {
        if (deep) launch_persistent<48, 256>(nm, &k_trace<48, 256>, L, work, ls, args...);
        else launch_persistent<32, 256>(nm, &k_trace<32, 256>, L, work, ls, args...);
    }
}

// The any-hit pair kernel, likewise.
template <typename... A>
void launch_any_pairs(const TraceLaunch& L, int64_t work, LaunchSink& ls, A... args) {
    const char* nm = "k_trace_any_pairs";
    const bool deep = L.lds_depth >= 48;
    if (L.block == 64) {
        if (deep) launch_persistent<48, 64>(nm, &k_trace_any_pairs<48, 64>, L, work, ls, args...);
        else launch_persistent<32, 64>(nm, &k_trace_any_pairs<32, 64>, L, work, ls, args...);
    } else if (L.block == 128) {
        if (deep) launch_persistent<48, 128>(nm, &k_trace_any_pairs<48, 128>, L, work, ls, args...);
        else launch_persistent<32, 128>(nm, &k_trace_any_pairs<32, 128>, L, work, ls, args...);
    } else {
        if (deep) launch_persistent<48, 256>(nm, &k_trace_any_pairs<48, 256>, L, work, ls, args...);
        else launch_persistent<32, 256>(nm, &k_trace_any_pairs<32, 256>, L, work, ls, args...);
    }
}

// The one-node any-hit kernel, likewise.
template <typename... A>
void launch_any_nodes(const TraceLaunch& L, int64_t work, LaunchSink& ls, A... args) {
    const char* nm = "k_trace_any";
    const bool deep = L.lds_depth >= 48;
    if (L.block == 64) {
        if (deep) launch_persistent<48, 64>(nm, &k_trace_any<48, 64>, L, work, ls, args...);
        else launch_persistent<32, 64>(nm, &k_trace_any<32, 64>, L, work, ls, args...);
    } else if (L.block == 128) {
        if (deep) launch_persistent<48, 128>(nm, &k_trace_any<48, 128>, L, work, ls, args...);
        else launch_persistent<32, 128>(nm, &k_trace_any<32, 128>, L, work, ls, args...);
    } else {
        if (deep) launch_persistent<48, 256>(nm, &k_trace_any<48, 256>, L, work, ls, args...);
        else launch_persistent<32, 256>(nm, &k_trace_any<32, 256>, L, work, ls, args...);
    }
}

// The any-hit kernel L.any_step names.
template <typename... A>
void launch_any(const TraceLaunch& L, int64_t work, LaunchSink& ls, A... args) {
    if (any_step_pairs(L)) launch_any_pairs(L, work, ls, args...);
    else launch_any_nodes(L, work, ls, args...);
}

// L.handoff: steps (any-hit: nodes or sibling pairs) a ray must have taken, once the
// rays have all been handed out, to go to a cooperative pass, so that the
// short rays still in flight finish where they are. A closest-hit ray's
// steps are dropped: its cooperative pass starts again from the root. From
// kNoHandoff on, no ray goes and the passes are not launched (small scenes,
// whose rays are all short, save the launches).
constexpr int kNoHandoff = 1 << 30;
// lanes of a phase-3 launch (the rays the cooperative passes send back)
constexpr int64_t kPhase3Lanes = 16384;

// Workgroups of a cooperative pass: as many one-wave workgroups as fit.
int coop_grid(const void* kernel) {
    int dev = 0, cus = 0, per = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus <= 0)
        cus = 256;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, kernel, 64, 0) != hipSuccess || per <= 0) per = 4;
    return cus * per;
}

// Any-hit in three launches: phase 1 (one lane per ray, long rays spilled),
// phase 2 (a wave per spilled ray), phase 3 (rays phase 2 could not hold,
// from their roots, one lane each; normally none). Words 20..25 of the
// counters: spill count, spill cursor, failure count, phase-3 dispenser.
void trace_any(const TRec* rec, int ntris, const float* rays, int64_t nrays, const uint32_t* nrays_dev,
               const float* limit, const int32_t* src, const int32_t* dst, uint8_t* visible, uint32_t* counters,
               const TraceLaunch& L, AnySpill* spill, uint32_t spill_cap, LaunchSink& ls) {
    const bool coop = spill && spill_cap > 0 && nrays < (int64_t(1) << 32) && L.handoff < kNoHandoff;
    auto* disp = reinterpret_cast<unsigned long long*>(counters + 4);
    launch_any(L, nrays, ls, rec, rays, nrays, nrays_dev, (const uint32_t*)nullptr, limit, src, dst, visible,
                     counters + 1, disp, fetch_word(counters, false), coop ? spill : (AnySpill*)nullptr,
                     counters + 20, coop ? spill_cap : 0u, coop ? L.handoff : 0, L.any_full ? 0 : 1,
                     2 * ntris - 1);
    if (!coop) return;
    uint32_t* fail_ids = reinterpret_cast<uint32_t*>(spill + spill_cap);
    static const int blocks = coop_grid(reinterpret_cast<const void*>(&k_any_coop));
    ls.launch("k_any_coop", &k_any_coop, dim3((unsigned)blocks), dim3(64), rec, rays, limit, src, dst, visible,
              counters + 1, fetch_word(counters, false), (const AnySpill*)spill, (const uint32_t*)(counters + 20),
              spill_cap, counters + 21, fail_ids, counters + 22, L.any_full ? 0 : 1);
    // phase 3 on a small grid: its rays are few (normally none), and a
    // full persistent grid costs tens of microseconds to launch and drain
    launch_any(L, std::min<int64_t>(spill_cap, kPhase3Lanes), ls, rec, rays, (int64_t)spill_cap,
                     (const uint32_t*)(counters + 22),
                     (const uint32_t*)fail_ids, limit, src, dst, visible, counters + 1,
                     reinterpret_cast<unsigned long long*>(counters + 24), fetch_word(counters, false),
                     (AnySpill*)nullptr, counters + 20, 0u, 0, L.any_full ? 0 : 1, 2 * ntris - 1);
}

// Closest-hit in three launches, like any-hit: phase 1 (one lane per ray,
// long rays handed over), phase 2 (a wave per long ray, k_closest_coop),
// phase 3 (the rays phase 2 sent back, one lane each, the reference's
// order). Words 26..30 of the counters: long rays, their cursor, phase-3
// dispenser (64-bit, words 28-29), phase-3 rays.
void trace_near(const TRec* rec, int ntris, const float* root, const float* rays, int64_t nrays,
                const uint32_t* nrays_dev, int32_t* prim, float* t, float* u, float* v, uint32_t* counters,
                const TraceLaunch& L, const TraceCoop* coop, LaunchSink& ls) {
    const bool on = coop && coop->longs && coop->long_cap > 0 && nrays < (int64_t(1) << 32) && L.handoff < kNoHandoff;
    LongRays lr;
    if (on) {
        lr.ids = coop->longs;
        lr.n = counters + 26;
        lr.cap = coop->long_cap;
        lr.after = L.handoff;
    }
    launch_trace(L, nrays, ls, rec, root, rays, nrays, nrays_dev, (const uint32_t*)nullptr, prim, t, u, v, counters,
                 reinterpret_cast<unsigned long long*>(counters + 2), fetch_word(counters, true), lr, 2 * ntris - 1);
    if (!on) return;
    uint32_t* fail_ids = coop->longs + coop->long_cap;
    static const int blocks = coop_grid(reinterpret_cast<const void*>(&k_closest_coop));
    ls.launch("k_closest_coop", &k_closest_coop, dim3((unsigned)blocks), dim3(64), rec, rays, prim, t, u, v,
              fetch_word(counters, true), (const uint32_t*)coop->longs, (const uint32_t*)(counters + 26),
              coop->long_cap, counters + 27, fail_ids, counters + 30);
    launch_trace(L, std::min<int64_t>(coop->long_cap, kPhase3Lanes), ls, rec, root, rays, (int64_t)coop->long_cap,
                 (const uint32_t*)(counters + 30), (const uint32_t*)fail_ids, prim, t, u, v, counters,
                 reinterpret_cast<unsigned long long*>(counters + 28), fetch_word(counters, true), LongRays{},
                 2 * ntris - 1);
}

}  // namespace

hipError_t trace_closest(const TRec* rec, int ntris, const float* root,
                         const float* rays, int64_t nrays, int32_t* prim, float* t, float* u,
                         float* v, uint32_t* counters, const TraceLaunch& L, hipStream_t s,
                         const TraceCoop* coop) {
    if (nrays <= 0) return hipSuccess;
    hipError_t e = hipMemsetAsync(counters + 2, 0, 8, s);
    if (e == hipSuccess) e = hipMemsetAsync(counters + 26, 0, 20, s);
    if (e != hipSuccess) return e;
    LaunchSink ls(s);
    trace_near(rec, ntris, root, rays, nrays, nullptr, prim, t, u, v, counters, L, coop, ls);
    return ls.err;
}

hipError_t trace_occluded(const TRec* rec, int ntris, const float* rays,
                          const float* limit, const int32_t* src, const int32_t* dst,
                          int64_t nrays, uint8_t* visible, uint32_t* counters, const TraceLaunch& L,
                          hipStream_t s, const TraceCoop* coop) {
    if (nrays <= 0) return hipSuccess;
    hipError_t e = hipMemsetAsync(counters + 4, 0, 8, s);
    if (e == hipSuccess) e = hipMemsetAsync(counters + 20, 0, 24, s);
    if (e != hipSuccess) return e;
    LaunchSink ls(s);
    trace_any(rec, ntris, rays, nrays, nullptr, limit, src, dst, visible, counters, L,
              coop ? coop->spill : nullptr, coop ? coop->spill_cap : 0u, ls);
    return ls.err;
}

// Queue variants: the ray count is read on the device (no host round trip),
// the grid is the full persistent size. counters as above. The dispensers
// and pass counters are cleared by a kernel, so the sequence can be recorded
// into a graph.
void trace_closest_queue(const TRec* rec, int ntris, const float* root, const float* rays,
                         const uint32_t* nrays_dev, int64_t capacity, int32_t* prim, float* t, float* u, float* v,
                         uint32_t* counters, const TraceLaunch& L, LaunchSink& ls, const TraceCoop* coop) {
    ls.launch("k_clear_words", &k_clear_words, dim3(1), dim3(64), counters + 2, 2);
    if (coop && coop->longs && L.handoff < kNoHandoff)
        ls.launch("k_clear_words", &k_clear_words, dim3(1), dim3(64), counters + 26, 5);
    trace_near(rec, ntris, root, rays, capacity, nrays_dev, prim, t, u, v, counters, L, coop, ls);
}

void trace_occluded_queue(const TRec* rec, int ntris, const float* rays, const uint32_t* nrays_dev,
                          int64_t capacity, const float* limit, const int32_t* src, const int32_t* dst,
                          uint8_t* visible, uint32_t* counters, const TraceLaunch& L, LaunchSink& ls,
                          const TraceCoop* coop) {
    ls.launch("k_clear_words", &k_clear_words, dim3(1), dim3(64), counters + 4, 2);
    if (coop && coop->spill && L.handoff < kNoHandoff)
        ls.launch("k_clear_words", &k_clear_words, dim3(1), dim3(64), counters + 20, 6);
    trace_any(rec, ntris, rays, capacity, nrays_dev, limit, src, dst, visible, counters, L,
              coop ? coop->spill : nullptr, coop ? coop->spill_cap : 0u, ls);
}

hipError_t trace_closest_queue(const TRec* rec, int ntris, const float* root,
                               const float* rays, const uint32_t* nrays_dev, int64_t capacity,
                               int32_t* prim, float* t, float* u, float* v, uint32_t* counters,
                               const TraceLaunch& L, hipStream_t s, const TraceCoop* coop) {
    LaunchSink ls(s);
    trace_closest_queue(rec, ntris, root, rays, nrays_dev, capacity, prim, t, u, v, counters, L, ls, coop);
    return ls.err;
}

hipError_t trace_occluded_queue(const TRec* rec, int ntris, const float* rays,
                                const uint32_t* nrays_dev, int64_t capacity, const float* limit,
                                const int32_t* src, const int32_t* dst, uint8_t* visible,
                                uint32_t* counters, const TraceLaunch& L, hipStream_t s, const TraceCoop* coop) {
    LaunchSink ls(s);
    trace_occluded_queue(rec, ntris, rays, nrays_dev, capacity, limit, src, dst, visible, counters, L, ls, coop);
    return ls.err;
}

}  // namespace cav
