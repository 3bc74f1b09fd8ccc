// Closest-hit and any-hit BVH traversal.
//
// Reference: BVH::intersect and BVH::testVisibility (BVH/BVH.h:23-189), the
// slab test of Core/BoundingBox.h:55-74 and Moller-Trumbore of
// Core/Triangle.h:7-52. The visiting order is the reference's exactly (both
// children tested at their parent, the far child stacked, the near one taken
// next, left first on ties), so hit ids and distances match bit for bit.
//
// Layout read: TRec[2N] (64-byte records in the sibling-pair layout: an
// internal node's two child boxes and child refs, or a leaf's v0/e1/e2 +
// original index + mesh id), rays float[R][6].
#include <cfloat>

#include "dev_common.h"
#include "launch.h"
#include "lbvh.h"

namespace cav {
namespace {

constexpr int kBlock = 256;
constexpr int kStack = 64;  // BVH/BVH.h:41,122

// Traversal stack of kStack entries per ray: the first D in LDS (one column
// per thread, so a wave's accesses to one depth hit 64 distinct banks), the
// rest in private memory, which only unusually deep traversals touch. The
// depth sp is a separate local so that it stays in a register: a depth held
// beside the dynamically indexed spill array would live in scratch memory
// with it, a scratch load and store on every push and pop.
template <int D>
__device__ __forceinline__ void stack_push(int (*lds)[kBlock], volatile int* spill, int& sp, int v) {
    if (sp < D) lds[sp][threadIdx.x] = v;
    else spill[sp - D] = v;
    ++sp;
}
// (the LDS read goes through an LDS-qualified pointer so that the compiler
// cannot merge it with the spill read into one flat load through a selected
// pointer)
typedef __attribute__((address_space(3))) int lds_int;
template <int D>
__device__ __forceinline__ int stack_pop(int (*lds)[kBlock], const volatile int* spill, int& sp) {
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

// Persistent traversal: a fixed grid whose lanes each keep one ray in flight
// and pull a new one as soon as theirs finishes, so a wave is not held by its
// slowest ray (nodes visited per ray vary by 1000x on scenes with many
// coincident Morton codes). Rays are handed out per wave from chunks taken
// with one atomic. Each ray's node order is the reference's, so results do not
// depend on which lane or wave traced it.
template <bool kClosest, int kLdsDepth>
__global__ __launch_bounds__(kBlock) void k_trace(const TRec* __restrict__ recs,
                                                  const float* __restrict__ root,
                                                  const float* __restrict__ rays, int64_t nrays_arg,
                                                  const uint32_t* __restrict__ nrays_dev,
                                                  const float* __restrict__ limit,
                                                  const int32_t* __restrict__ src,
                                                  const int32_t* __restrict__ dst,
                                                  int32_t* __restrict__ oprim,
                                                  float* __restrict__ ot, float* __restrict__ ou,
                                                  float* __restrict__ ov,
                                                  uint8_t* __restrict__ ovis,
                                                  uint32_t* __restrict__ overflow,
                                                  unsigned long long* __restrict__ next_ray,
                                                  unsigned long long* __restrict__ fetched,
                                                  int max_steps, int chunk) {
    __shared__ int lstack[kLdsDepth][kBlock];
    const int lane = threadIdx.x & 63;
    const uint64_t lt = (1ull << lane) - 1ull;
    const int64_t nrays = nrays_dev ? (int64_t)*nrays_dev : nrays_arg;

    int64_t ri = -1;       // ray held by this lane; -1 = needs one, -2 = pool drained
    v3 o = mk3(0, 0, 0), d = o, inv = o;
    float best = FLT_MAX, bu = 0.f, bv = 0.f, lim = 0.f;
    int bp = -1, sm = -1, dm = -1, cur = 0, steps = 0;
    bool hit = false, use_lim = false;
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
                    ri = mine;
                    load_ray(rays, ri, o, d);
                    inv = div3(mk3(1.f, 1.f, 1.f), d);
                    best = FLT_MAX; hit = false; bp = -1; bu = 0.f; bv = 0.f;
                    sp = 0; cur = 0; steps = 0;
                    if (kClosest) {
                        // BVH/BVH.h:34-38: a ray missing the root box has no hit
                        if (slab(root[0], root[1], root[2], root[3], root[4], root[5], o, inv) < 0.f) {
                            oprim[ri] = -1; ot[ri] = 0.f;
                            if (ou) ou[ri] = 0.f;
                            if (ov) ov[ri] = 0.f;
                            ri = -1;
                        }
                    } else {
                        lim = limit[ri];
                        use_lim = !(lim == INFINITY);
                        sm = src ? src[ri] : -1;
                        dm = dst ? dst[ri] : -1;
                    }
                }
            }
        }
        if (!__ballot(ri >= 0)) {
            if (!__ballot(ri == -1)) break;  // every lane drained
            continue;                        // lanes whose ray missed the root refill
        }
        if (ri < 0) continue;

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
            if (at_node) {
                const float dl = slab(q0.x, q0.y, q0.z, q0.w, q1.x, q1.y, o, inv);
                const float dr = slab(q1.z, q1.w, q2.x, q2.y, q2.z, q2.w, o, inv);
                const bool hl = kClosest ? (dl >= 0 && dl < best) : dl >= 0;
                const bool hr = kClosest ? (dr >= 0 && dr < best) : dr >= 0;
                bool next = false;
                if (hl && hr) {
                    const bool lfar = dl > dr;
                    const int far = lfar ? refs.x : refs.y;
                    const int near = lfar ? refs.y : refs.x;
                    if (sp < kStack) stack_push<kLdsDepth>(lstack, spill, sp, far); else ++ovf;
                    maxsp = max(maxsp, sp);
                    if (sp < kStack) { cur = near; next = true; } else ++ovf;
                } else if (hl || hr) {
                    if (sp < kStack) { cur = hl ? refs.x : refs.y; next = true; } else ++ovf;
                }
                if (!next) {
                    if (sp == 0) done = true;
                    else cur = stack_pop<kLdsDepth>(lstack, spill, sp);
                }
            } else {
                const TTri tr{q0, q1, q2};
                float t, u, v;
                if (kClosest) {
                    if (tri_hit(tr, o, d, t, u, v) && (!hit || t < best)) {
                        hit = true; best = t; bu = u; bv = v; bp = __float_as_int(tr.a.w);
                    }
                } else {
                    const int mesh = __float_as_int(tr.b.w);
                    if (mesh != sm && mesh != dm && tri_hit(tr, o, d, t, u, v) && (!use_lim || t < lim))
                        hit = true;  // occluded
                }
                if (!kClosest && hit) done = true;
                else if (sp == 0) done = true;
                else cur = stack_pop<kLdsDepth>(lstack, spill, sp);
            }
        }
        if (done) {
            if (kClosest) {
                oprim[ri] = hit ? bp : -1;
                ot[ri] = hit ? best : 0.f;
                if (ou) ou[ri] = hit ? bu : 0.f;
                if (ov) ov[ri] = hit ? bv : 0.f;
            } else {
                ovis[ri] = hit ? 0 : 1;
            }
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

// Workgroups of a persistent launch: max_per_cu per CU (0: as many as fit).
int persistent_blocks(const void* kernel, int max_per_cu) {
    // (initialised once, thread-safely: contexts on several threads launch here)
    static const int cus = [] {
        int dev = 0, n = 0;
        if (hipGetDevice(&dev) != hipSuccess || hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || n <= 0)
            n = 256;
        return n;
    }();
    int per = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, kernel, kBlock, 0) != hipSuccess || per <= 0) per = 4;
    if (max_per_cu > 0 && per > max_per_cu) per = max_per_cu;
    return cus * per;
}

// The kernel of the launch's stack depth on a persistent grid of its
// workgroups per CU, for `work` rays at most.
template <bool C, typename... A>
void launch_trace(const TraceLaunch& L, int64_t work, LaunchSink& ls, A... args) {
    const bool deep = L.lds_depth >= 48;
    const void* k = deep ? reinterpret_cast<const void*>(&k_trace<C, 48>)
                         : reinterpret_cast<const void*>(&k_trace<C, 32>);
    int64_t blocks = persistent_blocks(k, L.per_cu);
    const int64_t need = (work + kBlock - 1) / kBlock;
    if (blocks > need) blocks = need;
    if (blocks < 1) blocks = 1;
    const char* name = C ? "k_trace<closest>" : "k_trace<any>";
    if (deep) ls.launch(name, &k_trace<C, 48>, dim3((unsigned)blocks), dim3(kBlock), args..., L.chunk);
    else ls.launch(name, &k_trace<C, 32>, dim3((unsigned)blocks), dim3(kBlock), args..., L.chunk);
}

}  // namespace

hipError_t trace_closest(const TRec* rec, int ntris, const float* root,
                         const float* rays, int64_t nrays, int32_t* prim, float* t, float* u,
                         float* v, uint32_t* counters, const TraceLaunch& L, hipStream_t s) {
    if (nrays <= 0) return hipSuccess;
    hipError_t e = hipMemsetAsync(counters + 2, 0, 8, s);
    if (e != hipSuccess) return e;
    LaunchSink ls(s);
    launch_trace<true>(L, nrays, ls, rec, root, rays, nrays, (const uint32_t*)nullptr, (const float*)nullptr,
                       (const int32_t*)nullptr, (const int32_t*)nullptr, prim, t, u, v, (uint8_t*)nullptr,
                       counters, reinterpret_cast<unsigned long long*>(counters + 2), fetch_word(counters, true),
                       2 * ntris - 1);
    return ls.err;
}

hipError_t trace_occluded(const TRec* rec, int ntris, const float* rays,
                          const float* limit, const int32_t* src, const int32_t* dst,
                          int64_t nrays, uint8_t* visible, uint32_t* counters, const TraceLaunch& L,
                          hipStream_t s) {
    if (nrays <= 0) return hipSuccess;
    hipError_t e = hipMemsetAsync(counters + 4, 0, 8, s);
    if (e != hipSuccess) return e;
    LaunchSink ls(s);
    launch_trace<false>(L, nrays, ls, rec, (const float*)nullptr, rays, nrays, (const uint32_t*)nullptr, limit, src,
                        dst, (int32_t*)nullptr, (float*)nullptr, (float*)nullptr, (float*)nullptr, visible,
                        counters + 1, reinterpret_cast<unsigned long long*>(counters + 4), fetch_word(counters, false),
                        2 * ntris - 1);
    return ls.err;
}

// Queue variants: the ray count is read on the device (no host round trip),
// the grid is the full persistent size. counters as above. The dispenser is
// cleared by a kernel, so the sequence can be recorded into a graph.
void trace_closest_queue(const TRec* rec, int ntris, const float* root, const float* rays,
                         const uint32_t* nrays_dev, int64_t capacity, int32_t* prim, float* t, float* u, float* v,
                         uint32_t* counters, const TraceLaunch& L, LaunchSink& ls) {
    ls.launch("k_clear_words", &k_clear_words, dim3(1), dim3(64), counters + 2, 2);
    launch_trace<true>(L, capacity, ls, rec, root, rays, capacity, nrays_dev, (const float*)nullptr,
                       (const int32_t*)nullptr, (const int32_t*)nullptr, prim, t, u, v, (uint8_t*)nullptr,
                       counters, reinterpret_cast<unsigned long long*>(counters + 2), fetch_word(counters, true),
                       2 * ntris - 1);
}

void trace_occluded_queue(const TRec* rec, int ntris, const float* rays, const uint32_t* nrays_dev,
                          int64_t capacity, const float* limit, const int32_t* src, const int32_t* dst,
                          uint8_t* visible, uint32_t* counters, const TraceLaunch& L, LaunchSink& ls) {
    ls.launch("k_clear_words", &k_clear_words, dim3(1), dim3(64), counters + 4, 2);
    launch_trace<false>(L, capacity, ls, rec, (const float*)nullptr, rays, capacity, nrays_dev, limit, src, dst,
                        (int32_t*)nullptr, (float*)nullptr, (float*)nullptr, (float*)nullptr, visible,
                        counters + 1, reinterpret_cast<unsigned long long*>(counters + 4), fetch_word(counters, false),
                        2 * ntris - 1);
}

hipError_t trace_closest_queue(const TRec* rec, int ntris, const float* root,
                               const float* rays, const uint32_t* nrays_dev, int64_t capacity,
                               int32_t* prim, float* t, float* u, float* v, uint32_t* counters,
                               const TraceLaunch& L, hipStream_t s) {
    LaunchSink ls(s);
    trace_closest_queue(rec, ntris, root, rays, nrays_dev, capacity, prim, t, u, v, counters, L, ls);
    return ls.err;
}

hipError_t trace_occluded_queue(const TRec* rec, int ntris, const float* rays,
                                const uint32_t* nrays_dev, int64_t capacity, const float* limit,
                                const int32_t* src, const int32_t* dst, uint8_t* visible,
                                uint32_t* counters, const TraceLaunch& L, hipStream_t s) {
    LaunchSink ls(s);
    trace_occluded_queue(rec, ntris, rays, nrays_dev, capacity, limit, src, dst, visible, counters, L, ls);
    return ls.err;
}

}  // namespace cav
