// LBVH build entry points used by the C ABI layer; see lbvh.hip.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "dev_common.h"
#include "launch.h"
#include "radix_sort.h"

namespace cav {

// Optional per-kernel timing: an event is recorded on the stream after each
// launch, named after the kernel just launched. Durations are the gaps
// between consecutive events.
struct Marker {
    static constexpr int kMax = 160;
    hipStream_t s = nullptr;
    hipEvent_t ev[kMax];
    const char* name[kMax];
    int n = 0;
    bool on = false;
    void start(hipStream_t st) {
        s = st;
        n = 0;
        if (on) { hipEventRecord(ev[0], s); name[0] = "start"; n = 1; }
    }
    void mark(const char* nm) {
        if (on && n < kMax) { hipEventRecord(ev[n], s); name[n++] = nm; }
    }
};
inline void mark(Marker* m, const char* nm) { if (m) m->mark(nm); }

constexpr int kSortHistWords = 8 * 256;  // digit counts of up to 8 radix passes
constexpr int kMaxLevels = 4096;  // node heights supported by the level scheduler
// top levels holding at most this many nodes run in one workgroup
constexpr int kTreeletTailNodes = 256;   // treelet rounds

// Device buffers of one BVH build (library-owned; sized for n triangles).
struct LbvhBuffers {
    void* keys[2] = {nullptr, nullptr};       // u32 or u64 Morton codes
    uint32_t* vals[2] = {nullptr, nullptr};   // triangle indices
    uint32_t* hist = nullptr;                 // kSortHistWords digit counts, 8 tile counters, error word
    uint32_t* counts = nullptr;               // 8 * 256 * tiles sort look-back words
    uint32_t* scene_ord = nullptr;            // 6 order-preserving ints
    uint32_t* bounds_part = nullptr;          // per-workgroup partial bounds (6 each)
    NodeBox* nbox = nullptr;                  // 2n-1
    int4* topo = nullptr;                     // 2n-1
    uint32_t* visited = nullptr;              // 2n-1 (arrivals << 16 | height)
    unsigned long long* gate64 = nullptr;     // bottom_up_gate_words(n): split gates and edge arrivals (gates zero between builds)
    int32_t* ref = nullptr;                   // 2n-1 (restructured trees)
    uint32_t* partial = nullptr;              // scan partials
    uint32_t* leaf_ids = nullptr;             // n (treelet round start list)
    uint32_t* counters = nullptr;             // small device counters
    uint32_t* lvl_list = nullptr;             // 2n-1 internal node ids grouped by height
    uint32_t* lvl_cnt = nullptr;              // kMaxLevels + 2 (count per height, error flag)
    uint32_t* lvl_cursor = nullptr;           // kMaxLevels + 2
    uint32_t* lvl_off = nullptr;              // kMaxLevels + 2 (start of each height in lvl_list)
    uint32_t* h_lvl = nullptr;                // pinned host copy of lvl_cnt
    uint32_t* h_off = nullptr;                // host offsets (kMaxLevels + 2)
    int height = 0;                           // root height of the bucketed tree
    int sorted = 0;                           // which keys/vals buffer is sorted
    uint32_t* order = nullptr;                // = vals[sorted]
};

// tris must be 16-byte aligned
hipError_t lbvh_scene_bounds(const float* tris, int n, uint32_t* scene_ord, uint32_t* part,
                             hipStream_t s, Marker* mk);
hipError_t lbvh_build(LbvhBuffers& b, const float* tris, int n, int morton_bits, hipStream_t s,
                      Marker* mk);
hipError_t lbvh_refs(LbvhBuffers& b, int n, hipStream_t s, Marker* mk, const char* label);
hipError_t lbvh_flatten(LbvhBuffers& b, const float* tris, const int32_t* mesh, int n,
                        bool identity_refs, TRec* rec, float* root, hipStream_t s,
                        Marker* mk);
int64_t lbvh_partials(int64_t m);

// bottomup.hip: node ids, children, parents, boxes and heights of the
// radix tree over the sorted codes, in one climb from the leaves
int64_t bottom_up_gate_words(int n);
hipError_t lbvh_bottom_up(LbvhBuffers& b, const void* codes, int keybits, const float* tris, int n,
                          hipStream_t s, Marker* mk);

// levels.hip: height climb + bucketing by height (treelet rounds)
hipError_t levels_from_list(LbvhBuffers& b, int n, hipStream_t s, Marker* mk);
hipError_t levels_bucket(LbvhBuffers& b, int n, hipStream_t s, Marker* mk);
// deepest leaf (edges from the root) of the current tree into *out (device word)
hipError_t levels_max_depth(const LbvhBuffers& b, int n, uint32_t* out, hipStream_t s);

// treelet.hip: one restructuring round; `first` = leaves still at ids N-1..2N-2
hipError_t treelet_round_impl(LbvhBuffers& b, int n, bool first, hipStream_t s, Marker* mk);

// traverse.hip
// Traversal over a flattened BVH of ntris triangles. counters[0] (closest) and
// counters[1] (occluded) count stack overflows and step-cap exits (either
// means the results are partial); counters[2..5] hold the ray dispensers;
// counters[6] (closest) and counters[7] (occluded) keep the deepest stack any
// ray reached (atomic max, never reset by a launch); the 64-bit words at
// counters[16] (closest) and counters[18] (occluded) add up the 64-byte
// records every launch fetched (never reset by a launch); counters[20..25]
// are the cooperative any-hit pass's spill count, spill cursor, failure
// count and ray dispenser, counters[26..30] the cooperative closest-hit
// pass's long-ray count, cursor, phase-3 ray dispenser (64-bit, 28-29) and
// phase-3 count (all cleared by each launch).
// Launch shape of the persistent traversal kernels (and, for the render
// loops, of the per-sample shading kernels). Each ray's node order is fixed
// and each sample's shading is independent of where it runs, so results do
// not depend on it; the autotuner picks it.
struct TraceLaunch {
    int per_cu = 0;         // workgroups per CU (0: as many as fit)
    int chunk = 256;        // rays a wave takes from the dispenser at a time (64..4096)
    int lds_depth = 32;     // stack entries in LDS (32 or 48); deeper ones in scratch
    int block = 256;        // threads per traversal workgroup (64, 128 or 256)
    int shade_block = 256;  // threads per shading workgroup (64, 128 or 256)
    int handoff = 1024;     // steps before a ray still in flight at the end goes to a cooperative pass
    int any_step = 0;       // any-hit step: 0 or 1 one node, 2 a pair of sibling nodes
    int any_full = 0;       // 1: any-hit visits boxes beyond the ray's limit too, as the reference does
};
// Whether an any-hit step fetches a sibling pair (one 128-byte line, both
// nodes tested) instead of one node. One node per step measured as fast or
// faster from 2^12 to 2^24 triangles (2^22 shadow rays, 0.42 against 0.54 ms
// at 2^12, 59.7 against 63.7 ms at 2^24), so pairs are opt-in.
inline bool any_step_pairs(const TraceLaunch& L) { return L.any_step == 2; }

// Any-hit rays that outgrow a step budget hand their pending sibling pairs
// to a cooperative pass (one wave per ray): one record each.
constexpr int kSpillItems = 65;  // the current pair + a full stack
struct AnySpill {
    int64_t ray;
    int32_t n;                     // pending pairs
    int32_t pad;
    uint32_t item[kSpillItems];    // (base slot >> 1) << 2 | active mask
    uint8_t byte[kSpillItems];     // s of the far or only member | 128: near is the right one
};
// spill capacity for a batch of nrays any-hit rays. Rays spill only once
// every ray has been handed out, so about one per lane of the persistent grid
// (at most 8 workgroups x 256 lanes on 256 CUs) plus the few rays a wave
// still holds in its chunk; a full pool only means more rays finish on
// their lane. The closest-hit pass hands over ray ids only, as many.
inline uint32_t any_spill_capacity(int64_t nrays) {
    const int64_t cap = int64_t(1) << 20;
    return (uint32_t)(nrays < cap ? nrays : cap);
}

// Device buffers of the cooperative passes that finish the long rays still
// in flight once every ray has been handed out (one wave per ray instead of
// one lane); a null buffer leaves every ray on its lane.
struct TraceCoop {
    AnySpill* spill = nullptr;  // any-hit: spill_cap records, then spill_cap ray ids
    uint32_t spill_cap = 0;
    uint32_t* longs = nullptr;  // closest-hit: long_cap ray ids, then long_cap ray ids
    uint32_t long_cap = 0;
};
inline size_t trace_coop_bytes(uint32_t spill_cap, uint32_t long_cap) {
    return (sizeof(AnySpill) + sizeof(uint32_t)) * (size_t)spill_cap + 2 * sizeof(uint32_t) * (size_t)long_cap;
}
// the buffers of one allocation of trace_coop_bytes(spill_cap, long_cap) bytes
inline TraceCoop trace_coop_in(void* base, uint32_t spill_cap, uint32_t long_cap) {
    TraceCoop c;
    c.spill = spill_cap ? static_cast<AnySpill*>(base) : nullptr;
    c.spill_cap = spill_cap;
    c.longs = long_cap ? reinterpret_cast<uint32_t*>(static_cast<unsigned char*>(base) +
                                                     (sizeof(AnySpill) + sizeof(uint32_t)) * (size_t)spill_cap)
                       : nullptr;
    c.long_cap = long_cap;
    return c;
}

// coop: null (every ray traced by one lane) or the cooperative passes' buffers
hipError_t trace_closest(const TRec* rec, int ntris, const float* root,
                         const float* rays, int64_t nrays, int32_t* prim, float* t, float* u,
                         float* v, uint32_t* counters, const TraceLaunch& L, hipStream_t s,
                         const TraceCoop* coop = nullptr);
hipError_t trace_occluded(const TRec* rec, int ntris, const float* rays,
                          const float* limit, const int32_t* src, const int32_t* dst,
                          int64_t nrays, uint8_t* visible, uint32_t* counters, const TraceLaunch& L,
                          hipStream_t s, const TraceCoop* coop = nullptr);

hipError_t trace_closest_queue(const TRec* rec, int ntris, const float* root,
                               const float* rays, const uint32_t* nrays_dev, int64_t capacity,
                               int32_t* prim, float* t, float* u, float* v, uint32_t* counters,
                               const TraceLaunch& L, hipStream_t s, const TraceCoop* coop = nullptr);
hipError_t trace_occluded_queue(const TRec* rec, int ntris, const float* rays,
                                const uint32_t* nrays_dev, int64_t capacity, const float* limit,
                                const int32_t* src, const int32_t* dst, uint8_t* visible,
                                uint32_t* counters, const TraceLaunch& L, hipStream_t s,
                                const TraceCoop* coop = nullptr);
// the same, launched through a sink (a stream or a graph under construction)
void trace_closest_queue(const TRec* rec, int ntris, const float* root, const float* rays,
                         const uint32_t* nrays_dev, int64_t capacity, int32_t* prim, float* t, float* u, float* v,
                         uint32_t* counters, const TraceLaunch& L, LaunchSink& ls, const TraceCoop* coop = nullptr);
void trace_occluded_queue(const TRec* rec, int ntris, const float* rays, const uint32_t* nrays_dev,
                          int64_t capacity, const float* limit, const int32_t* src, const int32_t* dst,
                          uint8_t* visible, uint32_t* counters, const TraceLaunch& L, LaunchSink& ls,
                          const TraceCoop* coop = nullptr);

}  // namespace cav
