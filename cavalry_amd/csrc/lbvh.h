// LBVH build entry points used by the C ABI layer; see lbvh.hip.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "dev_common.h"
#include "radix_sort.h"

namespace cav {

struct StageEvents {
    hipEvent_t e[12];
};

// Device buffers of one BVH build (library-owned; sized for n triangles).
struct LbvhBuffers {
    void* keys[2] = {nullptr, nullptr};       // u32 or u64 Morton codes
    uint32_t* vals[2] = {nullptr, nullptr};   // triangle indices
    uint32_t* hist = nullptr;                 // 8 * 256
    uint32_t* counts = nullptr;               // 256 * tiles
    uint32_t* offsets = nullptr;              // 256 * tiles
    uint32_t* scene_ord = nullptr;            // 6 order-preserving ints
    NodeBox* nbox = nullptr;                  // 2n-1
    int4* topo = nullptr;                     // 2n-1
    uint32_t* visited = nullptr;              // 2n-1
    int32_t* ref = nullptr;                   // 2n-1 (restructured trees)
    uint32_t* partial = nullptr;              // scan partials
    uint32_t* leaf_ids = nullptr;             // n (treelet round start list)
    uint32_t* counters = nullptr;             // small device counters
    int sorted = 0;                           // which keys/vals buffer is sorted
    uint32_t* order = nullptr;                // = vals[sorted]
};

hipError_t lbvh_scene_bounds(const float* tris, int n, uint32_t* scene_ord, hipStream_t s);
hipError_t lbvh_build(LbvhBuffers& b, const float* tris, int n, int morton_bits, hipStream_t s,
                      StageEvents* ev);
hipError_t lbvh_refs(LbvhBuffers& b, int n, hipStream_t s);
hipError_t lbvh_flatten(LbvhBuffers& b, const float* tris, const int32_t* mesh, int n,
                        bool identity_refs, TNode* tn, TTri* tt, float* root, hipStream_t s);
int64_t lbvh_partials(int64_t m);

// treelet.hip: one restructuring round; `first` = leaves still at ids N-1..2N-2
hipError_t treelet_round_impl(LbvhBuffers& b, int n, bool first, hipStream_t s);

// traverse.hip
// Traversal over a flattened BVH of ntris triangles. overflow counts stack
// overflows and step-cap exits (either means the results are partial).
hipError_t trace_closest(const TNode* tn, const TTri* tt, int ntris, const float* root,
                         const float* rays, int64_t nrays, int32_t* prim, float* t, float* u,
                         float* v, uint32_t* overflow, hipStream_t s);
hipError_t trace_occluded(const TNode* tn, const TTri* tt, int ntris, const float* rays,
                          const float* limit, const int32_t* src, const int32_t* dst,
                          int64_t nrays, uint8_t* visible, uint32_t* overflow, hipStream_t s);

}  // namespace cav
