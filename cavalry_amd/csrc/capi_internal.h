// Internal definitions shared by the C ABI translation units.
#pragma once
#include <hip/hip_runtime.h>

#include <string>

#include "../../include/cavalry.h"
#include "lbvh.h"

struct cav_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    std::string err;
    void* render = nullptr;  // render-path state owned by the context (render.cpp)
    // traversal launch shapes: [CAV_TRACE_STANDALONE] for cav_trace_*, [CAV_TRACE_WAVEFRONT]
    // for the render loops and the server's pools. Standalone batches are usually large and
    // incoherent over trees far larger than L2, where fewer rays in flight thrash L2 less
    // (3 workgroups per CU: 4% faster on 2^24 random rays); the loops keep all that fit.
    cav::TraceLaunch launch[2] = {cav::TraceLaunch{3, 256, 32}, cav::TraceLaunch{0, 256, 32}};
};

struct cav_bvh {
    cav_ctx* ctx = nullptr;
    int n = 0;
    int morton_bits = 30;
    int opt_rounds = 0;
    cav::LbvhBuffers b;
    cav::TRec* rec = nullptr;   // 2n traversal records (sibling-pair layout)
    float* root = nullptr;      // 8 floats
    int32_t* mesh = nullptr;    // n, original order (may be null)
    // long-ray pass buffers (cav::trace_coop_bytes(spill_cap, spill_cap)),
    // grown to the largest batch traced
    mutable void* spill = nullptr;
    mutable uint32_t spill_cap = 0;
    cav_bvh_stats stats{};
    cav::Marker mk;
    bool events = false;
    int nk = 0;                          // kernel timings of the last build
    const char* kname[cav::Marker::kMax];
    float kms[cav::Marker::kMax];
};

namespace cav {
int fail(cav_ctx* ctx, int code, const std::string& msg);
int hip_fail(cav_ctx* ctx, hipError_t e, const char* what);
int dev_alloc(cav_ctx* ctx, void** p, size_t bytes, const char* what);
// frees context-owned state of other subsystems before the stream goes away
void ctx_release(cav_ctx* ctx);
// noalloc.hip: finds the no-alloc staging array on `device` (current), once per device
hipError_t noalloc_init(int device);
}  // namespace cav
