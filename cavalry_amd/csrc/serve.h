// Device side of the render server: persistent per-scene ray pools whose
// tasks belong to many requests; see serve.hip and server.cpp.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "launch.h"
#include "render.h"

namespace cav {

constexpr int kPoolSlots = 256;  // requests in flight per pool
constexpr int kPoolDetSlots = 32;  // of those, deterministic rlpath ones (slots 0..31)

// one admitted request, as the pool kernels see it
struct ReqDev {
    CameraDev cam;
    int spp, max_depth, integrator;  // CAV_INTEGRATOR_*
    float* res;                      // [spp * pixels][3]
    float* sxy;                      // [spp * pixels][2]
};

// Samples of a request admitted into the pool at one step: a contiguous run
// of sample ids (spp_round = 0), or an rlpath round, samples first ..
// first + spp_round - 1 of every pixel (n = pixels * spp_round)
struct Chunk {
    int32_t rid;
    int32_t spp_round;
    int64_t first;  // first sample id, or first sample of each pixel
    int64_t n;      // samples
    int64_t off;    // exclusive prefix of n over the chunks of one launch
};

// rlpath in a pool: per-slot learning state (RlReqState; t.Q null for slots
// that are not rlpath requests), the marks of requests with hits this bounce,
// the per-task rlpath state of both queues and of a bounce, and the update
// records of deterministic requests (request slot in the key's top bits).
struct PoolRl {
    RlReqState* req;   // [kPoolSlots]
    uint8_t* live;     // [kPoolSlots]
    RlQueue q[2];
    RlSlots sl;
    uint64_t* rec_keys[2];
    uint32_t* rec_idx[2];
    float* rec_val;
    uint32_t* rec_count;
    uint32_t *rec_hist, *rec_counts, *rec_offsets;
    int64_t rec_cap;
};

struct PoolDev {
    SceneDev scene;
    TraceLaunch launch;  // traversal launch shape
    HaltonDev halton;
    Queue q[2];
    int32_t* rid[2];   // request slot per task
    uint8_t* depth[2]; // bounce index per task
    Shadow sh;
    int32_t* sh_rid;
    uint8_t* vis;
    int32_t* hit_prim;
    float* hit_t;
    float* hit_u;              // hit barycentrics (when scene.needs_bary())
    float* hit_v;
    uint32_t* counts;          // [0], [1] ray queues, [2] shadow queue
    unsigned long long* rays;  // traced closest / shadow rays
    uint32_t* trace_counters;
    TraceCoop coop;      // buffers of the long-ray passes of the traversals
    const ReqDev* reqs;
    int64_t cap;
    PoolRl rl;
};

// Kinds of pool bounce: with or without rlpath tasks, and with deterministic
// rlpath updates (sorted records) on top.
enum { kBouncePlain = 0, kBounceRl = 1, kBounceRlDet = 2 };

// camera rays of the chunks, appended to queue qn
hipError_t pool_raygen(const PoolDev& p, const Chunk* chunks_dev, int nchunks, int64_t total, int qn,
                       hipStream_t s);
// one bounce of every task in queue qin: closest hits, shading (path or
// directlighting per task), shadow rays, their contributions; continuation
// rays go to queue qin ^ 1. Launched on a stream or recorded into a graph;
// ls.err / ls.failed name the first launch that failed.
void pool_bounce(const PoolDev& p, int qin, int kind, LaunchSink& ls);

// rlpath.hip: the rlpath kernels of a pool bounce. pool_rl_lighting after the
// closest hits (syncDimension, computeLighting up to the shadow ray),
// pool_rl_learn after the shadow rays (Q distributions and proposals, the
// table updates of every request with hits, next rays).
void pool_rl_lighting(const PoolDev& p, int qin, LaunchSink& ls);
void pool_rl_learn(const PoolDev& p, int qin, bool det, LaunchSink& ls);
// fresh tables (and cleared syncDimension words) for the rlpath request in
// request slot `slot`, whose state is st
hipError_t pool_rl_init(const PoolDev& p, const RlReqState& st, int nlights, hipStream_t s);

}  // namespace cav
