// Device side of the render server: persistent per-scene ray pools whose
// tasks belong to many requests; see serve.hip and server.cpp.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "launch.h"
#include "render.h"

namespace cav {

// one admitted request, as the pool kernels see it
struct ReqDev {
    CameraDev cam;
    int spp, max_depth, integrator;  // CAV_INTEGRATOR_PATH or _DIRECT
    float* res;                      // [spp * pixels][3]
    float* sxy;                      // [spp * pixels][2]
};

// a contiguous run of a request's sample ids admitted into the pool
struct Chunk {
    int32_t rid;
    int32_t pad;
    int64_t first;  // first sample id
    int64_t n;      // samples
    int64_t off;    // exclusive prefix of n over the chunks of one launch
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
    const ReqDev* reqs;
    int64_t cap;
};

// camera rays of the chunks, appended to queue qn
hipError_t pool_raygen(const PoolDev& p, const Chunk* chunks_dev, int nchunks, int64_t total, int qn,
                       hipStream_t s);
// one bounce of every task in queue qin: closest hits, shading (path or
// directlighting per task), shadow rays, their contributions; continuation
// rays go to queue qin ^ 1. Launched on a stream or recorded into a graph;
// ls.err / ls.failed name the first launch that failed.
void pool_bounce(const PoolDev& p, int qin, LaunchSink& ls);

}  // namespace cav
