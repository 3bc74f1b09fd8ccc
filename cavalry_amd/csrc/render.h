// Wavefront renderer: device-side views and the launcher; see render.hip.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/cavalry.h"
#include "dev_common.h"
#include "lbvh.h"

namespace cav {

struct HaltonDev {
    const uint32_t* primes;  // 1024
    const uint32_t* start;   // 1024
    const uint16_t* perm;    // sum of the primes
};

struct SceneDev {
    const TNode* tn;
    const TTri* tt;
    const float* root;
    int ntris;
    const float* soup;         // float[ntris][9], original order
    const int32_t* tri_mesh;   // ntris
    const cav_mesh* meshes;
    const float* mesh_area;
    const int32_t* lights;     // mesh ids of area lights
    int nmeshes, nlights;
};

struct CameraDev {
    float m[16];
    float depth;
    int width, height;
};

// ray queue, structure of arrays (rays stay 24-byte records for the tracer)
struct Queue {
    float* ray;      // [cap][6]
    uint32_t* sid;   // sample id
    float* mult;     // [cap][3] path throughput
    float* pdf;      // pdf of the BSDF sample that made the ray
    uint8_t* delta;  // made by a delta BSDF
};

struct Shadow {
    float* ray;      // [cap][6]
    float* limit;
    int32_t* src;
    int32_t* dst;
    uint32_t* sid;
    float* contrib;  // [cap][3] unoccluded contribution
};

struct RenderArgs {
    SceneDev scene;
    HaltonDev halton;
    CameraDev cam;
    int width, height, spp, max_depth;
    int64_t nsamp;
    Queue q[2];
    Shadow sh;
    uint8_t* vis;
    int32_t* hit_prim;
    float* hit_t;
    float* res;                // [nsamp][3]
    float* sxy;                // [nsamp][2]
    uint32_t* counts;          // queue counts (16 words)
    unsigned long long* rays;  // traced closest / shadow rays
    uint32_t* trace_counters;  // traversal counters (8 words)
    float* film_rgb;           // [w*h][3]
    float* film_w;             // [w*h]
};

hipError_t render_path(const RenderArgs& a, hipStream_t s);

}  // namespace cav
