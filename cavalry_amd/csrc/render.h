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

// one RGBA8 texture inside the scene's texel buffer
struct TexDev {
    int64_t off;  // first texel
    int32_t w, h;
};

// an environment-type light (cav_env_light)
struct EnvDev {
    float color[3];
    int32_t texture;
    float m[9];  // world -> light, row-major
};

struct SceneDev {
    const TRec* rec;
    const float* root;
    int ntris;
    const float* soup;         // float[ntris][9], original order
    const int32_t* tri_mesh;   // ntris
    const cav_mesh* meshes;
    const float* mesh_area;
    const int32_t* lights;     // the light list: area-light mesh id >= 0, or -1 - k for env[k]
    int nmeshes, nlights;
    const float* normals;        // float[ntris][9] per-vertex normals, or null
    const uint8_t* mesh_normals; // per mesh: 1 = shade with the interpolated normals
    const float* uv;             // float[ntris][6] per-vertex texture coordinates, or null (all zero)
    const uint32_t* texels;      // RGBA8 texels of every texture, row 0 = v 0; null when no mesh has one
    const TexDev* tex;           // indexed by cav_mesh::texture
    const EnvDev* env;           // environment-type lights, or null
    int env_miss;                // env light that rays which miss see, or -1
    // Halton dimensions a light sample takes after randInt + rand4: 3 (a
    // triangle and a point on it) when the scene has area lights, else 2 (a
    // sphere direction). syncDimension brings every task of a bounce to the
    // maximum over the queue, which is 3 as soon as any task picked an area
    // light; this fixes it per scene.
    int light_dims;
    // hit barycentrics are needed for vertex normals and texture coordinates
    __host__ __device__ bool needs_bary() const { return normals != nullptr || texels != nullptr; }
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
    TraceLaunch launch;      // traversal launch shape
    HaltonDev halton;
    CameraDev cam;
    int width, height, spp, max_depth;
    int64_t nsamp;           // samples of the whole render (width * height * spp)
    int64_t nq;              // samples of this round (width * height * spp_round): queue capacity
    int spp_round, spp_done; // this round's samples per pixel and those done before it
    Queue q[2];
    Shadow sh;
    uint8_t* vis;
    int32_t* hit_prim;
    float* hit_t;
    float* hit_u;              // hit barycentrics, written when scene.needs_bary()
    float* hit_v;
    float* res;                // [nsamp][3]
    float* sxy;                // [nsamp][2]
    uint32_t* counts;          // queue counts (16 words)
    unsigned long long* rays;  // traced closest / shadow rays
    uint32_t* trace_counters;  // traversal counters (32 words, see lbvh.h)
    TraceCoop coop;            // buffers of the long-ray passes of the traversals
    float* film_rgb;           // [w*h][3]
    float* film_w;             // [w*h]
    float4* film_partial;      // [9][w*h] per-source-pixel partial sums (room for double ones)
    bool film_f64;             // film sums in binary64 (CAV_RENDER_FILM_F64)
    uint32_t* sort_perm;       // [cap] hit queue order by material (null: queue order)
    uint32_t* sort_hist;       // sort_hist_words(cap)
};

// extra per-task state of the rlpath ray queues
struct RlQueue {
    int32_t* prev_entry;  // Q entry of the previous hit (-1: camera ray)
    int32_t* prev_cell;   // direction bin taken there
    int32_t* dim;         // next Halton dimension
};

// rlpath per-slot results, carried from one kernel of a bounce to the next
struct RlSlots {
    int32_t* entry;   // Q entry of the hit
    int32_t* light;   // chosen light
    float* sel;       // its selection probability
    float* lpdf;      // light sample pdf (solid angle)
    float* emis;      // luminance of MIS-weighted emission found by this ray
    float* ldir;      // [cap][3] direction to the light sample
    float* inc;       // [cap][3] unoccluded incident radiance
    int32_t* sh_idx;  // slot of the shadow task
    float* ndir;      // [cap][3] Q-sampled next direction
    float* npdf;
    int32_t* ncell;   // Q-sampled bin, -1 when the distribution was empty
};

// Q table and light statistics of one rlpath render (E = entries, L = lights)
struct RlTables {
    float* Q;        // [E][128]
    float* newQ;     // [E][128]
    int32_t* cnt;    // [E][128] proposals this bounce
    int32_t* total;  // [E][128]
    float* ldist;    // [E][L] light-selection CDF
    float* lres;     // [E][L]
    float* lnew;     // [E][L]
    int32_t* lcount; // [E][L]
    int32_t* lnewc;  // [E][L]
};

// findQEntry's grid (ReinforcementLearningPathTracing.cu:155-165): g entries
// per axis over the scene box
struct Grid {
    float lo[3], size[3];
    int g;
};

// Deterministic table updates of a bounce as (key, value) records
struct RlRecords {
    uint64_t* keys;
    float* val;
    uint32_t* count;
    uint32_t* idx;  // record number, the sort's value
};

// Everything one rlpath request's tasks read and update besides their own
// queue slots: its tables and grid, its per-sample results, its
// syncDimension words ([0] the bounce's synced dimension, [1] whether any
// task sampled a Q bin), its learning-rate scale, its index in the top bits
// of deterministic record keys and whether it uses them. A render has one; a server pool one
// per request slot (t.Q == null for slots that are not rlpath requests).
struct RlReqState {
    RlTables t;
    Grid gr;
    float* res;
    int32_t* dims;
    float alpha_scale;
    uint32_t rlx;
    int max_depth;  // continuation rays of the last bounce are not queued
    int det;        // updates go to the sorted records (CAV_RENDER_DETERMINISTIC), else float atomics
};

struct RlArgs {
    int grid;           // Q entries per axis
    float alpha_scale;  // 1: the reference's learning rate n/(T+n)
    float bounds[6];    // scene bounds (min xyz, max xyz)
    RlQueue q[2];
    RlSlots slots;
    float *Q, *newQ, *ldist, *lres, *lnew;
    int32_t *cnt, *total, *lcount, *lnewc;
    int32_t* dims;  // [0] synced dimension at the bounce start, [1] any Q sample
    // Deterministic table updates: each bounce's updates as records
    // (key, index) + value, sorted by key (table, bin or (entry, light),
    // sample, kind) and summed per key in that order. Null: atomics.
    uint64_t* rec_keys[2];  // rec_cap each
    uint32_t* rec_idx[2];   // rec_cap each
    float* rec_val;         // rec_cap
    uint32_t* rec_count;    // 1
    uint32_t *rec_hist, *rec_counts, *rec_offsets;  // radix sort work space
    int64_t rec_cap;
    // shading order (CAV_RENDER_SORT_BY_MATERIAL): (material, Q entry) keys
    // and queue slots of each bounce's hits, and their radix sort's work space
    uint32_t *skeys[2], *svals[2];
    uint32_t *shist, *scounts, *soffsets;
};

// one spp round (a.spp_done .. a.spp_done + a.spp_round) of each integrator;
// rlpath's Q table and light statistics persist across the rounds of a render
// and are initialised when init_tables is set
// stable counting sort of the first *count hits by material (misses last)
hipError_t sort_hits_by_material(const SceneDev& sc, const int32_t* hit_prim, const uint32_t* count, int64_t cap,
                                 uint32_t* hist, uint32_t* perm, hipStream_t s);
int64_t sort_hist_words(int64_t cap);

hipError_t render_path(const RenderArgs& a, hipStream_t s);
hipError_t render_direct(const RenderArgs& a, hipStream_t s);
hipError_t render_rlpath(const RenderArgs& a, const RlArgs& rl, bool init_tables, hipStream_t s);
// camera rays for every sample and the gathered box-filtered film, shared by
// both integrators
void render_raygen(const RenderArgs& a, hipStream_t s);
void render_film(const RenderArgs& a, hipStream_t s);
void film_launch(int width, int height, int spp, const float* res, const float* sxy, float4* partial, float* rgb,
                 float* wt, hipStream_t s);

}  // namespace cav
