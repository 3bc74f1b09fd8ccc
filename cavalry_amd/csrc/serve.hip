// Pool kernels of the render server: tasks of many requests share one
// wavefront per scene and advance one bounce per step.
//
// A task carries its request slot and its bounce index. Everything a sample
// computes depends only on its request (camera, spp, depth limit, result
// buffers) and on its own sample id, which is its Halton index exactly as in
// a render of that request alone (Core/Sampler.cu:29-51). Its `path`
// dimensions are a function of its own depth, as in the plain loop. So a
// request's film is bit-identical whether it is rendered alone or mixed with
// others, and whether its samples are admitted at once or in chunks.
#include "../../include/cavalry.h"
#include "dev_common.h"
#include "launch.h"
#include "serve.h"
#include "shade.h"

namespace cav {
namespace {

using namespace shade;

// first Halton dimension of bounce `depth` of the `path` integrator:
// 2 for the camera sample, then per bounce 5 + the light's dimensions + 2,
// plus 1 for Russian roulette after the fourth
__device__ inline int path_dim0(int depth, int light_dims) {
    return 2 + (7 + light_dims) * depth + (depth > 4 ? depth - 4 : 0);
}

__global__ __launch_bounds__(kBlock) void k_pool_raygen(PoolDev p, const Chunk* __restrict__ chunks, int nchunks,
                                                        int64_t total, int qn) {
    const int64_t g = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    const bool live = g < total;
    int rid = 0;
    int64_t si = 0;
    v3 ro = mk3(0, 0, 0), rd = ro;
    if (live) {
        int lo = 0, hi = nchunks - 1;  // last chunk with off <= g
        while (lo < hi) {
            const int mid = (lo + hi + 1) / 2;
            if (chunks[mid].off <= g) lo = mid;
            else hi = mid - 1;
        }
        const Chunk c = chunks[lo];
        rid = c.rid;
        const ReqDev& r = p.reqs[rid];
        if (c.spp_round > 0) {  // an rlpath round: c.spp_round samples of every pixel from sample c.first
            const int64_t k = g - c.off;
            si = (k / c.spp_round) * r.spp + c.first + k % c.spp_round;
        } else {
            si = c.first + (g - c.off);
        }
        const int64_t pix = si / r.spp;
        const int x = (int)(pix % r.cam.width), y = (int)(pix / r.cam.width);
        float fx = (float)x, fy = (float)y;
        fx += halton(p.halton, (uint32_t)si, 0);
        fy += halton(p.halton, (uint32_t)si, 1);
        r.sxy[si * 2] = fx;
        r.sxy[si * 2 + 1] = fy;
        r.res[si * 3] = 0.f; r.res[si * 3 + 1] = 0.f; r.res[si * 3 + 2] = 0.f;
        const v3 pl = mk3(fx - r.cam.width / 2.f, r.cam.height / 2.f - fy, r.cam.depth);
        ro = apply_m(r.cam.m, mk3(0, 0, 0));
        rd = nrm(sub3(apply_m(r.cam.m, pl), ro));
    }
    const uint32_t slot = block_push<kBlock>(live, &p.counts[qn]);
    if (live) {
        const Queue& q = p.q[qn];
        float* rr = q.ray + (int64_t)slot * 6;
        rr[0] = ro.x; rr[1] = ro.y; rr[2] = ro.z; rr[3] = rd.x; rr[4] = rd.y; rr[5] = rd.z;
        q.sid[slot] = (uint32_t)si;
        q.mult[(int64_t)slot * 3] = 1.f; q.mult[(int64_t)slot * 3 + 1] = 1.f; q.mult[(int64_t)slot * 3 + 2] = 1.f;
        q.pdf[slot] = 1.f;
        q.delta[slot] = 1;
        p.rid[qn][slot] = rid;
        p.depth[qn][slot] = 0;
        if (p.reqs[rid].integrator == CAV_INTEGRATOR_RLPATH) {
            // genInitialRays' QEntryInfo {-1, -1}; the camera sample used 2 dimensions
            p.rl.q[qn].prev_entry[slot] = -1;
            p.rl.q[qn].prev_cell[slot] = -1;
            p.rl.q[qn].dim[slot] = 2;
        }
    }
}

template <int B>
__global__ __launch_bounds__(B) void k_pool_shade(PoolDev p, int qin) {
    const int qout = qin ^ 1;
    const int64_t i = (int64_t)blockIdx.x * B + threadIdx.x;
    bool live = i < (int64_t)p.counts[qin];
    // rlpath tasks are shaded by the rlpath kernels
    live = live && p.reqs[p.rid[qin][i]].integrator != CAV_INTEGRATOR_RLPATH;
    const int prim = live ? p.hit_prim[i] : -1;
    if (live && prim < 0 && p.scene.env_miss >= 0) {  // a miss sees the background, weighted as the loops do
        const Queue& qi = p.q[qin];
        const ReqDev& r = p.reqs[p.rid[qin][i]];
        float* acc = r.res + (int64_t)qi.sid[i] * 3;
        v3 e = miss_radiance(p.scene, mk3(qi.ray[i * 6 + 3], qi.ray[i * 6 + 4], qi.ray[i * 6 + 5]));
        if (r.integrator != CAV_INTEGRATOR_DIRECT) e = mul3(e, mk3(qi.mult[i * 3], qi.mult[i * 3 + 1], qi.mult[i * 3 + 2]));
        acc[0] = acc[0] + e.x; acc[1] = acc[1] + e.y; acc[2] = acc[2] + e.z;
    }
    live = live && prim >= 0;
    ShadeOut o;
    uint32_t sid = 0;
    int rid = 0, depth = 0;
    if (live) {
        const Queue& qi = p.q[qin];
        sid = qi.sid[i];
        rid = p.rid[qin][i];
        depth = p.depth[qin][i];
        const ReqDev& r = p.reqs[rid];
        const float* ray = qi.ray + i * 6;
        const v3 ro = mk3(ray[0], ray[1], ray[2]), rd = mk3(ray[3], ray[4], ray[5]);
        float* acc = r.res + (int64_t)sid * 3;
        const float hu = p.scene.needs_bary() ? p.hit_u[i] : 0.f, hv = p.scene.needs_bary() ? p.hit_v[i] : 0.f;
        if (r.integrator == CAV_INTEGRATOR_DIRECT) {
            direct_shade(p.scene, p.halton, sid, ro, rd, prim, p.hit_t[i], hu, hv, acc, o);
        } else {
            path_shade(depth, path_dim0(depth, p.scene.light_dims), p.scene, p.halton, sid, ro, rd,
                       mk3(qi.mult[i * 3], qi.mult[i * 3 + 1], qi.mult[i * 3 + 2]), qi.pdf[i], qi.delta[i] != 0,
                       prim, p.hit_t[i], hu, hv, acc, o);
            // the plain loop's last bounce pushes rays that are never traced
            if (depth + 1 >= r.max_depth) o.push_next = false;
        }
    }
    const uint32_t ss = block_push<B>(o.push_sh, &p.counts[2]);
    if (o.push_sh) {
        write_shadow(p.sh, ss, o);
        p.sh.sid[ss] = sid;
        p.sh_rid[ss] = rid;
    }
    const uint32_t ns = block_push<B>(o.push_next, &p.counts[qout]);
    if (o.push_next) {
        write_next(p.q[qout], ns, o);
        p.q[qout].sid[ns] = sid;
        p.rid[qout][ns] = rid;
        p.depth[qout][ns] = (uint8_t)(depth + 1);
    }
}

__global__ __launch_bounds__(kBlock) void k_pool_nee(PoolDev p) {
    const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (i >= (int64_t)p.counts[2] || !p.vis[i]) return;
    const ReqDev& r = p.reqs[p.sh_rid[i]];
    if (r.integrator == CAV_INTEGRATOR_RLPATH) return;  // read by k_qdist
    const uint32_t sid = p.sh.sid[i];
    float* res = r.res + (int64_t)sid * 3;
    res[0] = res[0] + p.sh.contrib[i * 3];
    res[1] = res[1] + p.sh.contrib[i * 3 + 1];
    res[2] = res[2] + p.sh.contrib[i * 3 + 2];
}

__global__ void k_pool_step(PoolDev p, int qin) {
    if (threadIdx.x == 0) {
        p.rays[0] += p.counts[qin];
        p.rays[1] += p.counts[2];
        p.counts[qin] = 0;
        p.counts[2] = 0;
    }
}

}  // namespace

hipError_t pool_raygen(const PoolDev& p, const Chunk* chunks_dev, int nchunks, int64_t total, int qn,
                       hipStream_t s) {
    if (total <= 0) return hipSuccess;
    LaunchSink ls(s);
    ls.launch("k_pool_raygen", &k_pool_raygen, dim3((unsigned)((total + kBlock - 1) / kBlock)), dim3(kBlock), p,
              chunks_dev, nchunks, total, qn);
    return ls.err;
}

void pool_bounce(const PoolDev& p, int qin, int kind, LaunchSink& ls) {
    const unsigned nb = (unsigned)((p.cap + kBlock - 1) / kBlock);
    trace_closest_queue(p.scene.rec, p.scene.ntris, p.scene.root, p.q[qin].ray, p.counts + qin, p.cap, p.hit_prim,
                        p.hit_t, p.scene.needs_bary() ? p.hit_u : nullptr, p.scene.needs_bary() ? p.hit_v : nullptr,
                        p.trace_counters, p.launch, ls, &p.coop);
    const auto grid = [&p](int b) { return dim3((unsigned)((p.cap + b - 1) / b)); };
    if (p.launch.shade_block == 64) ls.launch("k_pool_shade", &k_pool_shade<64>, grid(64), dim3(64), p, qin);
    else if (p.launch.shade_block == 128) ls.launch("k_pool_shade", &k_pool_shade<128>, grid(128), dim3(128), p, qin);
    else ls.launch("k_pool_shade", &k_pool_shade<256>, grid(256), dim3(256), p, qin);
    if (kind != kBouncePlain) pool_rl_lighting(p, qin, ls);
    trace_occluded_queue(p.scene.rec, p.scene.ntris, p.sh.ray, p.counts + 2, p.cap, p.sh.limit, p.sh.src, p.sh.dst,
                         p.vis, p.trace_counters, p.launch, ls, &p.coop);
    ls.launch("k_pool_nee", &k_pool_nee, dim3(nb), dim3(kBlock), p);
    if (kind != kBouncePlain) pool_rl_learn(p, qin, kind == kBounceRlDet, ls);
    ls.launch("k_pool_step", &k_pool_step, dim3(1), dim3(64), p, qin);
}

}  // namespace cav
