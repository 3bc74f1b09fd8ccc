// Wavefront `path` integrator kernels.
//
// Reference: Integrators/PathTracingIntegrator.cu (genInitialRays,
// intersectScene, computeLighting, evalMaterial, genNextRay,
// addSamplesToFilm) and the device code it calls: HaltonSampler
// (Samplers/HaltonSampler.h), PerspectiveCamera::genRay, the matte / plastic /
// mirror / glass BSDFs (BSDFs/*.h), DiffuseAreaLight + TriangleMesh::sample
// (Lights/DiffuseAreaLight.h, Core/TriangleMesh.h:108-153), the box filter
// (Core/Film.h:25-49, Filters/BoxFilter.h).
//
// One bounce = closest-hit traversal of the ray queue, k_shade (emission with
// MIS, light sample, BSDF sample, Russian roulette; pushes a shadow task and a
// next-ray task with one atomic per wave), any-hit traversal of the shadow
// queue, k_nee (adds the contributions of unoccluded shadow tasks). Queues are
// structure-of-arrays; counts stay on the device, so the bounce loop never
// waits for the host. Every sample's Halton dimensions are a fixed function
// of the depth (the reference's syncDimension never changes them for this
// integrator), so each sample's result does not depend on its queue mates.
// The film is gathered per pixel in a fixed order, without float atomics.
#include <cfloat>

#include "../../include/cavalry.h"
#include "dev_common.h"
#include "render.h"
#include "shade.h"

namespace cav {
namespace {

using namespace shade;

// genCameraSampleSPPs (Core/Sampler.cu:29-51) + genInitialRays
// (PathTracingIntegrator.cu:252-266) for one round of `spp_round` samples per
// pixel starting at `spp_done`: queue slot j is sample k = spp_done +
// j % spp_round of pixel j / spp_round, whose Halton index (and sample id) is
// pixel * spp + k.
__global__ __launch_bounds__(kBlock) void k_raygen(int64_t nq, int width, int spp, int spp_round,
                                                   int spp_done, CameraDev cam, HaltonDev h, Queue q,
                                                   float* __restrict__ res, float* __restrict__ sxy,
                                                   uint32_t* counts) {
    const int64_t j = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (j == 0) counts[0] = (uint32_t)nq;
    if (j >= nq) return;
    const int64_t pix = j / spp_round;
    const int64_t si = pix * spp + spp_done + j % spp_round;
    const int x = (int)(pix % width), y = (int)(pix / width);
    float fx = (float)x, fy = (float)y;
    fx += halton(h, (uint32_t)si, 0);
    fy += halton(h, (uint32_t)si, 1);
    sxy[si * 2] = fx;
    sxy[si * 2 + 1] = fy;
    const v3 pl = mk3(fx - cam.width / 2.f, cam.height / 2.f - fy, cam.depth);
    const v3 ro = apply_m(cam.m, mk3(0, 0, 0));
    const v3 rd = nrm(sub3(apply_m(cam.m, pl), ro));
    float* r = q.ray + j * 6;
    r[0] = ro.x; r[1] = ro.y; r[2] = ro.z; r[3] = rd.x; r[4] = rd.y; r[5] = rd.z;
    q.sid[j] = (uint32_t)si;
    q.mult[j * 3] = 1.f; q.mult[j * 3 + 1] = 1.f; q.mult[j * 3 + 2] = 1.f;
    q.pdf[j] = 1.f;
    q.delta[j] = 1;
    res[si * 3] = 0.f; res[si * 3 + 1] = 0.f; res[si * 3 + 2] = 0.f;
}

// computeLighting + the shading half of evalMaterial + genNextRay for one
// bounce (PathTracingIntegrator.cu:124-249); the per-task work is
// shade::path_shade.
template <int B>
__global__ __launch_bounds__(B) void k_shade(int depth, int dim0, SceneDev sc, HaltonDev h,
                                                  Queue qi, Queue qo, Shadow sh,
                                                  const int32_t* __restrict__ hit_prim,
                                                  const float* __restrict__ hit_t,
                                                  const float* __restrict__ hit_u,
                                                  const float* __restrict__ hit_v,
                                                  float* __restrict__ res, uint32_t* counts,
                                                  int qin, int qout, const uint32_t* __restrict__ perm) {
    const int64_t j = (int64_t)blockIdx.x * B + threadIdx.x;
    const uint32_t n = counts[qin];
    bool live = j < (int64_t)n;
    const int64_t i = live ? task_slot(perm, j) : 0;
    int prim = live ? hit_prim[i] : -1;
    if (live && prim < 0 && sc.env_miss >= 0) {  // intersectScene's miss (PathTracingIntegrator.cu:112-116)
        float* acc = res + (int64_t)qi.sid[i] * 3;
        const v3 e = mul3(miss_radiance(sc, mk3(qi.ray[i * 6 + 3], qi.ray[i * 6 + 4], qi.ray[i * 6 + 5])),
                          mk3(qi.mult[i * 3], qi.mult[i * 3 + 1], qi.mult[i * 3 + 2]));
        acc[0] = acc[0] + e.x; acc[1] = acc[1] + e.y; acc[2] = acc[2] + e.z;
    }
    live = live && prim >= 0;
    ShadeOut o;
    uint32_t sid = 0;
    if (live) {
        sid = qi.sid[i];
        const float* r = qi.ray + i * 6;
        path_shade(depth, dim0, sc, h, sid, mk3(r[0], r[1], r[2]), mk3(r[3], r[4], r[5]),
                   mk3(qi.mult[i * 3], qi.mult[i * 3 + 1], qi.mult[i * 3 + 2]), qi.pdf[i], qi.delta[i] != 0, prim,
                   hit_t[i], hit_u ? hit_u[i] : 0.f, hit_u ? hit_v[i] : 0.f, res + (int64_t)sid * 3, o);
    }
    const uint32_t ss = block_push<B>(o.push_sh, &counts[2]);
    if (o.push_sh) {
        write_shadow(sh, ss, o);
        sh.sid[ss] = sid;
    }
    const uint32_t ns = block_push<B>(o.push_next, &counts[qout]);
    if (o.push_next) {
        write_next(qo, ns, o);
        qo.sid[ns] = sid;
    }
}

// DirectLighting::renderRay (Integrators/DirectLightingIntegrator.cu:22-63);
// the per-task work is shade::direct_shade.
__global__ __launch_bounds__(kBlock) void k_direct(SceneDev sc, HaltonDev h, Queue qi, Shadow sh,
                                                   const int32_t* __restrict__ hit_prim,
                                                   const float* __restrict__ hit_t,
                                                   const float* __restrict__ hit_u,
                                                   const float* __restrict__ hit_v, float* __restrict__ res,
                                                   uint32_t* counts, const uint32_t* __restrict__ perm) {
    const int64_t j = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    bool live = j < (int64_t)counts[0];
    const int64_t i = live ? task_slot(perm, j) : 0;
    const int prim = live ? hit_prim[i] : -1;
    if (live && prim < 0 && sc.env_miss >= 0) {  // DirectLightingIntegrator.cu:31-36: the background
        float* acc = res + (int64_t)qi.sid[i] * 3;
        const v3 e = miss_radiance(sc, mk3(qi.ray[i * 6 + 3], qi.ray[i * 6 + 4], qi.ray[i * 6 + 5]));
        acc[0] = acc[0] + e.x; acc[1] = acc[1] + e.y; acc[2] = acc[2] + e.z;
    }
    live = live && prim >= 0;
    ShadeOut o;
    uint32_t sid = 0;
    if (live) {
        sid = qi.sid[i];
        const float* r = qi.ray + i * 6;
        direct_shade(sc, h, sid, mk3(r[0], r[1], r[2]), mk3(r[3], r[4], r[5]), prim, hit_t[i],
                     hit_u ? hit_u[i] : 0.f, hit_u ? hit_v[i] : 0.f, res + (int64_t)sid * 3, o);
    }
    const uint32_t ss = block_push<kBlock>(o.push_sh, &counts[2]);
    if (o.push_sh) {
        write_shadow(sh, ss, o);
        sh.sid[ss] = sid;
    }
}

// evalMaterial's accumulation for unoccluded shadow rays.
__global__ __launch_bounds__(kBlock) void k_nee(Shadow sh, const uint8_t* __restrict__ vis,
                                                float* __restrict__ res, const uint32_t* counts) {
    const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (i >= (int64_t)counts[2] || !vis[i]) return;
    const uint32_t sid = sh.sid[i];
    res[(int64_t)sid * 3] = res[(int64_t)sid * 3] + sh.contrib[i * 3];
    res[(int64_t)sid * 3 + 1] = res[(int64_t)sid * 3 + 1] + sh.contrib[i * 3 + 1];
    res[(int64_t)sid * 3 + 2] = res[(int64_t)sid * 3 + 2] + sh.contrib[i * 3 + 2];
}

// end of a bounce: tally traced rays, empty the consumed queues
__global__ void k_step(uint32_t* counts, unsigned long long* rays, int qin) {
    if (threadIdx.x == 0) {
        rays[0] += counts[qin];
        rays[1] += counts[2];
        counts[qin] = 0;
        counts[2] = 0;
    }
}

// addSamplesToFilm with the box filter (Core/Film.h:25-49,
// Filters/BoxFilter.h:19-25) in a fixed order instead of float atomics, in
// two passes. k_film_partial: for source pixel p and each of the 9 target
// offsets (dy, dx) in row-major order, the sum over p's samples k = 0..spp-1
// (in order) of those whose filter support covers pixel p + (dx, dy).
// k_film_final: each target pixel adds the partials of its 3x3 neighbours in
// row-major neighbour order. Samples with a non-finite value are skipped.
// Partial and final sums in T: float (the reference's FP32 film) or double
// (CAV_RENDER_FILM_F64: the same sums in the same order in binary64, rounded
// to binary32 once when written out).
template <typename T>
struct alignas(4 * sizeof(T)) Sum4 {
    T x, y, z, w;
};

template <typename T>
__global__ __launch_bounds__(kBlock) void k_film_partial(int width, int height, int spp,
                                                         const float* __restrict__ res,
                                                         const float* __restrict__ sxy,
                                                         Sum4<T>* __restrict__ partial) {
    const int64_t pix = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    const int64_t npix = (int64_t)width * height;
    if (pix >= npix) return;
    const int sx0 = (int)(pix % width), sy0 = (int)(pix / width);
    Sum4<T> acc[9];
#pragma unroll
    for (int d = 0; d < 9; ++d) acc[d] = Sum4<T>{0, 0, 0, 0};
    for (int k = 0; k < spp; ++k) {
        const int64_t si = pix * spp + k;
        const float c0 = res[si * 3], c1 = res[si * 3 + 1], c2 = res[si * 3 + 2];
        if (!isfinite(c0) || !isfinite(c1) || !isfinite(c2)) continue;
        const float fx = sxy[si * 2], fy = sxy[si * 2 + 1];
        const int x_lo = (int)floorf(fx - 1.f), x_hi = (int)ceilf(fx + 1.f);
        const int y_lo = (int)floorf(fy - 1.f), y_hi = (int)ceilf(fy + 1.f);
#pragma unroll
        for (int d = 0; d < 9; ++d) {
            const int x = sx0 + d % 3 - 1, y = sy0 + d / 3 - 1;
            if (x < x_lo || x > x_hi || y < y_lo || y > y_hi) continue;
            if (fabsf((float)x - fx) > 1.f || fabsf((float)y - fy) > 1.f) continue;
            acc[d].x += (T)c0; acc[d].y += (T)c1; acc[d].z += (T)c2; acc[d].w += (T)1;
        }
    }
#pragma unroll
    for (int d = 0; d < 9; ++d) partial[d * npix + pix] = acc[d];
}

template <typename T>
__global__ __launch_bounds__(kBlock) void k_film_final(int width, int height,
                                                       const Sum4<T>* __restrict__ partial,
                                                       float* __restrict__ rgb, float* __restrict__ wt) {
    const int64_t pix = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    const int64_t npix = (int64_t)width * height;
    if (pix >= npix) return;
    const int x = (int)(pix % width), y = (int)(pix / width);
    T r = 0, g = 0, b = 0, w = 0;
    for (int py = y - 1; py <= y + 1; ++py)
        for (int px = x - 1; px <= x + 1; ++px) {
            if (px < 0 || py < 0 || px >= width || py >= height) continue;
            const int d = (y - py + 1) * 3 + (x - px + 1);  // offset from the source to this pixel
            const Sum4<T> v = partial[d * npix + (int64_t)py * width + px];
            r += v.x; g += v.y; b += v.z; w += v.w;
        }
    rgb[pix * 3] = (float)r;
    rgb[pix * 3 + 1] = (float)g;
    rgb[pix * 3 + 2] = (float)b;
    wt[pix] = (float)w;
}

template <typename T>
void film_sums(int width, int height, int spp, const float* res, const float* sxy, Sum4<T>* partial, float* rgb,
               float* wt, hipStream_t s) {
    const int64_t npix = (int64_t)width * height;
    const dim3 grid((unsigned)((npix + kBlock - 1) / kBlock));
    hipLaunchKernelGGL(k_film_partial<T>, grid, dim3(kBlock), 0, s, width, height, spp, res, sxy, partial);
    hipLaunchKernelGGL(k_film_final<T>, grid, dim3(kBlock), 0, s, width, height, partial, rgb, wt);
}

}  // namespace

void render_raygen(const RenderArgs& a, hipStream_t s) {
    const unsigned nb = (unsigned)((a.nq + kBlock - 1) / kBlock);
    hipLaunchKernelGGL(k_raygen, dim3(nb), dim3(kBlock), 0, s, a.nq, a.width, a.spp, a.spp_round,
                       a.spp_done, a.cam, a.halton, a.q[0], a.res, a.sxy, a.counts);
}

void film_launch(int width, int height, int spp, const float* res, const float* sxy, float4* partial, float* rgb,
                 float* wt, hipStream_t s) {
    film_sums(width, height, spp, res, sxy, reinterpret_cast<Sum4<float>*>(partial), rgb, wt, s);
}

void render_film(const RenderArgs& a, hipStream_t s) {
    if (a.film_f64)  // film_partial holds 9 x w*h double sums (twice the float ones)
        film_sums(a.width, a.height, a.spp, a.res, a.sxy, reinterpret_cast<Sum4<double>*>(a.film_partial),
                  a.film_rgb, a.film_w, s);
    else
        film_launch(a.width, a.height, a.spp, a.res, a.sxy, a.film_partial, a.film_rgb, a.film_w, s);
}

hipError_t render_direct(const RenderArgs& a, hipStream_t s) {
    const int64_t nq = a.nq;
    const unsigned nb = (unsigned)((nq + kBlock - 1) / kBlock);
    hipError_t e = hipMemsetAsync(a.counts, 0, 64, s);
    if (e != hipSuccess) return e;
    render_raygen(a, s);
    float* hu = a.scene.needs_bary() ? a.hit_u : nullptr;
    float* hv = a.scene.needs_bary() ? a.hit_v : nullptr;
    if ((e = trace_closest_queue(a.scene.rec, a.scene.ntris, a.scene.root, a.q[0].ray, a.counts, nq,
                                 a.hit_prim, a.hit_t, hu, hv, a.trace_counters, a.launch, s, &a.coop)) != hipSuccess)
        return e;
    if (a.sort_perm && (e = sort_hits_by_material(a.scene, a.hit_prim, a.counts, nq, a.sort_hist, a.sort_perm, s)) !=
                           hipSuccess)
        return e;
    hipLaunchKernelGGL(k_direct, dim3(nb), dim3(kBlock), 0, s, a.scene, a.halton, a.q[0], a.sh, a.hit_prim, a.hit_t,
                       hu, hv, a.res, a.counts, a.sort_perm);
    if ((e = trace_occluded_queue(a.scene.rec, a.scene.ntris, a.sh.ray, a.counts + 2, nq, a.sh.limit,
                                  a.sh.src, a.sh.dst, a.vis, a.trace_counters, a.launch, s, &a.coop)) != hipSuccess)
        return e;
    hipLaunchKernelGGL(k_nee, dim3(nb), dim3(kBlock), 0, s, a.sh, a.vis, a.res, a.counts);
    hipLaunchKernelGGL(k_step, dim3(1), dim3(64), 0, s, a.counts, a.rays, 0);
    return hipGetLastError();
}

hipError_t render_path(const RenderArgs& a, hipStream_t s) {
    const int64_t nsamp = a.nq;
    const unsigned nb = (unsigned)((nsamp + kBlock - 1) / kBlock);
    float* hu = a.scene.needs_bary() ? a.hit_u : nullptr;
    float* hv = a.scene.needs_bary() ? a.hit_v : nullptr;
    hipError_t e = hipMemsetAsync(a.counts, 0, 64, s);
    if (e != hipSuccess) return e;
    render_raygen(a, s);
    int qin = 0, dim0 = 2;
    for (int depth = 0; depth < a.max_depth; ++depth) {
        const int qout = qin ^ 1;
        if ((e = trace_closest_queue(a.scene.rec, a.scene.ntris, a.scene.root, a.q[qin].ray,
                                     a.counts + qin, nsamp, a.hit_prim, a.hit_t, hu, hv,
                                     a.trace_counters, a.launch, s, &a.coop)) != hipSuccess)
            return e;
        if (a.sort_perm &&
            (e = sort_hits_by_material(a.scene, a.hit_prim, a.counts + qin, nsamp, a.sort_hist, a.sort_perm, s)) !=
                hipSuccess)
            return e;
        {
            // each sample's shading is its own: the workgroup size changes only the speed
            LaunchSink ls(s);
            const auto grid = [nsamp](int b) { return dim3((unsigned)((nsamp + b - 1) / b)); };
            const int sb = a.launch.shade_block;
            if (sb == 64)
                ls.launch("k_shade", &k_shade<64>, grid(64), dim3(64), depth, dim0, a.scene, a.halton, a.q[qin],
                          a.q[qout], a.sh, a.hit_prim, a.hit_t, hu, hv, a.res, a.counts, qin, qout, a.sort_perm);
            else if (sb == 128)
                ls.launch("k_shade", &k_shade<128>, grid(128), dim3(128), depth, dim0, a.scene, a.halton, a.q[qin],
                          a.q[qout], a.sh, a.hit_prim, a.hit_t, hu, hv, a.res, a.counts, qin, qout, a.sort_perm);
            else
                ls.launch("k_shade", &k_shade<256>, grid(256), dim3(256), depth, dim0, a.scene, a.halton, a.q[qin],
                          a.q[qout], a.sh, a.hit_prim, a.hit_t, hu, hv, a.res, a.counts, qin, qout, a.sort_perm);
            if (ls.err != hipSuccess) return ls.err;
        }
        if ((e = trace_occluded_queue(a.scene.rec, a.scene.ntris, a.sh.ray, a.counts + 2,
                                      nsamp, a.sh.limit, a.sh.src, a.sh.dst, a.vis, a.trace_counters, a.launch,
                                      s, &a.coop)) != hipSuccess)
            return e;
        hipLaunchKernelGGL(k_nee, dim3(nb), dim3(kBlock), 0, s, a.sh, a.vis, a.res, a.counts);
        hipLaunchKernelGGL(k_step, dim3(1), dim3(64), 0, s, a.counts, a.rays, qin);
        dim0 += 7 + a.scene.light_dims + (depth > 3 ? 1 : 0);
        qin = qout;
    }
    return hipGetLastError();
}

}  // namespace cav
