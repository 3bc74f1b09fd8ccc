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
// bounce (PathTracingIntegrator.cu:124-249).
__global__ __launch_bounds__(kBlock) void k_shade(int depth, int dim0, SceneDev sc, HaltonDev h,
                                                  Queue qi, Queue qo, Shadow sh,
                                                  const int32_t* __restrict__ hit_prim,
                                                  const float* __restrict__ hit_t,
                                                  float* __restrict__ res, uint32_t* counts,
                                                  int qin, int qout, const uint32_t* __restrict__ perm) {
    const int64_t j = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    const uint32_t n = counts[qin];
    bool live = j < (int64_t)n;
    const int64_t i = live ? task_slot(perm, j) : 0;
    int prim = live ? hit_prim[i] : -1;
    live = live && prim >= 0;

    bool push_sh = false, push_next = false;
    uint32_t sid = 0;
    v3 sro = mk3(0, 0, 0), srd = sro, contrib = sro, nro = sro, nrd = sro, nmult = sro;
    float slim = 0.f, npdf = 0.f;
    int ssrc = 0, sdst = 0;
    bool ndelta = false;

    if (live) {
        sid = qi.sid[i];
        const float* r = qi.ray + i * 6;
        const v3 ro = mk3(r[0], r[1], r[2]), rd = mk3(r[3], r[4], r[5]);
        v3 mult = mk3(qi.mult[i * 3], qi.mult[i * 3 + 1], qi.mult[i * 3 + 2]);
        const float mat_pdf = qi.pdf[i];
        const bool from_delta = qi.delta[i] != 0;
        const float t = hit_t[i];
        const int mesh = sc.tri_mesh[prim];
        const cav_mesh M = sc.meshes[mesh];
        const v3 nn = tri_normal(sc, prim);
        const v3 pos = add3(ro, scale3(rd, t));
        const Bsdf bsdf = make_bsdf(M);
        v3 t0, t1;
        frame(nn, t0, t1);
        v3 acc = mk3(res[sid * 3], res[sid * 3 + 1], res[sid * 3 + 2]);
        if (M.is_light) {
            const v3 Le = dot3(rd, nn) >= 0 ? mk3(0, 0, 0) : mk3(M.emission[0], M.emission[1], M.emission[2]);
            if (from_delta) {
                acc = add3(acc, mul3(Le, mult));
            } else if (depth > 0) {
                const v3 l2s = sub3(ro, pos);
                const float c = fabsf(dot3(nn, nrm(l2s)));
                const float lpdf = dot3(l2s, l2s) / (c * sc.mesh_area[mesh]);
                const float w = power_heuristic(mat_pdf, lpdf);
                if (isfinite(w)) acc = add3(acc, scale3(mul3(Le, mult), w));
            }
            res[sid * 3] = acc.x; res[sid * 3 + 1] = acc.y; res[sid * 3 + 2] = acc.z;
        }
        const v3 wo = scale3(rd, -1);
        // light sample: DiffuseAreaLight::sampleRayToPoint / TriangleMesh::sample
        const int li = halton_int(h, sid, dim0, sc.nlights);
        const int lm = sc.lights[li];
        const cav_mesh L = sc.meshes[lm];
        const int tri = L.first_tri + halton_int(h, sid, dim0 + 5, L.num_tris);
        const float tmp = sqrtf(halton(h, sid, dim0 + 6));
        const float u = 1 - tmp;
        const float v = tmp * halton(h, sid, dim0 + 7);
        const v3 p0 = soup_v(sc.soup, tri, 0), p1 = soup_v(sc.soup, tri, 1), p2 = soup_v(sc.soup, tri, 2);
        const v3 lp = add3(add3(scale3(p0, u), scale3(p1, v)), scale3(p2, 1.f - u - v));
        const v3 ln = tri_normal(sc, tri);
        const v3 l2s = sub3(pos, lp);
        const float c = fabsf(dot3(ln, nrm(l2s)));
        const float prob = dot3(l2s, l2s) / (c * sc.mesh_area[lm]);
        const v3 dir = nrm(sub3(lp, pos));
        const float lim = len3(sub3(lp, pos));
        const v3 incident = dot3(dir, ln) >= 0 ? mk3(0, 0, 0) : mk3(L.emission[0], L.emission[1], L.emission[2]);
        if (isfinite(prob)) {
            const v3 em = divs3(scale3(mult, (float)sc.nlights), prob);
            const v3 il = to_local(dir, t0, t1, nn), ol = to_local(wo, t0, t1, nn);
            const float mp = bsdf.pdf(il, ol);
            const float w = power_heuristic(prob, mp);
            const float cosv = fabsf(dot3(dir, nn));
            const v3 e = scale3(mul3(incident, bsdf.eval(il, ol)), cosv);
            contrib = scale3(mul3(e, em), w);
            sro = pos; srd = dir; slim = lim; ssrc = mesh; sdst = lm;
            push_sh = true;
        }
        // genNextRay
        int dim = dim0 + 8;
        bool alive = true;
        if (depth > 3) {
            float q = 1;
            q = fmn(q, 1 - mult.x);
            q = fmn(q, 1 - mult.y);
            q = fmn(q, 1 - mult.z);
            q = fmx(q, 0.05f);
            if (halton(h, sid, dim) < q) alive = false;
            else mult = divs3(mult, 1.f - q);
            ++dim;
        }
        if (alive) {
            const float ux = halton(h, sid, dim), uy = halton(h, sid, dim + 1);
            v3 wl;
            float p = 0.f;
            const v3 f = bsdf.sample(ux, uy, wl, to_local(wo, t0, t1, nn), p);
            nrd = to_world(wl, t0, t1, nn);
            nro = add3(pos, scale3(nrd, 0.001f));
            if (zero3(f)) nmult = mk3(0, 0, 0);
            else nmult = divs3(scale3(mul3(mult, f), fabsf(dot3(nrd, nn))), p);
            npdf = p;
            ndelta = bsdf.delta();
            push_next = true;
        }
    }
    const uint32_t ss = block_push<kBlock>(push_sh, &counts[2]);
    if (push_sh) {
        float* r = sh.ray + (int64_t)ss * 6;
        r[0] = sro.x; r[1] = sro.y; r[2] = sro.z; r[3] = srd.x; r[4] = srd.y; r[5] = srd.z;
        sh.limit[ss] = slim; sh.src[ss] = ssrc; sh.dst[ss] = sdst; sh.sid[ss] = sid;
        sh.contrib[(int64_t)ss * 3] = contrib.x;
        sh.contrib[(int64_t)ss * 3 + 1] = contrib.y;
        sh.contrib[(int64_t)ss * 3 + 2] = contrib.z;
    }
    const uint32_t ns = block_push<kBlock>(push_next, &counts[qout]);
    if (push_next) {
        float* r = qo.ray + (int64_t)ns * 6;
        r[0] = nro.x; r[1] = nro.y; r[2] = nro.z; r[3] = nrd.x; r[4] = nrd.y; r[5] = nrd.z;
        qo.sid[ns] = sid;
        qo.mult[(int64_t)ns * 3] = nmult.x;
        qo.mult[(int64_t)ns * 3 + 1] = nmult.y;
        qo.mult[(int64_t)ns * 3 + 2] = nmult.z;
        qo.pdf[ns] = npdf;
        qo.delta[ns] = ndelta ? 1 : 0;
    }
}

// DirectLighting::renderRay (Integrators/DirectLightingIntegrator.cu:22-63):
// emission of the hit light without MIS, one light sample, and a shadow task
// for every hit carrying Material::eval * lightsCount / pdf (no MIS weight).
__global__ __launch_bounds__(kBlock) void k_direct(SceneDev sc, HaltonDev h, Queue qi, Shadow sh,
                                                   const int32_t* __restrict__ hit_prim,
                                                   const float* __restrict__ hit_t, float* __restrict__ res,
                                                   uint32_t* counts, const uint32_t* __restrict__ perm) {
    const int64_t j = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    bool live = j < (int64_t)counts[0];
    const int64_t i = live ? task_slot(perm, j) : 0;
    const int prim = live ? hit_prim[i] : -1;
    live = live && prim >= 0;
    uint32_t sid = 0;
    v3 pos = mk3(0, 0, 0), dir = pos, contrib = pos;
    float lim = 0.f;
    int mesh = 0, lm = 0;
    if (live) {
        sid = qi.sid[i];
        const float* r = qi.ray + i * 6;
        const v3 ro = mk3(r[0], r[1], r[2]), rd = mk3(r[3], r[4], r[5]);
        mesh = sc.tri_mesh[prim];
        const cav_mesh M = sc.meshes[mesh];
        const v3 nn = tri_normal(sc, prim);
        pos = add3(ro, scale3(rd, hit_t[i]));
        if (M.is_light) {
            const v3 Le = dot3(rd, nn) >= 0 ? mk3(0, 0, 0) : mk3(M.emission[0], M.emission[1], M.emission[2]);
            res[sid * 3] = res[sid * 3] + Le.x;
            res[sid * 3 + 1] = res[sid * 3 + 1] + Le.y;
            res[sid * 3 + 2] = res[sid * 3 + 2] + Le.z;
        }
        const int dim0 = 2;
        const int li = halton_int(h, sid, dim0, sc.nlights);
        lm = sc.lights[li];
        const cav_mesh L = sc.meshes[lm];
        const int tri = L.first_tri + halton_int(h, sid, dim0 + 5, L.num_tris);
        const float tmp = sqrtf(halton(h, sid, dim0 + 6));
        const float u = 1 - tmp;
        const float v = tmp * halton(h, sid, dim0 + 7);
        const v3 p0 = soup_v(sc.soup, tri, 0), p1 = soup_v(sc.soup, tri, 1), p2 = soup_v(sc.soup, tri, 2);
        const v3 lp = add3(add3(scale3(p0, u), scale3(p1, v)), scale3(p2, 1.f - u - v));
        const v3 ln = tri_normal(sc, tri);
        const v3 l2s = sub3(pos, lp);
        const float c = fabsf(dot3(ln, nrm(l2s)));
        const float prob = dot3(l2s, l2s) / (c * sc.mesh_area[lm]);
        dir = nrm(sub3(lp, pos));
        lim = len3(sub3(lp, pos));
        const v3 incident = dot3(dir, ln) >= 0 ? mk3(0, 0, 0) : mk3(L.emission[0], L.emission[1], L.emission[2]);
        const Bsdf bsdf = make_bsdf(M);
        v3 t0, t1;
        frame(nn, t0, t1);
        const v3 il = to_local(dir, t0, t1, nn), ol = to_local(scale3(rd, -1), t0, t1, nn);
        const float mult = (float)sc.nlights / prob;
        contrib = scale3(scale3(mul3(incident, bsdf.eval(il, ol)), fabsf(dot3(dir, nn))), mult);
    }
    const uint32_t ss = block_push<kBlock>(live, &counts[2]);
    if (live) {
        float* r = sh.ray + (int64_t)ss * 6;
        r[0] = pos.x; r[1] = pos.y; r[2] = pos.z; r[3] = dir.x; r[4] = dir.y; r[5] = dir.z;
        sh.limit[ss] = lim; sh.src[ss] = mesh; sh.dst[ss] = lm; sh.sid[ss] = sid;
        sh.contrib[(int64_t)ss * 3] = contrib.x;
        sh.contrib[(int64_t)ss * 3 + 1] = contrib.y;
        sh.contrib[(int64_t)ss * 3 + 2] = contrib.z;
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
__global__ __launch_bounds__(kBlock) void k_film_partial(int width, int height, int spp,
                                                         const float* __restrict__ res,
                                                         const float* __restrict__ sxy,
                                                         float4* __restrict__ partial) {
    const int64_t pix = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    const int64_t npix = (int64_t)width * height;
    if (pix >= npix) return;
    const int sx0 = (int)(pix % width), sy0 = (int)(pix / width);
    float4 acc[9];
#pragma unroll
    for (int d = 0; d < 9; ++d) acc[d] = make_float4(0.f, 0.f, 0.f, 0.f);
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
            acc[d].x += c0; acc[d].y += c1; acc[d].z += c2; acc[d].w += 1.f;
        }
    }
#pragma unroll
    for (int d = 0; d < 9; ++d) partial[d * npix + pix] = acc[d];
}

__global__ __launch_bounds__(kBlock) void k_film_final(int width, int height,
                                                       const float4* __restrict__ partial,
                                                       float* __restrict__ rgb, float* __restrict__ wt) {
    const int64_t pix = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    const int64_t npix = (int64_t)width * height;
    if (pix >= npix) return;
    const int x = (int)(pix % width), y = (int)(pix / width);
    float r = 0, g = 0, b = 0, w = 0;
    for (int py = y - 1; py <= y + 1; ++py)
        for (int px = x - 1; px <= x + 1; ++px) {
            if (px < 0 || py < 0 || px >= width || py >= height) continue;
            const int d = (y - py + 1) * 3 + (x - px + 1);  // offset from the source to this pixel
            const float4 v = partial[d * npix + (int64_t)py * width + px];
            r += v.x; g += v.y; b += v.z; w += v.w;
        }
    rgb[pix * 3] = r;
    rgb[pix * 3 + 1] = g;
    rgb[pix * 3 + 2] = b;
    wt[pix] = w;
}

}  // namespace

void render_raygen(const RenderArgs& a, hipStream_t s) {
    const unsigned nb = (unsigned)((a.nq + kBlock - 1) / kBlock);
    hipLaunchKernelGGL(k_raygen, dim3(nb), dim3(kBlock), 0, s, a.nq, a.width, a.spp, a.spp_round,
                       a.spp_done, a.cam, a.halton, a.q[0], a.res, a.sxy, a.counts);
}

void render_film(const RenderArgs& a, hipStream_t s) {
    const int64_t npix = (int64_t)a.width * a.height;
    const dim3 grid((unsigned)((npix + kBlock - 1) / kBlock));
    hipLaunchKernelGGL(k_film_partial, grid, dim3(kBlock), 0, s, a.width, a.height, a.spp, a.res, a.sxy,
                       a.film_partial);
    hipLaunchKernelGGL(k_film_final, grid, dim3(kBlock), 0, s, a.width, a.height, a.film_partial, a.film_rgb,
                       a.film_w);
}

hipError_t render_direct(const RenderArgs& a, hipStream_t s) {
    const int64_t nq = a.nq;
    const unsigned nb = (unsigned)((nq + kBlock - 1) / kBlock);
    hipError_t e = hipMemsetAsync(a.counts, 0, 64, s);
    if (e != hipSuccess) return e;
    render_raygen(a, s);
    if ((e = trace_closest_queue(a.scene.tn, a.scene.tt, a.scene.ntris, a.scene.root, a.q[0].ray, a.counts, nq,
                                 a.hit_prim, a.hit_t, nullptr, nullptr, a.trace_counters, s)) != hipSuccess)
        return e;
    if (a.sort_perm && (e = sort_hits_by_material(a.scene, a.hit_prim, a.counts, nq, a.sort_hist, a.sort_perm, s)) !=
                           hipSuccess)
        return e;
    hipLaunchKernelGGL(k_direct, dim3(nb), dim3(kBlock), 0, s, a.scene, a.halton, a.q[0], a.sh, a.hit_prim, a.hit_t,
                       a.res, a.counts, a.sort_perm);
    if ((e = trace_occluded_queue(a.scene.tn, a.scene.tt, a.scene.ntris, a.sh.ray, a.counts + 2, nq, a.sh.limit,
                                  a.sh.src, a.sh.dst, a.vis, a.trace_counters, s)) != hipSuccess)
        return e;
    hipLaunchKernelGGL(k_nee, dim3(nb), dim3(kBlock), 0, s, a.sh, a.vis, a.res, a.counts);
    hipLaunchKernelGGL(k_step, dim3(1), dim3(64), 0, s, a.counts, a.rays, 0);
    return hipGetLastError();
}

hipError_t render_path(const RenderArgs& a, hipStream_t s) {
    const int64_t nsamp = a.nq;
    const unsigned nb = (unsigned)((nsamp + kBlock - 1) / kBlock);
    hipError_t e = hipMemsetAsync(a.counts, 0, 64, s);
    if (e != hipSuccess) return e;
    render_raygen(a, s);
    int qin = 0, dim0 = 2;
    for (int depth = 0; depth < a.max_depth; ++depth) {
        const int qout = qin ^ 1;
        if ((e = trace_closest_queue(a.scene.tn, a.scene.tt, a.scene.ntris, a.scene.root, a.q[qin].ray,
                                     a.counts + qin, nsamp, a.hit_prim, a.hit_t, nullptr, nullptr,
                                     a.trace_counters, s)) != hipSuccess)
            return e;
        if (a.sort_perm &&
            (e = sort_hits_by_material(a.scene, a.hit_prim, a.counts + qin, nsamp, a.sort_hist, a.sort_perm, s)) !=
                hipSuccess)
            return e;
        hipLaunchKernelGGL(k_shade, dim3(nb), dim3(kBlock), 0, s, depth, dim0, a.scene, a.halton,
                           a.q[qin], a.q[qout], a.sh, a.hit_prim, a.hit_t, a.res, a.counts, qin, qout, a.sort_perm);
        if ((e = trace_occluded_queue(a.scene.tn, a.scene.tt, a.scene.ntris, a.sh.ray, a.counts + 2,
                                      nsamp, a.sh.limit, a.sh.src, a.sh.dst, a.vis, a.trace_counters,
                                      s)) != hipSuccess)
            return e;
        hipLaunchKernelGGL(k_nee, dim3(nb), dim3(kBlock), 0, s, a.sh, a.vis, a.res, a.counts);
        hipLaunchKernelGGL(k_step, dim3(1), dim3(64), 0, s, a.counts, a.rays, qin);
        dim0 += 10 + (depth > 3 ? 1 : 0);
        qin = qout;
    }
    return hipGetLastError();
}

}  // namespace cav
