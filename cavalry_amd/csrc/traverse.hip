// Closest-hit and any-hit BVH traversal.
//
// Reference: BVH::intersect and BVH::testVisibility (BVH/BVH.h:23-189), the
// slab test of Core/BoundingBox.h:55-74 and Moller-Trumbore of
// Core/Triangle.h:7-52. The visiting order is the reference's exactly (both
// children tested at their parent, the far child stacked, the near one taken
// next, left first on ties), so hit ids and distances match bit for bit.
//
// Layout read: TNode[N-1] (64 B, both child boxes + refs), TTri[N] (48 B,
// v0/e1/e2 + original index + mesh id), rays float[R][6].
#include <cfloat>

#include "dev_common.h"
#include "lbvh.h"

namespace cav {
namespace {

constexpr int kBlock = 256;
constexpr int kStack = 64;  // BVH/BVH.h:41,122

// Core/BoundingBox.h:55-74; returns the entry distance or -1.
__device__ __forceinline__ float slab(float lx, float ly, float lz, float hx, float hy, float hz,
                                      v3 o, v3 inv) {
    const float t0x = (lx - o.x) * inv.x, t0y = (ly - o.y) * inv.y, t0z = (lz - o.z) * inv.z;
    const float t1x = (hx - o.x) * inv.x, t1y = (hy - o.y) * inv.y, t1z = (hz - o.z) * inv.z;
    const float maxmin = fmin_fmax(t0x, t1x, fmin_fmax(t0y, t1y, fmn(t0z, t1z)));
    const float minmax = fmax_fmin(t0x, t1x, fmax_fmin(t0y, t1y, fmx(t0z, t1z)));
    if (maxmin <= minmax && minmax > 0) return fmx(0.f, maxmin);
    return -1.f;
}

// Core/Triangle.h:7-52
__device__ __forceinline__ bool tri_hit(const TTri& tr, v3 o, v3 d, float& t, float& u, float& v) {
    const v3 v0 = mk3(tr.a.x, tr.a.y, tr.a.z);
    const v3 e1 = mk3(tr.b.x, tr.b.y, tr.b.z);
    const v3 e2 = mk3(tr.c.x, tr.c.y, tr.c.z);
    const v3 tvec = sub3(o, v0);
    const v3 pvec = cross3(d, e2);
    float det = dot3(e1, pvec);
    if (det == 0) return false;
    det = 1.f / det;
    u = dot3(tvec, pvec) * det;
    if (u < 0.0f || u > 1.0f) return false;
    const v3 qvec = cross3(tvec, e1);
    v = dot3(d, qvec) * det;
    if (v < 0.0f || (u + v) > 1.0f) return false;
    t = dot3(e2, qvec) * det;
    return t > 0.f;
}

__device__ __forceinline__ void load_ray(const float* __restrict__ rays, int64_t i, v3& o, v3& d) {
    const float2* r = reinterpret_cast<const float2*>(rays + i * 6);
    const float2 a = r[0], b = r[1], c = r[2];
    o = mk3(a.x, a.y, b.x);
    d = mk3(b.y, c.x, c.y);
}

__global__ __launch_bounds__(kBlock) void k_closest(const TNode* __restrict__ tn,
                                                    const TTri* __restrict__ tt,
                                                    const float* __restrict__ root,
                                                    const float* __restrict__ rays, int64_t nrays,
                                                    int32_t* __restrict__ oprim,
                                                    float* __restrict__ ot, float* __restrict__ ou,
                                                    float* __restrict__ ov,
                                                    uint32_t* __restrict__ overflow,
                                                    int max_steps) {
    const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (i >= nrays) return;
    v3 o, d;
    load_ray(rays, i, o, d);
    const v3 inv = div3(mk3(1.f, 1.f, 1.f), d);
    bool hit = false;
    float best = FLT_MAX, bu = 0.f, bv = 0.f;
    int bp = -1;
    uint32_t ovf = 0;
    if (slab(root[0], root[1], root[2], root[3], root[4], root[5], o, inv) >= 0.f) {
        int stack[kStack];
        int sp = 0;
        int cur = 0;
        // every node is visited at most once in a valid tree
        for (int step = 0; step < max_steps; ++step) {
            if (cur >= 0) {
                const TNode nd = tn[cur];
                const float dl = slab(nd.l0.x, nd.l0.y, nd.l0.z, nd.l0.w, nd.l1.x, nd.l1.y, o, inv);
                const float dr = slab(nd.l1.z, nd.l1.w, nd.r1.x, nd.r1.y, nd.r1.z, nd.r1.w, o, inv);
                const bool hl = dl >= 0 && dl < best, hr = dr >= 0 && dr < best;
                if (hl && hr) {
                    const bool lfar = dl > dr;
                    const int far = lfar ? nd.refs.x : nd.refs.y;
                    const int near = lfar ? nd.refs.y : nd.refs.x;
                    if (sp < kStack) stack[sp++] = far; else ++ovf;
                    if (sp < kStack) { cur = near; continue; }
                    ++ovf;
                } else if (hl || hr) {
                    if (sp < kStack) { cur = hl ? nd.refs.x : nd.refs.y; continue; }
                    ++ovf;
                }
            } else {
                const TTri tr = tt[~cur];
                float t, u, v;
                if (tri_hit(tr, o, d, t, u, v) && (!hit || t < best)) {
                    hit = true; best = t; bu = u; bv = v; bp = __float_as_int(tr.a.w);
                }
            }
            if (sp == 0) break;
            cur = stack[--sp];
        }
    }
    oprim[i] = hit ? bp : -1;
    ot[i] = hit ? best : 0.f;
    if (ou) ou[i] = hit ? bu : 0.f;
    if (ov) ov[i] = hit ? bv : 0.f;
    if (ovf) atomicAdd(overflow, ovf);
}

__global__ __launch_bounds__(kBlock) void k_occluded(const TNode* __restrict__ tn,
                                                     const TTri* __restrict__ tt,
                                                     const float* __restrict__ rays,
                                                     const float* __restrict__ limit,
                                                     const int32_t* __restrict__ src,
                                                     const int32_t* __restrict__ dst,
                                                     int64_t nrays, uint8_t* __restrict__ visible,
                                                     uint32_t* __restrict__ overflow,
                                                     int max_steps) {
    const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (i >= nrays) return;
    v3 o, d;
    load_ray(rays, i, o, d);
    const v3 inv = div3(mk3(1.f, 1.f, 1.f), d);
    const float lim = limit[i];
    const bool use_lim = !(lim == INFINITY);
    const int sm = src ? src[i] : -1, dm = dst ? dst[i] : -1;
    uint32_t ovf = 0;
    bool vis = true;
    int stack[kStack];
    int sp = 0;
    int cur = 0;
    for (int step = 0; step < max_steps; ++step) {
        if (cur >= 0) {
            const TNode nd = tn[cur];
            const float dl = slab(nd.l0.x, nd.l0.y, nd.l0.z, nd.l0.w, nd.l1.x, nd.l1.y, o, inv);
            const float dr = slab(nd.l1.z, nd.l1.w, nd.r1.x, nd.r1.y, nd.r1.z, nd.r1.w, o, inv);
            const bool hl = dl >= 0, hr = dr >= 0;
            if (hl && hr) {
                const bool lfar = dl > dr;
                const int far = lfar ? nd.refs.x : nd.refs.y;
                const int near = lfar ? nd.refs.y : nd.refs.x;
                if (sp < kStack) stack[sp++] = far; else ++ovf;
                if (sp < kStack) { cur = near; continue; }
                ++ovf;
            } else if (hl || hr) {
                if (sp < kStack) { cur = hl ? nd.refs.x : nd.refs.y; continue; }
                ++ovf;
            }
        } else {
            const TTri tr = tt[~cur];
            const int mesh = __float_as_int(tr.b.w);
            if (mesh != sm && mesh != dm) {
                float t, u, v;
                if (tri_hit(tr, o, d, t, u, v) && (!use_lim || t < lim)) { vis = false; break; }
            }
        }
        if (sp == 0) break;
        cur = stack[--sp];
    }
    visible[i] = vis ? 1 : 0;
    if (ovf) atomicAdd(overflow, ovf);
}

}  // namespace

hipError_t trace_closest(const TNode* tn, const TTri* tt, int ntris, const float* root,
                         const float* rays, int64_t nrays, int32_t* prim, float* t, float* u,
                         float* v, uint32_t* overflow, hipStream_t s) {
    if (nrays <= 0) return hipSuccess;
    const int64_t nb = (nrays + kBlock - 1) / kBlock;
    hipLaunchKernelGGL(k_closest, dim3((unsigned)nb), dim3(kBlock), 0, s, tn, tt, root, rays,
                       nrays, prim, t, u, v, overflow, 2 * ntris - 1);
    return hipGetLastError();
}

hipError_t trace_occluded(const TNode* tn, const TTri* tt, int ntris, const float* rays,
                          const float* limit, const int32_t* src, const int32_t* dst,
                          int64_t nrays, uint8_t* visible, uint32_t* overflow, hipStream_t s) {
    if (nrays <= 0) return hipSuccess;
    const int64_t nb = (nrays + kBlock - 1) / kBlock;
    hipLaunchKernelGGL(k_occluded, dim3((unsigned)nb), dim3(kBlock), 0, s, tn, tt, rays, limit,
                       src, dst, nrays, visible, overflow, 2 * ntris - 1);
    return hipGetLastError();
}

}  // namespace cav
