// CPU oracle: closest-hit and any-hit BVH traversal.
//
// TEST INFRASTRUCTURE ONLY: a restatement of BVH/BVH.h:23-189,
// Core/BoundingBox.h:54-74 and Core/Triangle.h:7-92 used to check the HIP
// traversal kernels and as the timed CPU baseline. Not part of the product.
#include "ref_api.h"
#include "ref_math.h"

#include <cfloat>
#include <vector>

using namespace ref;

namespace {

struct View {
    const float* box; const int32_t* left; const int32_t* right; const int32_t* prim;
    const float* tris; const int32_t* mesh;
};

// Core/BoundingBox.h:55-74 (slab test on integer-ordered min/max)
static inline bool box_hit(const float* b, f3 o, f3 inv, float& mind) {
    f3 tmin = mul(mk(b[0], b[1], b[2]) - o, inv);
    f3 tmax = mul(mk(b[3], b[4], b[5]) - o, inv);
    float maxmin = fmin_fmax(tmin.x, tmax.x, fmin_fmax(tmin.y, tmax.y, fminr(tmin.z, tmax.z)));
    float minmax = fmax_fmin(tmin.x, tmax.x, fmax_fmin(tmin.y, tmax.y, fmaxr(tmin.z, tmax.z)));
    if (maxmin <= minmax) {
        if (minmax > 0) { mind = fmaxr(0.f, maxmin); return true; }
    }
    mind = -1;
    return false;
}

// Core/Triangle.h:7-52 (Moller-Trumbore, edges formed at call time)
static inline bool tri_hit(const float* t9, f3 o, f3 d, float& t, float& u, float& v) {
    f3 v0 = mk(t9[0], t9[1], t9[2]);
    f3 e1 = mk(t9[3], t9[4], t9[5]) - v0;
    f3 e2 = mk(t9[6], t9[7], t9[8]) - v0;
    f3 tvec = o - v0;
    f3 pvec = cross(d, e2);
    float det = dot(e1, pvec);
    if (det == 0) return false;
    det = 1.f / det;
    u = dot(tvec, pvec) * det;
    if (u < 0.0f || u > 1.0f) return false;
    f3 qvec = cross(tvec, e1);
    v = dot(d, qvec) * det;
    if (v < 0.0f || (u + v) > 1.0f) return false;
    t = dot(e2, qvec) * det;
    if (t <= 0.f) return false;
    return true;
}

}  // namespace

// BVH/BVH.h:23-107. hit_prim = original triangle index or -1.
extern "C" int ref_trace_closest(const float* box, const int32_t* left, const int32_t* right,
                                 const int32_t* prim, const float* tris, const float* rays,
                                 int64_t nrays, int32_t* hit_prim, float* hit_t, float* hit_u,
                                 float* hit_v, int32_t* visits, int32_t* max_stack) {
    int overflow = 0;
#pragma omp parallel for schedule(dynamic, 1024) reduction(+ : overflow)
    for (int64_t r = 0; r < nrays; ++r) {
        const float* ry = rays + r * 6;
        f3 o = mk(ry[0], ry[1], ry[2]), d = mk(ry[3], ry[4], ry[5]);
        f3 inv = divv(mk(1, 1, 1), d);
        bool hit = false;
        float best = FLT_MAX, bu = 0, bv = 0;
        int bp = -1, nvis = 0, depth = 0;
        float md;
        if (box_hit(box, o, inv, md)) {
            int stack[64];
            int top = 0;
            stack[0] = 0;
            while (top >= 0) {
                int cur = stack[top--];
                ++nvis;
                if (prim[cur] >= 0) {
                    float t, u, v;
                    int p = prim[cur];
                    if (tri_hit(tris + (int64_t)p * 9, o, d, t, u, v)) {
                        if (!hit || t < best) { hit = true; best = t; bp = p; bu = u; bv = v; }
                    }
                } else {
                    float dl, dr;
                    int l = left[cur], rr = right[cur];
                    box_hit(box + (int64_t)l * 6, o, inv, dl);
                    box_hit(box + (int64_t)rr * 6, o, inv, dr);
                    bool hl = dl >= 0 && dl < best, hr = dr >= 0 && dr < best;
                    int push[2], np = 0;
                    if (hl && hr) {
                        if (dl > dr) { push[0] = l; push[1] = rr; }
                        else { push[0] = rr; push[1] = l; }
                        np = 2;
                    } else if (hl) { push[0] = l; np = 1; }
                    else if (hr) { push[0] = rr; np = 1; }
                    for (int k = 0; k < np; ++k) {
                        if (top + 1 >= 64) { ++overflow; continue; }
                        stack[++top] = push[k];
                    }
                    if (top + 1 > depth) depth = top + 1;
                }
            }
        }
        hit_prim[r] = hit ? bp : -1;
        hit_t[r] = hit ? best : 0.f;
        if (hit_u) hit_u[r] = hit ? bu : 0.f;
        if (hit_v) hit_v[r] = hit ? bv : 0.f;
        if (visits) visits[r] = nvis;
        if (max_stack) max_stack[r] = depth;
    }
    return overflow;
}

// BVH/BVH.h:109-189. visible[r] = 1 when nothing blocks the segment.
// limit[r] = +inf means no distance limit. Triangles whose mesh id equals
// src_mesh[r] or dst_mesh[r] are skipped (mesh/src/dst may be null).
extern "C" int ref_trace_occluded(const float* box, const int32_t* left, const int32_t* right,
                                  const int32_t* prim, const float* tris, const int32_t* mesh,
                                  const float* rays, const float* limit, const int32_t* src_mesh,
                                  const int32_t* dst_mesh, int64_t nrays, uint8_t* visible) {
    int overflow = 0;
#pragma omp parallel for schedule(dynamic, 1024) reduction(+ : overflow)
    for (int64_t r = 0; r < nrays; ++r) {
        const float* ry = rays + r * 6;
        f3 o = mk(ry[0], ry[1], ry[2]), d = mk(ry[3], ry[4], ry[5]);
        f3 inv = divv(mk(1, 1, 1), d);
        float lim = limit[r];
        bool use_lim = !(lim == INFINITY);
        int sm = src_mesh ? src_mesh[r] : -1, dm = dst_mesh ? dst_mesh[r] : -1;
        // the root test only reports an error in the reference; traversal
        // starts from the root either way
        int stack[64];
        int top = 0;
        stack[0] = 0;
        bool vis = true;
        while (top >= 0 && vis) {
            int cur = stack[top--];
            if (prim[cur] >= 0) {
                int p = prim[cur];
                if (mesh && (mesh[p] == sm || mesh[p] == dm)) continue;
                float t, u, v;
                if (tri_hit(tris + (int64_t)p * 9, o, d, t, u, v)) {
                    if (!use_lim || t < lim) vis = false;
                }
            } else {
                float dl, dr;
                int l = left[cur], rr = right[cur];
                box_hit(box + (int64_t)l * 6, o, inv, dl);
                box_hit(box + (int64_t)rr * 6, o, inv, dr);
                int push[2], np = 0;
                if (dl >= 0 && dr >= 0) {
                    if (dl > dr) { push[0] = l; push[1] = rr; }
                    else { push[0] = rr; push[1] = l; }
                    np = 2;
                } else if (dl >= 0) { push[0] = l; np = 1; }
                else if (dr >= 0) { push[0] = rr; np = 1; }
                for (int k = 0; k < np; ++k) {
                    if (top + 1 >= 64) { ++overflow; continue; }
                    stack[++top] = push[k];
                }
            }
        }
        visible[r] = vis ? 1 : 0;
    }
    return overflow;
}
