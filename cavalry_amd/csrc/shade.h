// Device code shared by the wavefront integrators: the Halton sampler, frames,
// the matte / plastic / mirror / glass BSDFs, triangle helpers and the
// wave-aggregated queue append.
//
// Reference: Samplers/HaltonSampler.h, Utils/MathsCommons.h,
// Utils/RandomUtils.h, BSDFs/*.h, Materials/*.h, Core/TriangleMesh.h.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/cavalry.h"
#include "dev_common.h"
#include "render.h"

namespace cav {
namespace shade {

constexpr int kBlock = 256;
constexpr double kPi = 3.14159265358979323846;

// ---------------------------------------------------------------- Halton
__device__ inline float halton(const HaltonDev& h, uint32_t index, int dim) {
    const uint32_t base = h.primes[dim];
    const uint16_t* pm = h.perm + h.start[dim];
    float r = 0.f, f = 1.f;
    const float fb = (float)base;
    uint32_t i = index;
    while (i > 0) {
        const uint32_t q = i / base;
        f = f / fb;
        r = r + f * (float)pm[i - q * base];
        i = q;
    }
    return r;
}
__device__ inline int halton_int(const HaltonDev& h, uint32_t index, int dim, int n) {
    const float f = halton(h, index, dim);
    int r = (int)(f * (float)n);
    if (r >= n) r = n - 1;
    return r;
}

// ---------------------------------------------------------------- vectors
__device__ inline v3 nrm(v3 v) {
    const float inv = 1.0f / sqrtf(dot3(v, v));
    return scale3(v, inv);
}
__device__ inline float len3(v3 v) { return sqrtf(dot3(v, v)); }
__device__ inline v3 divs3(v3 a, float s) { return mk3(a.x / s, a.y / s, a.z / s); }
__device__ inline bool zero3(v3 c) { return c.x == 0 && c.y == 0 && c.z == 0; }
__device__ inline float clampf(float x, float lo, float hi) { return fmx(lo, fmn(x, hi)); }

__device__ inline void frame(v3 n, v3& t0, v3& t1) {  // Utils/MathsCommons.h:193-203
    if (fabsf(n.x) < 0.5) t0 = cross3(mk3(1, 0, 0), n);
    else t0 = cross3(mk3(0, 1, 0), n);
    t0 = nrm(t0);
    t1 = nrm(cross3(n, t0));
}
__device__ inline v3 to_local(v3 v, v3 t0, v3 t1, v3 n) {
    return mk3(dot3(t0, v), dot3(t1, v), dot3(n, v));
}
__device__ inline v3 to_world(v3 v, v3 t0, v3 t1, v3 n) {
    return mk3(t0.x * v.x + t1.x * v.y + n.x * v.z, t0.y * v.x + t1.y * v.y + n.y * v.z,
               t0.z * v.x + t1.z * v.y + n.z * v.z);
}

// Utils/MathsCommons.h:109-162; the reference's `sinZenith == 0` guard never
// fires, so the clamped quotient is always used.
__device__ inline float c2z(v3 w) { return w.z * w.z; }
__device__ inline float s2z(v3 w) { return fmx(0.f, 1.f - c2z(w)); }
__device__ inline float sz(v3 w) { return sqrtf(s2z(w)); }
__device__ inline float tz(v3 w) { return sz(w) / w.z; }
__device__ inline float t2z(v3 w) { return s2z(w) / c2z(w); }
__device__ inline float caz(v3 w) { return clampf(w.x / sz(w), -1.f, 1.f); }
__device__ inline float saz(v3 w) { return clampf(w.y / sz(w), -1.f, 1.f); }
__device__ inline bool same_hemi(v3 a, v3 b) { return a.z * b.z > 0; }
__device__ inline float pow5(float f) { return (f * f) * (f * f) * f; }
__device__ inline v3 reflect3(v3 i, v3 n) { return sub3(scale3(scale3(n, 2.0f), dot3(n, i)), i); }
__device__ inline bool refract3(v3 i, v3 n, float eta, v3& out) {
    const float c = dot3(n, i);
    const float s2 = fmx(0.f, 1.f - c * c);
    const float s2t = eta * eta * s2;
    if (s2t >= 1) return false;
    const float ct = sqrtf(1 - s2t);
    out = add3(scale3(scale3(i, -1.f), eta), scale3(n, eta * c - ct));
    return true;
}
__device__ inline bool eq3(v3 a, v3 b) { return a.x == b.x && a.y == b.y && a.z == b.z; }

// Utils/RandomUtils.h:20-41: theta is a double (M_PI), so are cos/sin.
__device__ inline v3 cos_hemi(float ux, float uy) {
    const float ox = 2.f * ux - 1.f, oy = 2.f * uy - 1.f;
    float dx = 0.f, dy = 0.f;
    if (!(ox == 0 && oy == 0)) {
        double th;
        float r;
        if (fabsf(ox) > fabsf(oy)) { r = ox; th = (kPi / 4.f) * (double)(oy / ox); }
        else { r = oy; th = (kPi / 2.f) - (kPi / 4.f) * (double)(ox / oy); }
        dx = r * (float)cos(th);
        dy = r * (float)sin(th);
    }
    return mk3(dx, dy, sqrtf(fmx(0.f, 1 - dx * dx - dy * dy)));
}
__device__ inline float cos_pdf(float z) { return (float)fabs((double)z / kPi); }

// ---------------------------------------------------------------- GGX
struct Ggx {
    float ax, ay;
    __device__ float D(v3 h) const {
        const float t2 = t2z(h);
        if (isinf(t2)) return 0;
        const float c4 = c2z(h) * c2z(h);
        const float e = (caz(h) * caz(h) / (ax * ax) + saz(h) * saz(h) / (ay * ay)) * t2;
        return (float)(1 / (kPi * ax * ay * c4 * (1 + e) * (1 + e)));
    }
    __device__ float lambda(v3 w) const {
        if (dot3(w, mk3(0, 0, 1)) == 0) return 0;
        const float at = fabsf(tz(w));
        const float a = sqrtf(caz(w) * caz(w) * ax * ax + saz(w) * saz(w) * ay * ay);
        const float a2 = (a * at) * (a * at);
        return (-1 + sqrtf(1.f + a2)) / 2;
    }
    __device__ float G1(v3 w) const { return 1 / (1 + lambda(w)); }
    __device__ float G(v3 i, v3 o) const { return 1.f / (1.f + lambda(i) + lambda(o)); }
    __device__ float pdf(v3 h, v3 o) const { return D(h) * G1(o) * fabsf(dot3(o, h)) / fabsf(o.z); }
    __device__ static void slopes(float cosT, float u1, float u2, float& sx, float& sy) {
        if (cosT > .9999) {
            const float r = sqrtf(u1 / (1 - u1));
            const float phi = (float)(6.28318530718 * u2);
            sx = r * (float)cos((double)phi);  // binary64, rounded once: equal on every host and device
            sy = r * (float)sin((double)phi);
            return;
        }
        const float sinT = sqrtf(fmx(0.f, 1.f - cosT * cosT));
        const float tanT = sinT / cosT;
        const float a = 1 / tanT;
        const float g1 = 2 / (1 + sqrtf(1.f + 1.f / (a * a)));
        const float A = 2 * u1 / g1 - 1;
        float tmp = 1.f / (A * A - 1.f);
        if (tmp > 1e10) tmp = 1e10;
        const float B = tanT;
        const float Dd = sqrtf(fmx(B * B * tmp * tmp - (A * A - B * B) * tmp, 0.f));
        const float s1 = B * tmp - Dd, s2 = B * tmp + Dd;
        sx = (A < 0 || s2 > 1.f / tanT) ? s1 : s2;
        float S;
        if (u2 > 0.5f) { S = 1.f; u2 = 2.f * (u2 - .5f); }
        else { S = -1.f; u2 = 2.f * (.5f - u2); }
        const float z = (u2 * (u2 * (u2 * 0.27385f - 0.73369f) + 0.46341f)) /
                        (u2 * (u2 * (u2 * 0.093073f + 0.309420f) - 1.000000f) + 0.597999f);
        sy = S * z * sqrtf(1.f + sx * sx);
    }
    __device__ v3 sample(float u1, float u2, v3 o) const {
        const bool flip = o.z < 0;
        if (flip) o = scale3(o, -1.f);
        const v3 os = nrm(mk3(ax * o.x, ay * o.y, o.z));
        float sx, sy;
        slopes(os.z, u1, u2, sx, sy);
        const float t = caz(os) * sx - saz(os) * sy;
        sy = saz(os) * sx + caz(os) * sy;
        sx = t;
        sx = ax * sx;
        sy = ay * sy;
        v3 r = nrm(mk3(-sx, -sy, 1.f));
        if (flip) r = scale3(r, -1.f);
        return r;
    }
};

__device__ inline v3 schlick(float c, v3 f0) {
    const float t = pow5(1 - fabsf(c));
    return add3(f0, scale3(sub3(mk3(1, 1, 1), f0), t));
}

// ---------------------------------------------------------------- BSDF
struct Bsdf {
    int type;
    v3 kd, ks, f0, refl, trans;
    Ggx ggx;
    float kr, ior;

    __device__ bool delta() const { return type == CAV_MAT_MIRROR || type == CAV_MAT_GLASS; }
    // MicrofacetDistribution.h:157-159 through MicrofacetBSDF / FresnelBlendBSDF;
    // the other BSDFs keep the base class's false (Core/BSDF.h:25)
    __device__ bool almost_delta() const {
        return (type == CAV_MAT_METAL || type == CAV_MAT_SUBSTRATE) && ((double)ggx.ax < 0.05 || (double)ggx.ay < 0.05);
    }
    __device__ v3 lam(v3 i, v3 o) const {
        if (!same_hemi(i, o)) return mk3(0, 0, 0);
        return divs3(kd, (float)kPi);
    }
    __device__ v3 mf(v3 i, v3 o) const {
        if (!same_hemi(i, o)) return mk3(0, 0, 0);
        const float co = fabsf(o.z), ci = fabsf(i.z);
        v3 h = add3(i, o);
        if (ci == 0 || co == 0) return mk3(0, 0, 0);
        if (h.x == 0 && h.y == 0 && h.z == 0) return mk3(0, 0, 0);
        h = nrm(h);
        const v3 F = schlick(dot3(o, h), f0);
        return divs3(mul3(scale3(scale3(ks, ggx.D(h)), ggx.G(o, i)), F), 4 * ci * co);
    }
    __device__ float mf_pdf(v3 i, v3 o) const {
        const v3 h = nrm(add3(i, o));
        return ggx.pdf(h, o) / (4.f * dot3(o, h));
    }
    // FresnelBlendBSDF (BSDFs/Fresnel.h:58-150): Ashikhmin-Shirley diffuse
    // + GGX specular, sampled half / half
    __device__ v3 fb_eval(v3 i, v3 o) const {
        if (!same_hemi(i, o)) return mk3(0, 0, 0);
        const float c = (float)(28.0 / (23.0 * kPi));
        v3 d = mul3(scale3(kd, c), sub3(mk3(1, 1, 1), ks));
        d = scale3(d, 1 - pow5(1 - .5f * fabsf(i.z)));
        d = scale3(d, 1 - pow5(1 - .5f * fabsf(o.z)));
        v3 h = add3(i, o);
        if (h.x == 0 && h.y == 0 && h.z == 0) return mk3(0, 0, 0);
        h = nrm(h);
        const v3 s = divs3(scale3(schlick(dot3(i, h), ks), ggx.D(h)),
                           4 * fabsf(dot3(i, h)) * fmx(fabsf(i.z), fabsf(o.z)));
        return add3(d, s);
    }
    __device__ float fb_pdf(v3 i, v3 o, v3 h) const {
        return (1.f - 0.5f) * cos_pdf(i.z) + 0.5f * ggx.pdf(h, o) / (4.f * dot3(o, h));
    }
    __device__ v3 srefl(v3 i, v3 o) const {
        if (!same_hemi(i, o)) return mk3(0, 0, 0);
        return divs3(mul3(refl, schlick(fabsf(i.z), f0)), fabsf(i.z));
    }
    __device__ v3 strans(v3 i, v3 o) const {
        if (same_hemi(i, o)) return mk3(0, 0, 0);
        return divs3(mul3(trans, sub3(mk3(1, 1, 1), schlick(fabsf(i.z), f0))), fabsf(i.z));
    }
    __device__ bool srefract(v3 o, v3& out) const {
        const float eta = o.z > 0 ? (1.f / ior) : (ior / 1.f);
        v3 n = mk3(0, 0, 1);
        if (o.z < 0) n.z = -1;
        return refract3(o, n, eta, out);
    }
    __device__ float sprob(v3 o) const {
        float p = schlick(fabsf(o.z), f0).x;
        if (zero3(refl)) p = 0;
        else if (zero3(trans)) p = 1;
        return p;
    }
    __device__ v3 eval(v3 i, v3 o) const {
        if (type == CAV_MAT_MATTE) return lam(i, o);
        if (type == CAV_MAT_PLASTIC) return add3(lam(i, o), mf(i, o));
        if (type == CAV_MAT_METAL) return mf(i, o);
        if (type == CAV_MAT_SUBSTRATE) return fb_eval(i, o);
        if (type == CAV_MAT_MIRROR) {
            if (!(i.x == -o.x && i.y == -o.y && i.z == o.z)) return mk3(0, 0, 0);
            return scale3(mk3(1, 1, 1), kr / fabsf(i.z));
        }
        if (same_hemi(i, o)) {
            if (!(i.x == -o.x && i.y == -o.y && i.z == o.z)) return mk3(0, 0, 0);
            return srefl(i, o);
        }
        v3 r;
        if (!srefract(o, r) || !eq3(r, i)) return mk3(0, 0, 0);
        return strans(i, o);
    }
    __device__ float pdf(v3 i, v3 o) const {
        if (type == CAV_MAT_MATTE) return cos_pdf(i.z);
        if (type == CAV_MAT_PLASTIC) return 0.5f * cos_pdf(i.z) + (1.f - 0.5f) * mf_pdf(i, o);
        if (type == CAV_MAT_METAL) return mf_pdf(i, o);
        if (type == CAV_MAT_SUBSTRATE) return fb_pdf(i, o, nrm(add3(i, o)));
        if (type == CAV_MAT_MIRROR) return (i.x == -o.x && i.y == -o.y && i.z == o.z) ? 1.f : 0.f;
        if (same_hemi(i, o)) {
            if (!(i.x == -o.x && i.y == -o.y && i.z == o.z)) return 0;
            return sprob(o);
        }
        v3 r;
        if (!srefract(o, r) || !eq3(r, i)) return 0;
        return 1.f - sprob(o);
    }
    __device__ v3 sample(float ux, float uy, v3& wi, v3 o, float& p) const {
        if (type == CAV_MAT_MATTE) {
            wi = cos_hemi(ux, uy);
            p = cos_pdf(wi.z);
            if (o.z < 0) wi.z *= -1;
            return lam(wi, o);
        }
        if (type == CAV_MAT_PLASTIC) {
            float q = ux;
            if (q < 0.5f) {
                q = q * (1.f / 0.5f);
                wi = cos_hemi(q, uy);
                if (o.z < 0) wi.z *= -1;
            } else {
                q = (q - 0.5f) * (1.f / (1.f - 0.5f));
                wi = reflect3(o, ggx.sample(q, uy, o));
            }
            p = pdf(wi, o);
            return eval(wi, o);
        }
        if (type == CAV_MAT_METAL) {  // MicrofacetBSDF::sampleReflection (BSDFs/Microfacet.h:96-105)
            const v3 h = ggx.sample(ux, uy, o);
            wi = reflect3(o, h);
            p = ggx.pdf(h, o) / (4.f * dot3(o, h));
            return mf(wi, o);
        }
        if (type == CAV_MAT_SUBSTRATE) {  // FresnelBlendBSDF::sample (BSDFs/Fresnel.h:101-128)
            v3 h;
            if (ux < 0.5f) {
                h = ggx.sample(ux * (1.f / 0.5f), uy, o);
                wi = reflect3(o, h);
                if (!same_hemi(wi, o)) {  // the reference leaves the pdf unset; the value is zero
                    p = 0;
                    return mk3(0, 0, 0);
                }
            } else {
                wi = cos_hemi((ux - 0.5f) * (1.f / (1.f - 0.5f)), uy);
                if (o.z < 0) wi.z *= -1;
                h = nrm(add3(wi, o));
            }
            p = fb_pdf(wi, o, h);
            return fb_eval(wi, o);
        }
        if (type == CAV_MAT_MIRROR) {
            wi = scale3(o, -1);
            wi.z = o.z;
            p = 1;
            return eval(wi, o);
        }
        const float pr = sprob(o);
        if (ux < pr) {
            wi = mk3(-o.x, -o.y, o.z);
            p = 1;
            const v3 r = srefl(wi, o);
            p *= pr;
            return r;
        }
        v3 r = mk3(0, 0, 0);
        if (srefract(o, wi)) {
            p = 1;
            r = strans(wi, o);
        } else {
            // total internal reflection: the reference leaves the direction and
            // pdf unset (BSDFs/Specular.h:70-82); defined here as the mirror
            // direction with pdf 0 (the value is 0 either way)
            wi = mk3(-o.x, -o.y, o.z);
            p = 0;
        }
        p *= (1.f - pr);
        return r;
    }
};

__device__ inline Bsdf make_bsdf(const cav_mesh& m) {
    Bsdf b;
    b.type = m.material;
    b.kd = mk3(m.kd[0], m.kd[1], m.kd[2]);
    b.ks = mk3(m.ks[0], m.ks[1], m.ks[2]);
    b.refl = mk3(m.kr[0], m.kr[1], m.kr[2]);
    b.trans = mk3(m.kt[0], m.kt[1], m.kt[2]);
    b.ggx = Ggx{m.roughness, m.roughness_v != 0.f ? m.roughness_v : m.roughness};
    b.kr = m.kr[0];
    b.ior = m.ior;
    const float eta = m.material == CAV_MAT_GLASS ? m.ior : 1.5f;  // Materials/Plastic.h:111, Glass.h:238
    const float t = (eta - 1) / (eta + 1);
    b.f0 = mk3(t * t, t * t, t * t);
    if (m.material == CAV_MAT_METAL) {  // Materials/Metal.h:24-27: colour 1, the metal's F0
        b.f0 = b.ks;
        b.ks = mk3(1, 1, 1);
    }
    return b;
}

// Texture2D::readTexture (Core/Texture.h:92-95, texture set up at :101-125):
// normalized coordinates, wrap addressing on both axes, bilinear filtering of
// RGBA8 texels read as value / 255. The filter is evaluated here in float
// with exact fractional weights: texel coordinate x = u W - 0.5, the weights
// are the fractions of x and y, texels outside the image wrap around.
__device__ inline float wrap01(float u) {
    float w = u - floorf(u);
    if (!(w >= 0.f && w < 1.f)) w = 0.f;  // non-finite input, or a tiny negative u that rounds to 1
    return w;
}
__device__ inline v3 tex_fetch(const SceneDev& sc, int ti, float u, float v) {
    const TexDev T = sc.tex[ti];
    const float x = wrap01(u) * (float)T.w - 0.5f, y = wrap01(v) * (float)T.h - 0.5f;
    const float fx = floorf(x), fy = floorf(y);
    const float a = x - fx, b = y - fy;
    int i0 = (int)fx, j0 = (int)fy;
    int i1 = i0 + 1, j1 = j0 + 1;
    if (i0 < 0) i0 += T.w;
    if (j0 < 0) j0 += T.h;
    if (i1 >= T.w) i1 -= T.w;
    if (j1 >= T.h) j1 -= T.h;
    const uint32_t* px = sc.texels + T.off;
    const uint32_t c00 = px[(int64_t)j0 * T.w + i0], c10 = px[(int64_t)j0 * T.w + i1];
    const uint32_t c01 = px[(int64_t)j1 * T.w + i0], c11 = px[(int64_t)j1 * T.w + i1];
    float r[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        const float t00 = (float)((c00 >> (8 * c)) & 255u) / 255.f, t10 = (float)((c10 >> (8 * c)) & 255u) / 255.f;
        const float t01 = (float)((c01 >> (8 * c)) & 255u) / 255.f, t11 = (float)((c11 >> (8 * c)) & 255u) / 255.f;
        r[c] = (1.f - b) * ((1.f - a) * t00 + a * t10) + b * ((1.f - a) * t01 + a * t11);
    }
    return mk3(r[0], r[1], r[2]);
}

__device__ inline v3 soup_v(const float* soup, int t, int j) {
    const float* p = soup + (int64_t)t * 9 + j * 3;
    return mk3(p[0], p[1], p[2]);
}
__device__ inline v3 tri_normal(const SceneDev& sc, int t) {
    const v3 v0 = soup_v(sc.soup, t, 0);
    v3 n = nrm(cross3(sub3(soup_v(sc.soup, t, 1), v0), sub3(soup_v(sc.soup, t, 2), v0)));
    if (sc.meshes[sc.tri_mesh[t]].reverse_orientation) n = scale3(n, -1.f);
    return n;
}

// Shading normal at barycentrics (u, v) of triangle t
// (Triangle::fillIntersectionInformation, Core/Triangle.h:94-136, and
// TriangleMesh::getNormal, Core/TriangleMesh.h:76-104): with vertex normals
// normalize(n0 (1 - u - v) + u n1 + v n2), else the geometric normal; flipped
// for ReverseOrientation.
__device__ inline v3 shading_normal(const SceneDev& sc, int t, float u, float v) {
    const int mesh = sc.tri_mesh[t];
    if (!(sc.normals && sc.mesh_normals[mesh])) return tri_normal(sc, t);
    const float* q = sc.normals + (int64_t)t * 9;
    const v3 n0 = mk3(q[0], q[1], q[2]), n1 = mk3(q[3], q[4], q[5]), n2 = mk3(q[6], q[7], q[8]);
    v3 n = nrm(add3(add3(scale3(n0, 1.f - u - v), scale3(n1, u)), scale3(n2, v)));
    if (sc.meshes[mesh].reverse_orientation) n = scale3(n, -1.f);
    return n;
}
__device__ inline float power_heuristic(float a, float b) {
    a *= a;
    b *= b;
    return a / (a + b);
}
// glm mat4 * vec4 and the divide (Utils/MathsCommons.h:57-60)
__device__ inline v3 apply_m(const float* m, v3 v) {
    float r[4];
#pragma unroll
    for (int k = 0; k < 4; ++k)
        r[k] = (m[0 * 4 + k] * v.x + m[1 * 4 + k] * v.y) + (m[2 * 4 + k] * v.z + m[3 * 4 + k] * 1.f);
    return mk3(r[0] / r[3], r[1] / r[3], r[2] / r[3]);
}

// The BSDF at a hit (Material::getBSDF): the mesh's constants, with a Kd
// texture read at the interpolated texture coordinates
// tex0 (1 - u - v) + u tex1 + v tex2 (Core/Triangle.h:128-131) multiplying a
// matte's colour (Materials/MatteMaterial.h:23-33) or replacing a plastic's
// diffuse colour (Materials/Plastic.h:38-50).
__device__ inline Bsdf surface_bsdf(const SceneDev& sc, int mesh, int prim, float u, float v) {
    const cav_mesh& M = sc.meshes[mesh];
    Bsdf b = make_bsdf(M);
    const int ti = M.texture;
    if (sc.texels && ti >= 0 &&
        (M.material == CAV_MAT_MATTE || M.material == CAV_MAT_PLASTIC || M.material == CAV_MAT_SUBSTRATE)) {
        float tu = 0.f, tv = 0.f;
        if (sc.uv) {
            const float* q = sc.uv + (int64_t)prim * 6;
            const float w = 1.f - u - v;
            tu = (q[0] * w + u * q[2]) + v * q[4];
            tv = (q[1] * w + u * q[3]) + v * q[5];
        }
        const v3 texel = tex_fetch(sc, ti, tu, tv);
        b.kd = M.material == CAV_MAT_MATTE ? mul3(b.kd, texel) : texel;
    }
    return b;
}

// EnvironmentMap::evaluateRay (Core/Light.h:91-113): the colour, times the
// texel at (phi / 2pi, theta / pi) of the direction in the light's frame.
// atan2 and acos are evaluated in binary64 and rounded once (the reference
// calls the float routines); the M_PI divisions are binary64 as written.
__device__ inline v3 env_eval(const SceneDev& sc, int e, v3 d) {
    const EnvDev& E = sc.env[e];
    v3 r = mk3(E.color[0], E.color[1], E.color[2]);
    if (E.texture >= 0 && sc.texels) {
        const float* m = E.m;
        const v3 l = mk3((m[0] * d.x + m[1] * d.y) + m[2] * d.z, (m[3] * d.x + m[4] * d.y) + m[5] * d.z,
                         (m[6] * d.x + m[7] * d.y) + m[8] * d.z);
        float phi = (float)atan2((double)l.y, (double)l.x);
        phi = phi < 0 ? (float)((double)phi + 2 * kPi) : phi;
        const float theta = (float)acos((double)clampf(l.z, -1.f, 1.f));
        const float tu = (float)((double)phi / (2.0 * kPi)), tv = (float)((double)theta / kPi);
        r = mul3(r, tex_fetch(sc, E.texture, tu, tv));
    }
    return r;
}

// Radiance a ray that misses every triangle carries back: the background
// environment map, or nothing.
__device__ inline v3 miss_radiance(const SceneDev& sc, v3 d) {
    return sc.env_miss >= 0 ? env_eval(sc, sc.env_miss, d) : mk3(0, 0, 0);
}

// uniformSampleSphere (Utils/RandomUtils.h:70-78): angles through binary64
// M_PI as written, cos / sin in binary64 rounded once.
__device__ inline v3 uniform_sphere(float x, float y) {
    const float u = (float)((double)(x * 2.f) * kPi);
    const float v = (float)(((double)y - 0.5) * kPi);
    const float cu = (float)cos((double)u), su = (float)sin((double)u);
    const float cv = (float)cos((double)v), sv = (float)sin((double)v);
    return mk3(cv * cu, sv, cv * su);
}

// TriangleMesh::sampleSphere (Core/TriangleMesh.h:155-211): a direction
// uniform within the cone the sphere (centre c, radius rad) subtends from pos,
// the point where it meets the sphere, its outward normal and the cone's
// solid-angle pdf. Small cones (sin^2 < sin^2 1.5 deg) draw sin^2 theta
// directly, as the reference does.
__device__ inline void sample_sphere(v3 pos, v3 c, float rad, float ux, float uy, v3& p, v3& n, float& pdf) {
    const float dc = len3(sub3(pos, c));
    const float inv_dc = 1 / dc;
    const v3 wc = scale3(sub3(c, pos), inv_dc);
    v3 wx, wy;
    frame(wc, wx, wy);
    const float st_max = rad * inv_dc;
    const float st_max2 = st_max * st_max;
    const float inv_st_max = 1 / st_max;
    const float ct_max = sqrtf(fmx(0.f, 1 - st_max2));
    float ct = (ct_max - 1) * ux + 1;
    float st2 = 1 - ct * ct;
    if (st_max2 < 0.00068523f) {
        st2 = st_max2 * ux;
        ct = sqrtf(1 - st2);
    }
    const float ca = st2 * inv_st_max + ct * sqrtf(fmx(0.f, 1.f - st2 * inv_st_max * inv_st_max));
    const float sa = sqrtf(fmx(0.f, 1.f - ca * ca));
    const float phi = (float)((double)(uy * 2.f) * kPi);
    const float cp = (float)cos((double)phi), sp = (float)sin((double)phi);
    n = scale3(add3(add3(scale3(wx, sa * cp), scale3(wy, sa * sp)), scale3(wc, ct)), -1.f);
    p = add3(c, scale3(n, rad));
    pdf = (float)(1.0 / (2.0 * kPi * (double)(1 - ct_max)));
}

// TriangleMesh::pdf (Core/TriangleMesh.h:214-232) of a ray from o that hit
// light mesh m at pos with normal nn: the cone pdf of a sphere seen from
// outside, else the area pdf converted to solid angle.
__device__ inline float light_pdf(const SceneDev& sc, int m, v3 o, v3 pos, v3 nn) {
    const cav_mesh& L = sc.meshes[m];
    if (L.sphere[3] > 0) {
        const v3 oc = sub3(o, mk3(L.sphere[0], L.sphere[1], L.sphere[2]));
        const float d2 = dot3(oc, oc);
        if (d2 > L.sphere[3] * L.sphere[3]) {
            const float st2 = L.sphere[3] * L.sphere[3] / d2;
            const float ct = sqrtf(fmx(0.f, 1.f - st2));
            return (float)(1.0 / (2.0 * kPi * (double)(1 - ct)));
        }
    }
    const v3 l2s = sub3(o, pos);
    const float c = fabsf(dot3(nn, nrm(l2s)));
    return dot3(l2s, l2s) / (c * sc.mesh_area[m]);
}

// LightObject::sampleRayToPoint for light-list entry lm seen from pos, with
// its dimensions from d on (after randInt + rand4): an area light
// (DiffuseAreaLight::sampleRayToPoint, TriangleMesh::sample: a uniform
// triangle, then a uniform point; Core/TriangleMesh.h:108-153) or an
// environment-type light (EnvironmentMap::sampleRayToPoint: a uniform sphere
// direction, pdf 1 / 4pi, no distance limit; Core/Light.h:75-84).
struct LightPick {
    v3 dir, incident;
    float prob, lim;
    int dst;  // target mesh the shadow ray ignores (-1: none)
};
__device__ inline LightPick sample_light(const SceneDev& sc, const HaltonDev& h, uint32_t sid, int d, v3 pos,
                                         int lm) {
    LightPick r;
    if (lm < 0) {
        const int e = -1 - lm;
        r.dir = uniform_sphere(halton(h, sid, d), halton(h, sid, d + 1));
        r.prob = (float)(1.0 / (4.0 * kPi));
        r.incident = env_eval(sc, e, r.dir);
        r.lim = __builtin_inff();
        r.dst = -1;
        return r;
    }
    const cav_mesh L = sc.meshes[lm];
    r.dst = lm;
    if (L.sphere[3] > 0) {  // TriangleMesh::sample (Core/TriangleMesh.h:112-118)
        const v3 c = mk3(L.sphere[0], L.sphere[1], L.sphere[2]);
        const v3 pc = sub3(pos, c);
        if (dot3(pc, pc) > L.sphere[3] * L.sphere[3]) {
            v3 lp, ln;
            sample_sphere(pos, c, L.sphere[3], halton(h, sid, d), halton(h, sid, d + 1), lp, ln, r.prob);
            if (L.reverse_orientation) ln = scale3(ln, -1.f);
            r.dir = nrm(sub3(lp, pos));
            r.lim = len3(sub3(lp, pos));
            r.incident = dot3(r.dir, ln) >= 0 ? mk3(0, 0, 0) : mk3(L.emission[0], L.emission[1], L.emission[2]);
            return r;
        }
    }
    const int tri = L.first_tri + halton_int(h, sid, d, L.num_tris);
    const float tmp = sqrtf(halton(h, sid, d + 1));
    const float u = 1 - tmp;
    const float v = tmp * halton(h, sid, d + 2);
    const v3 p0 = soup_v(sc.soup, tri, 0), p1 = soup_v(sc.soup, tri, 1), p2 = soup_v(sc.soup, tri, 2);
    const v3 lp = add3(add3(scale3(p0, u), scale3(p1, v)), scale3(p2, 1.f - u - v));
    const v3 ln = shading_normal(sc, tri, u, v);
    const v3 l2s = sub3(pos, lp);
    const float c = fabsf(dot3(ln, nrm(l2s)));
    r.prob = dot3(l2s, l2s) / (c * sc.mesh_area[lm]);
    r.dir = nrm(sub3(lp, pos));
    r.lim = len3(sub3(lp, pos));
    r.incident = dot3(r.dir, ln) >= 0 ? mk3(0, 0, 0) : mk3(L.emission[0], L.emission[1], L.emission[2]);
    return r;
}

// What shading one task produces: an optional shadow task and an optional
// continuation ray.
struct ShadeOut {
    bool push_sh = false, push_next = false;
    v3 sro, srd, contrib, nro, nrd, nmult;
    float slim = 0.f, npdf = 0.f;
    int ssrc = 0, sdst = 0;
    bool ndelta = false;
};

// One `path` bounce of one task (PathTracingIntegrator.cu:124-249):
// emission with MIS (added to acc3, the sample's radiance), the light sample
// (a shadow task carrying the unoccluded contribution), Russian roulette and
// the BSDF sample of the next ray. dim0 = the task's first Halton dimension
// of this bounce.
__device__ inline void path_shade(int depth, int dim0, const SceneDev& sc, const HaltonDev& h, uint32_t sid, v3 ro,
                                  v3 rd, v3 mult, float mat_pdf, bool from_delta, int prim, float t, float hu,
                                  float hv, float* acc3, ShadeOut& o) {
    const int mesh = sc.tri_mesh[prim];
    const cav_mesh M = sc.meshes[mesh];
    const v3 nn = shading_normal(sc, prim, hu, hv);
    const v3 pos = add3(ro, scale3(rd, t));
    const Bsdf bsdf = surface_bsdf(sc, mesh, prim, hu, hv);
    v3 t0, t1;
    frame(nn, t0, t1);
    if (M.is_light) {
        v3 acc = mk3(acc3[0], acc3[1], acc3[2]);
        const v3 Le = dot3(rd, nn) >= 0 ? mk3(0, 0, 0) : mk3(M.emission[0], M.emission[1], M.emission[2]);
        if (from_delta) {
            acc = add3(acc, mul3(Le, mult));
        } else if (depth > 0) {
            const float lpdf = light_pdf(sc, mesh, ro, pos, nn);
            const float w = power_heuristic(mat_pdf, lpdf);
            if (isfinite(w)) acc = add3(acc, scale3(mul3(Le, mult), w));
        }
        acc3[0] = acc.x; acc3[1] = acc.y; acc3[2] = acc.z;
    }
    const v3 wo = scale3(rd, -1);
    // the light sample (randInt, then rand4 that nothing reads, then the light's own)
    const int li = halton_int(h, sid, dim0, sc.nlights);
    const LightPick L = sample_light(sc, h, sid, dim0 + 5, pos, sc.lights[li]);
    if (isfinite(L.prob)) {
        const v3 em = divs3(scale3(mult, (float)sc.nlights), L.prob);
        const v3 il = to_local(L.dir, t0, t1, nn), ol = to_local(wo, t0, t1, nn);
        const float mp = bsdf.pdf(il, ol);
        const float w = power_heuristic(L.prob, mp);
        const float cosv = fabsf(dot3(L.dir, nn));
        const v3 e = scale3(mul3(L.incident, bsdf.eval(il, ol)), cosv);
        o.contrib = scale3(mul3(e, em), w);
        o.sro = pos; o.srd = L.dir; o.slim = L.lim; o.ssrc = mesh; o.sdst = L.dst;
        o.push_sh = true;
    }
    // genNextRay
    int dim = dim0 + 5 + sc.light_dims;
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
        o.nrd = to_world(wl, t0, t1, nn);
        o.nro = add3(pos, scale3(o.nrd, 0.001f));
        if (zero3(f)) o.nmult = mk3(0, 0, 0);
        else o.nmult = divs3(scale3(mul3(mult, f), fabsf(dot3(o.nrd, nn))), p);
        o.npdf = p;
        o.ndelta = bsdf.delta();
        o.push_next = true;
    }
}

// DirectLighting::renderRay for one camera-ray hit: the light's emission
// without MIS (added to acc3) and one light sample weighted by
// lightsCount / pdf without MIS; the shadow task carries it.
__device__ inline void direct_shade(const SceneDev& sc, const HaltonDev& h, uint32_t sid, v3 ro, v3 rd, int prim,
                                    float t, float hu, float hv, float* acc3, ShadeOut& o) {
    const int mesh = sc.tri_mesh[prim];
    const cav_mesh M = sc.meshes[mesh];
    const v3 nn = shading_normal(sc, prim, hu, hv);
    const v3 pos = add3(ro, scale3(rd, t));
    if (M.is_light) {
        const v3 Le = dot3(rd, nn) >= 0 ? mk3(0, 0, 0) : mk3(M.emission[0], M.emission[1], M.emission[2]);
        acc3[0] = acc3[0] + Le.x;
        acc3[1] = acc3[1] + Le.y;
        acc3[2] = acc3[2] + Le.z;
    }
    const int dim0 = 2;
    const int li = halton_int(h, sid, dim0, sc.nlights);
    const LightPick L = sample_light(sc, h, sid, dim0 + 5, pos, sc.lights[li]);
    const Bsdf bsdf = surface_bsdf(sc, mesh, prim, hu, hv);
    v3 t0, t1;
    frame(nn, t0, t1);
    const v3 il = to_local(L.dir, t0, t1, nn), ol = to_local(scale3(rd, -1), t0, t1, nn);
    const float m = (float)sc.nlights / L.prob;
    o.contrib = scale3(scale3(mul3(L.incident, bsdf.eval(il, ol)), fabsf(dot3(L.dir, nn))), m);
    o.sro = pos; o.srd = L.dir; o.slim = L.lim; o.ssrc = mesh; o.sdst = L.dst;
    o.push_sh = true;
}

__device__ inline void write_shadow(const Shadow& sh, uint32_t ss, const ShadeOut& o) {
    float* r = sh.ray + (int64_t)ss * 6;
    r[0] = o.sro.x; r[1] = o.sro.y; r[2] = o.sro.z; r[3] = o.srd.x; r[4] = o.srd.y; r[5] = o.srd.z;
    sh.limit[ss] = o.slim; sh.src[ss] = o.ssrc; sh.dst[ss] = o.sdst;
    sh.contrib[(int64_t)ss * 3] = o.contrib.x;
    sh.contrib[(int64_t)ss * 3 + 1] = o.contrib.y;
    sh.contrib[(int64_t)ss * 3 + 2] = o.contrib.z;
}

__device__ inline void write_next(const Queue& qo, uint32_t ns, const ShadeOut& o) {
    float* r = qo.ray + (int64_t)ns * 6;
    r[0] = o.nro.x; r[1] = o.nro.y; r[2] = o.nro.z; r[3] = o.nrd.x; r[4] = o.nrd.y; r[5] = o.nrd.z;
    qo.mult[(int64_t)ns * 3] = o.nmult.x;
    qo.mult[(int64_t)ns * 3 + 1] = o.nmult.y;
    qo.mult[(int64_t)ns * 3 + 2] = o.nmult.z;
    qo.pdf[ns] = o.npdf;
    qo.delta[ns] = o.ndelta ? 1 : 0;
}

// Queue slot handled by thread j: through the material order when the queue
// was sorted, else j itself.
__device__ inline int64_t task_slot(const uint32_t* perm, int64_t j) { return perm ? (int64_t)perm[j] : j; }

// Appends the threads of the block that have `push` set to a queue and
// returns each pushing thread's slot. One atomic per block: with one atomic
// per wave, the ~65K waves of a 4M-task launch serialise on the counter's L2
// channel for ~0.35 ms. Every thread of the block must call it (it contains
// barriers); slots within the block follow thread order.
template <int kThreads>
__device__ inline uint32_t block_push(bool push, uint32_t* count) {
    constexpr int kWaves = kThreads / 64;
    __shared__ uint32_t wave_base[kWaves];
    __shared__ uint32_t block_base;
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const uint64_t m = __ballot(push);
    if (lane == 0) wave_base[wv] = (uint32_t)__popcll(m);
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t total = 0;
        for (int w = 0; w < kWaves; ++w) {
            const uint32_t c = wave_base[w];
            wave_base[w] = total;
            total += c;
        }
        block_base = total ? atomicAdd(count, total) : 0u;
    }
    __syncthreads();
    const uint32_t slot = block_base + wave_base[wv] + (uint32_t)__popcll(m & ((1ull << lane) - 1ull));
    __syncthreads();  // the next call may reuse wave_base / block_base
    return slot;
}


}  // namespace shade
}  // namespace cav
