// CPU oracle: the `path` integrator, one sample at a time.
//
// TEST INFRASTRUCTURE ONLY: a restatement of Integrators/PathTracingIntegrator.cu
// and the device code it calls (Halton sampler, perspective camera, matte /
// plastic / mirror / glass BSDFs, diffuse area lights, box-filter film), used
// to check the HIP wavefront renderer and as the timed CPU baseline.
//
// In the reference every sample of a queue consumes the same number of
// Halton dimensions per bounce (8 for light selection and sampling, 2 for the
// BSDF, 1 more for Russian roulette after depth 3), so syncDimension never
// changes a dimension and each sample can be traced on its own.
//
// Float semantics: IEEE binary32 operations in the reference's order, no
// contraction; expressions the reference evaluates in double (through M_PI
// and double literals) are evaluated in double here too. normalize(v) is
// v * (1/sqrt(dot(v,v))) with a correctly rounded sqrt and division.
#include "ref_api.h"
#include "ref_math.h"

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstring>
#include <vector>

using namespace ref;

namespace {

const double kPi = 3.14159265358979323846;

// ---------------------------------------------------------------- Halton
// Samplers/HaltonSampler.cu:7-57: first 1024 primes, per-prime digit
// permutations from a Fisher-Yates shuffle driven by std::rand() after
// std::srand(533799). The table is frozen to the Microsoft C runtime's rand()
// (state = state * 214013 + 2531011, output bits 16..30), the runtime the
// reference was built with.
struct Halton {
    std::vector<uint32_t> primes;
    std::vector<uint32_t> start;
    std::vector<uint16_t> perm;
    Halton() {
        primes.push_back(2);
        for (uint32_t i = 3; primes.size() < 1024; ++i) {
            bool prime = true;
            for (uint32_t p : primes) {
                if (p * p > i) break;
                if (i % p == 0) { prime = false; break; }
            }
            if (prime) primes.push_back(i);
        }
        uint32_t total = 0;
        for (uint32_t p : primes) { start.push_back(total); total += p; }
        perm.resize(total);
        uint32_t state = 533799u;
        auto rnd = [&]() { state = state * 214013u + 2531011u; return (int)((state >> 16) & 0x7fffu); };
        for (size_t k = 0; k < primes.size(); ++k) {
            const int p = (int)primes[k];
            std::vector<uint16_t> v(p);
            for (int j = 0; j < p; ++j) v[j] = (uint16_t)j;
            for (int i = p - 1; i > 0; --i) std::swap(v[i], v[rnd() % (i + 1)]);
            for (int j = 0; j < p; ++j) perm[start[k] + j] = v[j];
        }
    }
    // HaltonSampler::rand1 / runHalton (HaltonSampler.h:13-23, 58-65)
    float rand1(uint64_t index, int& dim) const {
        const uint32_t base = primes[dim];
        const uint16_t* pm = perm.data() + start[dim];
        ++dim;
        float r = 0, f = 1;
        uint64_t i = index;
        while (i > 0) {
            f = f / (float)base;
            r = r + f * (float)pm[i % base];
            i = i / base;
        }
        return r;
    }
    int rand_int(int n, uint64_t index, int& dim) const {  // HaltonSampler.h:50-56
        float f = rand1(index, dim);
        int r = (int)(f * (float)n);
        if (r >= n) r = n - 1;
        return r;
    }
};

const Halton& halton() {
    static Halton h;
    return h;
}

// ---------------------------------------------------------------- vectors
f3 normalize(f3 v) {
    float inv = 1.0f / std::sqrt(dot(v, v));
    return v * inv;
}
float length(f3 v) { return std::sqrt(dot(v, v)); }
f3 neg(f3 v) { return mk(-v.x, -v.y, -v.z); }
bool all_zero(f3 c) { return c.x == 0 && c.y == 0 && c.z == 0; }
f3 mulv(f3 a, f3 b) { return mk(a.x * b.x, a.y * b.y, a.z * b.z); }
f3 divs(f3 a, float s) { return mk(a.x / s, a.y / s, a.z / s); }
float clampf(float x, float lo, float hi) { return fmaxr(lo, fminr(x, hi)); }

// Utils/MathsCommons.h:193-203
void coordinate_system(f3 n, f3& t0, f3& t1) {
    if (std::fabs(n.x) < 0.5) t0 = cross(mk(1, 0, 0), n);
    else t0 = cross(mk(0, 1, 0), n);
    t0 = normalize(t0);
    t1 = normalize(cross(n, t0));
}
// Core/IntersectionResult.h:28-52 (glm mat3 products, rows / columns)
f3 to_local(f3 v, f3 t0, f3 t1, f3 n) { return mk(dot(t0, v), dot(t1, v), dot(n, v)); }
f3 to_world(f3 v, f3 t0, f3 t1, f3 n) {
    return mk(t0.x * v.x + t1.x * v.y + n.x * v.z, t0.y * v.x + t1.y * v.y + n.y * v.z,
              t0.z * v.x + t1.z * v.y + n.z * v.z);
}

// Utils/MathsCommons.h:109-162 (the sinZenith == 0 test compares a function
// pointer and is always false, so the clamped quotient is always taken)
float cos2z(f3 w) { return w.z * w.z; }
float sin2z(f3 w) { return fmaxr(0.f, 1.f - cos2z(w)); }
float sinz(f3 w) { return std::sqrt(sin2z(w)); }
float tanz(f3 w) { return sinz(w) / w.z; }
float tan2z(f3 w) { return sin2z(w) / cos2z(w); }
float cos_az(f3 w) { return clampf(w.x / sinz(w), -1.f, 1.f); }
float sin_az(f3 w) { return clampf(w.y / sinz(w), -1.f, 1.f); }
bool same_hemi(f3 a, f3 b) { return a.z * b.z > 0; }
float pow5(float f) { return (f * f) * (f * f) * f; }
f3 reflect(f3 i, f3 n) { return (2.0f * n) * dot(n, i) - i; }
bool refract(f3 i, f3 n, float eta, f3& out) {
    float c = dot(n, i);
    float s2 = fmaxr(0.f, 1.f - c * c);
    float s2t = eta * eta * s2;
    if (s2t >= 1) return false;
    float ct = std::sqrt(1 - s2t);
    out = eta * (i * -1.f) + (eta * c - ct) * n;
    return true;
}
bool veq(f3 a, f3 b) { return a.x == b.x && a.y == b.y && a.z == b.z; }

// Utils/RandomUtils.h:20-41 (theta is double through M_PI)
f3 cosine_hemisphere(float ux, float uy) {
    float ox = 2.f * ux - 1.f, oy = 2.f * uy - 1.f;
    float dx = 0, dy = 0;
    if (!(ox == 0 && oy == 0)) {
        double theta;
        float r;
        if (std::fabs(ox) > std::fabs(oy)) { r = ox; theta = (kPi / 4.f) * (double)(oy / ox); }
        else { r = oy; theta = (kPi / 2.f) - (kPi / 4.f) * (double)(ox / oy); }
        dx = r * (float)std::cos(theta);
        dy = r * (float)std::sin(theta);
    }
    float z = std::sqrt(fmaxr(0.f, 1 - dx * dx - dy * dy));
    return mk(dx, dy, z);
}
float cos_pdf(float z) { return (float)std::fabs((double)z / kPi); }

// ---------------------------------------------------------------- BSDFs
struct GGX {
    float ax, ay;
    float D(f3 h) const {  // MicrofacetDistribution.h:8-22
        float t2 = tan2z(h);
        if (std::isinf(t2)) return 0;
        const float c4 = cos2z(h) * cos2z(h);
        float e = (cos_az(h) * cos_az(h) / (ax * ax) + sin_az(h) * sin_az(h) / (ay * ay)) * t2;
        return (float)(1 / (kPi * ax * ay * c4 * (1 + e) * (1 + e)));
    }
    float lambda(f3 w) const {
        if (dot(w, mk(0, 0, 1)) == 0) return 0;
        float at = std::fabs(tanz(w));
        float a = std::sqrt(cos_az(w) * cos_az(w) * ax * ax + sin_az(w) * sin_az(w) * ay * ay);
        float a2 = (a * at) * (a * at);
        return (-1 + std::sqrt(1.f + a2)) / 2;
    }
    float G1(f3 w) const { return 1 / (1 + lambda(w)); }
    float G(f3 i, f3 o) const { return 1.f / (1.f + lambda(i) + lambda(o)); }
    float pdf(f3 h, f3 o) const { return D(h) * G1(o) * std::fabs(dot(o, h)) / std::fabs(o.z); }
    static void sample11(float cosT, float u1, float u2, float* sx, float* sy) {
        if (cosT > .9999) {  // the reference's float cos/sin, here in binary64 rounded once
            float r = std::sqrt(u1 / (1 - u1));
            float phi = (float)(6.28318530718 * u2);
            *sx = r * (float)std::cos((double)phi);
            *sy = r * (float)std::sin((double)phi);
            return;
        }
        float sinT = std::sqrt(fmaxr(0.f, 1.f - cosT * cosT));
        float tanT = sinT / cosT;
        float a = 1 / tanT;
        float g1 = 2 / (1 + std::sqrt(1.f + 1.f / (a * a)));
        float A = 2 * u1 / g1 - 1;
        float tmp = 1.f / (A * A - 1.f);
        if (tmp > 1e10) tmp = 1e10;
        float B = tanT;
        float Dd = std::sqrt(fmaxr(B * B * tmp * tmp - (A * A - B * B) * tmp, 0.f));
        float s1 = B * tmp - Dd, s2 = B * tmp + Dd;
        *sx = (A < 0 || s2 > 1.f / tanT) ? s1 : s2;
        float S;
        if (u2 > 0.5f) { S = 1.f; u2 = 2.f * (u2 - .5f); }
        else { S = -1.f; u2 = 2.f * (.5f - u2); }
        float z = (u2 * (u2 * (u2 * 0.27385f - 0.73369f) + 0.46341f)) /
                  (u2 * (u2 * (u2 * 0.093073f + 0.309420f) - 1.000000f) + 0.597999f);
        *sy = S * z * std::sqrt(1.f + *sx * *sx);
    }
    f3 sample(float u1, float u2, f3 o) const {  // MicrofacetDistribution.h:288-317
        bool flip = o.z < 0;
        if (flip) o = o * -1.f;
        f3 os = normalize(mk(ax * o.x, ay * o.y, o.z));
        float sx, sy;
        sample11(os.z, u1, u2, &sx, &sy);
        float t = cos_az(os) * sx - sin_az(os) * sy;
        sy = sin_az(os) * sx + cos_az(os) * sy;
        sx = t;
        sx = ax * sx;
        sy = ay * sy;
        f3 r = normalize(mk(-sx, -sy, 1.f));
        if (flip) r = r * -1.f;
        return r;
    }
};

f3 schlick(float c, f3 f0) {  // BSDFs/Fresnel.h:12-17
    float t = pow5(1 - std::fabs(c));
    return f0 + (mk(1, 1, 1) - f0) * t;
}

enum { kMatte = 0, kMirror = 1, kSubstrate = 2, kMetal = 3, kGlass = 4, kPlastic = 5 };

struct Bsdf {
    int type;
    f3 kd;            // matte / plastic diffuse
    f3 ks;            // plastic / substrate specular colour (metal: 1)
    GGX ggx;          // plastic, metal, substrate
    f3 f0;            // Fresnel (plastic: from IOR 1.5; glass: from its IOR; metal: its F0)
    float kr;         // mirror reflectivity
    f3 refl, trans;   // glass
    float ior;        // glass below IOR (above = 1)

    bool is_delta() const { return type == kMirror || type == kGlass; }
    // GGX::isAlmostDelta (BSDFs/MicrofacetDistribution.h:157-159) for the
    // BSDFs that forward it (Microfacet.h:212, Fresnel.h:146); false otherwise
    bool is_almost_delta() const {
        return (type == kMetal || type == kSubstrate) && ((double)ggx.ax < 0.05 || (double)ggx.ay < 0.05);
    }

    // LambertianBSDF (BSDFs/Lambertian.h)
    f3 lam_eval(f3 i, f3 o) const {
        if (!same_hemi(i, o)) return mk(0, 0, 0);
        return divs(kd, (float)kPi);
    }
    // MicrofacetBSDF reflection only (BSDFs/Microfacet.h:36-53, 97-116)
    f3 mf_eval(f3 i, f3 o) const {
        if (!same_hemi(i, o)) return mk(0, 0, 0);
        float co = std::fabs(o.z), ci = std::fabs(i.z);
        f3 h = i + o;
        if (ci == 0 || co == 0) return mk(0, 0, 0);
        if (h.x == 0 && h.y == 0 && h.z == 0) return mk(0, 0, 0);
        h = normalize(h);
        f3 F = schlick(dot(o, h), f0);
        return divs(mulv(ks * ggx.D(h) * ggx.G(o, i), F), 4 * ci * co);
    }
    float mf_pdf(f3 i, f3 o, f3 h) const { return ggx.pdf(h, o) / (4.f * dot(o, h)); }
    float mf_pdf(f3 i, f3 o) const { return mf_pdf(i, o, normalize(i + o)); }
    // FresnelBlendBSDF (BSDFs/Fresnel.h:58-150)
    f3 fb_eval(f3 i, f3 o) const {
        if (!same_hemi(i, o)) return mk(0, 0, 0);
        const float c = (float)(28.0 / (23.0 * kPi));
        f3 diffuse = mulv(kd * c, mk(1, 1, 1) - ks) * (1 - pow5(1 - .5f * std::fabs(i.z))) *
                     (1 - pow5(1 - .5f * std::fabs(o.z)));
        f3 h = i + o;
        if (h.x == 0 && h.y == 0 && h.z == 0) return mk(0, 0, 0);
        h = normalize(h);
        f3 spec = divs(schlick(dot(i, h), ks) * ggx.D(h),
                       4 * std::fabs(dot(i, h)) * fmaxr(std::fabs(i.z), std::fabs(o.z)));
        return diffuse + spec;
    }
    float fb_pdf(f3 i, f3 o, f3 h) const {
        return (1.f - 0.5f) * cos_pdf(i.z) + 0.5f * ggx.pdf(h, o) / (4.f * dot(o, h));
    }

    f3 eval(f3 i, f3 o) const {
        switch (type) {
            case kMatte: return lam_eval(i, o);
            case kPlastic: return lam_eval(i, o) + mf_eval(i, o);
            case kMetal: return mf_eval(i, o);
            case kSubstrate: return fb_eval(i, o);
            case kMirror:  // BSDFs/MirrorBSDF.h:18-24
                if (!(i.x == -o.x && i.y == -o.y && i.z == o.z)) return mk(0, 0, 0);
                return mk(1, 1, 1) * kr / std::fabs(i.z);
            default: {  // SpecularBSDF with transmission (BSDFs/Specular.h:86-107)
                if (same_hemi(i, o)) {
                    if (!(i.x == -o.x && i.y == -o.y && i.z == o.z)) return mk(0, 0, 0);
                    return spec_refl(i, o);
                }
                f3 refr;
                if (!spec_refract(o, refr)) return mk(0, 0, 0);
                if (!veq(refr, i)) return mk(0, 0, 0);
                return spec_trans(i, o);
            }
        }
    }
    f3 spec_refl(f3 i, f3 o) const {
        if (!same_hemi(i, o)) return mk(0, 0, 0);
        f3 F = schlick(std::fabs(i.z), f0);
        return divs(mulv(refl, F), std::fabs(i.z));
    }
    f3 spec_trans(f3 i, f3 o) const {
        if (same_hemi(i, o)) return mk(0, 0, 0);
        f3 F = schlick(std::fabs(i.z), f0);
        return divs(mulv(trans, mk(1, 1, 1) - F), std::fabs(i.z));
    }
    bool spec_refract(f3 o, f3& out) const {
        float eta = o.z > 0 ? (1.f / ior) : (ior / 1.f);
        f3 n = mk(0, 0, 1);
        if (o.z < 0) n.z = -1;
        return refract(o, n, eta, out);
    }
    float spec_refl_prob(f3 o) const {
        float p = schlick(std::fabs(o.z), f0).x;
        if (all_zero(refl)) p = 0;
        else if (all_zero(trans)) p = 1;
        return p;
    }

    float pdf(f3 i, f3 o) const {
        switch (type) {
            case kMatte: return cos_pdf(i.z);
            case kPlastic: return 0.5f * cos_pdf(i.z) + (1.f - 0.5f) * mf_pdf(i, o);
            case kMetal: return mf_pdf(i, o);
            case kSubstrate: return fb_pdf(i, o, normalize(i + o));
            case kMirror:
                if (i.x == -o.x && i.y == -o.y && i.z == o.z) return 1;
                return 0;
            default: {
                if (same_hemi(i, o)) {
                    if (!(i.x == -o.x && i.y == -o.y && i.z == o.z)) return 0;
                    return spec_refl_prob(o);
                }
                f3 refr;
                if (!spec_refract(o, refr)) return 0;
                if (!veq(refr, i)) return 0;
                return 1.f - spec_refl_prob(o);
            }
        }
    }

    f3 sample(float ux, float uy, f3& wi, f3 o, float& pdfv) const {
        switch (type) {
            case kMatte: {
                wi = cosine_hemisphere(ux, uy);
                pdfv = cos_pdf(wi.z);
                if (o.z < 0) wi.z *= -1;
                return lam_eval(wi, o);
            }
            case kPlastic: {  // BSDFs/PlasticBSDF.h:34-49
                float p = ux;
                if (p < 0.5f) {
                    p = p * (1.f / 0.5f);
                    wi = cosine_hemisphere(p, uy);
                    if (o.z < 0) wi.z *= -1;
                } else {
                    p = (p - 0.5f) * (1.f / (1.f - 0.5f));
                    f3 h = ggx.sample(p, uy, o);
                    wi = reflect(o, h);
                }
                pdfv = pdf(wi, o);
                return eval(wi, o);
            }
            case kMetal: {  // MicrofacetBSDF::sampleReflection (BSDFs/Microfacet.h:96-105)
                f3 h = ggx.sample(ux, uy, o);
                wi = reflect(o, h);
                pdfv = mf_pdf(wi, o, h);
                return mf_eval(wi, o);
            }
            case kSubstrate: {  // FresnelBlendBSDF::sample (BSDFs/Fresnel.h:101-128)
                f3 h;
                if (ux < 0.5f) {
                    h = ggx.sample(ux * (1.f / 0.5f), uy, o);
                    wi = reflect(o, h);
                    if (!same_hemi(wi, o)) {  // unset pdf in the reference; the value is zero
                        pdfv = 0;
                        return mk(0, 0, 0);
                    }
                } else {
                    wi = cosine_hemisphere((ux - 0.5f) * (1.f / (1.f - 0.5f)), uy);
                    if (o.z < 0) wi.z *= -1;
                    h = normalize(wi + o);
                }
                pdfv = fb_pdf(wi, o, h);
                return fb_eval(wi, o);
            }
            case kMirror: {
                wi = o * -1;
                wi.z = o.z;
                pdfv = 1;
                return eval(wi, o);
            }
            default: {  // SpecularBSDF::sample (BSDFs/Specular.h:122-144)
                float pr = spec_refl_prob(o);
                if (ux < pr) {
                    wi = mk(-o.x, -o.y, o.z);
                    pdfv = 1;
                    f3 r = spec_refl(wi, o);
                    pdfv *= pr;
                    return r;
                }
                f3 r = mk(0, 0, 0);
                if (spec_refract(o, wi)) {
                    pdfv = 1;
                    r = spec_trans(wi, o);
                } else {
                    // total internal reflection: direction and pdf are unset in the
                    // reference (BSDFs/Specular.h:70-82); the mirror direction, pdf 0
                    wi = mk(-o.x, -o.y, o.z);
                    pdfv = 0;
                }
                pdfv *= (1.f - pr);
                return r;
            }
        }
    }
};

// ---------------------------------------------------------------- scene
struct Scene {
    const float* tris; int ntris;
    const cav_mesh_ref* meshes; int nmeshes;
    const int32_t* tri_mesh;
    std::vector<int> lights;  // the light list: area-light mesh id, or -1 - k for env[k]
    std::vector<float> area;  // per mesh
    const float* box; const int32_t* left; const int32_t* right; const int32_t* prim;
    const float* normals = nullptr;         // float[ntris][9] per-vertex normals, or null
    const uint8_t* mesh_normals = nullptr;  // per mesh: 1 = use them (null: every mesh)
    const ref_surface* surf = nullptr;      // texture coordinates and textures, or null
    const ref_env_light* env = nullptr;
    int env_miss = -1;
    int light_dims = 3;  // Halton dimensions of a light sample after randInt + rand4
};

// The light list in the scene's order (Scene::lightsHost) and the light
// sample's share of the Halton dimensions: syncDimension (PathTracingIntegrator
// .cu:281-315) takes each bounce to the maximum over the queue, 3 (triangle +
// point) as soon as any task sampled an area light's triangles, else 2 (a
// direction: environment lights, sphere lights seen from outside); fixed here
// per scene, 3 with any area light (hits on a sphere light's tessellated
// surface lie inside the sphere and sample its triangles).
void init_lights(Scene& s) {
    const ref_surface* sf = s.surf;
    s.lights.clear();
    if (sf && sf->light_order) {
        s.lights.assign(sf->light_order, sf->light_order + sf->nlight_order);
    } else {
        for (int m = 0; m < s.nmeshes; ++m)
            if (s.meshes[m].is_light) s.lights.push_back(m);
        for (int k = 0; sf && k < sf->nenv; ++k) s.lights.push_back(-1 - k);
    }
    s.env = sf ? sf->env : nullptr;
    s.env_miss = sf ? sf->env_background : -1;
    s.light_dims = 2;
    for (int e : s.lights)
        if (e >= 0) s.light_dims = 3;
}

f3 tri_v(const Scene& s, int t, int j) {
    const float* p = s.tris + (int64_t)t * 9 + j * 3;
    return mk(p[0], p[1], p[2]);
}

Bsdf make_bsdf(const cav_mesh_ref& m) {
    Bsdf b{};
    b.type = m.material;
    b.kd = mk(m.kd[0], m.kd[1], m.kd[2]);
    b.ks = mk(m.ks[0], m.ks[1], m.ks[2]);
    b.ggx = GGX{m.roughness, m.roughness_v != 0.f ? m.roughness_v : m.roughness};
    b.kr = m.kr[0];
    b.refl = mk(m.kr[0], m.kr[1], m.kr[2]);
    b.trans = mk(m.kt[0], m.kt[1], m.kt[2]);
    b.ior = m.ior;
    float eta = m.material == kGlass ? m.ior : 1.5f;  // Materials/Plastic.h:111, Glass.h:238
    float t = (eta - 1) / (eta + 1);
    b.f0 = mk(t * t, t * t, t * t);
    if (m.material == kMetal) {  // Materials/Metal.h:24-27
        b.f0 = b.ks;
        b.ks = mk(1, 1, 1);
    }
    return b;
}

// Texture2D::readTexture (Core/Texture.h:92-95; set up at :101-125):
// normalized coordinates wrapped to [0, 1) on both axes, bilinear filtering
// between the four texels around x = u W - 0.5, y = v H - 0.5 (indices wrap
// around the image), texels read as value / 255. The weights are the exact
// fractions of x and y.
float wrap_unit(float u) {
    float w = u - std::floor(u);
    if (!(w >= 0.f && w < 1.f)) w = 0.f;
    return w;
}
f3 tex_fetch(const uint8_t* rgba, int W, int H, float u, float v) {
    const float x = wrap_unit(u) * (float)W - 0.5f, y = wrap_unit(v) * (float)H - 0.5f;
    const float fx = std::floor(x), fy = std::floor(y);
    const float a = x - fx, b = y - fy;
    const int x0 = ((int)fx + W) % W, y0 = ((int)fy + H) % H;
    const int x1 = (x0 + 1) % W, y1 = (y0 + 1) % H;
    auto at = [&](int xx, int yy, int c) { return (float)rgba[((int64_t)yy * W + xx) * 4 + c] / 255.f; };
    float r[3];
    for (int c = 0; c < 3; ++c) {
        const float lo = (1.f - a) * at(x0, y0, c) + a * at(x1, y0, c);
        const float hi = (1.f - a) * at(x0, y1, c) + a * at(x1, y1, c);
        r[c] = (1.f - b) * lo + b * hi;
    }
    return mk(r[0], r[1], r[2]);
}

// Material::getBSDF at a hit: the mesh's constants; a Kd texture read at
// tex0 (1 - u - v) + u tex1 + v tex2 (Core/Triangle.h:128-131) multiplies a
// matte colour (Materials/MatteMaterial.h:23-33) and replaces a plastic's
// diffuse colour (Materials/Plastic.h:38-50).
Bsdf surface_bsdf(const Scene& s, int mesh, int prim, float u, float v) {
    const cav_mesh_ref& M = s.meshes[mesh];
    Bsdf b = make_bsdf(M);
    const ref_surface* sf = s.surf;
    const int ti = M.texture;
    if (!sf || !sf->texels || ti < 0 || (M.material != kMatte && M.material != kPlastic && M.material != kSubstrate))
        return b;
    float tu = 0.f, tv = 0.f;
    if (sf->tri_uv) {
        const float* q = sf->tri_uv + (int64_t)prim * 6;
        tu = (q[0] * (1.f - u - v) + u * q[2]) + v * q[4];
        tv = (q[1] * (1.f - u - v) + u * q[3]) + v * q[5];
    }
    const f3 texel = tex_fetch(sf->texels + sf->tex_offset[ti] * 4, sf->tex_size[2 * ti], sf->tex_size[2 * ti + 1],
                               tu, tv);
    b.kd = M.material == kMatte ? mulv(b.kd, texel) : texel;
    return b;
}

// EnvironmentMap::evaluateRay (Core/Light.h:91-113): colour times the texel
// at (phi / 2pi, theta / pi) of the direction in the light's frame; atan2 and
// acos in binary64 rounded once, the M_PI divisions in binary64 as written.
f3 env_eval(const Scene& s, int e, f3 d) {
    const ref_env_light& E = s.env[e];
    f3 r = mk(E.color[0], E.color[1], E.color[2]);
    const ref_surface* sf = s.surf;
    if (E.texture >= 0 && sf && sf->texels) {
        const float* m = E.world_to_light;
        const f3 l = mk((m[0] * d.x + m[1] * d.y) + m[2] * d.z, (m[3] * d.x + m[4] * d.y) + m[5] * d.z,
                        (m[6] * d.x + m[7] * d.y) + m[8] * d.z);
        float phi = (float)std::atan2((double)l.y, (double)l.x);
        phi = phi < 0 ? (float)((double)phi + 2 * kPi) : phi;
        const float theta = (float)std::acos((double)clampf(l.z, -1.f, 1.f));
        const float tu = (float)((double)phi / (2.0 * kPi)), tv = (float)((double)theta / kPi);
        const int ti = E.texture;
        r = mulv(r, tex_fetch(sf->texels + sf->tex_offset[ti] * 4, sf->tex_size[2 * ti], sf->tex_size[2 * ti + 1],
                              tu, tv));
    }
    return r;
}

// uniformSampleSphere (Utils/RandomUtils.h:70-78), cos / sin in binary64
f3 uniform_sphere(float x, float y) {
    const float u = (float)((double)(x * 2.f) * kPi);
    const float v = (float)(((double)y - 0.5) * kPi);
    const float cu = (float)std::cos((double)u), su = (float)std::sin((double)u);
    const float cv = (float)std::cos((double)v), sv = (float)std::sin((double)v);
    return mk(cv * cu, sv, cv * su);
}

// single-ray traversal on the reference node arrays (BVH/BVH.h)
bool box_hit1(const float* b, f3 o, f3 inv, float& mind) {
    f3 tmin = mul(mk(b[0], b[1], b[2]) - o, inv);
    f3 tmax = mul(mk(b[3], b[4], b[5]) - o, inv);
    float maxmin = fmin_fmax(tmin.x, tmax.x, fmin_fmax(tmin.y, tmax.y, fminr(tmin.z, tmax.z)));
    float minmax = fmax_fmin(tmin.x, tmax.x, fmax_fmin(tmin.y, tmax.y, fmaxr(tmin.z, tmax.z)));
    if (maxmin <= minmax && minmax > 0) { mind = fmaxr(0.f, maxmin); return true; }
    mind = -1;
    return false;
}
bool tri_hit1(const Scene& s, int p, f3 o, f3 d, float& t, float& u, float& v) {
    f3 v0 = tri_v(s, p, 0);
    f3 e1 = tri_v(s, p, 1) - v0, e2 = tri_v(s, p, 2) - v0;
    f3 tv = o - v0, pv = cross(d, e2);
    float det = dot(e1, pv);
    if (det == 0) return false;
    det = 1.f / det;
    u = dot(tv, pv) * det;
    if (u < 0.0f || u > 1.0f) return false;
    f3 qv = cross(tv, e1);
    v = dot(d, qv) * det;
    if (v < 0.0f || (u + v) > 1.0f) return false;
    t = dot(e2, qv) * det;
    return t > 0.f;
}
int closest(const Scene& s, f3 o, f3 d, float& tb, float* bu = nullptr, float* bv = nullptr) {
    f3 inv = divv(mk(1, 1, 1), d);
    float md;
    tb = FLT_MAX;
    int bp = -1;
    float best_u = 0, best_v = 0;
    if (!box_hit1(s.box, o, inv, md)) return -1;
    int st[64], top = 0;
    st[0] = 0;
    while (top >= 0) {
        int c = st[top--];
        if (s.prim[c] >= 0) {
            float t, u, v;
            if (tri_hit1(s, s.prim[c], o, d, t, u, v) && (bp < 0 || t < tb)) {
                tb = t;
                bp = s.prim[c];
                best_u = u;
                best_v = v;
            }
        } else {
            float dl, dr;
            int l = s.left[c], r = s.right[c];
            box_hit1(s.box + (int64_t)l * 6, o, inv, dl);
            box_hit1(s.box + (int64_t)r * 6, o, inv, dr);
            bool hl = dl >= 0 && dl < tb, hr = dr >= 0 && dr < tb;
            if (hl && hr) {
                if (dl > dr) { if (top < 63) st[++top] = l; if (top < 63) st[++top] = r; }
                else { if (top < 63) st[++top] = r; if (top < 63) st[++top] = l; }
            } else if (hl) { if (top < 63) st[++top] = l; }
            else if (hr) { if (top < 63) st[++top] = r; }
        }
    }
    if (bu) *bu = best_u;
    if (bv) *bv = best_v;
    return bp;
}
bool visible(const Scene& s, f3 o, f3 d, float lim, int sm, int dm) {
    f3 inv = divv(mk(1, 1, 1), d);
    int st[64], top = 0;
    st[0] = 0;
    while (top >= 0) {
        int c = st[top--];
        if (s.prim[c] >= 0) {
            int p = s.prim[c];
            int m = s.tri_mesh[p];
            if (m == sm || m == dm) continue;
            float t, u, v;
            if (tri_hit1(s, p, o, d, t, u, v) && t < lim) return false;
        } else {
            float dl, dr;
            int l = s.left[c], r = s.right[c];
            box_hit1(s.box + (int64_t)l * 6, o, inv, dl);
            box_hit1(s.box + (int64_t)r * 6, o, inv, dr);
            if (dl >= 0 && dr >= 0) {
                if (dl > dr) { if (top < 63) st[++top] = l; if (top < 63) st[++top] = r; }
                else { if (top < 63) st[++top] = r; if (top < 63) st[++top] = l; }
            } else if (dl >= 0) { if (top < 63) st[++top] = l; }
            else if (dr >= 0) { if (top < 63) st[++top] = r; }
        }
    }
    return true;
}

f3 tri_normal(const Scene& s, int p) {  // Core/Triangle.h:211-222 without vertex normals
    f3 v0 = tri_v(s, p, 0);
    f3 n = normalize(cross(tri_v(s, p, 1) - v0, tri_v(s, p, 2) - v0));
    if (s.meshes[s.tri_mesh[p]].reverse_orientation) n = n * -1.f;
    return n;
}

// Triangle::fillIntersectionInformation / TriangleMesh::getNormal
// (Core/Triangle.h:94-136, Core/TriangleMesh.h:76-104): the interpolated
// vertex normal normalize(n0 (1 - u - v) + u n1 + v n2) for meshes that have
// vertex normals, else the geometric normal; flipped for ReverseOrientation.
f3 shading_normal(const Scene& s, int p, float u, float v) {
    const int m = s.tri_mesh[p];
    if (!(s.normals && (!s.mesh_normals || s.mesh_normals[m]))) return tri_normal(s, p);
    const float* q = s.normals + (int64_t)p * 9;
    f3 n0 = mk(q[0], q[1], q[2]), n1 = mk(q[3], q[4], q[5]), n2 = mk(q[6], q[7], q[8]);
    f3 n = normalize((n0 * (1.f - u - v) + u * n1) + v * n2);
    if (s.meshes[m].reverse_orientation) n = n * -1.f;
    return n;
}

float power_heuristic(float a, float b) { a *= a; b *= b; return a / (a + b); }

// TriangleMesh::sampleSphere (Core/TriangleMesh.h:155-211): a direction
// uniform in the cone the sphere subtends from pos; the point it meets, the
// outward normal there, and the cone pdf. Trig in binary64 rounded once.
void sample_sphere(f3 pos, f3 c, float rad, float ux, float uy, f3& p, f3& n, float& pdf) {
    float dc = length(pos - c);
    float inv_dc = 1 / dc;
    f3 wc = (c - pos) * inv_dc;
    f3 wx, wy;
    coordinate_system(wc, wx, wy);
    float st_max = rad * inv_dc;
    float st_max2 = st_max * st_max;
    float inv_st_max = 1 / st_max;
    float ct_max = std::sqrt(std::max(0.f, 1 - st_max2));
    float ct = (ct_max - 1) * ux + 1;
    float st2 = 1 - ct * ct;
    if (st_max2 < 0.00068523f) {  // sin^2(1.5 deg): small-angle form
        st2 = st_max2 * ux;
        ct = std::sqrt(1 - st2);
    }
    float ca = st2 * inv_st_max + ct * std::sqrt(std::max(0.f, 1.f - st2 * inv_st_max * inv_st_max));
    float sa = std::sqrt(std::max(0.f, 1.f - ca * ca));
    float phi = (float)((double)(uy * 2.f) * kPi);
    float cp = (float)std::cos((double)phi), sp = (float)std::sin((double)phi);
    n = ((wx * (sa * cp) + wy * (sa * sp)) + wc * ct) * -1.f;
    p = c + n * rad;
    pdf = (float)(1.0 / (2.0 * kPi * (double)(1 - ct_max)));
}

// TriangleMesh::pdf (Core/TriangleMesh.h:214-232) for a ray from o that hit
// light mesh m at pos, normal n: the cone pdf for a sphere seen from outside,
// else the area pdf in solid angle.
float light_pdf(const Scene& s, int m, f3 o, f3 pos, f3 n) {
    const cav_mesh_ref& L = s.meshes[m];
    if (L.sphere[3] > 0) {
        f3 oc = o - mk(L.sphere[0], L.sphere[1], L.sphere[2]);
        float d2 = dot(oc, oc);
        if (d2 > L.sphere[3] * L.sphere[3]) {
            float st2 = L.sphere[3] * L.sphere[3] / d2;
            float ct = std::sqrt(std::max(0.f, 1.f - st2));
            return (float)(1.0 / (2.0 * kPi * (double)(1 - ct)));
        }
    }
    f3 l2s = o - pos;
    float c = std::fabs(dot(n, normalize(l2s)));
    return dot(l2s, l2s) / (c * s.area[m]);
}

// LightObject::sampleRayToPoint of light-list entry lm from pos, reading
// dimensions from dim on and leaving dim light_dims further: an area light
// (DiffuseAreaLight / TriangleMesh::sample, Core/TriangleMesh.h:108-153) or
// an environment-type light (uniform sphere direction, pdf 1 / 4pi, no
// distance limit, Core/Light.h:75-84).
struct LightPick { f3 dir, incident; float prob, lim; int dst; };
LightPick sample_light(const Scene& s, const Halton& H, uint64_t hidx, int& dim, f3 pos, int lm) {
    LightPick r;
    const int d0 = dim;
    if (lm < 0) {
        const float x = H.rand1(hidx, dim);
        const float y = H.rand1(hidx, dim);
        r.dir = uniform_sphere(x, y);
        r.prob = (float)(1.0 / (4.0 * kPi));
        r.incident = env_eval(s, -1 - lm, r.dir);
        r.lim = INFINITY;
        r.dst = -1;
    } else {
        const cav_mesh_ref& L = s.meshes[lm];
        r.dst = lm;
        const f3 sc = mk(L.sphere[0], L.sphere[1], L.sphere[2]);
        const f3 pc = pos - sc;
        if (L.sphere[3] > 0 && dot(pc, pc) > L.sphere[3] * L.sphere[3]) {  // Core/TriangleMesh.h:112-118
            const float ux = H.rand1(hidx, dim);
            const float uy = H.rand1(hidx, dim);
            f3 lp, ln;
            sample_sphere(pos, sc, L.sphere[3], ux, uy, lp, ln, r.prob);
            if (L.reverse_orientation) ln = ln * -1.f;
            r.dir = normalize(lp - pos);
            r.lim = length(lp - pos);
            r.incident = dot(r.dir, ln) >= 0 ? mk(0, 0, 0) : mk(L.emission[0], L.emission[1], L.emission[2]);
            dim = d0 + s.light_dims;
            return r;
        }
        int tri = L.first_tri + H.rand_int(L.num_tris, hidx, dim);
        float tmp = std::sqrt(H.rand1(hidx, dim));
        float u = 1 - tmp;
        float v = tmp * H.rand1(hidx, dim);
        f3 p0 = tri_v(s, tri, 0), p1 = tri_v(s, tri, 1), p2 = tri_v(s, tri, 2);
        f3 lp = (p0 * u + p1 * v) + p2 * (1.f - u - v);
        f3 ln = shading_normal(s, tri, u, v);
        f3 l2s = pos - lp;
        float c = std::fabs(dot(ln, normalize(l2s)));
        r.prob = dot(l2s, l2s) / (c * s.area[lm]);
        r.dir = normalize(lp - pos);
        r.lim = length(lp - pos);
        r.incident = dot(r.dir, ln) >= 0 ? mk(0, 0, 0) : mk(L.emission[0], L.emission[1], L.emission[2]);
    }
    dim = d0 + s.light_dims;
    return r;
}

// glm mat4 * vec4 followed by the divide (Utils/MathsCommons.h:57-60)
f3 apply(const float* m, f3 v) {
    float r[4];
    for (int k = 0; k < 4; ++k)
        r[k] = (m[0 * 4 + k] * v.x + m[1 * 4 + k] * v.y) + (m[2 * 4 + k] * v.z + m[3 * 4 + k] * 1.f);
    return mk(r[0] / r[3], r[1] / r[3], r[2] / r[3]);
}

}  // namespace

extern "C" void ref_halton_table(uint32_t* primes, uint32_t* start, uint16_t* perm, int perm_cap) {
    const Halton& h = halton();
    for (int i = 0; i < 1024; ++i) { primes[i] = h.primes[i]; start[i] = h.start[i]; }
    for (int i = 0; i < perm_cap && i < (int)h.perm.size(); ++i) perm[i] = h.perm[i];
}

extern "C" float ref_halton(uint64_t index, int dim) {
    int d = dim;
    return halton().rand1(index, d);
}

extern "C" void ref_mesh_areas(const float* tris, const cav_mesh_ref* meshes, int nmeshes, float* area) {
    for (int m = 0; m < nmeshes; ++m) {  // TriangleMesh::computeArea (Core/TriangleMesh.cu:189-199)
        float a = 0;
        for (int t = meshes[m].first_tri; t < meshes[m].first_tri + meshes[m].num_tris; ++t) {
            const float* p = tris + (int64_t)t * 9;
            f3 p0 = mk(p[0], p[1], p[2]), p1 = mk(p[3], p[4], p[5]), p2 = mk(p[6], p[7], p[8]);
            a += length(cross(p1 - p0, p2 - p0)) * 0.5f;
        }
        area[m] = a;
    }
}

// box-filtered film (Core/Film.h:25-49, Filters/BoxFilter.h:19-25) in a
// fixed order: for each pixel, its 3x3 neighbours in row-major order each
// contribute the sum (in sample order) of their samples whose filter support
// covers the pixel; non-finite samples are skipped
void gather_film(int width, int height, int spp, const std::vector<float>& res, const std::vector<float>& sx,
                 const std::vector<float>& sy, float* film_rgb, float* film_w, int x0, int y0, int x1, int y1) {
    // pixels of [x0, x1) x [y0, y1), stored row by row from (x0, y0)
    const int cw = x1 - x0;
#pragma omp parallel for schedule(static)
    for (int64_t cpix = 0; cpix < (int64_t)cw * (y1 - y0); ++cpix) {
        const int x = x0 + (int)(cpix % cw), y = y0 + (int)(cpix / cw);
        const int64_t pix = cpix;
        float r = 0, g = 0, b = 0, w = 0;
        for (int py = y - 1; py <= y + 1; ++py)
            for (int px = x - 1; px <= x + 1; ++px) {
                if (px < 0 || py < 0 || px >= width || py >= height) continue;
                float pr = 0, pg = 0, pb = 0, pw = 0;
                for (int k = 0; k < spp; ++k) {
                    const int64_t si = ((int64_t)py * width + px) * spp + k;
                    const float* c = &res[si * 3];
                    if (!std::isfinite(c[0]) || !std::isfinite(c[1]) || !std::isfinite(c[2])) continue;
                    const float fx = sx[si], fy = sy[si];
                    if (x < (int)std::floor(fx - 1.f) || x > (int)std::ceil(fx + 1.f)) continue;
                    if (y < (int)std::floor(fy - 1.f) || y > (int)std::ceil(fy + 1.f)) continue;
                    if (std::fabs((float)x - fx) > 1.f || std::fabs((float)y - fy) > 1.f) continue;
                    pr += c[0]; pg += c[1]; pb += c[2]; pw += 1.f;
                }
                r += pr; g += pg; b += pb; w += pw;
            }
        film_rgb[pix * 3 + 0] = r;
        film_rgb[pix * 3 + 1] = g;
        film_rgb[pix * 3 + 2] = b;
        film_w[pix] = w;
    }
}

// Renders with the `path` integrator. The BVH node arrays are the oracle's own
// (ref_lbvh_build + ref_treelet_optimize). Outputs per-pixel sum (rgb) and
// weight of the box-filtered film (Core/Film.h:25-49), accumulated in a fixed
// order: for each pixel, samples of pixels (y-1..y+1, x-1..x+1) in row-major
// pixel order, then sample order.
// The same restricted to the pixels of crop = {x0, y0, x1, y1} of the full
// image: only the samples that reach them (their pixels and the ring around)
// are traced, with the full image's sample indices, so the crop equals that
// part of the full film. samples_out (optional): per crop pixel and sample,
// radiance r, g, b and film position x, y.
extern "C" int ref_render_path_crop(const float* tris, int ntris, const int32_t* tri_mesh,
                                    const cav_mesh_ref* meshes, int nmeshes, const float* box,
                                    const int32_t* left, const int32_t* right, const int32_t* prim,
                                    const ref_surface* surf, const float* cam16, float cam_depth, int width,
                                    int height, int spp, int max_depth, const int32_t* crop, float* film_rgb,
                                    float* film_w, float* samples_out, int64_t* rays_traced) {
    Scene s{tris, ntris, meshes, nmeshes, tri_mesh, {}, {}, box, left, right, prim,
             surf ? surf->tri_normals : nullptr, surf ? surf->mesh_normals : nullptr, surf};
    init_lights(s);
    s.area.resize(nmeshes);
    ref_mesh_areas(tris, meshes, nmeshes, s.area.data());
    const Halton& H = halton();
    const int64_t nsamp = (int64_t)width * height * spp;
    std::vector<float> res((size_t)nsamp * 3), sx((size_t)nsamp), sy((size_t)nsamp);
    int64_t traced = 0;
    const int nl = (int)s.lights.size();
    const int cx0 = crop[0], cy0 = crop[1], cx1 = crop[2], cy1 = crop[3];
    if (cx0 < 0 || cy0 < 0 || cx1 > width || cy1 > height || cx0 >= cx1 || cy0 >= cy1) return -1;
    const int ex0 = std::max(0, cx0 - 1), ey0 = std::max(0, cy0 - 1);
    const int ex1 = std::min(width, cx1 + 1), ey1 = std::min(height, cy1 + 1);
    const int64_t ew = ex1 - ex0, nwork = ew * (ey1 - ey0) * spp;

#pragma omp parallel for schedule(dynamic, 16) reduction(+ : traced)
    for (int64_t j = 0; j < nwork; ++j) {
        const int64_t wp = j / spp;
        const int64_t si = ((int64_t)(ey0 + wp / ew) * width + (ex0 + wp % ew)) * spp + j % spp;
        const int64_t pix = si / spp;
        const int x = (int)(pix % width), y = (int)(pix / width);
        const uint64_t hidx = (uint64_t)pix * spp + (uint64_t)(si % spp);  // Core/Sampler.cu:29-51
        int dim = 0;
        float fx = (float)x, fy = (float)y;
        fx += H.rand1(hidx, dim);
        fy += H.rand1(hidx, dim);
        sx[si] = fx;
        sy[si] = fy;
        // PerspectiveCamera::genRay (Cameras/PerspectiveCamera.h:32-48)
        f3 pl = mk(fx - width / 2.f, height / 2.f - fy, cam_depth);
        f3 ro = apply(cam16, mk(0, 0, 0));
        f3 rd = normalize(apply(cam16, pl) - ro);
        f3 result = mk(0, 0, 0), mult = mk(1, 1, 1);
        float mat_pdf = 1;
        bool from_delta = true;
        for (int depth = 0; depth < max_depth; ++depth) {
            float t, hu, hv;
            ++traced;
            int p = closest(s, ro, rd, t, &hu, &hv);
            if (p < 0) {  // intersectScene's miss (PathTracingIntegrator.cu:112-116)
                if (s.env_miss >= 0) result = result + mulv(env_eval(s, s.env_miss, rd), mult);
                break;
            }
            const int mesh = tri_mesh[p];
            const cav_mesh_ref& M = meshes[mesh];
            f3 n = shading_normal(s, p, hu, hv);
            f3 pos = ro + t * rd;
            Bsdf bsdf = surface_bsdf(s, mesh, p, hu, hv);
            f3 t0, t1;
            coordinate_system(n, t0, t1);
            // computeLighting (PathTracingIntegrator.cu:174-225)
            if (M.is_light) {
                f3 Le = dot(rd, n) >= 0 ? mk(0, 0, 0) : mk(M.emission[0], M.emission[1], M.emission[2]);
                if (from_delta) {
                    result = result + mulv(Le, mult);
                } else if (depth > 0) {
                    float lpdf = light_pdf(s, mesh, ro, pos, n);
                    float w = power_heuristic(mat_pdf, lpdf);
                    if (std::isfinite(w)) result = result + mulv(Le, mult) * w;
                }
            }
            f3 wo = rd * -1;
            int li = H.rand_int(nl, hidx, dim);
            dim += 4;  // rand4, unused by the reference
            {
                const LightPick L = sample_light(s, H, hidx, dim, pos, s.lights[li]);
                const float prob = L.prob;
                const f3 dir = L.dir, incident = L.incident;
                // a shadow ray only matters (and is only counted) when prob is finite
                if (std::isfinite(prob) && (++traced, visible(s, pos, dir, L.lim, mesh, L.dst))) {
                    f3 em = mult * (float)nl / prob;
                    f3 il = to_local(dir, t0, t1, n), ol = to_local(wo, t0, t1, n);
                    float mp = bsdf.pdf(il, ol);
                    float w = power_heuristic(prob, mp);
                    float cosv = std::fabs(dot(dir, n));  // Material::eval (Core/Material.h:20-26)
                    f3 contrib = mulv(incident, bsdf.eval(il, ol)) * cosv;
                    result = result + mulv(contrib, em) * w;
                }
            }
            // genNextRay (PathTracingIntegrator.cu:124-171)
            if (depth > 3) {
                float q = 1;
                q = fminr(q, 1 - mult.x);
                q = fminr(q, 1 - mult.y);
                q = fminr(q, 1 - mult.z);
                q = fmaxr(q, 0.05f);
                if (H.rand1(hidx, dim) < q) break;
                mult = divs(mult, 1.f - q);
            }
            float ux = H.rand1(hidx, dim);
            float uy = H.rand1(hidx, dim);
            f3 wl;
            float npdf = 0;
            f3 f = bsdf.sample(ux, uy, wl, to_local(wo, t0, t1, n), npdf);
            rd = to_world(wl, t0, t1, n);
            ro = pos + rd * 0.001f;
            if (all_zero(f)) mult = mk(0, 0, 0);
            else mult = divs(mulv(mult, f) * std::fabs(dot(rd, n)), npdf);
            mat_pdf = npdf;
            from_delta = bsdf.is_delta();
        }
        res[si * 3 + 0] = result.x;
        res[si * 3 + 1] = result.y;
        res[si * 3 + 2] = result.z;
    }

    gather_film(width, height, spp, res, sx, sy, film_rgb, film_w, cx0, cy0, cx1, cy1);
    if (samples_out) {
        const int cw = cx1 - cx0;
        for (int y = cy0; y < cy1; ++y)
            for (int x = cx0; x < cx1; ++x)
                for (int k = 0; k < spp; ++k) {
                    const int64_t si = ((int64_t)y * width + x) * spp + k;
                    float* o = samples_out + (((int64_t)(y - cy0) * cw + (x - cx0)) * spp + k) * 5;
                    o[0] = res[si * 3]; o[1] = res[si * 3 + 1]; o[2] = res[si * 3 + 2];
                    o[3] = sx[si]; o[4] = sy[si];
                }
    }
    if (rays_traced) *rays_traced = traced;
    return 0;
}

extern "C" int ref_render_path(const float* tris, int ntris, const int32_t* tri_mesh,
                               const cav_mesh_ref* meshes, int nmeshes, const float* box,
                               const int32_t* left, const int32_t* right, const int32_t* prim, const ref_surface* surf,
                               const float* cam16, float cam_depth, int width, int height,
                               int spp, int max_depth, float* film_rgb, float* film_w,
                               int64_t* rays_traced) {
    const int32_t full[4] = {0, 0, width, height};
    return ref_render_path_crop(tris, ntris, tri_mesh, meshes, nmeshes, box, left, right, prim, surf, cam16,
                                cam_depth, width, height, spp, max_depth, full, film_rgb, film_w, nullptr,
                                rays_traced);
}


// Renders with the `directlighting` integrator
// (Integrators/DirectLightingIntegrator.cu:22-140): the camera ray's hit adds
// the light's emission (no MIS), then one light sample whose unoccluded
// contribution is Material::eval * lightsCount / pdf (no MIS weight).
extern "C" int ref_render_direct(const float* tris, int ntris, const int32_t* tri_mesh,
                                 const cav_mesh_ref* meshes, int nmeshes, const float* box,
                                 const int32_t* left, const int32_t* right, const int32_t* prim, const ref_surface* surf,
                                 const float* cam16, float cam_depth, int width, int height, int spp,
                                 float* film_rgb, float* film_w, int64_t* rays_traced) {
    Scene s{tris, ntris, meshes, nmeshes, tri_mesh, {}, {}, box, left, right, prim,
             surf ? surf->tri_normals : nullptr, surf ? surf->mesh_normals : nullptr, surf};
    init_lights(s);
    s.area.resize(nmeshes);
    ref_mesh_areas(tris, meshes, nmeshes, s.area.data());
    const Halton& H = halton();
    const int64_t nsamp = (int64_t)width * height * spp;
    std::vector<float> res((size_t)nsamp * 3), sx((size_t)nsamp), sy((size_t)nsamp);
    int64_t traced = 0;
    const int nl = (int)s.lights.size();
#pragma omp parallel for schedule(dynamic, 16) reduction(+ : traced)
    for (int64_t si = 0; si < nsamp; ++si) {
        const int64_t pix = si / spp;
        const int x = (int)(pix % width), y = (int)(pix / width);
        const uint64_t hidx = (uint64_t)si;
        int dim = 0;
        float fx = (float)x, fy = (float)y;
        fx += H.rand1(hidx, dim);
        fy += H.rand1(hidx, dim);
        sx[si] = fx;
        sy[si] = fy;
        f3 pl = mk(fx - width / 2.f, height / 2.f - fy, cam_depth);
        f3 ro = apply(cam16, mk(0, 0, 0));
        f3 rd = normalize(apply(cam16, pl) - ro);
        f3 result = mk(0, 0, 0);
        float t, hu, hv;
        ++traced;
        const int p = closest(s, ro, rd, t, &hu, &hv);
        if (p >= 0) {
            const int mesh = tri_mesh[p];
            const cav_mesh_ref& M = meshes[mesh];
            f3 n = shading_normal(s, p, hu, hv);
            f3 pos = ro + t * rd;
            if (M.is_light) {
                f3 Le = dot(rd, n) >= 0 ? mk(0, 0, 0) : mk(M.emission[0], M.emission[1], M.emission[2]);
                result = result + Le;
            }
            const int li = H.rand_int(nl, hidx, dim);
            dim += 4;  // rand4, unused by the reference
            const LightPick L = sample_light(s, H, hidx, dim, pos, s.lights[li]);
            const float prob = L.prob;
            const f3 dir = L.dir, incident = L.incident;
            ++traced;
            if (visible(s, pos, dir, L.lim, mesh, L.dst)) {
                Bsdf bsdf = surface_bsdf(s, mesh, p, hu, hv);
                f3 t0, t1;
                coordinate_system(n, t0, t1);
                f3 il = to_local(dir, t0, t1, n), ol = to_local(rd * -1, t0, t1, n);
                const float mult = (float)nl / prob;
                f3 contrib = (mulv(incident, bsdf.eval(il, ol)) * std::fabs(dot(dir, n))) * mult;
                result = result + contrib;
            }
        } else if (s.env_miss >= 0) {  // DirectLightingIntegrator.cu:31-36
            result = result + env_eval(s, s.env_miss, rd);
        }
        res[si * 3 + 0] = result.x;
        res[si * 3 + 1] = result.y;
        res[si * 3 + 2] = result.z;
    }
    gather_film(width, height, spp, res, sx, sy, film_rgb, film_w, 0, 0, width, height);
    if (rays_traced) *rays_traced = traced;
    return 0;
}


// ====================================================================== rlpath
// Integrators/ReinforcementLearningPathTracing.cu, bounce by bounce over the
// live samples (in sample order), because the Q-table and the per-cell light
// distributions learnt in one bounce are shared by every sample. The
// reference's syncDimension is applied as written: each sync sets every task
// of the queue to the largest Halton dimension in the queue.
namespace {

struct QTab {
    int g, E, L;
    std::vector<float> Q, newQ;
    std::vector<int> cnt, total;
    std::vector<float> ldist, lres, lnew;
    std::vector<int> lcount, lnewc;
};

const int kNumX = 16, kNumY = 8, kNumXY = 128;

// QEntry::dirToCellIndex (RL .cu:113-128); (int)NaN is taken as 0, the value
// the reference's float->int conversion gives.
int f2i(float v) { return std::isnan(v) ? 0 : (v >= 2147483648.f ? 2147483647 : (v <= -2147483648.f ? (-2147483647 - 1) : (int)v)); }
int clampi(int x, int lo, int hi) { return x < lo ? lo : (x > hi ? hi : x); }
int dir_to_cell(f3 d) {
    float u = d.z;
    float y = (u + 1.f) / 2.f;
    int ti = clampi(f2i(y * kNumY), 0, kNumY - 1);
    d.x /= std::sqrt(1.f - u * u);
    d.y /= std::sqrt(1.f - u * u);
    // acos / cos / sin are evaluated in binary64 and rounded once, so the
    // single-precision results are correctly rounded (the HIP kernels do the
    // same) rather than depending on a libm's float routines.
    float v = (float)std::acos((double)d.x);
    if (d.y < 0) v = (float)((2.0 * kPi) - (double)v);
    v = (float)((double)v / (2.0 * kPi));
    int pi = clampi(f2i(v * kNumX), 0, kNumX - 1);
    return ti * kNumX + pi;
}
// QEntry::sampleDirectionInCell (RL .cu:131-147)
f3 cell_dir(float rx, float ry, int cell) {
    int ti = cell / kNumX, pi = cell % kNumX;
    float u = ((float)ti + rx) * (1.f / (float)kNumY);
    u = u * 2 - 1.f;
    float v = ((float)pi + ry) * (1.f / (float)kNumX);
    float xy = std::sqrt(1.0f - u * u);
    float phi = (float)(2 * kPi * v);
    return mk(xy * (float)std::cos((double)phi), xy * (float)std::sin((double)phi), u);
}
// findQEntry (RL .cu:155-165)
int find_entry(const float* b6, int g, f3 p) {
    f3 size = mk(b6[3], b6[4], b6[5]) - mk(b6[0], b6[1], b6[2]);
    f3 es = divv(size, mk((float)g, (float)g, (float)g));
    p = p - mk(b6[0], b6[1], b6[2]);
    int x = clampi(f2i(p.x / es.x), 0, g - 1);
    int y = clampi(f2i(p.y / es.y), 0, g - 1);
    int z = clampi(f2i(p.z / es.z), 0, g - 1);
    return x * g * g + y * g + z;
}
float luminance(f3 c) { return dot(c, mk(0.212671f, 0.715160f, 0.072169f)); }

// Distribution1D::getCorrespondingIndex / sample (Utils/RandomUtils.h:120-160)
int dist_index(const float* cdf, int n, float f) {
    int l = 0, r = n - 1;
    while (l < r) {
        int m = (l + r) / 2;
        if (cdf[m] >= f) {
            if (m == 0) return 0;
            if (cdf[m - 1] < f) return m;
            r = m - 1;
        } else {
            l = m + 1;
        }
    }
    return l;
}
float dist_pdf(const float* cdf, int i) { return i == 0 ? cdf[0] : cdf[i] - cdf[i - 1]; }

struct Task {
    uint32_t sid;
    f3 ro, rd, mult;
    float pdf;
    bool delta;
    int prev_entry, prev_cell;
    int dim;
};

}  // namespace

// finishUpdateQ + averageQ (RL .cu:44-110, 738-767) over `nent` entries of
// 128 bins, then prepareForUpdateQ (newQ and the proposal counts cleared).
extern "C" void ref_rl_finish_q(float* Q, float* newQ, int32_t* cnt, int32_t* total, int nent,
                                float alpha_scale) {
    for (int64_t i = 0; i < (int64_t)nent * kNumXY; ++i) {
        if (cnt[i] != 0) {
            float a = (float)cnt[i] / (float)(total[i] + cnt[i]);
            a *= alpha_scale;
            if (a != 0) {
                Q[i] = Q[i] * (1.f - a) + a * newQ[i] / (float)cnt[i];
                total[i] += cnt[i];
            }
        }
        newQ[i] = 0.f;
        cnt[i] = 0;
    }
    for (int e = 0; e < nent; ++e) {
        float* Qe = Q + (int64_t)e * kNumXY;
        const int32_t* tot = total + (int64_t)e * kNumXY;
        float sq = 0;
        int c = 0;
        for (int i = 0; i < kNumXY; ++i)
            if (tot[i] > 0) { sq += Qe[i]; c += 1; }
        if (c == 0) continue;
        const float avg = sq / (float)c;
        for (int i = 0; i < kNumXY; ++i)
            if (tot[i] == 0) Qe[i] = avg;
    }
}

// finishLightingTraining (RL .cu:463-514) over `nent` entries of L lights;
// the new results and counts are cleared (prepareLightingTraining).
extern "C" void ref_rl_finish_light(float* ldist, float* lres, int32_t* lcount, float* lnew,
                                    int32_t* lnewc, int nent, int L) {
    for (int e = 0; e < nent; ++e) {
        float sum_r = 0, valid = 0;
        for (int i = 0; i < L; ++i) {
            const size_t j = (size_t)e * L + i;
            if (lnewc[j] > 0) {
                float tc = (float)(lcount[j] + lnewc[j]);
                lres[j] = ((float)lcount[j] / tc) * lres[j] + ((float)lnewc[j] / tc) * (lnew[j] / (float)lnewc[j]);
                lcount[j] += lnewc[j];
                sum_r += lres[j];
                valid += 1;
            }
            lnewc[j] = 0;
            lnew[j] = 0;
        }
        if (sum_r > 0) {
            float avg = sum_r / (float)valid;
            float dens = 0;
            for (int i = 0; i < L; ++i) {
                const size_t j = (size_t)e * L + i;
                if (lres[j] > 0) dens += lres[j];
                else if (lcount[j] > 0) dens += avg / (float)lcount[j];
                else dens += avg;
                ldist[j] = dens;
            }
            for (int i = 0; i < L; ++i) ldist[(size_t)e * L + i] /= dens;
        }
    }
}

// rlpath render. scene6 = scene bounds (union of triangle boxes); q_grid = Q
// entries per axis (the reference's Q_TABLE_SIZE is 32); alpha_scale scales
// the learning rate n/(T+n) (1 = the reference's rule).
extern "C" int ref_render_rlpath(const float* tris, int ntris, const int32_t* tri_mesh,
                                 const cav_mesh_ref* meshes, int nmeshes, const float* box,
                                 const int32_t* left, const int32_t* right, const int32_t* prim, const ref_surface* surf,
                                 const float* scene6, int q_grid, float alpha_scale,
                                 const float* cam16, float cam_depth, int width, int height,
                                 int spp, int max_depth, int spp_per_round, float* film_rgb,
                                 float* film_w, int64_t* rays_traced) {
    Scene s{tris, ntris, meshes, nmeshes, tri_mesh, {}, {}, box, left, right, prim,
             surf ? surf->tri_normals : nullptr, surf ? surf->mesh_normals : nullptr, surf};
    init_lights(s);
    s.area.resize(nmeshes);
    ref_mesh_areas(tris, meshes, nmeshes, s.area.data());
    const Halton& H = halton();
    const int L = (int)s.lights.size();
    QTab T;
    T.g = q_grid;
    T.E = q_grid * q_grid * q_grid;
    T.L = L;
    T.Q.assign((size_t)T.E * kNumXY, 1.f);  // QEntry(): Q = defaultQ = 1
    T.newQ.assign((size_t)T.E * kNumXY, 0.f);
    T.cnt.assign((size_t)T.E * kNumXY, 0);
    T.total.assign((size_t)T.E * kNumXY, 0);
    T.ldist.resize((size_t)T.E * L);
    for (int e = 0; e < T.E; ++e)  // initialiseLightSamplingDist
        for (int i = 0; i < L; ++i) T.ldist[(size_t)e * L + i] = ((float)i + 1.f) / (float)L;
    T.lres.assign((size_t)T.E * L, 0.f);
    T.lnew.assign((size_t)T.E * L, 0.f);
    T.lcount.assign((size_t)T.E * L, 0);
    T.lnewc.assign((size_t)T.E * L, 0);
    const float cell_sa = (float)(4.f * kPi / (float)kNumXY);

    const int64_t nsamp = (int64_t)width * height * spp;
    std::vector<float> res((size_t)nsamp * 3, 0.f), sx((size_t)nsamp), sy((size_t)nsamp);
    int64_t traced = 0;
    // spp rounds in the reference's sppBySpp order (Core/Sampler.cu:29-51,
    // 77-96): round slot j is sample done + j % round of pixel j / round; the
    // Q table and light statistics carry over from round to round.
    const int round_spp = spp_per_round > 0 && spp_per_round < spp ? spp_per_round : spp;
    for (int done = 0; done < spp; done += round_spp) {
        const int this_spp = std::min(round_spp, spp - done);
        std::vector<Task> q;
        q.reserve((size_t)width * height * this_spp);
        for (int64_t j = 0; j < (int64_t)width * height * this_spp; ++j) {
            const int64_t pix = j / this_spp;
            const int64_t si = pix * spp + done + j % this_spp;
            const int x = (int)(pix % width), y = (int)(pix / width);
            int dim = 0;
            float fx = (float)x + H.rand1((uint64_t)si, dim);
            float fy = (float)y + H.rand1((uint64_t)si, dim);
            sx[si] = fx;
            sy[si] = fy;
            f3 pl = mk(fx - width / 2.f, height / 2.f - fy, cam_depth);
            f3 ro = apply(cam16, mk(0, 0, 0));
            f3 rd = normalize(apply(cam16, pl) - ro);
            q.push_back(Task{(uint32_t)si, ro, rd, mk(1, 1, 1), 1.f, true, -1, -1, dim});
        }
        struct Hit { Task t; int prim; float dist, hu, hv; };
        for (int depth = 0; depth < max_depth && !q.empty(); ++depth) {
            std::vector<Hit> hits;
            hits.reserve(q.size());
            for (const Task& t : q) {
                float d, hu, hv;
                ++traced;
                int p = closest(s, t.ro, t.rd, d, &hu, &hv);
                if (p >= 0) {
                    hits.push_back(Hit{t, p, d, hu, hv});
                } else if (s.env_miss >= 0) {  // intersectScene's miss (RL .cu:292-296)
                    f3 a = mulv(env_eval(s, s.env_miss, t.rd), t.mult);
                    float* acc = &res[(size_t)t.sid * 3];
                    acc[0] += a.x; acc[1] += a.y; acc[2] += a.z;
                }
            }
            q.clear();
            if (hits.empty()) break;
            int mx = 0;
            for (auto& h : hits) mx = std::max(mx, h.t.dim);
            for (auto& h : hits) h.t.dim = mx;  // syncDimension

            struct LRes { f3 contrib; f3 dir; float emis, sampled, lpdf; };
            struct QRes { float sum, pdfq; };
            struct NInfo { f3 dir; float pdf; int cell; bool valid; };
            std::vector<LRes> lr(hits.size());
            std::vector<QRes> qr(hits.size());
            std::vector<NInfo> ni(hits.size());
            // computeLighting (RL .cu:388-460)
            for (size_t k = 0; k < hits.size(); ++k) {
                Hit& h = hits[k];
                Task& t = h.t;
                const int mesh = tri_mesh[h.prim];
                const cav_mesh_ref& M = meshes[mesh];
                f3 n = shading_normal(s, h.prim, h.hu, h.hv);
                f3 pos = t.ro + h.dist * t.rd;
                Bsdf bsdf = surface_bsdf(s, mesh, h.prim, h.hu, h.hv);
                f3 t0, t1;
                coordinate_system(n, t0, t1);
                float* acc = &res[(size_t)t.sid * 3];
                f3 emis = mk(0, 0, 0);
                if (M.is_light) {
                    f3 Le = dot(t.rd, n) >= 0 ? mk(0, 0, 0) : mk(M.emission[0], M.emission[1], M.emission[2]);
                    if (t.delta) {
                        f3 a = mulv(Le, t.mult);
                        acc[0] += a.x; acc[1] += a.y; acc[2] += a.z;
                    } else if (depth > 0) {
                        float lp = light_pdf(s, mesh, t.ro, pos, n);
                        float w = power_heuristic(t.pdf, lp);
                        if (std::isfinite(w)) {
                            emis = Le;
                            f3 a = mulv(Le, t.mult) * w;
                            acc[0] += a.x; acc[1] += a.y; acc[2] += a.z;
                        }
                    }
                }
                const int entry = find_entry(scene6, T.g, pos);
                const float* cdf = &T.ldist[(size_t)entry * L];
                const int li = dist_index(cdf, L, H.rand1(t.sid, t.dim));
                const float sel = dist_pdf(cdf, li);
                t.dim += 4;  // rand4
                const LightPick Lp = sample_light(s, H, t.sid, t.dim, pos, s.lights[li]);
                float lpdf = Lp.prob;
                const f3 dir = Lp.dir;
                f3 incident = Lp.incident;
                ++traced;
                if (!(visible(s, pos, dir, Lp.lim, mesh, Lp.dst) && std::isfinite(lpdf))) {
                    incident = mk(0, 0, 0);
                    lpdf = 1;
                }
                f3 wo = t.rd * -1;
                f3 il = to_local(dir, t0, t1, n), ol = to_local(wo, t0, t1, n);
                f3 contrib = mulv(incident, bsdf.eval(il, ol)) * std::fabs(dot(dir, n));
                contrib = contrib * (1.f / (lpdf * sel));
                T.lnew[(size_t)entry * L + li] += luminance(incident);
                T.lnewc[(size_t)entry * L + li] += 1;
                lr[k] = LRes{contrib, dir, luminance(emis), luminance(incident), lpdf};
            }
            // computeQDistributions (RL .cu:517-606)
            bool any_valid = false;
            for (size_t k = 0; k < hits.size(); ++k) {
                Hit& h = hits[k];
                Task& t = h.t;
                const int mesh = tri_mesh[h.prim];
                f3 n = shading_normal(s, h.prim, h.hu, h.hv);
                f3 pos = t.ro + h.dist * t.rd;
                Bsdf bsdf = surface_bsdf(s, mesh, h.prim, h.hu, h.hv);
                f3 t0, t1;
                coordinate_system(n, t0, t1);
                const int entry = find_entry(scene6, T.g, pos);
                float* Qe = &T.Q[(size_t)entry * kNumXY];
                f3 ol = to_local(t.rd * -1.f, t0, t1, n);
                float px = H.rand1(t.sid, t.dim), py = H.rand1(t.sid, t.dim);
                float cdf[kNumXY];
                float sum = 0.f;
                for (int c = 0; c < kNumXY; ++c) {
                    f3 d = cell_dir(px, py, c);
                    float sc = luminance(bsdf.eval(to_local(d, t0, t1, n), ol));
                    float qv = std::fabs(dot(d, n)) * sc * Qe[c];
                    sum += qv;
                    cdf[c] = sum;
                }
                for (int c = 0; c < kNumXY; ++c) cdf[c] /= sum;
                sum *= cell_sa;
                qr[k].sum = sum;
                qr[k].pdfq = 0.f;
                if (t.prev_entry != -1) {
                    const size_t idx = (size_t)t.prev_entry * kNumXY + t.prev_cell;
                    T.newQ[idx] += sum + lr[k].emis;
                    T.cnt[idx] += 1;
                }
                const int lc = dir_to_cell(lr[k].dir);
                if (sum != 0) qr[k].pdfq = dist_pdf(cdf, lc) / cell_sa;
                if (lr[k].sampled > 0) {
                    float sa = 1.f / lr[k].lpdf;
                    float ratio = fminr(1.f, 0.5f * sa / cell_sa);
                    float prop = lr[k].sampled * ratio;
                    if (prop > Qe[lc]) {
                        T.newQ[(size_t)entry * kNumXY + lc] += prop;
                        T.cnt[(size_t)entry * kNumXY + lc] += 1;
                    }
                }
                if (sum != 0) {
                    int cell = dist_index(cdf, kNumXY, H.rand1(t.sid, t.dim));
                    float p = dist_pdf(cdf, cell);
                    p = (float)(kNumXY * p / (4 * kPi));
                    ni[k] = NInfo{cell_dir(px, py, cell), p, cell, true};
                    any_valid = true;
                } else {
                    ni[k].valid = false;
                }
            }
            // accumulateLightingContribs (RL .cu:609-647)
            for (size_t k = 0; k < hits.size(); ++k) {
                Hit& h = hits[k];
                Task& t = h.t;
                const int mesh = tri_mesh[h.prim];
                f3 n = shading_normal(s, h.prim, h.hu, h.hv);
                Bsdf bsdf = surface_bsdf(s, mesh, h.prim, h.hu, h.hv);
                f3 t0, t1;
                coordinate_system(n, t0, t1);
                f3 il = to_local(lr[k].dir, t0, t1, n), ol = to_local(t.rd * -1.f, t0, t1, n);
                float mp = bsdf.pdf(il, ol);
                float sp = 0;
                if (qr[k].sum != 0) sp = qr[k].pdfq;
                if (qr[k].sum == 0 || sp == 0 || bsdf.is_delta() || bsdf.is_almost_delta()) sp = mp;
                float w = power_heuristic(lr[k].lpdf, sp);
                f3 a = mulv(lr[k].contrib, t.mult) * w;
                float* acc = &res[(size_t)t.sid * 3];
                acc[0] += a.x; acc[1] += a.y; acc[2] += a.z;
            }
            ref_rl_finish_q(T.Q.data(), T.newQ.data(), T.cnt.data(), T.total.data(), T.E, alpha_scale);
            ref_rl_finish_light(T.ldist.data(), T.lres.data(), T.lcount.data(), T.lnew.data(), T.lnewc.data(), T.E, L);
            (void)any_valid;
            mx = 0;
            for (auto& h : hits) mx = std::max(mx, h.t.dim);
            for (auto& h : hits) h.t.dim = mx;  // syncDimension
            // genNextRay (RL .cu:305-375)
            for (size_t k = 0; k < hits.size(); ++k) {
                Hit& h = hits[k];
                Task t = h.t;
                if (depth > 3) {
                    float qq = 1;
                    qq = fminr(qq, 1 - t.mult.x);
                    qq = fminr(qq, 1 - t.mult.y);
                    qq = fminr(qq, 1 - t.mult.z);
                    qq = fmaxr(qq, 0.05f);
                    if (H.rand1(t.sid, t.dim) < qq) continue;
                    t.mult = t.mult / (1.f - qq);
                }
                const int mesh = tri_mesh[h.prim];
                f3 n = shading_normal(s, h.prim, h.hu, h.hv);
                f3 pos = t.ro + h.dist * t.rd;
                Bsdf bsdf = surface_bsdf(s, mesh, h.prim, h.hu, h.hv);
                f3 t0, t1;
                coordinate_system(n, t0, t1);
                f3 wo = t.rd * -1.f;
                const int entry = find_entry(scene6, T.g, pos);
                f3 f, nd;
                float npdf;
                int cell;
                if (bsdf.is_delta() || bsdf.is_almost_delta() || !ni[k].valid) {
                    float ux = H.rand1(t.sid, t.dim), uy = H.rand1(t.sid, t.dim);
                    f3 wl;
                    npdf = 0;
                    f = bsdf.sample(ux, uy, wl, to_local(wo, t0, t1, n), npdf);
                    nd = to_world(wl, t0, t1, n);
                    cell = dir_to_cell(nd);
                } else {
                    cell = ni[k].cell;
                    npdf = ni[k].pdf;
                    nd = ni[k].dir;
                    f = bsdf.eval(to_local(nd, t0, t1, n), to_local(wo, t0, t1, n));
                }
                f3 no = pos + nd * 0.001f;
                if (all_zero(f)) t.mult = mk(0, 0, 0);
                else t.mult = mulv(t.mult, f) * std::fabs(dot(nd, n)) / npdf;
                t.ro = no;
                t.rd = nd;
                t.pdf = npdf;
                t.delta = bsdf.is_delta();
                t.prev_entry = entry;
                t.prev_cell = cell;
                q.push_back(t);
            }
        }
    }  // rounds

    gather_film(width, height, spp, res, sx, sy, film_rgb, film_w, 0, 0, width, height);
    if (rays_traced) *rays_traced = traced;
    return 0;
}

extern "C" int ref_rl_dir_to_cell(const float* dir3) { return dir_to_cell(mk(dir3[0], dir3[1], dir3[2])); }
extern "C" void ref_rl_cell_dir(float rx, float ry, int cell, float* out3) {
    f3 d = cell_dir(rx, ry, cell);
    out3[0] = d.x; out3[1] = d.y; out3[2] = d.z;
}

extern "C" void ref_tex_fetch(const uint8_t* rgba, int w, int h, float u, float v, float* out3) {
    const f3 r = tex_fetch(rgba, w, h, u, v);
    out3[0] = r.x;
    out3[1] = r.y;
    out3[2] = r.z;
}

// BSDF of one mesh in its local frame (z = shading normal), for unit checks:
// f(wi, wo) and pdf(wi, wo) of n direction pairs.
extern "C" void ref_bsdf_eval(const cav_mesh_ref* m, const float* wi, const float* wo, int n, float* f_out,
                              float* pdf_out) {
    const Bsdf b = make_bsdf(*m);
    for (int k = 0; k < n; ++k) {
        const f3 i = mk(wi[3 * k], wi[3 * k + 1], wi[3 * k + 2]), o = mk(wo[3 * k], wo[3 * k + 1], wo[3 * k + 2]);
        const f3 f = b.eval(i, o);
        f_out[3 * k] = f.x;
        f_out[3 * k + 1] = f.y;
        f_out[3 * k + 2] = f.z;
        pdf_out[k] = b.pdf(i, o);
    }
}

// BSDF::sample for n (u, wo) pairs: sampled wi, f and pdf.
extern "C" void ref_bsdf_sample(const cav_mesh_ref* m, const float* u2, const float* wo, int n, float* wi_out,
                                float* f_out, float* pdf_out) {
    const Bsdf b = make_bsdf(*m);
    for (int k = 0; k < n; ++k) {
        const f3 o = mk(wo[3 * k], wo[3 * k + 1], wo[3 * k + 2]);
        f3 i = mk(0, 0, 0);
        float p = 0;
        const f3 f = b.sample(u2[2 * k], u2[2 * k + 1], i, o, p);
        wi_out[3 * k] = i.x;
        wi_out[3 * k + 1] = i.y;
        wi_out[3 * k + 2] = i.z;
        f_out[3 * k] = f.x;
        f_out[3 * k + 1] = f.y;
        f_out[3 * k + 2] = f.z;
        pdf_out[k] = p;
    }
}

// EnvironmentMap::evaluateRay of environment light e of surf for n world directions.
extern "C" void ref_env_eval(const ref_surface* surf, int e, const float* dirs, int n, float* out) {
    Scene s{};
    s.surf = surf;
    s.env = surf->env;
    for (int k = 0; k < n; ++k) {
        const f3 r = env_eval(s, e, mk(dirs[3 * k], dirs[3 * k + 1], dirs[3 * k + 2]));
        out[3 * k] = r.x;
        out[3 * k + 1] = r.y;
        out[3 * k + 2] = r.z;
    }
}
