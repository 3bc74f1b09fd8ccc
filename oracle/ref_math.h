// CPU oracle: scalar float semantics shared by every oracle routine.
//
// TEST INFRASTRUCTURE ONLY. Nothing under oracle/ is part of the product; only
// tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it.
//
// Every routine is compiled with -ffp-contract=off: each float expression is
// evaluated as written, one IEEE-754 binary32 (or binary64) operation at a time,
// round-to-nearest-even, with correctly rounded '/' and sqrtf. The HIP kernels
// are built with the same rule, so LBVH and traversal results can be compared
// bit for bit.
#pragma once
#include <cmath>
#include <cstdint>
#include <cstring>

namespace ref {

struct f3 { float x, y, z; };

static inline f3 mk(float x, float y, float z) { return f3{x, y, z}; }
static inline f3 operator+(f3 a, f3 b) { return mk(a.x + b.x, a.y + b.y, a.z + b.z); }
static inline f3 operator-(f3 a, f3 b) { return mk(a.x - b.x, a.y - b.y, a.z - b.z); }
static inline f3 operator*(f3 a, float s) { return mk(a.x * s, a.y * s, a.z * s); }
static inline f3 operator*(float s, f3 a) { return mk(s * a.x, s * a.y, s * a.z); }
static inline f3 operator/(f3 a, float s) { return mk(a.x / s, a.y / s, a.z / s); }
static inline f3 mul(f3 a, f3 b) { return mk(a.x * b.x, a.y * b.y, a.z * b.z); }
static inline f3 divv(f3 a, f3 b) { return mk(a.x / b.x, a.y / b.y, a.z / b.z); }
// dot and cross in the operand order of the reference's float3 helpers
static inline float dot(f3 a, f3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
static inline f3 cross(f3 a, f3 b) {
    return mk(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}

// fminf / fmaxf as the gfx950 v_min_f32 / v_max_f32 instructions compute them
// (IEEE-754-2019 minimumNumber / maximumNumber): a NaN operand is ignored and
// -0 orders below +0. With that total order min/max are associative and
// commutative, so box unions and reductions are order independent.
static inline float fminr(float a, float b) {
    if (std::isnan(a)) return b;
    if (std::isnan(b)) return a;
    if (a < b) return a;
    if (b < a) return b;
    return std::signbit(a) ? a : b;
}
static inline float fmaxr(float a, float b) {
    if (std::isnan(a)) return b;
    if (std::isnan(b)) return a;
    if (a > b) return a;
    if (b > a) return b;
    return std::signbit(a) ? b : a;
}

static inline int32_t fbits(float f) { int32_t i; std::memcpy(&i, &f, 4); return i; }
static inline float bitsf(int32_t i) { float f; std::memcpy(&f, &i, 4); return f; }

// Utils/MathsCommons.h:93-100: min/max of float *bit patterns* compared as
// signed 32-bit integers (the reference's PTX vmin/vmax helpers).
// fmin_fmax(a,b,c) = imax(imin(a,b),c); fmax_fmin(a,b,c) = imin(imax(a,b),c).
static inline float fmin_fmax(float a, float b, float c) {
    int32_t ia = fbits(a), ib = fbits(b), ic = fbits(c);
    int32_t m = ia < ib ? ia : ib;
    return bitsf(m > ic ? m : ic);
}
static inline float fmax_fmin(float a, float b, float c) {
    int32_t ia = fbits(a), ib = fbits(b), ic = fbits(c);
    int32_t m = ia > ib ? ia : ib;
    return bitsf(m < ic ? m : ic);
}

struct Box { f3 lo, hi; };

// Core/BoundingBox.h:13-21
static inline f3 centroid(const Box& b) { return (b.lo + b.hi) * 0.5f; }
static inline f3 extent(const Box& b) { return b.hi - b.lo; }

// Core/BoundingBox.h:77-80: 2*(x*y + y*z + x*z), summed left to right.
static inline float surface_area(const Box& b) {
    f3 e = b.hi - b.lo;
    return 2.f * (e.x * e.y + e.y * e.z + e.x * e.z);
}

// Core/BoundingBox.h:84-96
static inline Box union_box(const Box& a, const Box& b) {
    return Box{mk(fminr(a.lo.x, b.lo.x), fminr(a.lo.y, b.lo.y), fminr(a.lo.z, b.lo.z)),
               mk(fmaxr(a.hi.x, b.hi.x), fmaxr(a.hi.y, b.hi.y), fmaxr(a.hi.z, b.hi.z))};
}

// BVH/SAH.h:4-22. The cost constants are double literals, so the arithmetic is
// carried out in binary64 and rounded to binary32 on return.
static inline float leaf_cost(float area, int nprims) {
    return (float)((double)area * 1.0 * (double)nprims);
}
static inline float internal_cost3(float area, float cl, float cr) {
    return (float)((double)area * 1.2 + (double)cl + (double)cr);
}
static inline float internal_cost2(float area, float children) {
    return (float)((double)area * 1.2 + (double)children);
}

// Core/Triangle.h:242-268: per-triangle box, fminf(v0, fminf(v1, v2)).
static inline Box tri_box(f3 v0, f3 v1, f3 v2) {
    return Box{mk(fminr(v0.x, fminr(v1.x, v2.x)), fminr(v0.y, fminr(v1.y, v2.y)),
                  fminr(v0.z, fminr(v1.z, v2.z))),
               mk(fmaxr(v0.x, fmaxr(v1.x, v2.x)), fmaxr(v0.y, fmaxr(v1.y, v2.y)),
                  fmaxr(v0.z, fmaxr(v1.z, v2.z)))};
}

}  // namespace ref
