// Device-side scalar helpers shared by the HIP kernels.
//
// All kernels are compiled with -ffp-contract=off: every float expression is
// evaluated one IEEE-754 operation at a time in the written order, so results
// match the C++ restatement of the reference bit for bit.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace cav {

struct v3 { float x, y, z; };

__device__ __forceinline__ v3 mk3(float x, float y, float z) { return v3{x, y, z}; }
__device__ __forceinline__ v3 add3(v3 a, v3 b) { return mk3(a.x + b.x, a.y + b.y, a.z + b.z); }
__device__ __forceinline__ v3 sub3(v3 a, v3 b) { return mk3(a.x - b.x, a.y - b.y, a.z - b.z); }
__device__ __forceinline__ v3 scale3(v3 a, float s) { return mk3(a.x * s, a.y * s, a.z * s); }
__device__ __forceinline__ v3 mul3(v3 a, v3 b) { return mk3(a.x * b.x, a.y * b.y, a.z * b.z); }
__device__ __forceinline__ v3 div3(v3 a, v3 b) { return mk3(a.x / b.x, a.y / b.y, a.z / b.z); }
// dot/cross in the operand order of the reference's float3 helpers
__device__ __forceinline__ float dot3(v3 a, v3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ __forceinline__ v3 cross3(v3 a, v3 b) {
    return mk3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}

// v_min_f32 / v_max_f32: NaN operands ignored, -0 below +0.
__device__ __forceinline__ float fmn(float a, float b) { return __builtin_fminf(a, b); }
__device__ __forceinline__ float fmx(float a, float b) { return __builtin_fmaxf(a, b); }

// Utils/MathsCommons.h:93-100: min/max on float bit patterns compared as
// signed integers. imax(imin(a,b),c) and imin(imax(a,b),c).
__device__ __forceinline__ float fmin_fmax(float a, float b, float c) {
    int ia = __float_as_int(a), ib = __float_as_int(b), ic = __float_as_int(c);
    return __int_as_float(max(min(ia, ib), ic));
}
__device__ __forceinline__ float fmax_fmin(float a, float b, float c) {
    int ia = __float_as_int(a), ib = __float_as_int(b), ic = __float_as_int(c);
    return __int_as_float(min(max(ia, ib), ic));
}

// Core/BoundingBox.h:77-80
__device__ __forceinline__ float surface_area(v3 lo, v3 hi) {
    v3 e = sub3(hi, lo);
    return 2.f * (e.x * e.y + e.y * e.z + e.x * e.z);
}

// BVH/SAH.h:4-22: double-precision cost arithmetic, rounded to float.
__device__ __forceinline__ float leaf_cost(float area) { return (float)((double)area * 1.0 * 1.0); }
__device__ __forceinline__ float internal_cost3(float area, float cl, float cr) {
    return (float)((double)area * 1.2 + (double)cl + (double)cr);
}
__device__ __forceinline__ float internal_cost2(float area, float children) {
    return (float)((double)area * 1.2 + (double)children);
}

// Order-preserving float -> u32 map (-0 orders below +0) used for min/max
// reductions through integer atomics.
__device__ __forceinline__ uint32_t f2ord(float f) {
    uint32_t u = __float_as_uint(f);
    return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}
__device__ __forceinline__ float ord2f(uint32_t u) {
    return __uint_as_float((u & 0x80000000u) ? (u & 0x7fffffffu) : ~u);
}

// Build-layout node, two 16-byte halves in the unified 2N-1 id space:
//   box  = {lo.x, lo.y, lo.z, sa}, {hi.x, hi.y, hi.z, cost}
//   topo = {left, right, parent, prim}  (prim = original triangle index for
//          leaves, -1 for internal nodes; left/right = -1 for leaves)
struct NodeBox { float4 a, b; };

// Traversal-layout internal node (64 bytes): both child boxes and the
// children's refs, ref >= 0: an internal node's slot, ref < 0: ~(a leaf's slot).
struct TNode {
    float4 l0;  // left lo.xyz, left hi.x
    float4 l1;  // left hi.yz, right lo.xy
    float4 r1;  // right lo.z, right hi.xyz
    int4 refs;  // left ref, right ref, 0, 0 (ref = slot, or ~slot for a leaf)
};

// A leaf's triangle: v0 | prim, e1 | mesh, e2 | 0 (the first 48 bytes of a
// leaf's record; the last 16 are zero).
struct TTri { float4 a, b, c; };

// Traversal records, 64 bytes each (a TNode or a padded TTri), in the
// sibling-pair layout: the root at slot 0, the children of the internal node
// of rank p at slots 2p + 2 (left) and 2p + 3 (right). A pair of siblings
// fills one 128-byte L2 line, and every record sits in one line.
struct TRec { float4 q0, q1, q2, q3; };

}  // namespace cav
