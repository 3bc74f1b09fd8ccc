// Host 4x4 transforms with glm's conventions, for the scene front end and
// the camera: column-major storage (element row r, column c at m[c * 4 + r]),
// transforms compose on the right (translate(m, v) = m * T(v)) and points are
// mapped with the homogeneous divide.
//
// Reference: the glm calls of SceneLoading/Parsing.cu:119-169 (translate,
// rotate, scale, a raw "Transform" pre-multiplied), :196-204 (lookAtLH) and
// Utils/MathsCommons.h:74-82 (apply, getTransformForNormal).
#pragma once
#include <cmath>

namespace cav {
namespace m4 {

struct V3 { float x, y, z; };

inline V3 sub(V3 a, V3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline float dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline V3 cross(V3 a, V3 b) { return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
inline V3 normalize(V3 v) {
    const float inv = 1.0f / std::sqrt(dot(v, v));
    return {v.x * inv, v.y * inv, v.z * inv};
}

struct M4 {
    float m[16];
    float& at(int col, int row) { return m[col * 4 + row]; }
    float at(int col, int row) const { return m[col * 4 + row]; }
};

inline M4 identity() {
    M4 r{};
    for (int i = 0; i < 4; ++i) r.at(i, i) = 1.f;
    return r;
}

// a * b: column j of the product is a's columns weighted by b's column j,
// summed in column order
inline M4 mul(const M4& a, const M4& b) {
    M4 r{};
    for (int j = 0; j < 4; ++j)
        for (int k = 0; k < 4; ++k)
            r.at(j, k) = ((a.at(0, k) * b.at(j, 0) + a.at(1, k) * b.at(j, 1)) + a.at(2, k) * b.at(j, 2)) +
                         a.at(3, k) * b.at(j, 3);
    return r;
}

// m * (v, 1) with the homogeneous divide; the columns are summed as
// (c0 x + c1 y) + (c2 z + c3)
inline V3 apply(const M4& m, V3 v) {
    float r[4];
    for (int k = 0; k < 4; ++k)
        r[k] = (m.at(0, k) * v.x + m.at(1, k) * v.y) + (m.at(2, k) * v.z + m.at(3, k) * 1.f);
    return {r[0] / r[3], r[1] / r[3], r[2] / r[3]};
}

// the upper 3x3 block times v (directions)
inline V3 apply3(const M4& m, V3 v) {
    float r[3];
    for (int k = 0; k < 3; ++k) r[k] = (m.at(0, k) * v.x + m.at(1, k) * v.y) + m.at(2, k) * v.z;
    return {r[0], r[1], r[2]};
}

inline M4 translate(const M4& m, V3 v) {
    M4 r = m;
    for (int k = 0; k < 4; ++k)
        r.at(3, k) = ((m.at(0, k) * v.x + m.at(1, k) * v.y) + m.at(2, k) * v.z) + m.at(3, k);
    return r;
}

inline M4 scale(const M4& m, V3 v) {
    M4 r = m;
    for (int k = 0; k < 4; ++k) {
        r.at(0, k) = m.at(0, k) * v.x;
        r.at(1, k) = m.at(1, k) * v.y;
        r.at(2, k) = m.at(2, k) * v.z;
    }
    return r;
}

// rotation by `radians` about `axis` (Rodrigues), composed on the right
inline M4 rotate(const M4& m, float radians, V3 axis) {
    const float c = std::cos(radians), s = std::sin(radians);
    const V3 a = normalize(axis);
    const V3 t{(1.f - c) * a.x, (1.f - c) * a.y, (1.f - c) * a.z};
    float R[3][3];  // R[col][row]
    R[0][0] = c + t.x * a.x;
    R[0][1] = t.x * a.y + s * a.z;
    R[0][2] = t.x * a.z - s * a.y;
    R[1][0] = t.y * a.x - s * a.z;
    R[1][1] = c + t.y * a.y;
    R[1][2] = t.y * a.z + s * a.x;
    R[2][0] = t.z * a.x + s * a.y;
    R[2][1] = t.z * a.y - s * a.x;
    R[2][2] = c + t.z * a.z;
    M4 r = m;
    for (int j = 0; j < 3; ++j)
        for (int k = 0; k < 4; ++k)
            r.at(j, k) = (m.at(0, k) * R[j][0] + m.at(1, k) * R[j][1]) + m.at(2, k) * R[j][2];
    return r;
}

// left-handed view matrix: rows s, u, f and the translation -dot(., eye)
inline M4 look_at_lh(V3 eye, V3 center, V3 up) {
    const V3 f = normalize(sub(center, eye));
    const V3 s = normalize(cross(up, f));
    const V3 u = cross(f, s);
    M4 r = identity();
    r.at(0, 0) = s.x; r.at(1, 0) = s.y; r.at(2, 0) = s.z;
    r.at(0, 1) = u.x; r.at(1, 1) = u.y; r.at(2, 1) = u.z;
    r.at(0, 2) = f.x; r.at(1, 2) = f.y; r.at(2, 2) = f.z;
    r.at(3, 0) = -dot(s, eye);
    r.at(3, 1) = -dot(u, eye);
    r.at(3, 2) = -dot(f, eye);
    return r;
}

// 4x4 inverse in glm::inverse's operation order (glm 0.9.8,
// detail/func_matrix.inl:297-356, which the reference calls at
// SceneLoading/Parsing.cu:257 and Core/Light.h:58): the 2x2 minors of the
// lower rows, the cofactors as (a f0 - b f1) + c f2, and the determinant as
// the first column times the first cofactor row summed pairwise, its
// reciprocal multiplied into every element. m(c, r) = column c, row r.
// False when singular.
inline bool invert(const float* m, float* out) {
    auto at = [&](int c, int r) { return m[c * 4 + r]; };
    const float c00 = at(2, 2) * at(3, 3) - at(3, 2) * at(2, 3);
    const float c02 = at(1, 2) * at(3, 3) - at(3, 2) * at(1, 3);
    const float c03 = at(1, 2) * at(2, 3) - at(2, 2) * at(1, 3);
    const float c04 = at(2, 1) * at(3, 3) - at(3, 1) * at(2, 3);
    const float c06 = at(1, 1) * at(3, 3) - at(3, 1) * at(1, 3);
    const float c07 = at(1, 1) * at(2, 3) - at(2, 1) * at(1, 3);
    const float c08 = at(2, 1) * at(3, 2) - at(3, 1) * at(2, 2);
    const float c10 = at(1, 1) * at(3, 2) - at(3, 1) * at(1, 2);
    const float c11 = at(1, 1) * at(2, 2) - at(2, 1) * at(1, 2);
    const float c12 = at(2, 0) * at(3, 3) - at(3, 0) * at(2, 3);
    const float c14 = at(1, 0) * at(3, 3) - at(3, 0) * at(1, 3);
    const float c15 = at(1, 0) * at(2, 3) - at(2, 0) * at(1, 3);
    const float c16 = at(2, 0) * at(3, 2) - at(3, 0) * at(2, 2);
    const float c18 = at(1, 0) * at(3, 2) - at(3, 0) * at(1, 2);
    const float c19 = at(1, 0) * at(2, 2) - at(2, 0) * at(1, 2);
    const float c20 = at(2, 0) * at(3, 1) - at(3, 0) * at(2, 1);
    const float c22 = at(1, 0) * at(3, 1) - at(3, 0) * at(1, 1);
    const float c23 = at(1, 0) * at(2, 1) - at(2, 0) * at(1, 1);
    const float f0[4] = {c00, c00, c02, c03}, f1[4] = {c04, c04, c06, c07}, f2[4] = {c08, c08, c10, c11};
    const float f3[4] = {c12, c12, c14, c15}, f4[4] = {c16, c16, c18, c19}, f5[4] = {c20, c20, c22, c23};
    const float v0[4] = {at(1, 0), at(0, 0), at(0, 0), at(0, 0)};
    const float v1[4] = {at(1, 1), at(0, 1), at(0, 1), at(0, 1)};
    const float v2[4] = {at(1, 2), at(0, 2), at(0, 2), at(0, 2)};
    const float v3[4] = {at(1, 3), at(0, 3), at(0, 3), at(0, 3)};
    float inv[4][4];  // inv[column][row]
    for (int k = 0; k < 4; ++k) {
        const float sa = (k & 1) ? -1.f : 1.f;  // (+1, -1, +1, -1) on columns 0 and 2
        const float sb = -sa;                   // (-1, +1, -1, +1) on columns 1 and 3
        inv[0][k] = ((v1[k] * f0[k] - v2[k] * f1[k]) + v3[k] * f2[k]) * sa;
        inv[1][k] = ((v0[k] * f0[k] - v2[k] * f3[k]) + v3[k] * f4[k]) * sb;
        inv[2][k] = ((v0[k] * f1[k] - v1[k] * f3[k]) + v3[k] * f5[k]) * sa;
        inv[3][k] = ((v0[k] * f2[k] - v1[k] * f4[k]) + v2[k] * f5[k]) * sb;
    }
    const float d0 = at(0, 0) * inv[0][0], d1 = at(0, 1) * inv[1][0];
    const float d2 = at(0, 2) * inv[2][0], d3 = at(0, 3) * inv[3][0];
    const float det = (d0 + d1) + (d2 + d3);
    if (det == 0 || !std::isfinite(det)) return false;
    const float r = 1.0f / det;
    for (int c = 0; c < 4; ++c)
        for (int k = 0; k < 4; ++k) out[c * 4 + k] = inv[c][k] * r;
    return true;
}

// transpose(inverse(upper 3x3 of m)) (getTransformForNormal,
// Utils/MathsCommons.h:80-82), as a row-major 3x3 r[row][col], with glm's
// 3x3 inverse (detail/func_matrix.inl:272-292): the determinant by the first
// column's cofactors, its reciprocal multiplied into each cofactor. False
// when singular.
inline bool normal_matrix(const M4& m, float r[3][3]) {
    auto g = [&](int c, int w) { return m.at(c, w); };  // glm m[c][w] of the upper 3x3
    const float det = (+g(0, 0) * (g(1, 1) * g(2, 2) - g(2, 1) * g(1, 2)) -
                       g(1, 0) * (g(0, 1) * g(2, 2) - g(2, 1) * g(0, 2))) +
                      g(2, 0) * (g(0, 1) * g(1, 2) - g(1, 1) * g(0, 2));
    if (det == 0 || !std::isfinite(det)) return false;
    const float o = 1.0f / det;
    float inv[3][3];  // inv[column][row], as glm's Inverse[c][w]
    inv[0][0] = +(g(1, 1) * g(2, 2) - g(2, 1) * g(1, 2)) * o;
    inv[1][0] = -(g(1, 0) * g(2, 2) - g(2, 0) * g(1, 2)) * o;
    inv[2][0] = +(g(1, 0) * g(2, 1) - g(2, 0) * g(1, 1)) * o;
    inv[0][1] = -(g(0, 1) * g(2, 2) - g(2, 1) * g(0, 2)) * o;
    inv[1][1] = +(g(0, 0) * g(2, 2) - g(2, 0) * g(0, 2)) * o;
    inv[2][1] = -(g(0, 0) * g(2, 1) - g(2, 0) * g(0, 1)) * o;
    inv[0][2] = +(g(0, 1) * g(1, 2) - g(1, 1) * g(0, 2)) * o;
    inv[1][2] = -(g(0, 0) * g(1, 2) - g(1, 0) * g(0, 2)) * o;
    inv[2][2] = +(g(0, 0) * g(1, 1) - g(1, 0) * g(0, 1)) * o;
    // the transpose, row-major: r[row][col] = transpose[col][row] = inv[row][col]
    for (int row = 0; row < 3; ++row)
        for (int col = 0; col < 3; ++col) r[row][col] = inv[row][col];
    return true;
}

inline V3 mul3(const float r[3][3], V3 v) {
    return {(r[0][0] * v.x + r[0][1] * v.y) + r[0][2] * v.z, (r[1][0] * v.x + r[1][1] * v.y) + r[1][2] * v.z,
            (r[2][0] * v.x + r[2][1] * v.y) + r[2][2] * v.z};
}

// glm::radians for float
inline float radians(float deg) { return deg * 0.01745329251994329576923690768489f; }

}  // namespace m4
}  // namespace cav
