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

// General 4x4 inverse by cofactors; false when singular.
inline bool invert(const float* m, float* out) {
    float inv[16];
    inv[0] = m[5] * m[10] * m[15] - m[5] * m[11] * m[14] - m[9] * m[6] * m[15] + m[9] * m[7] * m[14] + m[13] * m[6] * m[11] - m[13] * m[7] * m[10];
    inv[4] = -m[4] * m[10] * m[15] + m[4] * m[11] * m[14] + m[8] * m[6] * m[15] - m[8] * m[7] * m[14] - m[12] * m[6] * m[11] + m[12] * m[7] * m[10];
    inv[8] = m[4] * m[9] * m[15] - m[4] * m[11] * m[13] - m[8] * m[5] * m[15] + m[8] * m[7] * m[13] + m[12] * m[5] * m[11] - m[12] * m[7] * m[9];
    inv[12] = -m[4] * m[9] * m[14] + m[4] * m[10] * m[13] + m[8] * m[5] * m[14] - m[8] * m[6] * m[13] - m[12] * m[5] * m[10] + m[12] * m[6] * m[9];
    inv[1] = -m[1] * m[10] * m[15] + m[1] * m[11] * m[14] + m[9] * m[2] * m[15] - m[9] * m[3] * m[14] - m[13] * m[2] * m[11] + m[13] * m[3] * m[10];
    inv[5] = m[0] * m[10] * m[15] - m[0] * m[11] * m[14] - m[8] * m[2] * m[15] + m[8] * m[3] * m[14] + m[12] * m[2] * m[11] - m[12] * m[3] * m[10];
    inv[9] = -m[0] * m[9] * m[15] + m[0] * m[11] * m[13] + m[8] * m[1] * m[15] - m[8] * m[3] * m[13] - m[12] * m[1] * m[11] + m[12] * m[3] * m[9];
    inv[13] = m[0] * m[9] * m[14] - m[0] * m[10] * m[13] - m[8] * m[1] * m[14] + m[8] * m[2] * m[13] + m[12] * m[1] * m[10] - m[12] * m[2] * m[9];
    inv[2] = m[1] * m[6] * m[15] - m[1] * m[7] * m[14] - m[5] * m[2] * m[15] + m[5] * m[3] * m[14] + m[13] * m[2] * m[7] - m[13] * m[3] * m[6];
    inv[6] = -m[0] * m[6] * m[15] + m[0] * m[7] * m[14] + m[4] * m[2] * m[15] - m[4] * m[3] * m[14] - m[12] * m[2] * m[7] + m[12] * m[3] * m[6];
    inv[10] = m[0] * m[5] * m[15] - m[0] * m[7] * m[13] - m[4] * m[1] * m[15] + m[4] * m[3] * m[13] + m[12] * m[1] * m[7] - m[12] * m[3] * m[5];
    inv[14] = -m[0] * m[5] * m[14] + m[0] * m[6] * m[13] + m[4] * m[1] * m[14] - m[4] * m[2] * m[13] - m[12] * m[1] * m[6] + m[12] * m[2] * m[5];
    inv[3] = -m[1] * m[6] * m[11] + m[1] * m[7] * m[10] + m[5] * m[2] * m[11] - m[5] * m[3] * m[10] - m[9] * m[2] * m[7] + m[9] * m[3] * m[6];
    inv[7] = m[0] * m[6] * m[11] - m[0] * m[7] * m[10] - m[4] * m[2] * m[11] + m[4] * m[3] * m[10] + m[8] * m[2] * m[7] - m[8] * m[3] * m[6];
    inv[11] = -m[0] * m[5] * m[11] + m[0] * m[7] * m[9] + m[4] * m[1] * m[11] - m[4] * m[3] * m[9] - m[8] * m[1] * m[7] + m[8] * m[3] * m[5];
    inv[15] = m[0] * m[5] * m[10] - m[0] * m[6] * m[9] - m[4] * m[1] * m[10] + m[4] * m[2] * m[9] + m[8] * m[1] * m[6] - m[8] * m[2] * m[5];
    const float det = m[0] * inv[0] + m[1] * inv[4] + m[2] * inv[8] + m[3] * inv[12];
    if (det == 0 || !std::isfinite(det)) return false;
    const float r = 1.0f / det;
    for (int i = 0; i < 16; ++i) out[i] = inv[i] * r;
    return true;
}

// transpose(inverse(upper 3x3 of m)) (getTransformForNormal,
// Utils/MathsCommons.h:80-82), as a row-major 3x3 r[row][col]; false when
// singular
inline bool normal_matrix(const M4& m, float r[3][3]) {
    // a[row][col] of the upper 3x3
    float a[3][3];
    for (int row = 0; row < 3; ++row)
        for (int col = 0; col < 3; ++col) a[row][col] = m.at(col, row);
    const float c00 = a[1][1] * a[2][2] - a[1][2] * a[2][1];
    const float c01 = a[1][2] * a[2][0] - a[1][0] * a[2][2];
    const float c02 = a[1][0] * a[2][1] - a[1][1] * a[2][0];
    const float det = a[0][0] * c00 + a[0][1] * c01 + a[0][2] * c02;
    if (det == 0 || !std::isfinite(det)) return false;
    const float inv = 1.0f / det;
    // inverse(a)[i][j] = cofactor[j][i] / det, so its transpose is cofactor[i][j] / det
    r[0][0] = c00 * inv;
    r[0][1] = c01 * inv;
    r[0][2] = c02 * inv;
    r[1][0] = (a[0][2] * a[2][1] - a[0][1] * a[2][2]) * inv;
    r[1][1] = (a[0][0] * a[2][2] - a[0][2] * a[2][0]) * inv;
    r[1][2] = (a[0][1] * a[2][0] - a[0][0] * a[2][1]) * inv;
    r[2][0] = (a[0][1] * a[1][2] - a[0][2] * a[1][1]) * inv;
    r[2][1] = (a[0][2] * a[1][0] - a[0][0] * a[1][2]) * inv;
    r[2][2] = (a[0][0] * a[1][1] - a[0][1] * a[1][0]) * inv;
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
