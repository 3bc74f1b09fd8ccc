// Perspective camera set-up on the host: LookAt + fov -> camera-to-world
// matrix and film depth.
//
// Reference: the LookAt handling of SceneLoading/Parsing.cu:196-204 (glm
// lookAtLH, then the camera gets the inverse of the accumulated transform,
// Parsing.cu:257) and PerspectiveCamera's constructor
// (Cameras/PerspectiveCamera.cu:6-20). Matrices are column-major like glm:
// element (row r, column c) is m[c * 4 + r].
#include <cmath>

#include "../../include/cavalry.h"

namespace {

struct V { float x, y, z; };
V sub(V a, V b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
float dot(V a, V b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
V cross(V a, V b) { return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
V normalize(V v) {
    const float inv = 1.0f / std::sqrt(dot(v, v));
    return {v.x * inv, v.y * inv, v.z * inv};
}

// General 4x4 inverse by cofactors (column-major in and out).
bool invert4(const float* m, float* out) {
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

}  // namespace

extern "C" int cav_camera_look_at(const float* eye3, const float* center3, const float* up3,
                                  float fov_degrees, int width, int height, cav_camera* out) {
    if (!eye3 || !center3 || !up3 || !out || width <= 0 || height <= 0) return CAV_ERR_INVALID;
    const V eye{eye3[0], eye3[1], eye3[2]}, center{center3[0], center3[1], center3[2]};
    const V up{up3[0], up3[1], up3[2]};
    // glm::lookAtLH: rows s, u, f and the translation -dot(., eye)
    const V f = normalize(sub(center, eye));
    const V s = normalize(cross(up, f));
    const V u = cross(f, s);
    float view[16] = {s.x, u.x, f.x, 0, s.y, u.y, f.y, 0, s.z, u.z, f.z, 0,
                      -dot(s, eye), -dot(u, eye), -dot(f, eye), 1};
    if (!invert4(view, out->cam_to_world)) return CAV_ERR_INVALID;
    // glm::radians and PerspectiveCamera's film depth for unit-size pixels
    const float fov = fov_degrees * 0.01745329251994329576923690768489f;
    const float half = (float)(height <= width ? height : width) / 2.f;
    out->depth = half / std::tan(fov / 2.f);
    out->width = width;
    out->height = height;
    return CAV_OK;
}
