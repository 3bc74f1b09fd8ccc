// Perspective camera set-up on the host: LookAt + fov -> camera-to-world
// matrix and film depth.
//
// Reference: the LookAt handling of SceneLoading/Parsing.cu:196-204 (glm
// lookAtLH, then the camera gets the inverse of the accumulated transform,
// Parsing.cu:257) and PerspectiveCamera's constructor
// (Cameras/PerspectiveCamera.cu:6-20). Matrices are column-major like glm:
// element (row r, column c) is m[c * 4 + r].
#include <cmath>
#include <cstring>

#include "../../include/cavalry.h"
#include "mat4.h"

using namespace cav::m4;

extern "C" int cav_camera_from_matrix(const float* cam_to_world16, float fov_degrees, int width,
                                      int height, cav_camera* out) {
    if (!cam_to_world16 || !out || width <= 0 || height <= 0) return CAV_ERR_INVALID;
    std::memcpy(out->cam_to_world, cam_to_world16, sizeof out->cam_to_world);
    // glm::radians and PerspectiveCamera's film depth for unit-size pixels
    const float fov = radians(fov_degrees);
    const float half = (float)(height <= width ? height : width) / 2.f;
    out->depth = half / std::tan(fov / 2.f);
    out->width = width;
    out->height = height;
    return CAV_OK;
}

extern "C" int cav_camera_look_at(const float* eye3, const float* center3, const float* up3,
                                  float fov_degrees, int width, int height, cav_camera* out) {
    if (!eye3 || !center3 || !up3 || !out || width <= 0 || height <= 0) return CAV_ERR_INVALID;
    const M4 view = look_at_lh({eye3[0], eye3[1], eye3[2]}, {center3[0], center3[1], center3[2]},
                               {up3[0], up3[1], up3[2]});
    float c2w[16];
    if (!invert(view.m, c2w)) return CAV_ERR_INVALID;
    return cav_camera_from_matrix(c2w, fov_degrees, width, height, out);
}
