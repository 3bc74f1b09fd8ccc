// Test infrastructure only (never linked into the library): the reference's
// own camera arithmetic, compiled from the glm 0.9.8 headers the reference
// vendors (Dependencies/include/glm), for golden vectors of the host camera.
//
// For each line "ex ey ez cx cy cz ux uy uz" on stdin it prints the 16 floats
// (column-major, as glm stores them) of inverse(lookAtLH(eye, center, up)),
// the camera-to-world matrix SceneLoading/Parsing.cu:196-204,257 gives the
// camera for a LookAt with no other transform, as 32-bit hex patterns.
#include <cstdint>
#include <cstdio>
#include <cstring>

#include <glm/glm.hpp>
#include <glm/gtc/matrix_transform.hpp>

int main() {
    float v[9];
    while (std::scanf("%f %f %f %f %f %f %f %f %f", &v[0], &v[1], &v[2], &v[3], &v[4], &v[5], &v[6], &v[7],
                      &v[8]) == 9) {
        const glm::mat4 view = glm::lookAtLH(glm::vec3(v[0], v[1], v[2]), glm::vec3(v[3], v[4], v[5]),
                                             glm::vec3(v[6], v[7], v[8]));
        const glm::mat4 c2w = glm::inverse(view);
        for (int c = 0; c < 4; ++c)
            for (int r = 0; r < 4; ++r) {
                uint32_t u;
                std::memcpy(&u, &c2w[c][r], 4);
                std::printf("%08x%c", u, (c == 3 && r == 3) ? '\n' : ' ');
            }
    }
    return 0;
}
