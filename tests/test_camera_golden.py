"""The host camera against the reference's own arithmetic and against what
the geometry fixes.

* Bit for bit: cav_camera_look_at's camera-to-world matrix equals
  inverse(glm::lookAtLH(eye, center, up)) from the glm headers the reference
  vendors (tests/golden/camera_glm.json, written by make_camera_golden.py).
* Convention: the ray through the image centre points at `center`; pixels
  above the centre see along +up, pixels right of it along
  normalize(cross(up, forward)) (lookAtLH's left-handed frame), so a
  mirrored or upside-down image fails."""
import ctypes
import json
import os

import numpy as np
import pytest

from cavalry_amd.camera import look_at

HERE = os.path.dirname(os.path.abspath(__file__))


def _cases():
    with open(os.path.join(HERE, "golden", "camera_glm.json")) as f:
        return json.load(f)["cases"]


def _f32(hexes):
    return np.array([int(h, 16) for h in hexes], np.uint32).view(np.float32)


def test_look_at_matches_glm_bit_for_bit():
    cases = _cases()
    assert len(cases) >= 100
    for k, c in enumerate(cases):
        v = _f32(c["eye_center_up_bits"])
        cam = look_at(v[0:3], v[3:6], v[6:9], 45.0, 64, 64)
        got = np.ctypeslib.as_array(cam.cam_to_world).astype(np.float32).view(np.uint32)
        want = np.array([int(h, 16) for h in c["cam_to_world_bits"]], np.uint32)
        assert (got == want).all(), f"case {k}: got {got} want {want}"


def _ray_dir(cam, x, y):
    """Direction of the ray through film point (x, y) (pixel units, y down)."""
    m = np.array(cam.cam_to_world, np.float64).reshape(4, 4).T
    p = np.array([x - cam.width / 2, cam.height / 2 - y, cam.depth])
    d = m[:3, :3] @ p
    return d / np.linalg.norm(d)


@pytest.mark.parametrize("eye,center,up", [((0, 0, 0), (0, 0, 1), (0, 1, 0)),
                                           ((0.5, 0.5, -1.3), (0.5, 0.5, 0.5), (0, 1, 0)),
                                           ((3, -1, 2), (-1, 2, 0.5), (0.2, 0.1, 1.0))])
def test_camera_frame_convention(eye, center, up):
    cam = look_at(eye, center, up, 40.0, 96, 64)
    f = np.subtract(center, eye, dtype=np.float64)
    f /= np.linalg.norm(f)
    u = np.asarray(up, np.float64)
    s = np.cross(u, f)
    s /= np.linalg.norm(s)
    uu = np.cross(f, s)  # the image's up direction
    c = _ray_dir(cam, 48, 32)
    np.testing.assert_allclose(c, f, atol=1e-6)       # centre ray at `center`
    right, top = _ray_dir(cam, 90, 32), _ray_dir(cam, 48, 2)
    assert right @ s > 0.1 and abs(right @ uu) < 1e-6   # image right = lookAtLH's s
    assert top @ uu > 0.1 and abs(top @ s) < 1e-6       # image up = the component of `up`
    # the origin of every ray is the eye
    m = np.array(cam.cam_to_world, np.float64).reshape(4, 4).T
    np.testing.assert_allclose(m[:3, 3], eye, atol=1e-5)
    # fov: the shorter image side (64 px) spans 40 degrees
    edge = _ray_dir(cam, 48, 0)
    np.testing.assert_allclose(np.degrees(np.arccos(edge @ f)), 20.0, atol=1e-4)
