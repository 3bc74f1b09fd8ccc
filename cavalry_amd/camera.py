"""Host-side camera set-up (cav_camera_look_at): LookAt + fov -> camera."""
from __future__ import annotations

import ctypes

import numpy as np

from ._lib import Camera, CavError, lib


def look_at(eye, center, up, fov_degrees: float, width: int, height: int) -> Camera:
    """The reference's LookAt + PerspectiveCamera set-up
    (SceneLoading/Parsing.cu:196-204, Cameras/PerspectiveCamera.cu:6-20)."""
    cam = Camera()
    e = np.ascontiguousarray(eye, np.float32)
    c = np.ascontiguousarray(center, np.float32)
    u = np.ascontiguousarray(up, np.float32)
    rc = lib().cav_camera_look_at(e.ctypes.data_as(ctypes.c_void_p), c.ctypes.data_as(ctypes.c_void_p),
                                  u.ctypes.data_as(ctypes.c_void_p), float(fov_degrees),
                                  int(width), int(height), ctypes.byref(cam))
    if rc != 0:
        raise CavError(rc, "cav_camera_look_at failed")
    return cam


def for_scene(scene) -> Camera:
    return look_at(scene.eye, scene.center, scene.up, scene.fov, scene.width, scene.height)
