"""Host-side scene and renderer objects over the C ABI.

Mirrors the reference's render path: ``Scene.prepare_for_render`` builds the
device scene and its BVH (Core/Scene.cu:117-151) and ``Renderer.render`` runs
the integrator named in the parameters (Core/Renderer.cu:3-6,
Integrators/CreateIntegrator.cu:10-25). Films come back as per-pixel sums and
weights; ``Film.image`` divides them.
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _lib
from ._lib import check, lib
from .bvh import Context
from .camera import for_scene

INTEGRATORS = {"path": 0, "rlpath": 1}


def default_bvh_rounds(spp: int) -> int:
    """Core/Scene.cu:141-146: 3 rounds, or 10 when spp >= 64."""
    return 10 if spp >= 64 else 3


class Film:
    def __init__(self, rgb: np.ndarray, weight: np.ndarray):
        self.rgb = rgb
        self.weight = weight

    @property
    def image(self) -> np.ndarray:
        return self.rgb / np.maximum(self.weight, 1e-30)[..., None]


class Scene:
    """Device scene (cav_scene): soup, meshes, areas, lights and BVH."""

    def __init__(self, ctx: Context, arrays, bvh_rounds: int = 0, morton_bits: int = 30):
        self.ctx = ctx
        self.arrays = arrays
        h = ctypes.c_void_p()
        tris = np.ascontiguousarray(arrays.tris, np.float32)
        meshes = np.ascontiguousarray(arrays.meshes)
        check(ctx.handle, lib().cav_scene_create(
            ctx.handle, tris.ctypes.data_as(ctypes.c_void_p), len(tris),
            meshes.ctypes.data_as(ctypes.c_void_p), len(meshes), None, int(bvh_rounds),
            int(morton_bits), ctypes.byref(h)))
        self.handle = h

    def close(self) -> None:
        if self.handle:
            lib().cav_scene_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def render(ctx: Context, scene: Scene, spp: int, integrator: str = "path", camera=None,
           max_depth: int | None = None, q_grid: int = 32, alpha_scale: float = 1.0):
    """Renders the scene; returns (Film, stats dict). q_grid and alpha_scale
    apply to rlpath (Q entries per axis; scale of the n/(T+n) learning rate)."""
    a = scene.arrays
    cam = camera if camera is not None else for_scene(a)
    if integrator not in INTEGRATORS:
        raise ValueError(f"unknown integrator {integrator!r}")
    prm = _lib.RenderParams(INTEGRATORS[integrator], int(spp),
                            int(max_depth if max_depth is not None else a.max_depth), 0,
                            int(q_grid), float(alpha_scale))
    rgb = np.zeros((cam.height, cam.width, 3), np.float32)
    w = np.zeros((cam.height, cam.width), np.float32)
    st = _lib.RenderStats()
    check(ctx.handle, lib().cav_render(ctx.handle, scene.handle, ctypes.byref(cam), ctypes.byref(prm),
                                       rgb.ctypes.data_as(ctypes.c_void_p),
                                       w.ctypes.data_as(ctypes.c_void_p), ctypes.byref(st)))
    return Film(rgb, w), {k: getattr(st, k) for k, _ in st._fields_}
