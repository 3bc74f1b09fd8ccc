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

INTEGRATORS = {"path": 0, "rlpath": 1, "directlighting": 2}
SORT_BY_MATERIAL = 1  # CAV_RENDER_SORT_BY_MATERIAL
DETERMINISTIC = 2     # CAV_RENDER_DETERMINISTIC
FILM_F64 = 4          # CAV_RENDER_FILM_F64


def default_bvh_rounds(spp: int) -> int:
    """Core/Scene.cu:141-146: 3 rounds, or 10 when spp >= 64."""
    return 10 if spp >= 64 else 3


class Film:
    """Per-pixel sample sums (rgb, [H, W, 3]) and filter weights ([H, W])."""

    def __init__(self, rgb: np.ndarray, weight: np.ndarray):
        self.rgb = np.ascontiguousarray(rgb, np.float32)
        self.weight = np.ascontiguousarray(weight, np.float32)

    @property
    def image(self) -> np.ndarray:
        return self.rgb / np.maximum(self.weight, 1e-30)[..., None]

    def to_srgb8(self) -> np.ndarray:
        """Film::readCurrentResult: clamp to [0, 1], sRGB gamma, u8."""
        h, w = self.weight.shape
        out = np.empty((h, w, 3), np.uint8)
        check(None, lib().cav_film_to_srgb8(self.rgb.ctypes.data_as(ctypes.c_void_p),
                                            self.weight.ctypes.data_as(ctypes.c_void_p), w, h,
                                            out.ctypes.data_as(ctypes.c_void_p)))
        return out

    def save_png(self, path: str) -> None:
        """RenderResult::saveToPNG of the resolved film."""
        img = np.ascontiguousarray(self.to_srgb8())
        h, w = self.weight.shape
        check(None, lib().cav_write_png(path.encode(), img.ctypes.data_as(ctypes.c_void_p), w, h))

    def save_pfm(self, path: str) -> None:
        """FP32 colour PFM of rgb / weight."""
        h, w = self.weight.shape
        check(None, lib().cav_write_pfm(path.encode(), self.rgb.ctypes.data_as(ctypes.c_void_p),
                                        self.weight.ctypes.data_as(ctypes.c_void_p), w, h))


def read_pfm(path: str) -> np.ndarray:
    """Colour PFM -> float32 [H, W, 3], top row first."""
    with open(path, "rb") as f:
        if f.readline().strip() != b"PF":
            raise ValueError(f"{path}: not a colour PFM")
        w, h = (int(v) for v in f.readline().split())
        scale = float(f.readline())
        data = np.frombuffer(f.read(), "<f4" if scale < 0 else ">f4", count=w * h * 3)
    return data.reshape(h, w, 3)[::-1].astype(np.float32)


def rel_mse(img: np.ndarray, ref: np.ndarray, eps: float = 1e-2) -> float:
    """Relative MSE, mean over pixels and channels of (img - ref)^2 / (ref^2 + eps)."""
    img = np.asarray(img, np.float64)
    ref = np.asarray(ref, np.float64)
    return float(np.mean((img - ref) ** 2 / (ref * ref + eps)))


class Scene:
    """Device scene (cav_scene): soup, meshes, areas, lights and BVH."""

    def __init__(self, ctx: Context, arrays, bvh_rounds: int = 0, morton_bits: int = 30):
        self.ctx = ctx
        self.arrays = arrays
        self.handle = None
        h = ctypes.c_void_p()
        tris = np.ascontiguousarray(arrays.tris, np.float32)
        meshes = np.ascontiguousarray(arrays.meshes)
        normals = getattr(arrays, "normals", None)
        flags = getattr(arrays, "mesh_normals", None)
        uv = getattr(arrays, "uv", None)
        textures = getattr(arrays, "textures", None) or []
        tex = (_lib.Texture * max(1, len(textures)))()
        for k, t in enumerate(textures):
            if t.ndim != 3 or t.shape[2] != 4 or t.dtype != np.uint8 or not t.flags.c_contiguous:
                raise ValueError("textures must be contiguous uint8 [H, W, 4] arrays")
            tex[k] = _lib.Texture(t.ctypes.data, t.shape[1], t.shape[0])
        env = getattr(arrays, "env_lights", None)
        env = None if env is None or len(env) == 0 else np.ascontiguousarray(env)
        order = getattr(arrays, "light_order", None)
        self._keep = (tris, meshes, normals, flags, uv, textures, tex, env, order)
        d = _lib.SceneDesc(tris.ctypes.data, len(tris), meshes.ctypes.data, len(meshes),
                           None if uv is None else uv.ctypes.data,
                           None if normals is None else normals.ctypes.data,
                           None if flags is None else flags.ctypes.data, int(bvh_rounds), int(morton_bits),
                           ctypes.addressof(tex) if textures else None, len(textures),
                           None if env is None else env.ctypes.data, 0 if env is None else len(env),
                           None if order is None else order.ctypes.data, 0 if order is None else len(order),
                           int(getattr(arrays, "env_background", -1)))
        check(ctx.handle, lib().cav_scene_create_desc(ctx.handle, ctypes.byref(d), ctypes.byref(h)))
        self.handle = h
        ctx.own(self)

    @classmethod
    def from_pbrt(cls, ctx: Context, path: str, **overrides):
        """Scene::prepareForRender of a parsed .pbrt file (overrides as
        cavalry_amd.pbrt.overrides). Returns (Scene, RenderSetup)."""
        from . import pbrt
        self = cls.__new__(cls)
        self.ctx = ctx
        self.arrays = None
        ov = pbrt.overrides(**overrides)
        st = pbrt.RenderSetup()
        h = ctypes.c_void_p()
        check(ctx.handle, lib().cav_scene_create_from_pbrt(ctx.handle, path.encode(), ctypes.byref(ov),
                                                           ctypes.byref(h), ctypes.byref(st)))
        self.handle = h
        self.setup = st
        ctx.own(self)
        return self, st

    def bvh(self):
        """The scene's BVH (owned by the scene) as a BVH object that does not
        destroy it: export(), trace counters, max_depth()."""
        from .bvh import BVH
        view = BVH.__new__(BVH)
        view.ctx = self.ctx
        view.handle = ctypes.c_void_p(lib().cav_scene_bvh(self.handle))
        view.n = int(len(self.arrays.tris)) if self.arrays is not None else int(self.setup.ntris)
        view.close = lambda: None  # the scene frees it
        return view

    close_rank = 1  # after the servers that render it (Context.close)

    def close(self) -> None:
        if self.handle:
            check(self.ctx.handle, lib().cav_scene_destroy(self.handle))
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def render(ctx: Context, scene: Scene, spp: int, integrator: str = "path", camera=None,
           max_depth: int | None = None, q_grid: int = 32, alpha_scale: float = 1.0,
           spp_per_round: int = 0, sort_by_material: bool = False, deterministic: bool = False,
           film_f64: bool = False):
    """Renders the scene; returns (Film, stats dict). q_grid and alpha_scale
    apply to rlpath (Q entries per axis; scale of the n/(T+n) learning rate);
    spp_per_round fixes the samples per pixel of each round (0: as many as
    device memory holds); sort_by_material shades each bounce's hits in
    material order (same film, different speed); deterministic makes rlpath
    sum its table updates in sample order instead of with float atomics (the
    same film on every run, and the oracle's); film_f64 adds the film sums in
    binary64 (high-spp reference images)."""
    a = scene.arrays
    if a is None:  # a scene parsed from a file: its setup gives the defaults
        st = scene.setup
        cam = camera if camera is not None else st.camera
        max_depth = max_depth if max_depth is not None else st.max_depth
    else:
        cam = camera if camera is not None else for_scene(a)
    if integrator not in INTEGRATORS:
        raise ValueError(f"unknown integrator {integrator!r}")
    prm = _lib.RenderParams(INTEGRATORS[integrator], int(spp),
                            int(max_depth if max_depth is not None else a.max_depth),
                            (SORT_BY_MATERIAL if sort_by_material else 0) | (DETERMINISTIC if deterministic else 0)
                            | (FILM_F64 if film_f64 else 0),
                            int(q_grid), float(alpha_scale), int(spp_per_round))
    rgb = np.zeros((cam.height, cam.width, 3), np.float32)
    w = np.zeros((cam.height, cam.width), np.float32)
    st = _lib.RenderStats()
    check(ctx.handle, lib().cav_render(ctx.handle, scene.handle, ctypes.byref(cam), ctypes.byref(prm),
                                       rgb.ctypes.data_as(ctypes.c_void_p),
                                       w.ctypes.data_as(ctypes.c_void_p), ctypes.byref(st)))
    return Film(rgb, w), {k: getattr(st, k) for k, _ in st._fields_}
