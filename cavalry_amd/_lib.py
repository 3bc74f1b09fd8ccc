"""ctypes binding of libcavalry_amd.so (the C ABI declared in include/cavalry.h).

The library is built in-tree by ``python -m cavalry_amd.build``. There is no
fallback: if the shared object is missing or fails to load, every entry point
raises.
"""
from __future__ import annotations

import ctypes
import os
import re

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_lib", "libcavalry_amd.so")
HEADER = os.path.join(os.path.dirname(HERE), "include", "cavalry.h")

CAV_OK = 0
STATUS = {
    0: "CAV_OK", 1: "CAV_ERR_INVALID", 2: "CAV_ERR_HIP", 3: "CAV_ERR_OOM",
    4: "CAV_ERR_OVERFLOW", 5: "CAV_ERR_PARSE", 6: "CAV_ERR_BUSY", 7: "CAV_ERR_NOT_FOUND",
}


class CavError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"{STATUS.get(code, code)}: {msg}")
        self.code = code


class BvhStats(ctypes.Structure):
    _fields_ = [(k, ctypes.c_float) for k in (
        "ms_bounds", "ms_morton", "ms_sort", "ms_karras", "ms_refit", "ms_treelet",
        "ms_flatten", "ms_total", "root_cost_lbvh", "root_cost_final")] + [
        ("n", ctypes.c_int), ("morton_bits", ctypes.c_int), ("opt_rounds", ctypes.c_int)]


class Camera(ctypes.Structure):
    _fields_ = [("cam_to_world", ctypes.c_float * 16), ("depth", ctypes.c_float),
                ("width", ctypes.c_int32), ("height", ctypes.c_int32)]


class RenderParams(ctypes.Structure):
    _fields_ = [("integrator", ctypes.c_int32), ("spp", ctypes.c_int32),
                ("max_depth", ctypes.c_int32), ("flags", ctypes.c_int32),
                ("q_grid", ctypes.c_int32), ("alpha_scale", ctypes.c_float),
                ("spp_per_round", ctypes.c_int32)]


class RenderStats(ctypes.Structure):
    _fields_ = [("ms_total", ctypes.c_double), ("closest_rays", ctypes.c_int64),
                ("shadow_rays", ctypes.c_int64), ("mrays_per_s", ctypes.c_double),
                ("rounds", ctypes.c_int32)]


class Texture(ctypes.Structure):
    _fields_ = [("rgba", ctypes.c_void_p), ("width", ctypes.c_int32), ("height", ctypes.c_int32)]


class SceneDesc(ctypes.Structure):
    _fields_ = [("tris", ctypes.c_void_p), ("ntris", ctypes.c_int32), ("meshes", ctypes.c_void_p),
                ("nmeshes", ctypes.c_int32), ("tri_uv", ctypes.c_void_p), ("tri_normals", ctypes.c_void_p),
                ("mesh_normals", ctypes.c_void_p), ("bvh_opt_rounds", ctypes.c_int32),
                ("morton_bits", ctypes.c_int32), ("textures", ctypes.c_void_p), ("ntextures", ctypes.c_int32),
                ("env_lights", ctypes.c_void_p), ("nenv", ctypes.c_int32), ("light_order", ctypes.c_void_p),
                ("nlight_order", ctypes.c_int32), ("env_background", ctypes.c_int32)]


class ServeConfig(ctypes.Structure):
    _fields_ = [("arena_bytes", ctypes.c_int64), ("host_arena_bytes", ctypes.c_int64),
                ("pool_tasks", ctypes.c_int64), ("chunk_samples", ctypes.c_int64),
                ("max_steps_ahead", ctypes.c_int32), ("use_graphs", ctypes.c_int32),
                ("rl_workers", ctypes.c_int32)]


class TraceLaunch(ctypes.Structure):
    _fields_ = [("workgroups_per_cu", ctypes.c_int32), ("ray_chunk", ctypes.c_int32),
                ("lds_stack", ctypes.c_int32), ("workgroup_size", ctypes.c_int32),
                ("shade_workgroup_size", ctypes.c_int32), ("handoff_steps", ctypes.c_int32),
                ("any_hit_step", ctypes.c_int32), ("any_hit_whole_line", ctypes.c_int32)]


class ServeResult(ctypes.Structure):
    _fields_ = [("status", ctypes.c_int32), ("error", ctypes.c_int32), ("submit_ms", ctypes.c_double),
                ("start_ms", ctypes.c_double), ("done_ms", ctypes.c_double),
                ("closest_rays", ctypes.c_int64), ("shadow_rays", ctypes.c_int64)]


class ServeStats(ctypes.Structure):
    _fields_ = [("pool_closest_rays", ctypes.c_int64), ("pool_shadow_rays", ctypes.c_int64),
                ("rl_closest_rays", ctypes.c_int64), ("rl_shadow_rays", ctypes.c_int64),
                ("steps", ctypes.c_int64), ("pools", ctypes.c_int32), ("graph_pools", ctypes.c_int32)]


_lib = None

P = ctypes.c_void_p
I = ctypes.c_int
I64 = ctypes.c_int64

_SIGS = {
    "cav_ctx_create": ([I, ctypes.POINTER(P)], I),
    "cav_ctx_destroy": ([P], I),
    "cav_last_error": ([P], ctypes.c_char_p),
    "cav_ctx_stream": ([P], P),
    "cav_ctx_synchronize": ([P], I),
    "cav_version": ([], ctypes.c_char_p),
    "cav_bvh_build": ([P, P, I, P, P, I, I, ctypes.POINTER(P)], I),
    "cav_bvh_build_device": ([P, P, I, P, P, I, I, ctypes.POINTER(P)], I),
    "cav_bvh_rebuild_device": ([P, P, P, P, I, I], I),
    "cav_bvh_destroy": ([P], I),
    "cav_bvh_stats_get": ([P, ctypes.POINTER(BvhStats)], I),
    "cav_bvh_kernel_times": ([P, P, P, I], I),
    "cav_bvh_export": ([P, P, P, P, P, P, P, P, P, P, P], I),
    "cav_trace_closest": ([P, P, P, I64, P, P, P, P], I),
    "cav_trace_closest_device": ([P, P, P, I64, P, P, P, P], I),
    "cav_trace_occluded": ([P, P, P, P, P, P, I64, P], I),
    "cav_trace_occluded_device": ([P, P, P, P, P, P, I64, P], I),
    "cav_trace_closest_async": ([P, P, P, I64, P, P], I),
    "cav_trace_occluded_async": ([P, P, P, P, I64, P], I),
    "cav_trace_counters": ([P, P, P, ctypes.c_int32], I),
    "cav_trace_fetches": ([P, P, P, ctypes.c_int32], I),
    "cav_trace_closest_noalloc": ([P, P, P, ctypes.c_int64, P, P, P, P], I),
    "cav_bvh_max_depth": ([P, P, ctypes.POINTER(ctypes.c_int32)], I),
    "cav_camera_look_at": ([P, P, P, ctypes.c_float, I, I, ctypes.POINTER(Camera)], I),
    "cav_scene_create": ([P, P, I, P, I, P, I, I, ctypes.POINTER(P)], I),
    "cav_scene_create_desc": ([P, ctypes.POINTER(SceneDesc), ctypes.POINTER(P)], I),
    "cav_scene_destroy": ([P], I),
    "cav_scene_bvh": ([P], P),
    "cav_render": ([P, P, ctypes.POINTER(Camera), ctypes.POINTER(RenderParams), P, P,
                    ctypes.POINTER(RenderStats)], I),
    "cav_camera_from_matrix": ([P, ctypes.c_float, I, I, ctypes.POINTER(Camera)], I),
    "cav_pbrt_parse": ([ctypes.c_char_p, P, ctypes.POINTER(P), ctypes.c_char_p, I], I),
    "cav_pbrt_get": ([P, P, P, P, P], I),
    "cav_pbrt_get_normals": ([P, P, P], I),
    "cav_pbrt_get_textures": ([P, P, P, P], I),
    "cav_pbrt_get_lights": ([P, P, P, P, P, P], I),
    "cav_pbrt_destroy": ([P], None),
    "cav_scene_create_from_pbrt": ([P, ctypes.c_char_p, P, ctypes.POINTER(P), P], I),
    "cav_trace_set_launch": ([P, ctypes.c_int32, ctypes.POINTER(TraceLaunch)], I),
    "cav_trace_get_launch": ([P, ctypes.c_int32, ctypes.POINTER(TraceLaunch)], I),
    "cav_serve_open": ([P, ctypes.POINTER(ServeConfig), ctypes.POINTER(P)], I),
    "cav_serve_submit": ([P, P, ctypes.POINTER(Camera), ctypes.POINTER(RenderParams), P, P,
                          ctypes.POINTER(I64)], I),
    "cav_serve_poll": ([P, I64, ctypes.c_int32, ctypes.POINTER(ServeResult)], I),
    "cav_serve_cancel": ([P, I64], I),
    "cav_serve_close": ([P, ctypes.POINTER(ServeStats)], I),
    "cav_film_to_srgb8": ([P, P, I, I, P], I),
    "cav_write_png": ([ctypes.c_char_p, P, I, I], I),
    "cav_write_pfm": ([ctypes.c_char_p, P, P, I, I], I),
}


def lib() -> ctypes.CDLL:
    """Load the HIP library, raising if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(
                f"{LIB_PATH} is missing: build it with `python -m cavalry_amd.build` "
                "(there is no CPU fallback)")
        l = ctypes.CDLL(LIB_PATH)
        for name, (args, res) in _SIGS.items():
            f = getattr(l, name, None)
            if f is None:
                continue
            f.argtypes = args
            f.restype = res
        _lib = l
    return _lib


def header_symbols() -> list[str]:
    """Function names declared in include/cavalry.h."""
    with open(HEADER) as f:
        text = f.read()
    return sorted(set(re.findall(r"\b(cav_[a-z0-9_]+)\s*\(", text)))


def check(ctx, rc: int) -> None:
    if rc != CAV_OK:
        msg = lib().cav_last_error(ctx) if ctx else b""
        raise CavError(rc, (msg or b"").decode(errors="replace"))
