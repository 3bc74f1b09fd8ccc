"""Autotuner for the traversal launch shape (cav_trace_set_launch).

The persistent traversal kernels can run with any number of workgroups per
CU, any ray-dispenser chunk and 32 or 48 stack entries in LDS; every ray's
node order is the reference's either way (BVH/BVH.h:23-189), so the choice
changes only the speed. Which shape is fastest depends on the batch and the
tree: big incoherent batches over trees much larger than L2 prefer fewer rays
in flight (less L2 thrash), small or coherent ones prefer more (latency).

`tune_trace` times candidate shapes on the caller's own rays (closest-hit and
any-hit, best of `reps`), by coordinate descent from the current shape over
workgroups per CU, then chunk, then LDS depth, sets the fastest on the
context and records it in a JSON cache keyed by device, tree size and batch
size. A later call with the same key reuses the cached shape without timing
(`use_cache`). `tune_render` does the same for the render loops (WAVEFRONT)
on one render. Treelet rounds are not tuned: they change the tree, and with
it which of two equally distant triangles a ray reports.
"""
from __future__ import annotations

import ctypes
import json
import math
import os
import time

from ._lib import TraceLaunch, check, lib

STANDALONE, WAVEFRONT = 0, 1
PER_CU = (1, 2, 3, 4, 5)
CHUNKS = (128, 256, 512, 1024)
LDS = (32, 48)


def default_cache_path() -> str:
    return os.path.join(os.path.expanduser("~"), ".cache", "cavalry_amd", "tune.json")


def get_launch(ctx, which: int) -> dict:
    t = TraceLaunch()
    check(ctx.handle, lib().cav_trace_get_launch(ctx.handle, int(which), ctypes.byref(t)))
    return {"workgroups_per_cu": t.workgroups_per_cu, "ray_chunk": t.ray_chunk, "lds_stack": t.lds_stack}


def set_launch(ctx, which: int, shape: dict) -> None:
    t = TraceLaunch(int(shape["workgroups_per_cu"]), int(shape["ray_chunk"]), int(shape["lds_stack"]))
    check(ctx.handle, lib().cav_trace_set_launch(ctx.handle, int(which), ctypes.byref(t)))


def _device_name() -> str:
    try:
        import torch
        if torch.cuda.is_available():
            p = torch.cuda.get_device_properties(0)
            return f"{p.name}|{getattr(p, 'gcnArchName', '')}|{p.multi_processor_count}cu"
    except Exception:
        pass
    return "unknown"


def _size_class(n: int) -> str:
    return f"2^{max(0, int(round(math.log2(max(1, n)))))}"


def _load(path: str) -> dict:
    try:
        with open(path) as f:
            return json.load(f)
    except (OSError, ValueError):
        return {}


def _store(path: str, key: str, entry: dict) -> None:
    data = _load(path)
    data.setdefault("entries", {})[key] = entry
    os.makedirs(os.path.dirname(os.path.abspath(path)), exist_ok=True)
    tmp = path + ".tmp"
    with open(tmp, "w") as f:
        json.dump(data, f, indent=1, sort_keys=True)
    os.replace(tmp, path)


def _descend(ctx, which: int, timer, start: dict, reps: int) -> tuple[dict, dict]:
    """Coordinate descent over (workgroups per CU, chunk, LDS depth) from
    `start`; returns the best shape and every timing taken (ms)."""
    timings = {}

    def measure(shape):
        k = f"{shape['workgroups_per_cu']}/{shape['ray_chunk']}/{shape['lds_stack']}"
        if k not in timings:
            set_launch(ctx, which, shape)
            timings[k] = min(timer() for _ in range(reps))
        return timings[k]

    best = dict(start)
    best_ms = measure(best)
    for field, values in (("workgroups_per_cu", PER_CU), ("ray_chunk", CHUNKS), ("lds_stack", LDS)):
        for v in values:
            cand = dict(best, **{field: v})
            ms = measure(cand)
            if ms < best_ms:
                best, best_ms = cand, ms
    set_launch(ctx, which, best)
    return best, timings


def tune_trace(ctx, bvh, rays, nrays: int, shadow_rays=None, limit=None, reps: int = 2,
               cache_path: str | None = None, use_cache: bool = True) -> dict:
    """Tunes the STANDALONE launch shape on device rays (torch tensors or
    device pointers as the bvh entry points take them: `rays` float[n][6],
    optional `shadow_rays` and `limit` for the any-hit half). Returns
    {"shape", "key", "ms", "timings", "cached"}."""
    import torch

    path = cache_path or default_cache_path()
    key = f"{_device_name()}|standalone|tris={_size_class(bvh.n)}|rays={_size_class(nrays)}"
    if use_cache:
        hit = _load(path).get("entries", {}).get(key)
        if hit:
            set_launch(ctx, STANDALONE, hit["shape"])
            return {"shape": hit["shape"], "key": key, "ms": hit["ms"], "timings": hit.get("timings", {}),
                    "cached": True}
    dev = torch.device("cuda:0")
    prim = torch.empty(nrays, dtype=torch.int32, device=dev)
    tt = torch.empty(nrays, dtype=torch.float32, device=dev)
    vis = torch.empty(nrays, dtype=torch.uint8, device=dev)
    ptr = lambda x: x.data_ptr() if hasattr(x, "data_ptr") else int(x)

    def timer():
        ctx.synchronize()
        t0 = time.perf_counter()
        check(ctx.handle, lib().cav_trace_closest_async(ctx.handle, bvh.handle, ptr(rays), nrays,
                                                       prim.data_ptr(), tt.data_ptr()))
        if shadow_rays is not None:
            check(ctx.handle, lib().cav_trace_occluded_async(ctx.handle, bvh.handle, ptr(shadow_rays), ptr(limit),
                                                            nrays, vis.data_ptr()))
        ctx.synchronize()
        return (time.perf_counter() - t0) * 1e3

    best, timings = _descend(ctx, STANDALONE, timer, get_launch(ctx, STANDALONE), reps)
    entry = {"shape": best, "ms": timings[f"{best['workgroups_per_cu']}/{best['ray_chunk']}/{best['lds_stack']}"],
             "timings": timings, "when": time.strftime("%Y-%m-%d %H:%M:%S")}
    _store(path, key, entry)
    return {"shape": best, "key": key, "ms": entry["ms"], "timings": timings, "cached": False}


def tune_render(ctx, scene, spp: int, integrator: str = "path", camera=None, max_depth=None, reps: int = 2,
                cache_path: str | None = None, use_cache: bool = True) -> dict:
    """Tunes the WAVEFRONT launch shape on one render (device time of
    cav_render). Returns {"shape", "key", "ms", "timings", "cached"}."""
    from .render import render

    path = cache_path or default_cache_path()
    cam = camera if camera is not None else (scene.setup.camera if scene.arrays is None else None)
    from .camera import for_scene
    if cam is None:
        cam = for_scene(scene.arrays)
    ntris = len(scene.arrays.tris) if scene.arrays is not None else 0
    key = (f"{_device_name()}|wavefront|{integrator}|tris={_size_class(ntris)}|"
           f"samples={_size_class(cam.width * cam.height * spp)}")
    if use_cache:
        hit = _load(path).get("entries", {}).get(key)
        if hit:
            set_launch(ctx, WAVEFRONT, hit["shape"])
            return {"shape": hit["shape"], "key": key, "ms": hit["ms"], "timings": hit.get("timings", {}),
                    "cached": True}

    def timer():
        _, st = render(ctx, scene, spp, integrator=integrator, camera=cam, max_depth=max_depth)
        return float(st["ms_total"])

    best, timings = _descend(ctx, WAVEFRONT, timer, get_launch(ctx, WAVEFRONT), reps)
    entry = {"shape": best, "ms": timings[f"{best['workgroups_per_cu']}/{best['ray_chunk']}/{best['lds_stack']}"],
             "timings": timings, "when": time.strftime("%Y-%m-%d %H:%M:%S")}
    _store(path, key, entry)
    return {"shape": best, "key": key, "ms": entry["ms"], "timings": timings, "cached": False}
