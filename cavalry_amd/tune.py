"""Autotuner for the launch shapes (cav_trace_set_launch) and the treelet
round count.

The persistent traversal kernels can run with any workgroup size (64, 128 or
256 threads), any number of workgroups per CU, any ray-dispenser chunk and
32 or 48 stack entries in LDS; the per-sample shading kernels of the render
loops (path shading, the RL integrator's Q-table sampling) with any of the
same workgroup sizes. Every ray's node order is the reference's
(BVH/BVH.h:23-189) and every sample's shading is its own either way, so the
choice changes only the speed (the reference fixes 256 threads at compile
time, Utils/GpuCommons.h:63). Which shape is fastest depends on the batch and
the tree: big incoherent batches over trees much larger than L2 prefer fewer
rays in flight (less L2 thrash), small or coherent ones prefer more
(latency).

`tune_trace` times candidate shapes on the caller's own rays (closest-hit and
any-hit, best of `reps`), by coordinate descent from the current shape over
workgroup size (at the same waves per CU), workgroups per CU, chunk, LDS
depth, the steps after which a ray still in flight at the end of a launch
goes to a cooperative pass (a whole wave per ray) and whether an any-hit
step fetches one node or a pair of siblings, sets the fastest on the
context and records it in a JSON cache keyed by device, tree size and batch
size. A later call with the same key reuses
the cached shape without timing (`use_cache`). `tune_render` does the same
for the render loops (WAVEFRONT) on one render, and also descends over the
shading workgroup size.

Treelet rounds are different: they change the tree, and with it which of two
equally distant triangles a ray reports, exactly as the reference's
--bvh-opt-rounds does. `tune_rounds` therefore only recommends a count: it
times build + rounds + the caller's traces for each candidate and returns
the count with the least total time (cached like the shapes); the caller
decides whether to build with it.
"""
from __future__ import annotations

import ctypes
import json
import math
import os
import time

from ._lib import TraceLaunch, check, lib

STANDALONE, WAVEFRONT = 0, 1
SIZES = (64, 128, 256)          # threads per workgroup (traversal and shading)
WAVES = (4, 8, 12, 16, 20)      # traversal waves per CU tried (1..5 workgroups of 256)
PER_CU = tuple(sorted({w * 64 // b for w in WAVES for b in SIZES}))
CHUNKS = (128, 256, 512, 1024)
HANDOFFS = (256, 1024, 4096, 1 << 30)  # steps before a ray still in flight at the end goes to a
                                       # whole wave (1 << 30: never; the passes are not launched)
LDS = (32, 48)
ANY_STEPS = (1, 2)              # any-hit step: one node, or a pair of sibling nodes (0: one node)
ROUNDS = (0, 1, 2, 3, 4, 6, 8)


def default_cache_path() -> str:
    return os.path.join(os.path.expanduser("~"), ".cache", "cavalry_amd", "tune.json")


def get_launch(ctx, which: int) -> dict:
    t = TraceLaunch()
    check(ctx.handle, lib().cav_trace_get_launch(ctx.handle, int(which), ctypes.byref(t)))
    return {"workgroups_per_cu": t.workgroups_per_cu, "ray_chunk": t.ray_chunk, "lds_stack": t.lds_stack,
            "workgroup_size": t.workgroup_size, "shade_workgroup_size": t.shade_workgroup_size,
            "handoff_steps": t.handoff_steps, "any_hit_step": t.any_hit_step,
            "any_hit_whole_line": t.any_hit_whole_line}


def set_launch(ctx, which: int, shape: dict) -> None:
    """Sets a shape; missing workgroup sizes mean 256, a missing hand-off
    1024 steps, a missing any-hit step one node (shapes cached before
    they were tunable) and a missing any_hit_whole_line 0 (any-hit rays
    skip boxes that start past their limit; 1 visits them as the reference
    does, same answers)."""
    t = TraceLaunch(int(shape["workgroups_per_cu"]), int(shape["ray_chunk"]), int(shape["lds_stack"]),
                    int(shape.get("workgroup_size", 0) or 0), int(shape.get("shade_workgroup_size", 0) or 0),
                    int(shape.get("handoff_steps", 0) or 0), int(shape.get("any_hit_step", 0) or 0),
                    int(shape.get("any_hit_whole_line", 0) or 0))
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


def _key(shape: dict) -> str:
    return (f"{shape['workgroups_per_cu']}/{shape['ray_chunk']}/{shape['lds_stack']}/"
            f"{shape.get('workgroup_size', 256) or 256}/{shape.get('shade_workgroup_size', 256) or 256}/"
            f"{shape.get('handoff_steps', 1024) or 1024}/{shape.get('any_hit_step', 0) or 0}")


def _normal(shape: dict) -> dict:
    s = dict(shape)
    s["workgroup_size"] = s.get("workgroup_size", 0) or 256
    s["shade_workgroup_size"] = s.get("shade_workgroup_size", 0) or 256
    s["handoff_steps"] = s.get("handoff_steps", 0) or 1024
    s["any_hit_step"] = s.get("any_hit_step", 0) or 0
    s["any_hit_whole_line"] = s.get("any_hit_whole_line", 0) or 0
    return s


def _per_cu_for(size: int) -> tuple:
    """Workgroups per CU tried at one workgroup size: the same waves per CU
    for every size."""
    return tuple(w * 64 // size for w in WAVES)


def _descend(ctx, which: int, timer, start: dict, reps: int, shade: bool = False) -> tuple[dict, dict]:
    """Coordinate descent from `start` over the workgroup size (keeping the
    waves per CU), workgroups per CU, chunk, LDS depth, hand-off steps, the
    any-hit step and, with `shade`, the shading workgroup size; returns the best shape and
    every timing taken (ms)."""
    timings = {}

    def measure(shape):
        k = _key(shape)
        if k not in timings:
            set_launch(ctx, which, shape)
            timings[k] = min(timer() for _ in range(reps))
        return timings[k]

    best = _normal(start)
    best_ms = measure(best)

    def waves(s):
        return (s["workgroups_per_cu"] or 0) * s["workgroup_size"] // 64

    for b in SIZES:
        w = waves(best)
        per = max(1, w * 64 // b) if w else 0
        cand = dict(best, workgroup_size=b, workgroups_per_cu=per)
        ms = measure(cand)
        if ms < best_ms:
            best, best_ms = cand, ms
    fields = [("workgroups_per_cu", None), ("ray_chunk", CHUNKS), ("lds_stack", LDS), ("handoff_steps", HANDOFFS),
              ("any_hit_step", ANY_STEPS)]
    if shade:
        fields.append(("shade_workgroup_size", SIZES))
    for field, values in fields:
        for v in (values if values is not None else _per_cu_for(best["workgroup_size"])):
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
    entry = {"shape": best, "ms": timings[_key(best)], "timings": timings,
             "when": time.strftime("%Y-%m-%d %H:%M:%S")}
    _store(path, key, entry)
    return {"shape": best, "key": key, "ms": entry["ms"], "timings": timings, "cached": False}


def tune_render(ctx, scene, spp: int, integrator: str = "path", camera=None, max_depth=None, reps: int = 2,
                cache_path: str | None = None, use_cache: bool = True, **render_kw) -> dict:
    """Tunes the WAVEFRONT launch shape, shading workgroup size included, on
    one render (device time of cav_render; render_kw go to render.render,
    e.g. q_grid). Returns {"shape", "key", "ms", "timings", "cached"}."""
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
        _, st = render(ctx, scene, spp, integrator=integrator, camera=cam, max_depth=max_depth, **render_kw)
        return float(st["ms_total"])

    best, timings = _descend(ctx, WAVEFRONT, timer, get_launch(ctx, WAVEFRONT), reps, shade=True)
    entry = {"shape": best, "ms": timings[_key(best)], "timings": timings,
             "when": time.strftime("%Y-%m-%d %H:%M:%S")}
    _store(path, key, entry)
    return {"shape": best, "key": key, "ms": entry["ms"], "timings": timings, "cached": False}


def tune_rounds(ctx, triangles, rays, shadow_rays=None, limit=None, traces: int = 1,
                candidates=ROUNDS, morton_bits: int = 30, cache_path: str | None = None,
                use_cache: bool = True) -> dict:
    """Recommends a treelet round count for a triangle set and a ray batch:
    for each candidate, the device time of the build with that many rounds
    plus `traces` closest-hit (and, given `shadow_rays` and `limit`, any-hit)
    passes over the rays (host arrays); the count with the least total wins.
    Nothing is set: the rounds change the tree, as the reference's
    --bvh-opt-rounds does. Returns {"rounds", "key", "ms", "timings",
    "cached"} with timings[r] = {"build_ms", "trace_ms", "total_ms",
    "sah"}."""
    import numpy as np
    import torch

    from .bvh import BVH

    path = cache_path or default_cache_path()
    n, nrays = len(triangles), len(rays)
    key = (f"{_device_name()}|rounds|tris={_size_class(n)}|rays={_size_class(nrays)}|traces={traces}|"
           f"any={shadow_rays is not None}|cands={','.join(str(c) for c in candidates)}")
    if use_cache:
        hit = _load(path).get("entries", {}).get(key)
        if hit:
            return {"rounds": hit["rounds"], "key": key, "ms": hit["ms"], "timings": hit.get("timings", {}),
                    "cached": True}
    dev = torch.device("cuda:0")
    tris_d = torch.from_numpy(np.ascontiguousarray(triangles, np.float32)).to(dev)
    rays_d = torch.from_numpy(np.ascontiguousarray(rays, np.float32)).to(dev)
    srays_d = lim_d = None
    if shadow_rays is not None:
        srays_d = torch.from_numpy(np.ascontiguousarray(shadow_rays, np.float32)).to(dev)
        lim_d = torch.from_numpy(np.ascontiguousarray(limit, np.float32)).to(dev)
    prim = torch.empty(nrays, dtype=torch.int32, device=dev)
    tt = torch.empty(nrays, dtype=torch.float32, device=dev)
    vis = torch.empty(nrays, dtype=torch.uint8, device=dev)
    timings = {}
    for r in candidates:
        bvh = BVH.build(ctx, tris_d, optimization_rounds=int(r), morton_bits=morton_bits, device=True)
        try:
            build_ms = float(bvh.stats()["ms_total"])
            ctx.synchronize()
            t0 = time.perf_counter()
            for _ in range(traces):
                check(ctx.handle, lib().cav_trace_closest_async(ctx.handle, bvh.handle, rays_d.data_ptr(), nrays,
                                                               prim.data_ptr(), tt.data_ptr()))
                if srays_d is not None:
                    check(ctx.handle, lib().cav_trace_occluded_async(ctx.handle, bvh.handle, srays_d.data_ptr(),
                                                                    lim_d.data_ptr(), nrays, vis.data_ptr()))
            ctx.synchronize()
            trace_ms = (time.perf_counter() - t0) * 1e3
            sah = float(bvh.stats()["root_cost_final"])
        finally:
            bvh.close()
        timings[str(r)] = {"build_ms": build_ms, "trace_ms": trace_ms, "total_ms": build_ms + trace_ms, "sah": sah}
    best = min(candidates, key=lambda r: timings[str(r)]["total_ms"])
    entry = {"rounds": int(best), "ms": timings[str(best)]["total_ms"], "timings": timings,
             "when": time.strftime("%Y-%m-%d %H:%M:%S")}
    _store(path, key, entry)
    return {"rounds": int(best), "key": key, "ms": entry["ms"], "timings": timings, "cached": False}
