"""Host-side BVH and context objects over the C ABI.

Mirrors the reference's interfaces for this path:
``BVH.build(triangles, trianglesCount, sceneBounds, optimizationRounds)``
(BVH/BVH.h:19), ``BVH::intersect`` (BVH/BVH.h:23-107) and
``BVH::testVisibility`` (BVH/BVH.h:109-189), with the triangles given as a
float32[N, 9] soup and rays as float32[R, 6].
"""
from __future__ import annotations

import atexit
import ctypes
import weakref

import numpy as np

from . import _lib
from ._lib import check, lib


def _ptr(a):
    if a is None:
        return None
    if isinstance(a, int):
        return ctypes.c_void_p(a)
    if hasattr(a, "data_ptr"):  # torch tensor on the device
        return ctypes.c_void_p(a.data_ptr())
    return a.ctypes.data_as(ctypes.c_void_p)


_LIVE = weakref.WeakSet()


@atexit.register
def _close_all() -> None:
    # while the HIP runtime is still up: interpreter teardown may otherwise
    # free device objects after it
    for c in list(_LIVE):
        c.close()


class Context:
    """One device, one HIP stream (cav_ctx). Objects made on it (BVHs,
    scenes, servers) are closed before the context itself: the library
    requires them to be destroyed first."""

    def __init__(self, device: int = 0):
        h = ctypes.c_void_p()
        rc = lib().cav_ctx_create(device, ctypes.byref(h))
        if rc != 0:
            raise _lib.CavError(rc, f"cav_ctx_create(device={device}) failed")
        self.handle = h
        self._children = weakref.WeakSet()
        _LIVE.add(self)

    def own(self, obj) -> None:
        """Registers an object whose close() must run before this context's."""
        self._children.add(obj)

    @property
    def stream(self) -> int:
        return lib().cav_ctx_stream(self.handle) or 0

    def synchronize(self) -> None:
        check(self.handle, lib().cav_ctx_synchronize(self.handle))

    def close(self) -> None:
        if self.handle:
            # servers first (they hold scenes), then scenes (they own BVHs), then the rest
            for child in sorted(self._children, key=lambda c: getattr(c, "close_rank", 2)):
                child.close()
            lib().cav_ctx_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class BVH:
    """A built BVH living in device memory (cav_bvh)."""

    def __init__(self, ctx: Context, handle, n: int):
        self.ctx = ctx
        self.handle = handle
        self.n = n
        ctx.own(self)

    @staticmethod
    def build(ctx: Context, triangles, scene_bounds=None, optimization_rounds: int = 0,
              morton_bits: int = 30, tri_mesh=None, device: bool = False) -> "BVH":
        """LBVH build + ``optimization_rounds`` treelet rounds. ``triangles`` is a
        float32[N,9] array (host) or, with device=True, a device pointer/tensor."""
        if device:
            n = int(triangles.shape[0]) if hasattr(triangles, "shape") else None
            if n is None:
                raise ValueError("device build needs a tensor with a shape")
        else:
            triangles = np.ascontiguousarray(triangles, np.float32)
            n = len(triangles)
            if tri_mesh is not None:
                tri_mesh = np.ascontiguousarray(tri_mesh, np.int32)
        sb = None if scene_bounds is None else np.ascontiguousarray(scene_bounds, np.float32)
        h = ctypes.c_void_p()
        fn = lib().cav_bvh_build_device if device else lib().cav_bvh_build
        check(ctx.handle, fn(ctx.handle, _ptr(triangles), n, _ptr(sb), _ptr(tri_mesh),
                             morton_bits, optimization_rounds, ctypes.byref(h)))
        return BVH(ctx, h, n)

    def rebuild(self, triangles_dev, scene_bounds=None, optimization_rounds: int = 0,
                morton_bits: int = 30) -> None:
        sb = None if scene_bounds is None else np.ascontiguousarray(scene_bounds, np.float32)
        check(self.ctx.handle, lib().cav_bvh_rebuild_device(
            self.ctx.handle, self.handle, _ptr(triangles_dev), _ptr(sb), morton_bits,
            optimization_rounds))

    def stats(self) -> dict:
        s = _lib.BvhStats()
        check(self.ctx.handle, lib().cav_bvh_stats_get(self.handle, ctypes.byref(s)))
        return {k: getattr(s, k) for k, _ in s._fields_}

    def kernel_times(self) -> list:
        """[(kernel name, ms)] of the last build, in launch order."""
        k = lib().cav_bvh_kernel_times(self.handle, None, None, 0)
        names = (ctypes.c_char_p * max(k, 1))()
        ms = np.zeros(max(k, 1), np.float32)
        lib().cav_bvh_kernel_times(self.handle, names, _ptr(ms), k)
        return [(names[i].decode(), float(ms[i])) for i in range(k)]

    def export(self) -> dict:
        """Node arrays in the unified 2N-1 id space plus sorted codes and order."""
        m = 2 * self.n - 1
        out = {
            "box": np.empty((m, 6), np.float32), "sa": np.empty(m, np.float32),
            "cost": np.empty(m, np.float32), "left": np.empty(m, np.int32),
            "right": np.empty(m, np.int32), "parent": np.empty(m, np.int32),
            "prim": np.empty(m, np.int32), "codes": np.empty(self.n, np.uint64),
            "order": np.empty(self.n, np.int32),
        }
        check(self.ctx.handle, lib().cav_bvh_export(
            self.ctx.handle, self.handle, *[_ptr(out[k]) for k in (
                "box", "sa", "cost", "left", "right", "parent", "prim", "codes", "order")]))
        return out

    def intersect(self, rays) -> dict:
        """Closest hit per ray: prim (-1 = miss), t, u, v."""
        rays = np.ascontiguousarray(rays, np.float32)
        r = len(rays)
        out = {k: np.empty(r, np.float32) for k in ("t", "u", "v")}
        out["prim"] = np.empty(r, np.int32)
        check(self.ctx.handle, lib().cav_trace_closest(
            self.ctx.handle, self.handle, _ptr(rays), r, _ptr(out["prim"]), _ptr(out["t"]),
            _ptr(out["u"]), _ptr(out["v"])))
        return out

    def test_visibility(self, rays, limit, src_mesh=None, dst_mesh=None) -> np.ndarray:
        """1 where nothing blocks the ray before its limit (+inf = no limit)."""
        rays = np.ascontiguousarray(rays, np.float32)
        limit = np.ascontiguousarray(limit, np.float32)
        r = len(rays)
        vis = np.empty(r, np.uint8)
        src = None if src_mesh is None else np.ascontiguousarray(src_mesh, np.int32)
        dst = None if dst_mesh is None else np.ascontiguousarray(dst_mesh, np.int32)
        check(self.ctx.handle, lib().cav_trace_occluded(
            self.ctx.handle, self.handle, _ptr(rays), _ptr(limit), _ptr(src), _ptr(dst), r,
            _ptr(vis)))
        return vis

    def trace_counters(self, reset: bool = False) -> dict:
        """Stack overflows and deepest stacks of this BVH's standalone traces."""
        out = np.zeros(4, np.uint32)
        check(self.ctx.handle, lib().cav_trace_counters(self.ctx.handle, self.handle, _ptr(out), 1 if reset else 0))
        return {"overflow_closest": int(out[0]), "overflow_occluded": int(out[1]),
                "max_stack_closest": int(out[2]), "max_stack_occluded": int(out[3])}

    def trace_fetches(self, reset: bool = False) -> dict:
        """64-byte records (node visits + triangle tests) fetched by the
        closest-hit and any-hit launches since the last reset, and the rays
        of the last any-hit / closest-hit launch finished by the cooperative
        passes (and of those, traced again on one lane)."""
        out = np.zeros(6, np.uint64)
        check(self.ctx.handle, lib().cav_trace_fetches(self.ctx.handle, self.handle, _ptr(out), 1 if reset else 0))
        return {"closest": int(out[0]), "occluded": int(out[1]), "spilled_rays": int(out[2]),
                "respilled_rays": int(out[3]), "long_closest_rays": int(out[4]),
                "long_closest_retraced": int(out[5])}

    def max_depth(self) -> int:
        """Deepest leaf, in edges from the root."""
        d = ctypes.c_int32()
        check(self.ctx.handle, lib().cav_bvh_max_depth(self.ctx.handle, self.handle, ctypes.byref(d)))
        return int(d.value)

    def close(self) -> None:
        if self.handle:
            lib().cav_bvh_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
