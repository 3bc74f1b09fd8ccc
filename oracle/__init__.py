"""CPU oracle loader (test infrastructure only).

Only tests/, ``__graft_entry__.smoke()`` and bench.py's cpu_baseline leg may
import this package. It restates the reference algorithms on the CPU so the HIP
kernels can be checked against it; the product never calls it.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_SO = os.path.join(_HERE, "_build", "libcavalry_oracle.so")
_lib = None


def build() -> None:
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib() -> ctypes.CDLL:
    global _lib
    if _lib is None:
        if not os.path.exists(_SO):
            build()
        _lib = ctypes.CDLL(_SO)
    return _lib


def _p(a):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


def scene_bounds(tris: np.ndarray) -> np.ndarray:
    out = np.empty(6, np.float32)
    lib().ref_scene_bounds(_p(tris), ctypes.c_int(len(tris)), _p(out))
    return out


def morton(tris: np.ndarray, bits: int, scene6=None) -> np.ndarray:
    tris = np.ascontiguousarray(tris, np.float32)
    if scene6 is None:
        scene6 = scene_bounds(tris)
    out = np.empty(len(tris), np.uint64)
    lib().ref_morton(_p(tris), ctypes.c_int(len(tris)), _p(np.asarray(scene6, np.float32)),
                     ctypes.c_int(bits), _p(out))
    return out


class Nodes:
    """BVH nodes in the reference's unified id space (2N-1 nodes)."""

    def __init__(self, n: int):
        m = 2 * n - 1
        self.n = n
        self.box = np.zeros((m, 6), np.float32)
        self.sa = np.zeros(m, np.float32)
        self.cost = np.zeros(m, np.float32)
        self.left = np.zeros(m, np.int32)
        self.right = np.zeros(m, np.int32)
        self.parent = np.zeros(m, np.int32)
        self.prim = np.zeros(m, np.int32)

    def copy(self) -> "Nodes":
        c = Nodes(self.n)
        for k in ("box", "sa", "cost", "left", "right", "parent", "prim"):
            setattr(c, k, getattr(self, k).copy())
        return c


def lbvh_build(tris: np.ndarray, bits: int = 30, scene6=None):
    """Returns (sorted_codes, order, Nodes) before any treelet round."""
    tris = np.ascontiguousarray(tris, np.float32)
    n = len(tris)
    nd = Nodes(n)
    codes = np.empty(n, np.uint64)
    order = np.empty(n, np.int32)
    rc = lib().ref_lbvh_build(
        _p(tris), ctypes.c_int(n), ctypes.c_int(bits),
        None if scene6 is None else _p(np.asarray(scene6, np.float32)),
        _p(codes), _p(order), _p(nd.box), _p(nd.sa), _p(nd.cost), _p(nd.left), _p(nd.right),
        _p(nd.parent), _p(nd.prim))
    if rc != 0:
        raise ValueError(f"ref_lbvh_build failed ({rc})")
    return codes, order, nd


def treelet_optimize(nd: Nodes, rounds: int) -> int:
    done = np.zeros(1, np.int64)
    rc = lib().ref_treelet_optimize(ctypes.c_int(nd.n), ctypes.c_int(rounds), _p(nd.box),
                                    _p(nd.sa), _p(nd.cost), _p(nd.left), _p(nd.right),
                                    _p(nd.parent), _p(nd.prim), _p(done))
    if rc != 0:
        raise ValueError(f"ref_treelet_optimize failed ({rc})")
    return int(done[0])


def trace_closest(nd: Nodes, tris: np.ndarray, rays: np.ndarray, stats: bool = False):
    rays = np.ascontiguousarray(rays, np.float32)
    r = len(rays)
    prim = np.empty(r, np.int32)
    t = np.empty(r, np.float32)
    u = np.empty(r, np.float32)
    v = np.empty(r, np.float32)
    vis = np.empty(r, np.int32) if stats else None
    dep = np.empty(r, np.int32) if stats else None
    overflow = lib().ref_trace_closest(
        _p(nd.box), _p(nd.left), _p(nd.right), _p(nd.prim), _p(tris), _p(rays),
        ctypes.c_int64(r), _p(prim), _p(t), _p(u), _p(v), _p(vis), _p(dep))
    out = {"prim": prim, "t": t, "u": u, "v": v, "overflow": overflow}
    if stats:
        out["visits"] = vis
        out["depth"] = dep
    return out


def trace_occluded(nd: Nodes, tris: np.ndarray, rays: np.ndarray, limit: np.ndarray,
                   mesh=None, src=None, dst=None) -> np.ndarray:
    rays = np.ascontiguousarray(rays, np.float32)
    r = len(rays)
    vis = np.empty(r, np.uint8)
    lib().ref_trace_occluded(
        _p(nd.box), _p(nd.left), _p(nd.right), _p(nd.prim), _p(tris), _p(mesh), _p(rays),
        _p(np.ascontiguousarray(limit, np.float32)), _p(src), _p(dst), ctypes.c_int64(r),
        _p(vis))
    return vis


def halton(index: int, dim: int) -> float:
    f = lib().ref_halton
    f.restype = ctypes.c_float
    return float(f(ctypes.c_uint64(index), ctypes.c_int(dim)))


def halton_table():
    primes = np.empty(1024, np.uint32)
    start = np.empty(1024, np.uint32)
    lib().ref_halton_table(_p(primes), _p(start), None, ctypes.c_int(0))
    total = int(start[-1] + primes[-1])
    perm = np.empty(total, np.uint16)
    lib().ref_halton_table(_p(primes), _p(start), _p(perm), ctypes.c_int(total))
    return primes, start, perm


def mesh_areas(tris, meshes) -> np.ndarray:
    out = np.empty(len(meshes), np.float32)
    lib().ref_mesh_areas(_p(np.ascontiguousarray(tris, np.float32)), _p(meshes),
                         ctypes.c_int(len(meshes)), _p(out))
    return out


class _Surface(ctypes.Structure):
    """ref_surface of ref_api.h"""
    _fields_ = [("tri_normals", ctypes.c_void_p), ("mesh_normals", ctypes.c_void_p), ("tri_uv", ctypes.c_void_p),
                ("texels", ctypes.c_void_p), ("tex_offset", ctypes.c_void_p), ("tex_size", ctypes.c_void_p),
                ("env", ctypes.c_void_p), ("nenv", ctypes.c_int32), ("light_order", ctypes.c_void_p),
                ("nlight_order", ctypes.c_int32), ("env_background", ctypes.c_int32)]


def _surface(scene):
    """A ref_surface pointer for the scene's normals, texture coordinates and
    textures (the arrays are kept alive on the scene object)."""
    def arr(name, dt):
        a = getattr(scene, name, None)
        return None if a is None else np.ascontiguousarray(a, dt)

    nrm, flags, uv = arr("normals", np.float32), arr("mesh_normals", np.uint8), arr("uv", np.float32)
    textures = getattr(scene, "textures", None) or []
    if int(np.max(scene.meshes["texture"], initial=-1)) >= len(textures):
        raise ValueError("a mesh names a missing texture")
    texels = sizes = offs = None
    if textures:
        texels = np.concatenate([np.ascontiguousarray(t, np.uint8).ravel() for t in textures])
        sizes = np.array([[t.shape[1], t.shape[0]] for t in textures], np.int32)
        offs = np.concatenate([[0], np.cumsum(sizes[:, 0].astype(np.int64) * sizes[:, 1])[:-1]]).astype(np.int64)
    env = getattr(scene, "env_lights", None)
    env = None if env is None or len(env) == 0 else np.ascontiguousarray(env)
    if env is not None and int(np.max(env["texture"], initial=-1)) >= len(textures):
        raise ValueError("an environment light names a missing texture")
    order = arr("light_order", np.int32)
    keep = (nrm, flags, uv, texels, offs, sizes, env, order)
    ptr = [None if a is None else a.ctypes.data for a in keep]
    s = _Surface(*ptr[:6], ptr[6], 0 if env is None else len(env), ptr[7], 0 if order is None else len(order),
                 int(getattr(scene, "env_background", -1)))
    scene._oracle_keep = (keep, s)
    return ctypes.byref(s)


def tex_fetch(rgba, u: float, v: float) -> np.ndarray:
    """Bilinear, wrap-addressed read of an RGBA8 texture [H, W, 4] at (u, v)."""
    t = np.ascontiguousarray(rgba, np.uint8)
    out = np.zeros(3, np.float32)
    lib().ref_tex_fetch(_p(t), ctypes.c_int(t.shape[1]), ctypes.c_int(t.shape[0]), ctypes.c_float(u),
                        ctypes.c_float(v), _p(out))
    return out


def render_path(scene, spp: int, camera, bvh_rounds: int = 0, morton_bits: int = 30, crop=None,
                max_depth: int | None = None, samples: bool = False, bvh=None):
    """`path` integrator on the CPU. Returns (film_rgb [H,W,3], film_w [H,W],
    rays_traced). camera is a cavalry_amd Camera (cav_camera); the image size
    is the camera's. crop = (x0, y0, x1, y1) renders only those pixels of the
    full image (same sample indices, so the crop equals that part of the full
    film) and returns films of the crop's size; samples=True adds, per crop
    pixel and sample, radiance and film position [h, w, spp, 5]. bvh: node
    arrays from an earlier lbvh_build + treelet_optimize of the same triangles."""
    tris = np.ascontiguousarray(scene.tris, np.float32)
    if bvh is None:
        _, _, nd = lbvh_build(tris, morton_bits)
        if bvh_rounds:
            treelet_optimize(nd, bvh_rounds)
    else:
        nd = bvh
    w, h = int(camera.width), int(camera.height)
    x0, y0, x1, y1 = crop if crop is not None else (0, 0, w, h)
    rgb = np.zeros((y1 - y0, x1 - x0, 3), np.float32)
    wt = np.zeros((y1 - y0, x1 - x0), np.float32)
    smp = np.zeros((y1 - y0, x1 - x0, spp, 5), np.float32) if samples else None
    rays = np.zeros(1, np.int64)
    cam16 = np.ctypeslib.as_array(camera.cam_to_world).astype(np.float32)
    tri_mesh = scene.tri_mesh
    rect = np.int32([x0, y0, x1, y1])
    depth = scene.max_depth if max_depth is None else max_depth
    rc = lib().ref_render_path_crop(
        _p(tris), ctypes.c_int(len(tris)), _p(tri_mesh), _p(scene.meshes), ctypes.c_int(len(scene.meshes)),
        _p(nd.box), _p(nd.left), _p(nd.right), _p(nd.prim), _surface(scene), _p(cam16),
        ctypes.c_float(camera.depth), ctypes.c_int(w), ctypes.c_int(h), ctypes.c_int(spp),
        ctypes.c_int(depth), _p(rect), _p(rgb), _p(wt), _p(smp) if samples else None, _p(rays))
    if rc != 0:
        raise ValueError(f"ref_render_path failed ({rc})")
    if samples:
        return rgb, wt, int(rays[0]), smp
    return rgb, wt, int(rays[0])


def render_direct(scene, spp: int, camera, bvh_rounds: int = 0, morton_bits: int = 30):
    """`directlighting` integrator on the CPU. Returns (film_rgb, film_w, rays)."""
    tris = np.ascontiguousarray(scene.tris, np.float32)
    _, _, nd = lbvh_build(tris, morton_bits)
    if bvh_rounds:
        treelet_optimize(nd, bvh_rounds)
    w, h = scene.width, scene.height
    rgb = np.zeros((h, w, 3), np.float32)
    wt = np.zeros((h, w), np.float32)
    rays = np.zeros(1, np.int64)
    cam16 = np.ctypeslib.as_array(camera.cam_to_world).astype(np.float32)
    rc = lib().ref_render_direct(
        _p(tris), ctypes.c_int(len(tris)), _p(scene.tri_mesh), _p(scene.meshes), ctypes.c_int(len(scene.meshes)),
        _p(nd.box), _p(nd.left), _p(nd.right), _p(nd.prim), _surface(scene), _p(cam16),
        ctypes.c_float(camera.depth), ctypes.c_int(w), ctypes.c_int(h), ctypes.c_int(spp), _p(rgb), _p(wt),
        _p(rays))
    if rc != 0:
        raise ValueError(f"ref_render_direct failed ({rc})")
    return rgb, wt, int(rays[0])


def render_rlpath(scene, spp: int, camera, bvh_rounds: int = 0, morton_bits: int = 30,
                  q_grid: int = 32, alpha_scale: float = 1.0, max_depth: int | None = None,
                  spp_per_round: int = 0, bvh=None):
    """`rlpath` integrator on the CPU, bounce by bounce over the samples in
    sample order, in rounds of spp_per_round samples per pixel (0: one round).
    bvh: node arrays from an earlier lbvh_build + treelet_optimize of the same
    triangles. The image size is the camera's.
    Returns (film_rgb [H,W,3], film_w [H,W], rays_traced)."""
    tris = np.ascontiguousarray(scene.tris, np.float32)
    if bvh is None:
        _, _, nd = lbvh_build(tris, morton_bits)
        if bvh_rounds:
            treelet_optimize(nd, bvh_rounds)
    else:
        nd = bvh
    w, h = int(camera.width), int(camera.height)
    rgb = np.zeros((h, w, 3), np.float32)
    wt = np.zeros((h, w), np.float32)
    rays = np.zeros(1, np.int64)
    cam16 = np.ctypeslib.as_array(camera.cam_to_world).astype(np.float32)
    s6 = scene_bounds(tris)
    depth = scene.max_depth if max_depth is None else max_depth
    rc = lib().ref_render_rlpath(
        _p(tris), ctypes.c_int(len(tris)), _p(scene.tri_mesh), _p(scene.meshes), ctypes.c_int(len(scene.meshes)),
        _p(nd.box), _p(nd.left), _p(nd.right), _p(nd.prim), _surface(scene), _p(s6), ctypes.c_int(q_grid),
        ctypes.c_float(alpha_scale), _p(cam16), ctypes.c_float(camera.depth), ctypes.c_int(w), ctypes.c_int(h),
        ctypes.c_int(spp), ctypes.c_int(depth), ctypes.c_int(spp_per_round), _p(rgb), _p(wt), _p(rays))
    if rc != 0:
        raise ValueError(f"ref_render_rlpath failed ({rc})")
    return rgb, wt, int(rays[0])


def rl_dir_to_cell(d) -> int:
    """QEntry::dirToCellIndex of one direction."""
    return int(lib().ref_rl_dir_to_cell(_p(np.ascontiguousarray(d, np.float32))))


def rl_cell_dir(rx: float, ry: float, cell: int) -> np.ndarray:
    """QEntry::sampleDirectionInCell."""
    out = np.empty(3, np.float32)
    lib().ref_rl_cell_dir(ctypes.c_float(rx), ctypes.c_float(ry), ctypes.c_int(cell), _p(out))
    return out


def rl_finish_q(Q, newQ, cnt, total, alpha_scale: float = 1.0) -> None:
    """finishUpdateQ + averageQ in place on [E, 128] arrays (float32, float32,
    int32, int32); newQ and cnt are cleared."""
    for a, dt in ((Q, np.float32), (newQ, np.float32), (cnt, np.int32), (total, np.int32)):
        assert a.dtype == dt and a.flags.c_contiguous and a.shape == Q.shape and a.shape[1] == 128
    lib().ref_rl_finish_q(_p(Q), _p(newQ), _p(cnt), _p(total), ctypes.c_int(Q.shape[0]),
                          ctypes.c_float(alpha_scale))


def rl_finish_light(ldist, lres, lcount, lnew, lnewc) -> None:
    """finishLightingTraining in place on [E, L] arrays."""
    for a, dt in ((ldist, np.float32), (lres, np.float32), (lcount, np.int32), (lnew, np.float32),
                  (lnewc, np.int32)):
        assert a.dtype == dt and a.flags.c_contiguous and a.shape == ldist.shape
    lib().ref_rl_finish_light(_p(ldist), _p(lres), _p(lcount), _p(lnew), _p(lnewc),
                              ctypes.c_int(ldist.shape[0]), ctypes.c_int(ldist.shape[1]))


def bsdf_eval(mesh, wi, wo):
    """f(wi, wo) [n, 3] and pdf [n] of one mesh's BSDF in its local frame."""
    m = np.ascontiguousarray(mesh).copy()
    wi = np.ascontiguousarray(wi, np.float32).reshape(-1, 3)
    wo = np.ascontiguousarray(np.broadcast_to(np.asarray(wo, np.float32), wi.shape))
    f = np.zeros_like(wi)
    p = np.zeros(len(wi), np.float32)
    lib().ref_bsdf_eval(_p(m), _p(wi), _p(wo), ctypes.c_int(len(wi)), _p(f), _p(p))
    return f, p


def bsdf_sample(mesh, u, wo):
    """BSDF::sample for uniforms u [n, 2] and exitant directions wo: (wi, f, pdf)."""
    m = np.ascontiguousarray(mesh).copy()
    u = np.ascontiguousarray(u, np.float32).reshape(-1, 2)
    wo = np.ascontiguousarray(np.broadcast_to(np.asarray(wo, np.float32), (len(u), 3)))
    wi = np.zeros((len(u), 3), np.float32)
    f = np.zeros((len(u), 3), np.float32)
    p = np.zeros(len(u), np.float32)
    lib().ref_bsdf_sample(_p(m), _p(u), _p(wo), ctypes.c_int(len(u)), _p(wi), _p(f), _p(p))
    return wi, f, p


def env_eval(scene, e: int, dirs) -> np.ndarray:
    """EnvironmentMap::evaluateRay of scene.env_lights[e] for world directions [n, 3]."""
    d = np.ascontiguousarray(dirs, np.float32).reshape(-1, 3)
    out = np.zeros_like(d)
    lib().ref_env_eval(_surface(scene), ctypes.c_int(e), _p(d), ctypes.c_int(len(d)), _p(out))
    return out
