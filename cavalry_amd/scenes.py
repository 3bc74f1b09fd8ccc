"""Seeded synthetic workloads for the LBVH, traversal and render paths.

Every generator draws from ``numpy.random.Generator(PCG64(seed))`` so the same
seed gives the same float32 arrays on every machine with this numpy.

Triangle soups are ``float32[N, 9]`` rows ``v0.xyz, v1.xyz, v2.xyz``; rays are
``float32[R, 6]`` rows ``origin.xyz, direction.xyz``.
"""
from __future__ import annotations

import numpy as np


def _rng(seed: int, stream: int = 0) -> np.random.Generator:
    return np.random.Generator(np.random.PCG64([seed, stream]))


def _unit_sphere(rng: np.random.Generator, n: int) -> np.ndarray:
    """Uniform directions on S^2 (z uniform in [-1,1], phi uniform)."""
    z = rng.uniform(-1.0, 1.0, n)
    phi = rng.uniform(0.0, 2.0 * np.pi, n)
    r = np.sqrt(np.maximum(0.0, 1.0 - z * z))
    return np.stack([r * np.cos(phi), r * np.sin(phi), z], axis=1)


def gmm_triangles(n: int, seed: int = 1, components: int = 8, edge: float = 1e-3) -> np.ndarray:
    """Random triangle soup: centroids from a Gaussian mixture in the unit cube,
    vertex offsets of log-normal length (mu = ln edge, sigma = 0.7; edge = 1e-3
    by default) and uniform orientation. Returns float32[n, 9]."""
    rng = _rng(seed, 0)
    w = rng.uniform(0.5, 1.5, components)
    w /= w.sum()
    means = rng.uniform(0.15, 0.85, (components, 3))
    sig = rng.uniform(0.03, 0.12, components)
    comp = rng.choice(components, size=n, p=w)
    c = means[comp] + rng.standard_normal((n, 3)) * sig[comp, None]
    np.clip(c, 0.0, 1.0, out=c)
    out = np.empty((n, 9), dtype=np.float32)
    for j in range(3):
        length = rng.lognormal(np.log(edge), 0.7, n)
        u = _unit_sphere(rng, n)
        out[:, 3 * j:3 * j + 3] = (c + length[:, None] * u).astype(np.float32)
    return out


def random_rays(n: int, seed: int = 1, stream: int = 1, lo: float = -0.5,
                hi: float = 1.5) -> np.ndarray:
    """Origins uniform in [lo,hi]^3, directions uniform on S^2. float32[n, 6]."""
    rng = _rng(seed, stream)
    o = rng.uniform(lo, hi, (n, 3))
    d = _unit_sphere(rng, n)
    return np.concatenate([o, d], axis=1).astype(np.float32)


def shadow_limits(n: int, seed: int = 1, stream: int = 3) -> np.ndarray:
    """Shadow-ray distance limits uniform in (0, 2*sqrt(3)). float32[n]."""
    rng = _rng(seed, stream)
    return rng.uniform(0.0, 2.0 * np.sqrt(3.0), n).astype(np.float32)


# ---------------------------------------------------------------------------
# Scenes for the renderer

MESH_DTYPE = np.dtype([
    ("first_tri", "<i4"), ("num_tris", "<i4"), ("material", "<i4"), ("texture", "<i4"),
    ("kd", "<f4", 3), ("ks", "<f4", 3), ("kr", "<f4", 3), ("kt", "<f4", 3),
    ("roughness", "<f4"), ("ior", "<f4"), ("emission", "<f4", 3),
    ("is_light", "<i4"), ("reverse_orientation", "<i4"),
])
assert MESH_DTYPE.itemsize == 92  # cav_mesh in include/cavalry.h

MATTE, MIRROR, GLASS, PLASTIC = 0, 1, 4, 5


class SceneArrays:
    """A renderable scene: triangle soup, meshes (contiguous triangle ranges)
    and a LookAt camera."""

    def __init__(self, tris, meshes, eye, center, up, fov, width, height, max_depth=8):
        self.tris = np.ascontiguousarray(tris, np.float32)
        self.meshes = meshes
        self.eye = np.asarray(eye, np.float32)
        self.center = np.asarray(center, np.float32)
        self.up = np.asarray(up, np.float32)
        self.fov = float(fov)
        self.width = int(width)
        self.height = int(height)
        self.max_depth = int(max_depth)

    @property
    def tri_mesh(self) -> np.ndarray:
        out = np.empty(len(self.tris), np.int32)
        for i, m in enumerate(self.meshes):
            out[m["first_tri"]:m["first_tri"] + m["num_tris"]] = i
        return out


class _Builder:
    def __init__(self):
        self.tris = []
        self.meshes = []

    def mesh(self, tris, material=MATTE, kd=(0.5, 0.5, 0.5), ks=(0.25, 0.25, 0.25),
             kr=(1, 1, 1), kt=(1, 1, 1), roughness=0.1, ior=1.5, emission=None):
        tris = np.asarray(tris, np.float32).reshape(-1, 9)
        m = np.zeros((), MESH_DTYPE)
        m["first_tri"] = sum(len(t) for t in self.tris)
        m["num_tris"] = len(tris)
        m["material"] = material
        m["texture"] = -1
        m["kd"], m["ks"], m["kr"], m["kt"] = kd, ks, kr, kt
        m["roughness"], m["ior"] = roughness, ior
        if emission is not None:
            m["emission"] = emission
            m["is_light"] = 1
        self.tris.append(tris)
        self.meshes.append(m)

    def arrays(self):
        return np.concatenate(self.tris), np.array(self.meshes, MESH_DTYPE)


def quad(p0, p1, p2, p3):
    """Two triangles (p0,p1,p2), (p0,p2,p3); normal = (p1-p0) x (p2-p0)."""
    return [list(p0) + list(p1) + list(p2), list(p0) + list(p2) + list(p3)]


def split_quad(o, u, v, n=2):
    """Parallelogram o + s*u + t*v split into n quads along u."""
    o, u, v = (np.asarray(a, np.float64) for a in (o, u, v))
    out = []
    for k in range(n):
        a = o + u * (k / n)
        b = o + u * ((k + 1) / n)
        out += quad(a, b, b + v, a + v)
    return out


def box_tris(lo, hi):
    """Axis-aligned box, 12 triangles, outward normals."""
    x0, y0, z0 = lo
    x1, y1, z1 = hi
    t = []
    t += quad((x0, y0, z0), (x1, y0, z0), (x1, y0, z1), (x0, y0, z1))  # -y
    t += quad((x0, y1, z0), (x0, y1, z1), (x1, y1, z1), (x1, y1, z0))  # +y
    t += quad((x0, y0, z0), (x0, y1, z0), (x1, y1, z0), (x1, y0, z0))  # -z
    t += quad((x0, y0, z1), (x1, y0, z1), (x1, y1, z1), (x0, y1, z1))  # +z
    t += quad((x0, y0, z0), (x0, y0, z1), (x0, y1, z1), (x0, y1, z0))  # -x
    t += quad((x1, y0, z0), (x1, y1, z0), (x1, y1, z1), (x1, y0, z1))  # +x
    return t


def uv_sphere(center, r, nphi=64, ntheta=32):
    """UV sphere with pole fans: 2*nphi*(ntheta-2) + 2*nphi triangles."""
    c = np.asarray(center, np.float64)

    def p(i, j):
        th = np.pi * i / ntheta
        ph = 2 * np.pi * j / nphi
        return c + r * np.array([np.sin(th) * np.cos(ph), np.cos(th), np.sin(th) * np.sin(ph)])

    t = []
    for j in range(nphi):
        t.append(list(p(0, 0)) + list(p(1, j + 1)) + list(p(1, j)))
    for i in range(1, ntheta - 1):
        for j in range(nphi):
            a, b, cc, d = p(i, j), p(i, j + 1), p(i + 1, j + 1), p(i + 1, j)
            t.append(list(a) + list(b) + list(cc))
            t.append(list(a) + list(cc) + list(d))
    for j in range(nphi):
        t.append(list(p(ntheta, 0)) + list(p(ntheta - 1, j)) + list(p(ntheta - 1, j + 1)))
    return t


def cornell_box(seed: int = 1, width: int = 512, height: int = 512, max_depth: int = 8,
                materials: str = "matte") -> SceneArrays:
    """Cornell-style box: unit box open at z=0 (5 walls x 2 quads), a tall and a
    short box, a 64x32 UV sphere (3968 triangles) and a 0.25x0.25 ceiling light
    of radiance 15 facing down. Albedos U[0.1,0.9] per mesh from the seed.
    materials="mixed" makes the boxes plastic/mirror and the sphere glass."""
    rng = _rng(seed, 7)
    b = _Builder()

    def albedo():
        return tuple(rng.uniform(0.1, 0.9, 3))

    b.mesh(split_quad((0, 0, 1), (1, 0, 0), (0, 0, -1)), kd=albedo())     # floor, +y
    b.mesh(split_quad((0, 1, 0), (1, 0, 0), (0, 0, 1)), kd=albedo())      # ceiling, -y
    b.mesh(split_quad((0, 0, 1), (1, 0, 0), (0, 1, 0)), kd=albedo())      # back, -z
    b.mesh(split_quad((0, 0, 0), (0, 0, 1), (0, 1, 0)), kd=albedo())      # left, +x
    b.mesh(split_quad((1, 0, 1), (0, 0, -1), (0, 1, 0)), kd=albedo())     # right, -x
    tall = box_tris((0.53, 0.0, 0.45), (0.83, 0.6, 0.75))
    short = box_tris((0.17, 0.0, 0.2), (0.47, 0.3, 0.5))
    sph = uv_sphere((0.5, 0.72, 0.3), 0.12)
    if materials == "mixed":
        b.mesh(tall, material=PLASTIC, kd=albedo(), ks=(0.3, 0.3, 0.3), roughness=0.2)
        b.mesh(short, material=MIRROR, kr=(0.9, 0.9, 0.9))
        b.mesh(sph, material=GLASS, kr=(1, 1, 1), kt=(1, 1, 1), ior=1.5)
    else:
        b.mesh(tall, kd=albedo())
        b.mesh(short, kd=albedo())
        b.mesh(sph, kd=albedo())
    light = quad((0.375, 0.999, 0.375), (0.625, 0.999, 0.375), (0.625, 0.999, 0.625),
                 (0.375, 0.999, 0.625))
    b.mesh(light, kd=albedo(), emission=(15.0, 15.0, 15.0))
    tris, meshes = b.arrays()
    return SceneArrays(tris, meshes, (0.5, 0.5, -1.3), (0.5, 0.5, 0.5), (0, 1, 0), 40.0,
                       width, height, max_depth)
