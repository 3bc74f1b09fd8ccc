"""Seeded synthetic workloads for the LBVH, traversal and render paths.

Every generator draws from ``numpy.random.Generator(PCG64(seed))`` so the same
seed gives the same float32 arrays on every machine with this numpy.

Triangle soups are ``float32[N, 9]`` rows ``v0.xyz, v1.xyz, v2.xyz``; rays are
``float32[R, 6]`` rows ``origin.xyz, direction.xyz``.
"""
from __future__ import annotations

import functools

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
    ("is_light", "<i4"), ("reverse_orientation", "<i4"), ("roughness_v", "<f4"),
    ("sphere", "<f4", 4),
])
assert MESH_DTYPE.itemsize == 112  # cav_mesh in include/cavalry.h

MATTE, MIRROR, SUBSTRATE, METAL, GLASS, PLASTIC = 0, 1, 2, 3, 4, 5

# an environment-type light (cav_env_light): radiance, texture index (-1:
# none) and the world -> light rotation, row-major
ENV_DTYPE = np.dtype([("color", "<f4", 3), ("texture", "<i4"), ("world_to_light", "<f4", 9)])
assert ENV_DTYPE.itemsize == 52


def env_light(color, texture: int = -1, world_to_light=None) -> np.ndarray:
    e = np.zeros((), ENV_DTYPE)
    e["color"] = color
    e["texture"] = texture
    e["world_to_light"] = np.eye(3).ravel() if world_to_light is None else np.asarray(world_to_light).ravel()
    return e


class SceneArrays:
    """A renderable scene: triangle soup, meshes (contiguous triangle ranges)
    and a LookAt camera."""

    def __init__(self, tris, meshes, eye, center, up, fov, width, height, max_depth=8, normals=None,
                 mesh_normals=None, uv=None, textures=None, env_lights=None, light_order=None,
                 env_background=-1):
        self.tris = np.ascontiguousarray(tris, np.float32)
        # optional per-vertex shading normals float32[N, 9] and per-mesh use flags
        self.normals = None if normals is None else np.ascontiguousarray(normals, np.float32)
        self.mesh_normals = None if mesh_normals is None else np.ascontiguousarray(mesh_normals, np.uint8)
        # optional per-vertex texture coordinates float32[N, 6] and RGBA8 textures
        # (uint8 [H, W, 4], row 0 at v = 0) that meshes["texture"] indexes (-1: none)
        self.uv = None if uv is None else np.ascontiguousarray(uv, np.float32)
        self.textures = [np.ascontiguousarray(t, np.uint8) for t in (textures or [])]
        # environment-type lights (ENV_DTYPE), the light list order (mesh id >= 0 or
        # -1 - k for env_lights[k]; None: area lights in mesh order, then these) and
        # the one that rays which miss see (-1: none)
        self.env_lights = (np.zeros(0, ENV_DTYPE) if env_lights is None
                           else np.ascontiguousarray(np.atleast_1d(np.asarray(env_lights, ENV_DTYPE))))
        self.light_order = None if light_order is None else np.ascontiguousarray(light_order, np.int32)
        self.env_background = int(env_background)
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
        self.uv = []

    def mesh(self, tris, material=MATTE, kd=(0.5, 0.5, 0.5), ks=(0.25, 0.25, 0.25),
             kr=(1, 1, 1), kt=(1, 1, 1), roughness=0.1, ior=1.5, emission=None, texture=-1, uv=None,
             roughness_v=0.0, sphere=None):
        tris = np.asarray(tris, np.float32).reshape(-1, 9)
        if callable(uv):
            uv = uv()
        self.uv.append(np.zeros((len(tris), 6), np.float32) if uv is None else np.asarray(uv, np.float32))
        m = np.zeros((), MESH_DTYPE)
        m["first_tri"] = sum(len(t) for t in self.tris)
        m["num_tris"] = len(tris)
        m["material"] = material
        m["texture"] = texture
        m["kd"], m["ks"], m["kr"], m["kt"] = kd, ks, kr, kt
        m["roughness"], m["ior"], m["roughness_v"] = roughness, ior, roughness_v
        if emission is not None:
            m["emission"] = emission
            m["is_light"] = 1
        if sphere is not None:
            m["sphere"] = sphere
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


def octa_sphere(center, radius) -> tuple[np.ndarray, np.ndarray, np.ndarray]:
    """The sphere shape of the scene format (ShapeMesh/Sphere.cu:3-90): an
    octahedron subdivided four times (2048 triangles), midpoints normalised,
    scaled by the radius and translated to the centre, with the unit points as
    vertex normals, in float32 as the .pbrt parser computes it for a
    `Translate` + `Shape "sphere"`. Returns (tris [2048, 9], normals [2048, 9],
    sphere = centre and the radius the reference derives from the transform)."""
    f = np.float32

    def normalize(v):
        d = (v[0] * v[0] + v[1] * v[1]) + v[2] * v[2]
        inv = f(1.0) / np.sqrt(d)
        return np.array([v[0] * inv, v[1] * inv, v[2] * inv], f)

    pts = [np.array(v, f) for v in ((1, 0, 0), (-1, 0, 0), (0, 1, 0), (0, -1, 0), (0, 0, 1), (0, 0, -1))]
    cur = [(0, 2, 4), (0, 2, 5), (0, 3, 4), (0, 3, 5), (1, 2, 4), (1, 2, 5), (1, 3, 4), (1, 3, 5)]
    for _ in range(4):
        nxt = []
        for i0, i1, i2 in cur:
            ids = []
            for a, b in ((i0, i1), (i0, i2), (i1, i2)):
                pts.append(normalize((pts[a] + pts[b]) * f(0.5)))
                ids.append(len(pts) - 1)
            i01, i02, i12 = ids
            nxt += [(i01, i02, i12), (i0, i01, i02), (i1, i01, i12), (i2, i02, i12)]
        cur = nxt
    unit = np.array(pts, f)
    c, r = np.asarray(center, f), f(radius)
    pos = unit * r + c
    nrm = np.array([normalize(v) for v in unit], f)
    idx = np.array(cur, np.int64)
    # shapeParams: |centre - T(radius, 0, 0)|, rounded as the transform does
    d = c - np.array([r + c[0], c[1], c[2]], f)
    rad = np.sqrt((d[0] * d[0] + d[1] * d[1]) + d[2] * d[2])
    return pos[idx].reshape(-1, 9), nrm[idx].reshape(-1, 9), np.array([c[0], c[1], c[2], rad], f)


def planar_uv(tris: np.ndarray, scale: float = 1.0) -> np.ndarray:
    """Texture coordinates by planar projection along each triangle's
    dominant normal axis: the other two coordinates times `scale`."""
    t = np.asarray(tris, np.float64).reshape(-1, 3, 3)
    n = np.abs(np.cross(t[:, 1] - t[:, 0], t[:, 2] - t[:, 0]))
    ax = np.argmax(n, axis=1)
    keep = np.array([[1, 2], [0, 2], [0, 1]])[ax]                 # [N, 2]
    uv = np.take_along_axis(t, keep[:, None, :].repeat(3, 1), axis=2) * scale
    return uv.reshape(-1, 6).astype(np.float32)


def spherical_uv(tris: np.ndarray, center) -> np.ndarray:
    """Texture coordinates by spherical projection about `center`:
    u = atan2(z, x) / 2pi + 1/2, v = asin(y / r) / pi + 1/2."""
    d = np.asarray(tris, np.float64).reshape(-1, 3, 3) - np.asarray(center, np.float64)
    r = np.linalg.norm(d, axis=2)
    u = np.arctan2(d[..., 2], d[..., 0]) / (2 * np.pi) + 0.5
    v = np.arcsin(np.clip(d[..., 1] / np.maximum(r, 1e-30), -1, 1)) / np.pi + 0.5
    return np.stack([u, v], 2).reshape(-1, 6).astype(np.float32)


@functools.lru_cache(maxsize=16)
def procedural_texture(kind: str, size: int, seed: int) -> np.ndarray:
    """RGBA8 [size, size, 4] pattern: "checker" (8 x 8 squares of two seeded
    colours), "stripes" (diagonal bands) or "noise" (value-noise marble).
    Cached: do not modify the returned array."""
    rng = _rng(seed, 13)
    c0 = rng.uniform(0.1, 0.9, 3)
    c1 = rng.uniform(0.1, 0.9, 3)
    y, x = np.mgrid[0:size, 0:size].astype(np.float64) / size
    if kind == "checker":
        w = ((np.floor(x * 8) + np.floor(y * 8)) % 2)
    elif kind == "stripes":
        w = 0.5 + 0.5 * np.sin(2 * np.pi * 6 * (x + y))
    elif kind == "noise":
        p = np.stack([x.ravel() * 8, y.ravel() * 8, np.zeros(x.size)], 1)
        w = (0.5 + 0.5 * np.sin(6 * x + 4 * _value_noise(p, seed).reshape(size, size))) ** 2
    else:
        raise ValueError(kind)
    rgb = c0 * (1 - w[..., None]) + c1 * w[..., None]
    out = np.empty((size, size, 4), np.uint8)
    out[..., :3] = np.clip(np.round(rgb * 255), 0, 255)
    out[..., 3] = 255
    return out


def sky_texture(w: int = 32, h: int = 16) -> np.ndarray:
    """RGBA8 [h, w, 4] sky for environment lights: a vertical blue-to-white
    gradient with a warm band at one azimuth (row 0 at theta = 0)."""
    y, x = np.mgrid[0:h, 0:w].astype(np.float64)
    t = (y + 0.5) / h
    rgb = np.stack([0.3 + 0.7 * t, 0.5 + 0.5 * t, np.ones_like(t)], -1)
    rgb[..., 0] += 0.5 * np.exp(-((x + 0.5) / w - 0.25) ** 2 / 0.005)
    out = np.empty((h, w, 4), np.uint8)
    out[..., :3] = np.clip(np.round(rgb / rgb.max() * 255), 0, 255)
    out[..., 3] = 255
    return out


def rotation_y(degrees: float) -> np.ndarray:
    a = np.radians(degrees)
    return np.array([[np.cos(a), 0, np.sin(a)], [0, 1, 0], [-np.sin(a), 0, np.cos(a)]])


def sky_plane(seed: int = 1, width: int = 64, height: int = 48, albedo: float = 0.5, sky: float = 1.0,
              max_depth: int = 8, black_lights: int = 0) -> SceneArrays:
    """A 20 x 20 matte floor (albedo `albedo`) under a constant environment
    light of radiance `sky`, nothing else; the camera looks down at the floor
    and the horizon. black_lights adds that many environment lights of
    radiance 0 (how the reference treats a LightSource it does not know)."""
    b = _Builder()
    b.mesh(quad((-10, 0, -10), (-10, 0, 10), (10, 0, 10), (10, 0, -10)), kd=(albedo, albedo, albedo))
    env = [env_light((sky, sky, sky))] + [env_light((0, 0, 0)) for _ in range(black_lights)]
    tris, meshes = b.arrays()
    return SceneArrays(tris, meshes, (0, 2, -3), (0, 0.3, 2), (0, 1, 0), 60.0, width, height, max_depth,
                       env_lights=env, env_background=0)


def sphere_light_plane(width: int = 64, height: int = 48, albedo: float = 0.5, radiance: float = 2.0,
                       radius: float = 0.7, max_depth: int = 2, light_kd: float = 0.0) -> SceneArrays:
    """A 20 x 20 matte floor (albedo `albedo`, normal +y) under one sphere
    light of radius `radius` centred at height 1 (the octahedral sphere
    shape, radiance `radiance`, diffuse albedo `light_kd`); the camera looks
    at the floor under the sphere from the side."""
    b = _Builder()
    b.mesh(quad((-10, 0, -10), (-10, 0, 10), (10, 0, 10), (10, 0, -10)), kd=(albedo, albedo, albedo))
    st, sn, ss = octa_sphere((0, 1, 0), radius)
    b.mesh(st, kd=(light_kd,) * 3, emission=(radiance,) * 3, sphere=ss)
    tris, meshes = b.arrays()
    normals = np.zeros_like(tris)
    normals[len(tris) - len(st):] = sn
    return SceneArrays(tris, meshes, (0, 0.6, -4), (0, 0, 0.5), (0, 1, 0), 50.0, width, height, max_depth,
                       normals=normals, mesh_normals=np.array([0, 1], np.uint8))


def cornell_box(seed: int = 1, width: int = 512, height: int = 512, max_depth: int = 8,
                materials: str = "matte", smooth: bool = False, textured: bool = False,
                env: bool = False, lights: str = "quad") -> SceneArrays:
    """Cornell-style box: unit box open at z=0 (5 walls x 2 quads), a tall and a
    short box, a 64x32 UV sphere (3968 triangles) and a 0.25x0.25 ceiling light
    of radiance 15 facing down. Albedos U[0.1,0.9] per mesh from the seed.
    materials="mixed" makes the boxes plastic/mirror and the sphere glass;
    materials="metal" makes the tall box anisotropic substrate, the short box
    copper (GGX alpha 0.2) and the sphere almost-mirror copper (alpha 0.04).
    smooth=True gives the sphere radial per-vertex normals (smooth shading).
    textured=True puts a checker on the floor, stripes on the back wall and
    value-noise marble on the tall box (64 x 64 textures, planar UVs; the
    matte ones with kd = 1), and spherical UVs on the sphere. env=True adds a
    textured sky (rotated 30 degrees about y, seen by rays leaving through the
    open front) and a black light, ordered sky, ceiling light, black.
    lights="sphere" replaces the ceiling quad with a sphere light (the
    octahedral sphere shape, radius 0.08 at (0.5, 0.85, 0.5), radiance 20,
    sampled within the cone it subtends); lights="both" keeps the quad too."""
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
    elif materials == "metal":
        copper = (0.955, 0.638, 0.538)
        b.mesh(tall, material=SUBSTRATE, kd=albedo(), ks=(0.2, 0.2, 0.2), roughness=0.3, roughness_v=0.1)
        b.mesh(short, material=METAL, ks=copper, roughness=0.2)
        b.mesh(sph, material=METAL, ks=copper, roughness=0.04)
    else:
        b.mesh(tall, kd=albedo())
        b.mesh(short, kd=albedo())
        b.mesh(sph, kd=albedo())
    light = quad((0.375, 0.999, 0.375), (0.625, 0.999, 0.375), (0.625, 0.999, 0.625),
                 (0.375, 0.999, 0.625))
    if lights not in ("quad", "sphere", "both"):
        raise ValueError(f"lights must be quad, sphere or both, not {lights!r}")
    ka = albedo()
    if lights != "sphere":
        b.mesh(light, kd=ka, emission=(15.0, 15.0, 15.0))
    sl_n = None
    if lights != "quad":
        sl, sl_n, sl_s = octa_sphere((0.5, 0.85, 0.5), 0.08)
        b.mesh(sl, kd=ka, emission=(20.0, 20.0, 20.0), sphere=sl_s)
    tris, meshes = b.arrays()
    normals = flags = None
    if smooth or sl_n is not None:
        normals = np.zeros_like(tris)
        flags = np.zeros(len(meshes), np.uint8)
    if sl_n is not None:
        m = meshes[-1]
        normals[m["first_tri"]:m["first_tri"] + m["num_tris"]] = sl_n
        flags[-1] = 1
    if smooth:
        m = meshes[7]  # the sphere
        t = tris[m["first_tri"]:m["first_tri"] + m["num_tris"]].reshape(-1, 3, 3).astype(np.float64)
        n = t - np.array([0.5, 0.72, 0.3])
        n /= np.linalg.norm(n, axis=2, keepdims=True)
        normals[m["first_tri"]:m["first_tri"] + m["num_tris"]] = n.reshape(-1, 9)
        flags[7] = 1
    uv = textures = None
    if textured:
        uv = planar_uv(tris, 2.0)
        m = meshes[7]
        uv[m["first_tri"]:m["first_tri"] + m["num_tris"]] = spherical_uv(
            tris[m["first_tri"]:m["first_tri"] + m["num_tris"]], (0.5, 0.72, 0.3))
        textures = [procedural_texture(k, 64, seed + i) for i, k in enumerate(("checker", "stripes", "noise"))]
        meshes["texture"][[0, 2, 5]] = [0, 1, 2]
        # a textured matte takes its colour from the texture alone, as a
        # .pbrt Kd is either a colour or a texture
        meshes["kd"][[k for k in (0, 2, 5) if meshes["material"][k] == MATTE]] = 1.0
    envs = order = None
    background = -1
    if env:
        textures = (textures or []) + [sky_texture()]
        envs = [env_light((0.8, 0.9, 1.0), len(textures) - 1, rotation_y(30).T), env_light((0, 0, 0))]
        order = [-1] + [i for i in range(len(meshes)) if meshes[i]["is_light"]] + [-2]
        background = 0
    return SceneArrays(tris, meshes, (0.5, 0.5, -1.3), (0.5, 0.5, 0.5), (0, 1, 0), 40.0,
                       width, height, max_depth, normals=normals, mesh_normals=flags, uv=uv, textures=textures,
                       env_lights=envs, light_order=order, env_background=background)


@functools.lru_cache(maxsize=4)
def icosphere(level: int) -> tuple[np.ndarray, np.ndarray]:
    """Unit icosphere subdivided `level` times: (vertices [V, 3] float64,
    faces [20 * 4^level, 3] int64), faces wound outward. Cached: do not
    modify the returned arrays."""
    t = (1.0 + 5 ** 0.5) / 2
    v = np.array([[-1, t, 0], [1, t, 0], [-1, -t, 0], [1, -t, 0], [0, -1, t], [0, 1, t], [0, -1, -t], [0, 1, -t],
                  [t, 0, -1], [t, 0, 1], [-t, 0, -1], [-t, 0, 1]], np.float64)
    v /= np.linalg.norm(v, axis=1, keepdims=True)
    f = np.array([[0, 11, 5], [0, 5, 1], [0, 1, 7], [0, 7, 10], [0, 10, 11], [1, 5, 9], [5, 11, 4], [11, 10, 2],
                  [10, 7, 6], [7, 1, 8], [3, 9, 4], [3, 4, 2], [3, 2, 6], [3, 6, 8], [3, 8, 9], [4, 9, 5],
                  [2, 4, 11], [6, 2, 10], [8, 6, 7], [9, 8, 1]], np.int64)
    for _ in range(level):
        e = np.concatenate([f[:, [0, 1]], f[:, [1, 2]], f[:, [2, 0]]])
        e.sort(axis=1)
        key = e[:, 0] * (len(v) + 1) + e[:, 1]
        uk, first, inv = np.unique(key, return_index=True, return_inverse=True)
        uniq = e[first]
        mid = v[uniq[:, 0]] + v[uniq[:, 1]]
        mid /= np.linalg.norm(mid, axis=1, keepdims=True)
        base = len(v)
        v = np.concatenate([v, mid])
        inv = inv.reshape(3, -1) + base
        a, b, c = f[:, 0], f[:, 1], f[:, 2]
        ab, bc, ca = inv[0], inv[1], inv[2]
        f = np.concatenate([np.stack([a, ab, ca], 1), np.stack([b, bc, ab], 1), np.stack([c, ca, bc], 1),
                            np.stack([ab, bc, ca], 1)])
    return v, f


def _value_noise(p: np.ndarray, seed: int) -> np.ndarray:
    """Trilinear value noise in [-1, 1] on the integer lattice (smoothstep
    weights, lattice values from a seeded hash)."""
    i = np.floor(p).astype(np.int64)
    fr = p - i
    w = fr * fr * (3 - 2 * fr)

    def h(ix, iy, iz):
        x = (ix * 73856093) ^ (iy * 19349663) ^ (iz * 83492791) ^ (seed * 2654435761)
        x = (x ^ (x >> 13)) * 1274126177
        return ((x ^ (x >> 16)) & 0xFFFF) / 32767.5 - 1.0

    out = 0.0
    for dx in (0, 1):
        for dy in (0, 1):
            for dz in (0, 1):
                wt = (w[:, 0] if dx else 1 - w[:, 0]) * (w[:, 1] if dy else 1 - w[:, 1]) * (
                    w[:, 2] if dz else 1 - w[:, 2])
                out = out + wt * h(i[:, 0] + dx, i[:, 1] + dy, i[:, 2] + dz)
    return out


def interior(n_tris: int, seed: int = 1, width: int = 1920, height: int = 1080, max_depth: int = 16,
             textures: bool = True, tex_size: int = 1024) -> SceneArrays:
    """Procedural interior (the wavefront-at-scale scene family): a 10 x 4 x 10
    room, 64 ceiling quad lights (two equal triangles each, L ~ U[5, 20]),
    displaced icospheres (level 6-8, radial value-noise displacement of
    amplitude U[0.02, 0.1] r) and boxes, about n_tris triangles in total.
    Materials per object: 40% matte, 30% plastic (alpha ~ U[0.05, 0.5]),
    15% mirror, 15% glass (index 1.5). textures=True: four procedural RGBA8
    textures of tex_size^2 (checker, stripes, two marbles) on the floor, the
    far wall and half of the matte / plastic objects (spherical UVs on the
    spheres, planar elsewhere; textured matte has kd = 1)."""
    rng = _rng(seed, 7)
    trng = _rng(seed, 11)  # texture choices: the geometry and materials do not depend on them
    b = _Builder()
    X, Y, Z = 10.0, 4.0, 10.0
    textures = [procedural_texture(k, tex_size, seed + i)
                for i, k in enumerate(("checker", "stripes", "noise", "noise"))] if textures else []

    def textured(tris, mat, uv):
        if textures and mat["material"] in (MATTE, PLASTIC) and trng.uniform() < 0.5:
            mat = dict(mat, texture=int(trng.integers(len(textures))), uv=uv)
            if mat["material"] == MATTE:
                mat["kd"] = (1.0, 1.0, 1.0)
        return mat

    def wall(tris, kd, tex):
        kw = dict(kd=kd)
        if textures and tex >= 0:
            kw = dict(kd=(1.0, 1.0, 1.0), texture=tex, uv=planar_uv(np.asarray(tris, np.float32), 0.5))
        b.mesh(tris, **kw)

    # the room's faces, normals pointing inward
    wall(split_quad((0, 0, 0), (0, 0, Z), (X, 0, 0), 4), (0.6, 0.6, 0.6), 0)          # floor, +y
    b.mesh(split_quad((0, Y, 0), (X, 0, 0), (0, 0, Z), 4), kd=(0.7, 0.7, 0.7))        # ceiling, -y
    b.mesh(split_quad((0, 0, 0), (X, 0, 0), (0, Y, 0), 2), kd=(0.5, 0.3, 0.3))        # z = 0, +z
    wall(split_quad((0, 0, Z), (0, Y, 0), (X, 0, 0), 2), (0.3, 0.5, 0.3), 1)          # z = Z, -z
    b.mesh(split_quad((0, 0, 0), (0, Y, 0), (0, 0, Z), 2), kd=(0.3, 0.3, 0.5))        # x = 0, +x
    b.mesh(split_quad((X, 0, 0), (0, 0, Z), (0, Y, 0), 2), kd=(0.5, 0.5, 0.3))        # x = X, -x
    for i in range(8):
        for j in range(8):
            cx, cz = (i + 0.5) * X / 8, (j + 0.5) * Z / 8
            s = 0.2
            L = float(rng.uniform(5, 20))
            b.mesh(quad((cx - s, Y - 0.01, cz - s), (cx + s, Y - 0.01, cz - s), (cx + s, Y - 0.01, cz + s),
                        (cx - s, Y - 0.01, cz + s)), kd=(0.5, 0.5, 0.5), emission=(L, L, L))

    def material():
        u = float(rng.uniform())
        kd = tuple(float(x) for x in rng.uniform(0.1, 0.9, 3))
        if u < 0.4:
            return dict(material=MATTE, kd=kd)
        if u < 0.7:
            return dict(material=PLASTIC, kd=kd, ks=(0.25, 0.25, 0.25), roughness=float(rng.uniform(0.05, 0.5)))
        if u < 0.85:
            return dict(material=MIRROR, kr=(0.9, 0.9, 0.9))
        return dict(material=GLASS, kr=(1, 1, 1), kt=(1, 1, 1), ior=1.5)

    budget = n_tris - 20 - 128
    k = 0
    while budget > 20 * 4 ** 6 or k == 0:
        level = 8 if budget >= 20 * 4 ** 8 else (7 if budget >= 20 * 4 ** 7 else 6)
        v, f = icosphere(level)
        r = float(rng.uniform(0.3, 0.8))
        amp = float(rng.uniform(0.02, 0.1)) * r
        d = 1.0 + amp / r * _value_noise(v * 3.0 + k * 17.0, seed + k)
        c = np.array([rng.uniform(1.5, X - 1.5), r + 0.05, rng.uniform(2.5, Z - 1.5)])
        pts = c + v * (r * d)[:, None]
        tris = pts[f].reshape(-1, 9)
        b.mesh(tris, **textured(tris, material(), lambda: spherical_uv(tris, c)))
        budget -= len(f)
        k += 1
    for _ in range(max(4, budget // 12)):
        lo = np.array([rng.uniform(0.5, X - 1.5), 0.0, rng.uniform(1.5, Z - 1.5)])
        hi = lo + rng.uniform(0.2, 1.0, 3)
        tris = box_tris(lo, hi)
        b.mesh(tris, **textured(tris, material(), lambda: planar_uv(np.asarray(tris, np.float32), 2.0)))
        if len(b.meshes) > 4000:
            break
    tris, meshes = b.arrays()
    uv = np.concatenate(b.uv) if textures else None
    return SceneArrays(tris, meshes, (X / 2, 2.0, 0.3), (X / 2, 1.2, Z / 2), (0, 1, 0), 60.0, width, height,
                       max_depth=max_depth, uv=uv, textures=textures)
