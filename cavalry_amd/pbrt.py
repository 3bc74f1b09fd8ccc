"""PBRT scene files: the host parser over the C ABI (cav_pbrt_parse,
cav_scene_create_from_pbrt) and a writer for the generated scenes.

The parser mirrors SceneLoading/Lexing.cu + Parsing.cu and the CLI overrides
of main.cu; ``write`` emits the subset it reads back exactly (trianglemesh
shapes, matte / plastic / mirror / glass materials, diffuse area lights,
LookAt camera), with float32 values printed so they round-trip bit for bit.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

from . import _lib
from ._lib import Camera, check, lib
from .scenes import (ENV_DTYPE, GLASS, MATTE, MESH_DTYPE, METAL, MIRROR, PLASTIC, SUBSTRATE, SceneArrays,
                     octa_sphere)


class Overrides(ctypes.Structure):
    _fields_ = [("output", ctypes.c_char_p), ("integrator", ctypes.c_char_p), ("spp", ctypes.c_int32),
                ("bvh_opt_rounds", ctypes.c_int32), ("morton_bits", ctypes.c_int32)]


class RenderSetup(ctypes.Structure):
    _fields_ = [("camera", Camera), ("width", ctypes.c_int32), ("height", ctypes.c_int32),
                ("integrator", ctypes.c_int32), ("max_depth", ctypes.c_int32), ("spp", ctypes.c_int32),
                ("bvh_opt_rounds", ctypes.c_int32), ("morton_bits", ctypes.c_int32),
                ("warnings", ctypes.c_int32), ("ntris", ctypes.c_int64), ("nmeshes", ctypes.c_int32),
                ("nlights", ctypes.c_int32), ("output", ctypes.c_char * 1024)]


class ParseError(Exception):
    pass


def overrides(output=None, integrator=None, spp=0, bvh_opt_rounds=-1, morton_bits=0) -> Overrides:
    return Overrides(output.encode() if output else None, integrator.encode() if integrator else None,
                     int(spp), int(bvh_opt_rounds), int(morton_bits))


def _parse(path: str, **kw):
    """(tris, meshes, setup, messages, normals or None, mesh flags or None,
    uv or None, textures list)"""
    ov = overrides(**kw)
    h = ctypes.c_void_p()
    err = ctypes.create_string_buffer(1024)
    rc = lib().cav_pbrt_parse(path.encode(), ctypes.byref(ov), ctypes.byref(h), err, len(err))
    if rc != _lib.CAV_OK:
        raise ParseError(err.value.decode(errors="replace"))
    try:
        tp, mp, msg = ctypes.c_void_p(), ctypes.c_void_p(), ctypes.c_char_p()
        st = RenderSetup()
        check(None, lib().cav_pbrt_get(h, ctypes.byref(tp), ctypes.byref(mp), ctypes.byref(st), ctypes.byref(msg)))
        n, m = int(st.ntris), int(st.nmeshes)
        tris = np.ctypeslib.as_array((ctypes.c_float * (9 * n)).from_address(tp.value)).reshape(n, 9).copy()
        raw = (ctypes.c_char * (MESH_DTYPE.itemsize * m)).from_address(mp.value)
        meshes = np.frombuffer(bytes(raw), MESH_DTYPE).copy()
        messages = [s for s in (msg.value or b"").decode().split("\n") if s]
        nptr, fptr = ctypes.c_void_p(), ctypes.c_void_p()
        check(None, lib().cav_pbrt_get_normals(h, ctypes.byref(nptr), ctypes.byref(fptr)))
        normals = flags = None
        if nptr.value:
            normals = np.ctypeslib.as_array((ctypes.c_float * (9 * n)).from_address(nptr.value)).reshape(n, 9).copy()
            flags = np.ctypeslib.as_array((ctypes.c_uint8 * m).from_address(fptr.value)).copy()
        uptr, tptr, ntex = ctypes.c_void_p(), ctypes.c_void_p(), ctypes.c_int32()
        check(None, lib().cav_pbrt_get_textures(h, ctypes.byref(uptr), ctypes.byref(tptr), ctypes.byref(ntex)))
        uv = None
        if uptr.value:
            uv = np.ctypeslib.as_array((ctypes.c_float * (6 * n)).from_address(uptr.value)).reshape(n, 6).copy()
        textures = []
        if ntex.value:
            views = (_lib.Texture * ntex.value).from_address(tptr.value)
            for t in views:
                raw = (ctypes.c_uint8 * (4 * t.width * t.height)).from_address(t.rgba)
                textures.append(np.frombuffer(bytes(raw), np.uint8).reshape(t.height, t.width, 4).copy())
        eptr, nenv, optr, nord, bg = (ctypes.c_void_p(), ctypes.c_int32(), ctypes.c_void_p(), ctypes.c_int32(),
                                      ctypes.c_int32())
        check(None, lib().cav_pbrt_get_lights(h, ctypes.byref(eptr), ctypes.byref(nenv), ctypes.byref(optr),
                                              ctypes.byref(nord), ctypes.byref(bg)))
        envs = np.zeros(0, ENV_DTYPE)
        if nenv.value:
            raw = (ctypes.c_char * (ENV_DTYPE.itemsize * nenv.value)).from_address(eptr.value)
            envs = np.frombuffer(bytes(raw), ENV_DTYPE).copy()
        order = np.ctypeslib.as_array((ctypes.c_int32 * nord.value).from_address(optr.value)).copy() \
            if nord.value else np.zeros(0, np.int32)
        return tris, meshes, st, messages, normals, flags, uv, textures, envs, order, int(bg.value)
    finally:
        lib().cav_pbrt_destroy(h)


def parse(path: str, **kw):
    """Parses a scene on the host. Returns (tris [n, 9] float32, meshes
    (MESH_DTYPE), setup (RenderSetup), messages list)."""
    return _parse(path, **kw)[:4]


def parse_normals(path: str, **kw):
    """Per-vertex normals of a parsed scene, gathered per triangle corner:
    (normals [n, 9] float32 or None, one flag per mesh or None). None when
    no shape gave normals ("normal N", PLY nx/ny/nz, spheres)."""
    return _parse(path, **kw)[4:6]


def parse_textures(path: str, **kw):
    """Texture coordinates and textures of a parsed scene: (uv [n, 6] float32
    or None, list of RGBA8 [H, W, 4] textures, row 0 at v = 0) that
    meshes["texture"] indexes."""
    return _parse(path, **kw)[6:8]


def parse_lights(path: str, **kw):
    """The light list of a parsed scene: (environment lights (ENV_DTYPE),
    light order (mesh id >= 0 or -1 - k for environment light k),
    background environment light or -1)."""
    return _parse(path, **kw)[8:11]


def _f(x) -> str:
    """float32 value printed so that strtof reads it back exactly."""
    return np.format_float_positional(np.float32(x), unique=True, trim="-")


def _list(v) -> str:
    return "[ " + " ".join(_f(x) for x in np.asarray(v, np.float32).ravel()) + " ]"


def _srgb8(linear8: np.ndarray) -> np.ndarray:
    """Bytes that the loader's inverse sRGB curve maps back to (about) the
    given linear bytes."""
    x = linear8.astype(np.float64) / 255
    g = np.where(x <= 0.0031308, 12.92 * x, 1.055 * x ** (1 / 2.4) - 0.055)
    return np.clip(np.round(g * 255), 0, 255).astype(np.uint8)


def write(arr: SceneArrays, path: str, spp: int = 16, integrator: str = "path", filename: str = "output.png",
          sampler: str = "halton") -> None:
    """Writes a generated scene as a .pbrt file the parser reads back into the
    same triangles, meshes, camera, vertex normals and texture coordinates.
    Textures go to PNG files next to it (<name>_tex<k>.png), sRGB-encoded: the
    loader's inverse curve returns each byte to within one or two units.
    Environment lights become "infinite" LightSources (a Transform for their
    rotation; black ones a "point" light, which the reader lists as black)
    placed so the file order is the light order; the reader numbers them in
    that order."""
    out = [
        "# generated scene",
        f"LookAt {' '.join(_f(x) for x in arr.eye)} {' '.join(_f(x) for x in arr.center)} "
        f"{' '.join(_f(x) for x in arr.up)}",
        f'Camera "perspective" "float fov" [ {_f(arr.fov)} ]',
        f'Film "image" "integer xresolution" [ {arr.width} ] "integer yresolution" [ {arr.height} ] '
        f'"string filename" [ "{filename}" ]',
        f'Sampler "{sampler}" "integer pixelsamples" [ {int(spp)} ]',
        f'Integrator "{integrator}" "integer maxdepth" [ {int(arr.max_depth)} ]',
        'PixelFilter "box"',
        "WorldBegin",
    ]
    stem = os.path.splitext(os.path.basename(path))[0]
    envs = getattr(arr, "env_lights", None)
    envs = [] if envs is None else list(envs)
    order = getattr(arr, "light_order", None)
    if order is None:
        order = [i for i, m in enumerate(arr.meshes) if m["is_light"]] + [-1 - k for k in range(len(envs))]
    order = [int(e) for e in order]
    area_order = [e for e in order if e >= 0]
    if area_order != sorted(area_order):
        raise ValueError("area lights must appear in the light order in mesh order to have a .pbrt form")
    # the parser numbers environment lights in file order, which is light order,
    # and makes the last "infinite" one (a lit one; black ones are written as
    # another type) the background
    lit = [-1 - e for e in order if e < 0 and (envs[-1 - e]["color"].any() or envs[-1 - e]["texture"] >= 0)]
    if int(getattr(arr, "env_background", -1)) != (lit[-1] if lit else -1):
        raise ValueError("the background must be the last lit environment light in light order")
    for k, tex in enumerate(getattr(arr, "textures", None) or []):
        name = f"{stem}_tex{k}.png"
        img = np.ascontiguousarray(_srgb8(tex[::-1, :, :3]))  # files list rows top first
        check(None, lib().cav_write_png(os.path.join(os.path.dirname(os.path.abspath(path)), name).encode(),
                                        img.ctypes.data_as(ctypes.c_void_p), tex.shape[1], tex.shape[0]))
        if (arr.meshes["texture"] == k).any():  # a map of a light is read through its own "mapname"
            out.append(f'Texture "tex{k}" "spectrum" "imagemap" "string filename" [ "{name}" ]')
    pos = 0  # next entry of the light order to write

    def env_lines(k):
        e = envs[k]
        if not (e["color"].any() or e["texture"] >= 0):
            return ['LightSource "point" "rgb L" [ 0 0 0 ]']  # any unknown type is a black light
        lines = ["AttributeBegin"]
        R = np.asarray(e["world_to_light"], np.float64).reshape(3, 3)
        if not np.allclose(R, np.eye(3)):
            L = np.eye(4)
            L[:3, :3] = np.linalg.inv(R)  # light to world; Transform lists it column by column
            lines.append(f"  Transform {_list(L.T)}")
        tex = f' "string mapname" [ "{stem}_tex{int(e["texture"])}.png" ]' if e["texture"] >= 0 else ""
        lines.append(f'  LightSource "infinite" "rgb L" {_list(e["color"])}{tex}')
        lines.append("AttributeEnd")
        return lines

    for mi, m in enumerate(arr.meshes):
        if m["is_light"]:
            while order[pos] < 0:
                out += env_lines(-1 - order[pos])
                pos += 1
            pos += 1
        out.append("AttributeBegin")
        mat = int(m["material"])
        tex = int(m["texture"])
        kd = f'"texture Kd" "tex{tex}"' if tex >= 0 and mat in (MATTE, PLASTIC, SUBSTRATE) \
            else f'"rgb Kd" {_list(m["kd"])}'
        if mat == MATTE:
            if tex >= 0 and np.any(m["kd"] != 1):
                raise ValueError("a textured matte mesh needs kd = 1 to have a .pbrt form")
            out.append(f'  Material "matte" {kd}')
        elif mat == PLASTIC:
            out.append(f'  Material "plastic" {kd} "rgb Ks" {_list(m["ks"])} '
                       f'"float roughness" [ {_f(m["roughness"])} ] "bool remaproughness" "false"')
        elif mat == SUBSTRATE:
            rv = m["roughness_v"] if m["roughness_v"] != 0 else m["roughness"]
            out.append(f'  Material "substrate" {kd} "rgb Ks" {_list(m["ks"])} '
                       f'"float uroughness" [ {_f(m["roughness"])} ] "float vroughness" [ {_f(rv)} ] '
                       '"bool remaproughness" "false"')
        elif mat == METAL:
            if not np.array_equal(m["ks"], np.float32([0.955, 0.638, 0.538])):
                raise ValueError("only copper (the default F0) metal has an exact .pbrt form")
            rv = m["roughness_v"] if m["roughness_v"] != 0 else m["roughness"]
            out.append(f'  Material "metal" "float uroughness" [ {_f(m["roughness"])} ] '
                       f'"float vroughness" [ {_f(rv)} ] "bool remaproughness" "false"')
        elif mat == MIRROR:
            out.append(f'  Material "mirror" "float Kr" [ {_f(m["kr"][0])} ]')
        elif mat == GLASS:
            out.append(f'  Material "glass" "rgb Kr" {_list(m["kr"])} "rgb Kt" {_list(m["kt"])} '
                       f'"float eta" [ {_f(m["ior"])} ]')
        else:
            raise ValueError(f"material {mat} has no .pbrt form")
        if m["is_light"]:
            out.append(f'  AreaLightSource "diffuse" "rgb L" {_list(m["emission"])}')
        if m["reverse_orientation"]:
            out.append("  ReverseOrientation")
        t = arr.tris[m["first_tri"]:m["first_tri"] + m["num_tris"]]
        if m["sphere"][3] > 0:
            # a sphere shape: written as one when the triangles are those the
            # parser will make again from the centre and radius
            # (the stored radius went through the transform's rounding, so
            # the radius to write is searched among its float neighbours)
            c = m["sphere"][:3]
            r = None
            for k in range(-8, 9):
                cand = np.float32(m["sphere"][3])
                step = np.float32(np.inf if k > 0 else 0)
                for _ in range(abs(k)):
                    cand = np.nextafter(cand, step)
                st, _, ss = octa_sphere(c, cand)
                if np.array_equal(st, t) and np.array_equal(ss, m["sphere"]):
                    r = cand
                    break
            if r is None:
                raise ValueError("a sphere mesh that is not the sphere shape of its centre and radius "
                                 "has no .pbrt form")
            out.append(f"  Translate {_f(c[0])} {_f(c[1])} {_f(c[2])}")
            out.append(f'  Shape "sphere" "float radius" [ {_f(r)} ]')
            out.append("AttributeEnd")
            continue
        idx = " ".join(str(i) for i in range(3 * len(t)))
        normals = ""
        if getattr(arr, "normals", None) is not None and (arr.mesh_normals is None or arr.mesh_normals[mi]):
            normals = f' "normal N" {_list(arr.normals[m["first_tri"]:m["first_tri"] + m["num_tris"]])}'
        uv = getattr(arr, "uv", None)
        if uv is not None:
            normals += f' "float uv" {_list(uv[m["first_tri"]:m["first_tri"] + m["num_tris"]])}'
        out.append(f'  Shape "trianglemesh" "integer indices" [ {idx} ] "point P" {_list(t)}{normals}')
        out.append("AttributeEnd")
    for e in order[pos:]:
        out += env_lines(-1 - e)
    out.append("WorldEnd")
    with open(path, "w") as f:
        f.write("\n".join(out) + "\n")
