"""PBRT scene files: the host parser over the C ABI (cav_pbrt_parse,
cav_scene_create_from_pbrt) and a writer for the generated scenes.

The parser mirrors SceneLoading/Lexing.cu + Parsing.cu and the CLI overrides
of main.cu; ``write`` emits the subset it reads back exactly (trianglemesh
shapes, matte / plastic / mirror / glass materials, diffuse area lights,
LookAt camera), with float32 values printed so they round-trip bit for bit.
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _lib
from ._lib import Camera, check, lib
from .scenes import GLASS, MATTE, MESH_DTYPE, MIRROR, PLASTIC, SceneArrays


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
    """(tris, meshes, setup, messages, normals or None, mesh flags or None)"""
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
        return tris, meshes, st, messages, normals, flags
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
    return _parse(path, **kw)[4:]


def _f(x) -> str:
    """float32 value printed so that strtof reads it back exactly."""
    return np.format_float_positional(np.float32(x), unique=True, trim="-")


def _list(v) -> str:
    return "[ " + " ".join(_f(x) for x in np.asarray(v, np.float32).ravel()) + " ]"


def write(arr: SceneArrays, path: str, spp: int = 16, integrator: str = "path", filename: str = "output.png",
          sampler: str = "halton") -> None:
    """Writes a generated scene as a .pbrt file the parser reads back into the
    same triangles, meshes, camera and vertex normals."""
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
    for mi, m in enumerate(arr.meshes):
        out.append("AttributeBegin")
        mat = int(m["material"])
        if mat == MATTE:
            out.append(f'  Material "matte" "rgb Kd" {_list(m["kd"])}')
        elif mat == PLASTIC:
            out.append(f'  Material "plastic" "rgb Kd" {_list(m["kd"])} "rgb Ks" {_list(m["ks"])} '
                       f'"float roughness" [ {_f(m["roughness"])} ] "bool remaproughness" "false"')
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
        idx = " ".join(str(i) for i in range(3 * len(t)))
        normals = ""
        if getattr(arr, "normals", None) is not None and (arr.mesh_normals is None or arr.mesh_normals[mi]):
            normals = f' "normal N" {_list(arr.normals[m["first_tri"]:m["first_tri"] + m["num_tris"]])}'
        out.append(f'  Shape "trianglemesh" "integer indices" [ {idx} ] "point P" {_list(t)}{normals}')
        out.append("AttributeEnd")
    out.append("WorldEnd")
    with open(path, "w") as f:
        f.write("\n".join(out) + "\n")
