"""PBRT front end on the host (no GPU): lexer, parser semantics of
SceneLoading/Parsing.cu, shape builders, PLY, overrides and the writer's
round trip.

Expectations come from the file contents and closed forms: transformed
vertices (translation, scale, a 90-degree rotation, a raw matrix applied
before the current transform), sphere vertices on the radius, the
roughness-to-alpha polynomial of GGX evaluated in binary64, the defaults the
reference's factories state, and the CPU camera of cav_camera_look_at.
"""
import os
import struct

import numpy as np
import pytest

from cavalry_amd import pbrt, scenes
from cavalry_amd.camera import for_scene
from cavalry_amd.scenes import GLASS, MATTE, METAL, MIRROR, PLASTIC, SUBSTRATE

HEAD = """LookAt 0 0 -5  0 0 0  0 1 0
Camera "perspective" "float fov" [ 45 ]
Film "image" "integer xresolution" [ 32 ] "integer yresolution" [ 24 ]
Sampler "halton" "integer pixelsamples" [ 4 ]
Integrator "path" "integer maxdepth" [ 5 ]
"""

QUAD = '"integer indices" [ 0 1 2 0 2 3 ] "point P" [ 0 0 0  1 0 0  1 1 0  0 1 0 ]'


def write(tmp_path, body, head=HEAD, name="s.pbrt"):
    p = tmp_path / name
    p.write_text(head + body)
    return str(p)


def tri_vertices(tris):
    return tris.reshape(-1, 3)


def relevant_equal(a, b):
    """Fields each material type reads."""
    assert a["material"] == b["material"]
    for f in ("first_tri", "num_tris", "is_light", "reverse_orientation"):
        assert a[f] == b[f], f
    if a["is_light"]:
        np.testing.assert_array_equal(a["emission"], b["emission"])
    mat = int(a["material"])
    fields = {MATTE: ["kd"], PLASTIC: ["kd", "ks", "roughness"], MIRROR: [], GLASS: ["kr", "kt", "ior"],
              SUBSTRATE: ["kd", "ks", "roughness"], METAL: ["ks", "roughness"]}[mat]
    for f in fields:
        np.testing.assert_array_equal(a[f], b[f], err_msg=f)
    if mat == MIRROR:
        assert a["kr"][0] == b["kr"][0]


@pytest.mark.parametrize("materials", ["matte", "mixed"])
def test_writer_round_trip(tmp_path, materials):
    sc = scenes.cornell_box(seed=2, width=40, height=30, materials=materials)
    p = str(tmp_path / "c.pbrt")
    pbrt.write(sc, p, spp=16)
    tris, meshes, st, msgs = pbrt.parse(p)
    assert msgs == []
    np.testing.assert_array_equal(tris, sc.tris)
    assert len(meshes) == len(sc.meshes)
    for a, b in zip(meshes, sc.meshes):
        relevant_equal(a, b)
    cam = for_scene(sc)
    np.testing.assert_array_equal(np.ctypeslib.as_array(st.camera.cam_to_world), np.ctypeslib.as_array(cam.cam_to_world))
    assert st.camera.depth == cam.depth
    assert (st.width, st.height, st.spp, st.max_depth, st.integrator) == (40, 30, 16, sc.max_depth, 0)
    assert st.output == b"output.png" and st.morton_bits == 63 and st.bvh_opt_rounds == 3


def test_overrides_and_bvh_round_rule(tmp_path):
    p = write(tmp_path, 'WorldBegin\nShape "trianglemesh" ' + QUAD + '\nWorldEnd\n')
    _, _, st, _ = pbrt.parse(p, output="x.png", integrator="rlpath", spp=64)
    assert st.output == b"x.png" and st.integrator == 1 and st.spp == 64 and st.bvh_opt_rounds == 10
    _, _, st, _ = pbrt.parse(p, bvh_opt_rounds=0, morton_bits=30)
    assert st.bvh_opt_rounds == 0 and st.morton_bits == 30 and st.spp == 4 and st.max_depth == 5


def test_integrator_factory(tmp_path):
    for name, expect, warn in (("directlighting", (2, 1), 0), ("bdpt", (0, 8), 1)):
        head = HEAD.replace('Integrator "path" "integer maxdepth" [ 5 ]', f'Integrator "{name}"')
        p = write(tmp_path, 'WorldBegin\nShape "trianglemesh" ' + QUAD + '\nWorldEnd\n', head=head)
        _, _, st, msgs = pbrt.parse(p)
        assert (st.integrator, st.max_depth) == expect and st.warnings == warn and len(msgs) == warn


def test_transforms(tmp_path):
    body = f"""WorldBegin
AttributeBegin
  Translate 1 2 3
  Scale 2 2 2
  Shape "trianglemesh" {QUAD}
AttributeEnd
AttributeBegin
  Rotate 90 0 0 1
  Shape "trianglemesh" {QUAD}
AttributeEnd
AttributeBegin
  Translate 5 0 0
  Transform [ 1 0 0 0  0 1 0 0  0 0 1 0  0 0 7 1 ]
  Shape "trianglemesh" {QUAD}
AttributeEnd
Shape "trianglemesh" {QUAD}
WorldEnd
"""
    tris, meshes, _, _ = pbrt.parse(write(tmp_path, body))
    quad = np.array([[0, 0, 0], [1, 0, 0], [1, 1, 0], [0, 0, 0], [1, 1, 0], [0, 1, 0]], np.float64)
    v = tri_vertices(tris).astype(np.float64)
    np.testing.assert_allclose(v[0:6], quad * 2 + [1, 2, 3], atol=1e-6)
    np.testing.assert_allclose(v[6:12], np.stack([-quad[:, 1], quad[:, 0], quad[:, 2]], 1), atol=1e-6)
    # the raw matrix multiplies from the left: M * T(5,0,0) -> (x + 5, y, z + 7)
    np.testing.assert_allclose(v[12:18], quad + [5, 0, 7], atol=1e-6)
    np.testing.assert_array_equal(v[18:24], quad)  # the section's transform did not leak
    assert [int(m["num_tris"]) for m in meshes] == [2, 2, 2, 2]


def test_attribute_scoping_lights_and_orientation(tmp_path):
    body = f"""WorldBegin
Material "plastic" "rgb Kd" [ 0.1 0.2 0.3 ] "float roughness" [ 0.1 ]
AttributeBegin
  Material "mirror" "float Kr" [ 0.7 ]
  AreaLightSource "diffuse" "rgb L" [ 4 5 6 ]
  Shape "trianglemesh" {QUAD}
  Shape "trianglemesh" {QUAD}
AttributeEnd
ReverseOrientation
Shape "trianglemesh" {QUAD}
AttributeBegin
  Shape "trianglemesh" {QUAD}
AttributeEnd
MakeNamedMaterial "g" "string type" [ "glass" ] "float index" [ 1.33 ]
NamedMaterial "g"
Shape "trianglemesh" {QUAD}
Shape "sphere" "float radius" [ 2 ]
WorldEnd
"""
    tris, m, st, _ = pbrt.parse(write(tmp_path, body))
    assert list(m["material"]) == [MIRROR, MIRROR, PLASTIC, PLASTIC, GLASS, GLASS]
    assert m["kr"][0][0] == np.float32(0.7)
    # the light binds to the next shape only
    assert list(m["is_light"]) == [1, 0, 0, 0, 0, 0] and list(m["emission"][0]) == [4, 5, 6]
    # ReverseOrientation holds for the rest of its own section, not nested ones
    assert list(m["reverse_orientation"]) == [0, 0, 1, 0, 1, 1]
    # plastic roughness remapped by GGX::roughnessToAlpha
    x = np.log(0.1)
    alpha = 1.62142 + 0.819955 * x + 0.1734 * x * x + 0.0171201 * x ** 3 + 0.000640711 * x ** 4
    assert m["roughness"][2] == pytest.approx(alpha, rel=1e-5)
    np.testing.assert_array_equal(m["kd"][2], np.float32([0.1, 0.2, 0.3]))
    assert m["ior"][4] == np.float32(1.33) and list(m["kt"][4]) == [1, 1, 1]
    # sphere: octahedron subdivided four times, vertices on the radius
    sph = tris[m["first_tri"][5]:].reshape(-1, 3)
    assert m["num_tris"][5] == 8 * 4 ** 4
    np.testing.assert_allclose(np.linalg.norm(sph.astype(np.float64), axis=1), 2.0, rtol=1e-6)
    assert st.nlights == 1
    # its shape parameters (ShapeMesh/Sphere.cu:86-88); other meshes have none
    assert list(m["sphere"][5]) == [0, 0, 0, 2] and not m["sphere"][:5].any()


def test_sphere_shape_parameters(tmp_path):
    """centre = T(0), radius = |centre - T(r, 0, 0)| through the current
    transform; the triangles are the octahedral sphere shape."""
    body = """WorldBegin
AttributeBegin
  AreaLightSource "diffuse" "rgb L" [ 3 3 3 ]
  Translate 1 2 3
  Scale 2 2 2
  Shape "sphere" "float radius" [ 0.5 ]
AttributeEnd
AttributeBegin
  Translate 0.5 0.85 0.5
  Shape "sphere" "float radius" [ 0.08 ]
AttributeEnd
WorldEnd
"""
    tris, m, st, _ = pbrt.parse(write(tmp_path, body))
    assert list(m["sphere"][0]) == [1, 2, 3, 1] and m["is_light"][0] == 1
    t1, n1, s1 = scenes.octa_sphere((0.5, 0.85, 0.5), 0.08)
    np.testing.assert_array_equal(m["sphere"][1], s1)
    np.testing.assert_array_equal(tris[m["first_tri"][1]:], t1)
    normals, _ = pbrt.parse_normals(write(tmp_path, body))
    np.testing.assert_allclose(normals[m["first_tri"][1]:], n1, atol=2.5e-7)


def test_writer_round_trip_sphere_lights(tmp_path):
    sc = scenes.cornell_box(seed=3, width=40, height=30, lights="both")
    p = str(tmp_path / "s.pbrt")
    pbrt.write(sc, p)
    tris, meshes, st, msgs = pbrt.parse(p)
    np.testing.assert_array_equal(tris, sc.tris)
    np.testing.assert_array_equal(meshes["sphere"], sc.meshes["sphere"])
    for a, b in zip(meshes, sc.meshes):
        relevant_equal(a, b)
    assert st.nlights == 2


def test_material_defaults(tmp_path):
    body = f"""WorldBegin
Shape "trianglemesh" {QUAD}
Material "mirror"
Shape "trianglemesh" {QUAD}
Material "plastic"
Shape "trianglemesh" {QUAD}
Material "glass"
Shape "trianglemesh" {QUAD}
Material "metal"
Shape "trianglemesh" {QUAD}
Material "substrate"
Shape "trianglemesh" {QUAD}
Material "metal" "rgb eta" [ 0.2 0.9 1.1 ] "rgb k" [ 3.9 2.4 2.2 ] "float uroughness" [ 0.2 ] "float vroughness" [ 0.05 ] "bool remaproughness" "false"
Shape "trianglemesh" {QUAD}
Material "rubber"
Shape "trianglemesh" {QUAD}
WorldEnd
"""
    _, m, st, msgs = pbrt.parse(write(tmp_path, body))
    assert list(m["material"]) == [MATTE, MIRROR, PLASTIC, GLASS, METAL, SUBSTRATE, METAL, MATTE]
    assert list(m["kd"][0]) == [1, 1, 1]                 # default matte (Parsing.cu:290-292)
    assert m["kr"][1][0] == np.float32(0.9)               # MirrorMaterial()
    assert list(m["kd"][2]) == [np.float32(0.25)] * 3     # plastic defaults
    assert list(m["ks"][2]) == [np.float32(0.25)] * 3
    assert m["ior"][3] == np.float32(1.5)
    # metal: copper F0 by default, GGX alpha of roughness 0.1 on both axes (Metal.h:36-64)
    a01 = m["roughness"][2]  # plastic's default roughness is also 0.1, remapped
    np.testing.assert_array_equal(m["ks"][4], np.float32([0.955, 0.638, 0.538]))
    assert m["roughness"][4] == a01 and m["roughness_v"][4] == a01
    # substrate: Kd = Ks = 0.5 (Substrate.h:50-83)
    assert list(m["kd"][5]) == [0.5] * 3 and list(m["ks"][5]) == [0.5] * 3 and m["roughness_v"][5] == a01
    # F0 of a complex IOR: |((eta + ik - 1) / (eta + ik + 1))^2|
    n = np.array([0.2, 0.9, 1.1]) + 1j * np.array([3.9, 2.4, 2.2])
    np.testing.assert_allclose(m["ks"][6], np.abs(((n - 1) / (n + 1)) ** 2), rtol=1e-6)
    assert m["roughness"][6] == np.float32(0.2) and m["roughness_v"][6] == np.float32(0.05)
    assert st.warnings == 1 and "rubber" in msgs[0]


def test_include_and_lexer(tmp_path):
    (tmp_path / "sub").mkdir()
    (tmp_path / "sub" / "inc.pbrt").write_text(
        "# comment line\r\nShape \"trianglemesh\" \"integer indices\" [0 1 2] \"point P\" [ .5 -1e-3 0  1 0 0  0 1.5e+1 0 ]\r\n")
    tris, _, _, _ = pbrt.parse(write(tmp_path, 'WorldBegin\nInclude "sub/inc.pbrt"\nWorldEnd\n'))
    np.testing.assert_array_equal(tris[0], np.float32([0.5, -1e-3, 0, 1, 0, 0, 0, 15, 0]))


def test_ply(tmp_path):
    verts = np.array([[0, 0, 0], [1, 0, 0], [1, 1, 0], [0, 1, 0], [0, 0, 1]], np.float32)
    faces = [[0, 1, 2, 3], [0, 1, 4]]  # a quad keeps its first three indices
    hdr = ("ply\nformat binary_little_endian 1.0\nelement vertex 5\nproperty float x\nproperty float y\n"
           "property float z\nproperty float nx\nelement face 2\nproperty list uchar int vertex_indices\nend_header\n")
    body = b"".join(struct.pack("<4f", *v, 0.0) for v in verts)
    body += b"".join(struct.pack("<B", len(f)) + struct.pack(f"<{len(f)}i", *f) for f in faces)
    (tmp_path / "m.ply").write_bytes(hdr.encode() + body)
    asc = ("ply\nformat ascii 1.0\nelement vertex 5\nproperty float x\nproperty float y\nproperty float z\n"
           "element face 2\nproperty list uchar int vertex_indices\nend_header\n"
           + "\n".join(" ".join(str(float(c)) for c in v) for v in verts) + "\n"
           + "\n".join(f"{len(f)} " + " ".join(map(str, f)) for f in faces) + "\n")
    (tmp_path / "a.ply").write_text(asc)
    body = ('WorldBegin\nShape "plymesh" "string filename" "m.ply"\n'
            'Translate 0 0 2\nShape "plymesh" "string filename" [ "a.ply" ]\nWorldEnd\n')
    tris, m, _, _ = pbrt.parse(write(tmp_path, body))
    expect = verts[[0, 1, 2, 0, 1, 4]].reshape(2, 9)
    np.testing.assert_array_equal(tris[:2], expect)
    np.testing.assert_array_equal(tris[2:], expect + np.float32([0, 0, 2] * 3))


@pytest.mark.parametrize("body,needle", [
    ("WorldBegin\nBogusKeyword\nWorldEnd\n", "Unrecognized Keyword"),
    ('WorldBegin\nShape "trianglemesh" "point P" [ 0 0 0 ]\nWorldEnd\n', "No valid inputs"),
    ('WorldBegin\nNamedMaterial "nope"\nWorldEnd\n', "nope"),
    ('WorldBegin\nAttributeBegin\nWorldEnd\n', "Mismatching"),
    ('WorldBegin\nInclude "missing.pbrt"\nWorldEnd\n', "cannot open"),
])
def test_parse_errors(tmp_path, body, needle):
    with pytest.raises(pbrt.ParseError, match=needle):
        pbrt.parse(write(tmp_path, body))


def test_incomplete_options(tmp_path):
    head = HEAD.replace('Camera "perspective" "float fov" [ 45 ]\n', "")
    with pytest.raises(pbrt.ParseError, match="incomplete"):
        pbrt.parse(write(tmp_path, 'WorldBegin\nShape "trianglemesh" ' + QUAD + '\nWorldEnd\n', head=head))
    with pytest.raises(pbrt.ParseError, match="cannot open"):
        pbrt.parse(os.path.join(str(tmp_path), "none.pbrt"))


def test_vertex_normals(tmp_path):
    """"normal N" maps by transpose(inverse(upper 3x3)) and is normalised
    (Core/TriangleMesh.cu:152-166); PLY nx/ny/nz the same way (:81-90);
    sphere normals are the unit octahedron points mapped likewise
    (ShapeMesh/Sphere.cu:67-72); meshes without normals get a zero flag."""
    n = "[ 1 1 0  1 1 0  1 1 0  1 1 0 ]"
    verts = np.array([[0, 0, 0], [1, 0, 0], [1, 1, 0]], np.float32)
    hdr = ("ply\nformat ascii 1.0\nelement vertex 3\nproperty float x\nproperty float y\nproperty float z\n"
           "property float nx\nproperty float ny\nproperty float nz\n"
           "element face 1\nproperty list uchar int vertex_indices\nend_header\n")
    (tmp_path / "n.ply").write_text(hdr + "0 0 0 0 0 1\n1 0 0 0 1 1\n1 1 0 0 0 2\n3 0 1 2\n")
    body = f"""WorldBegin
AttributeBegin
  Scale 2 1 1
  Shape "trianglemesh" {QUAD} "normal N" {n}
AttributeEnd
AttributeBegin
  Rotate 90 0 0 1
  Shape "trianglemesh" {QUAD} "normal N" {n}
AttributeEnd
Shape "trianglemesh" {QUAD}
AttributeBegin
  Scale 1 1 4
  Shape "plymesh" "string filename" "n.ply"
AttributeEnd
AttributeBegin
  Translate 3 0 0
  Scale 1 2 1
  Shape "sphere" "float radius" [ 0.5 ]
AttributeEnd
WorldEnd
"""
    p = write(tmp_path, body)
    tris, m, _, _ = pbrt.parse(p)
    normals, flags = pbrt.parse_normals(p)
    assert list(flags) == [1, 1, 0, 1, 1]
    nv = normals.reshape(-1, 3).astype(np.float64)

    def unit(v):
        v = np.asarray(v, np.float64)
        return v / np.linalg.norm(v)

    np.testing.assert_allclose(nv[0:6], np.tile(unit([0.5, 1, 0]), (6, 1)), atol=1e-6)   # diag(1/2, 1, 1)
    np.testing.assert_allclose(nv[6:12], np.tile(unit([-1, 1, 0]), (6, 1)), atol=1e-6)   # rotated like points
    np.testing.assert_array_equal(nv[12:18], 0)                                           # no normals
    np.testing.assert_allclose(nv[18:21], [[0, 0, 1], unit([0, 1, 0.25]), [0, 0, 1]], atol=1e-6)  # diag(1,1,1/4)
    # sphere: the normal at position p of an ellipsoid x^2 + (y/2)^2 + z^2 = r^2 is (x, y/4, z) normalised
    s0 = int(m["first_tri"][4])
    pos = tris[s0:].reshape(-1, 3).astype(np.float64) - [3, 0, 0]
    expect = pos * [1, 0.25, 1]
    expect /= np.linalg.norm(expect, axis=1, keepdims=True)
    np.testing.assert_allclose(nv[3 * s0:], expect, atol=2e-6)
    # a scene without any normals reports none
    assert pbrt.parse_normals(write(tmp_path, f'WorldBegin\nShape "trianglemesh" {QUAD}\nWorldEnd\n', name="b.pbrt")) \
        == (None, None)


def test_vertex_normals_errors(tmp_path):
    with pytest.raises(pbrt.ParseError, match="one normal per vertex"):
        pbrt.parse(write(tmp_path, f'WorldBegin\nShape "trianglemesh" {QUAD} "normal N" [ 0 0 1 ]\nWorldEnd\n'))
    with pytest.raises(pbrt.ParseError, match="singular"):
        pbrt.parse(write(tmp_path, f'WorldBegin\nScale 1 0 1\nShape "trianglemesh" {QUAD} "normal N" '
                                   '[ 0 0 1  0 0 1  0 0 1  0 0 1 ]\nWorldEnd\n'))


def test_writer_round_trip_normals(tmp_path):
    sc = scenes.cornell_box(seed=5, width=40, height=30, materials="mixed", smooth=True)
    p = str(tmp_path / "s.pbrt")
    pbrt.write(sc, p)
    tris, _, _, _ = pbrt.parse(p)
    normals, flags = pbrt.parse_normals(p)
    np.testing.assert_array_equal(tris, sc.tris)
    np.testing.assert_array_equal(flags, sc.mesh_normals)
    on = sc.mesh_normals[np.repeat(np.arange(len(sc.meshes)), sc.meshes["num_tris"])].astype(bool)
    # the parser normalises the written unit normals again, which may move them by an ulp or two
    np.testing.assert_allclose(normals[on], sc.normals[on], atol=2.5e-7, rtol=0)


# ------------------------------------------------------------------ textures
def _png(path, img, ctype, depth=8, filters=(0,), palette=None, trns=None):
    """PNG encoder after the specification: rows top first, per-row filter
    types cycled from `filters`, one zlib stream."""
    import zlib
    chans = {0: 1, 2: 3, 3: 1, 4: 2, 6: 4}[ctype]
    h, w = img.shape[:2]
    data = img.reshape(h, w * chans)
    if depth == 16:
        data = data.astype(">u2").view(np.uint8).reshape(h, -1)
    bpp = chans * depth // 8
    raw = b""
    prev = np.zeros(data.shape[1], np.int64)
    for y in range(h):
        row = data[y].astype(np.int64)
        ft = filters[y % len(filters)]
        a = np.concatenate([np.zeros(bpp, np.int64), row[:-bpp]])
        c = np.concatenate([np.zeros(bpp, np.int64), prev[:-bpp]])
        if ft == 0:
            out = row
        elif ft == 1:
            out = row - a
        elif ft == 2:
            out = row - prev
        elif ft == 3:
            out = row - (a + prev) // 2
        else:
            p = a + prev - c
            pa, pb, pc = np.abs(p - a), np.abs(p - prev), np.abs(p - c)
            pred = np.where((pa <= pb) & (pa <= pc), a, np.where(pb <= pc, prev, c))
            out = row - pred
        raw += bytes([ft]) + bytes((out % 256).astype(np.uint8))
        prev = row

    def chunk(t, b):
        return struct.pack(">I", len(b)) + t + b + struct.pack(">I", zlib.crc32(t + b) & 0xffffffff)

    body = chunk(b"IHDR", struct.pack(">IIBBBBB", w, h, depth, ctype, 0, 0, 0))
    if palette is not None:
        body += chunk(b"PLTE", bytes(np.asarray(palette, np.uint8).ravel()))
    if trns is not None:
        body += chunk(b"tRNS", bytes(np.asarray(trns, np.uint8)))
    body += chunk(b"IDAT", zlib.compress(raw)) + chunk(b"IEND", b"")
    path.write_bytes(b"\x89PNG\r\n\x1a\n" + body)


def _inv_gamma8(v):
    """(uint8)(255 * inverseGammaCorrect(v / 255)) in float64; the library
    computes it in float, so it may differ by one."""
    x = np.asarray(v, np.float64) / 255
    return np.floor(255 * np.where(x <= 0.04045, x / 12.92, ((x + 0.055) / 1.055) ** 2.4)).astype(np.int64)


def _tex_scene(tmp_path, tex_lines, kd='"texture Kd" "t0"', material="matte", uv=' "float uv" [ 0 0  1 0  1 1  0 1 ]'):
    body = tex_lines + f'\nMaterial "{material}" {kd}\nShape "trianglemesh" {QUAD}{uv}\nWorldEnd\n'
    return write(tmp_path, "WorldBegin\n" + body)


def _check_texture(got, rgba_top_first):
    """got: parsed texture (bottom-up, linearised rgb); expect flipped rows
    and the inverse sRGB curve on rgb within one unit, alpha as is."""
    want = rgba_top_first[::-1].astype(np.int64)
    assert got.shape == want.shape
    assert np.abs(got[..., :3].astype(np.int64) - _inv_gamma8(want[..., :3])).max() <= 1
    np.testing.assert_array_equal(got[..., 3], want[..., 3])


@pytest.mark.parametrize("ctype,depth", [(0, 8), (2, 8), (4, 8), (6, 8), (6, 16), (0, 16), (3, 8)])
def test_png_textures(tmp_path, ctype, depth):
    """Every colour type and all five row filters decode to the pixels
    written, as RGBA8 with rows bottom-up and the sRGB curve removed."""
    rng = np.random.default_rng(ctype * 100 + depth)
    h, w = 7, 5
    if ctype == 3:
        pal = rng.integers(0, 256, (6, 3))
        idx = rng.integers(0, 6, (h, w))
        trns = [10, 250, 0]
        _png(tmp_path / "t.png", idx.astype(np.uint8), 3, palette=pal, trns=trns, filters=(0, 1, 2, 3, 4))
        a = np.where(idx < 3, np.array(trns + [255, 255, 255])[np.minimum(idx, 5)], 255)
        rgba = np.concatenate([pal[idx], a[..., None]], 2)
    else:
        chans = {0: 1, 2: 3, 4: 2, 6: 4}[ctype]
        img = rng.integers(0, 2 ** depth, (h, w, chans))
        _png(tmp_path / "t.png", img.astype(np.uint16 if depth == 16 else np.uint8), ctype, depth,
             filters=(0, 1, 2, 3, 4))
        hi = img >> 8 if depth == 16 else img
        g = hi[..., :1]
        rgba = {0: lambda: np.concatenate([g, g, g, np.full_like(g, 255)], 2),
                2: lambda: np.concatenate([hi, np.full_like(g, 255)], 2),
                4: lambda: np.concatenate([g, g, g, hi[..., 1:2]], 2),
                6: lambda: hi}[ctype]()
    p = _tex_scene(tmp_path, 'Texture "t0" "spectrum" "imagemap" "string filename" [ "t.png" ]')
    uv, textures = pbrt.parse_textures(p)
    assert len(textures) == 1
    _check_texture(textures[0], rgba)
    _, m, _, _ = pbrt.parse(p)
    assert m["texture"][0] == 0


def _tga(path, img, rle=False, top_first=False, grey=False):
    h, w = img.shape[:2]
    bpp = 1 if grey else img.shape[2]
    hdr = bytes([0, 0, (11 if rle else 3) if grey else (10 if rle else 2)]) + bytes(9) + \
        struct.pack("<HHBB", w, h, 8 * bpp, 32 if top_first else 0)
    rows = img if top_first else img[::-1]
    px = rows.reshape(h * w, -1)
    if not grey:
        px = px[:, [2, 1, 0, 3][:bpp]]   # b, g, r (, a)
    body = b""
    if rle:  # pixels in runs: repeat packets for equal neighbours, else literal packets of one
        k = 0
        while k < len(px):
            n = 1
            while k + n < len(px) and n < 128 and np.array_equal(px[k + n], px[k]):
                n += 1
            body += bytes([(n - 1) | (128 if n > 1 else 0)]) + bytes(px[k].astype(np.uint8))
            k += n
    else:
        body = bytes(px.astype(np.uint8).ravel())
    path.write_bytes(hdr + body)


@pytest.mark.parametrize("rle,top_first,chans", [(False, False, 3), (True, True, 4), (True, False, 1),
                                                 (False, True, 4)])
def test_tga_textures(tmp_path, rle, top_first, chans):
    rng = np.random.default_rng(7)
    img = rng.integers(0, 4, (6, 9, chans)).astype(np.uint8) * 60  # few values: runs for RLE
    _tga(tmp_path / "t.tga", img, rle=rle, top_first=top_first, grey=chans == 1)
    p = _tex_scene(tmp_path, 'Texture "t0" "spectrum" "imagemap" "string filename" "t.tga"')
    _, textures = pbrt.parse_textures(p)
    if chans == 1:
        rgba = np.concatenate([img, img, img, np.full_like(img, 255)], 2)
    elif chans == 3:
        rgba = np.concatenate([img, np.full_like(img[..., :1], 255)], 2)
    else:
        rgba = img
    _check_texture(textures[0], rgba)


def test_texture_binding_and_uv(tmp_path):
    """Kd textures bind by name to matte and plastic (the texel then decides
    the colour), the texture list keeps definition order, a Texture without
    a file is one transparent black texel, "float uv" and PLY u/v give the
    per-vertex texture coordinates and meshes without any read (0, 0)."""
    img = np.full((2, 2, 3), 128, np.uint8)
    _png(tmp_path / "a.png", img, 2)
    hdr = ("ply\nformat ascii 1.0\nelement vertex 3\nproperty float x\nproperty float y\nproperty float z\n"
           "property float u\nproperty float v\nelement face 1\nproperty list uchar int vertex_indices\nend_header\n")
    (tmp_path / "uv.ply").write_text(hdr + "0 0 0 0.25 0.5\n1 0 0 0.75 0.5\n1 1 0 0.75 1\n3 0 1 2\n")
    body = f"""WorldBegin
Texture "empty" "spectrum" "imagemap"
Texture "a" "spectrum" "imagemap" "string filename" "a.png"
Material "plastic" "texture Kd" "a" "float roughness" [ 0.3 ]
Shape "trianglemesh" {QUAD} "float uv" [ 0 0  2 0  2 2  0 2 ]
Material "matte" "texture Kd" "empty"
Shape "plymesh" "string filename" "uv.ply"
Material "mirror"
Shape "trianglemesh" {QUAD}
WorldEnd
"""
    p = write(tmp_path, body)
    _, m, _, _ = pbrt.parse(p)
    uv, textures = pbrt.parse_textures(p)
    assert list(m["texture"]) == [1, 0, -1]
    assert list(m["material"]) == [PLASTIC, MATTE, MIRROR]
    np.testing.assert_array_equal(textures[0], np.zeros((1, 1, 4), np.uint8))
    assert textures[1].shape == (2, 2, 4)
    np.testing.assert_array_equal(uv[0], [0, 0, 2, 0, 2, 2])   # quad triangles 0 1 2 and 0 2 3
    np.testing.assert_array_equal(uv[1], [0, 0, 2, 2, 0, 2])
    np.testing.assert_array_equal(uv[2], [0.25, 0.5, 0.75, 0.5, 0.75, 1])
    np.testing.assert_array_equal(uv[3:], 0)
    # a scene without textures reports none
    plain = write(tmp_path, f'WorldBegin\nShape "trianglemesh" {QUAD}\nWorldEnd\n', name="p.pbrt")
    assert pbrt.parse_textures(plain) == (None, [])


@pytest.mark.parametrize("body,needle", [
    ('Material "matte" "texture Kd" "nope"\n', "texture not found"),
    ('Texture "t" "spectrum" "imagemap" "string filename" "x.exr"\n', "not supported"),
    ('Texture "t" "spectrum" "imagemap" "string filename" "missing.png"\n', "cannot open"),
    (f'Shape "trianglemesh" {QUAD} "float uv" [ 0 0 1 ]\n', "two numbers per vertex"),
])
def test_texture_errors(tmp_path, body, needle):
    with pytest.raises(pbrt.ParseError, match=needle):
        pbrt.parse(write(tmp_path, "WorldBegin\n" + body + "WorldEnd\n"))
    (tmp_path / "bad.png").write_bytes(b"\x89PNG\r\n\x1a\nnot really")
    with pytest.raises(pbrt.ParseError, match="PNG"):
        pbrt.parse(write(tmp_path, 'WorldBegin\nTexture "t" "spectrum" "imagemap" "string filename" "bad.png"\n'
                                   'WorldEnd\n', name="b.pbrt"))


def test_writer_round_trip_textures(tmp_path):
    sc = scenes.cornell_box(seed=5, width=40, height=30, materials="mixed", textured=True)
    p = str(tmp_path / "t.pbrt")
    pbrt.write(sc, p)
    tris, m, _, _ = pbrt.parse(p)
    uv, textures = pbrt.parse_textures(p)
    np.testing.assert_array_equal(tris, sc.tris)
    np.testing.assert_array_equal(uv, sc.uv)
    np.testing.assert_array_equal(m["texture"], sc.meshes["texture"])
    # a textured plastic's Kd is not written (the texel replaces it): the parser keeps its default
    want = sc.meshes.copy()
    want["kd"][(want["texture"] >= 0) & (want["material"] == PLASTIC)] = np.float32(0.25)
    for a, b in zip(m, want):
        relevant_equal(a, b)
    assert len(textures) == len(sc.textures)
    for a, b in zip(textures, sc.textures):
        assert np.abs(a[..., :3].astype(int) - b[..., :3]).max() <= 2   # sRGB encode, then decode


def test_light_sources_and_order(tmp_path):
    """LightSource "infinite" (L, scale, mapname, the inverse of its
    transform), other types as black lights, AreaLightSource listed where it
    appears (Parsing.cu:320-342, Lights/LightObject.h:89-126)."""
    _png(tmp_path / "sky.png", np.full((2, 4, 3), 200, np.uint8), 2)
    body = f"""WorldBegin
LightSource "point" "rgb I" [ 5 5 5 ]
AttributeBegin
  AreaLightSource "diffuse" "rgb L" [ 3 3 3 ]
  Shape "trianglemesh" {QUAD}
AttributeEnd
Shape "trianglemesh" {QUAD}
AttributeBegin
  Rotate 90 0 1 0
  LightSource "infinite" "rgb L" [ 0.5 0.6 0.7 ] "string mapname" "sky.png"
AttributeEnd
LightSource "infinite" "rgb scale" [ 2 2 2 ]
WorldEnd
"""
    p = write(tmp_path, body)
    envs, order, background = pbrt.parse_lights(p)
    _, m, st, msgs = pbrt.parse(p)
    assert list(order) == [-1, 0, -2, -3] and background == 2 and st.nlights == 4
    np.testing.assert_array_equal(envs["color"], np.float32([[0, 0, 0], [0.5, 0.6, 0.7], [2, 2, 2]]))
    assert list(envs["texture"]) == [-1, 0, -1]
    # world -> light = inverse of a 90-degree turn about y, row-major
    np.testing.assert_allclose(envs["world_to_light"][1].reshape(3, 3), scenes.rotation_y(90).T, atol=1e-6)
    np.testing.assert_array_equal(envs["world_to_light"][2], np.eye(3, dtype=np.float32).ravel())
    assert list(m["is_light"]) == [1, 0] and any("point" in s for s in msgs)
    _, textures = pbrt.parse_textures(p)
    assert len(textures) == 1 and textures[0].shape == (2, 4, 4)


def test_writer_round_trip_lights(tmp_path):
    sc = scenes.cornell_box(seed=6, width=40, height=30, materials="mixed", env=True)
    p = str(tmp_path / "e.pbrt")
    pbrt.write(sc, p)
    envs, order, background = pbrt.parse_lights(p)
    np.testing.assert_array_equal(order, sc.light_order)
    assert background == sc.env_background
    np.testing.assert_array_equal(envs["color"], sc.env_lights["color"])
    np.testing.assert_allclose(envs["world_to_light"], sc.env_lights["world_to_light"], atol=1e-6)
    # the sky map is read again through "mapname": same image, its own index
    _, textures = pbrt.parse_textures(p)
    assert envs["texture"][1] == -1 and len(textures) == 1
    got, want = textures[envs["texture"][0]], sc.textures[sc.env_lights["texture"][0]]
    assert np.abs(got[..., :3].astype(int) - want[..., :3]).max() <= 2
