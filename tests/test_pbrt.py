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
from cavalry_amd.scenes import GLASS, MATTE, MIRROR, PLASTIC

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
    fields = {MATTE: ["kd"], PLASTIC: ["kd", "ks", "roughness"], MIRROR: [], GLASS: ["kr", "kt", "ior"]}[mat]
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
WorldEnd
"""
    _, m, st, msgs = pbrt.parse(write(tmp_path, body))
    assert list(m["material"]) == [MATTE, MIRROR, PLASTIC, GLASS, MATTE]
    assert list(m["kd"][0]) == [1, 1, 1]                 # default matte (Parsing.cu:290-292)
    assert m["kr"][1][0] == np.float32(0.9)               # MirrorMaterial()
    assert list(m["kd"][2]) == [np.float32(0.25)] * 3     # plastic defaults
    assert list(m["ks"][2]) == [np.float32(0.25)] * 3
    assert m["ior"][3] == np.float32(1.5)
    assert st.warnings == 1 and "metal" in msgs[0]


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
