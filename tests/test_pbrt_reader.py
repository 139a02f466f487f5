"""The .pbrt / PLY reader against the reference's own parser: every scene is
read here and by the reference (oracle/_ref/ref_driver --describe), and the
film size, samples, depth, camera matrix and fov, skybox, and per mesh the
transforms, light or material and the packed vertices, normals and faces must
be identical bit for bit.  Scenes: the reference's own scene files and crafted
ones that exercise its quirks (shared transforms and lights, doubled vertices
of binary PLY files, the sphere tessellation, named materials, instancing).
Without the reference build those comparisons are skipped; the round trip with
the writer and the number parsing run everywhere."""
import json
import os
import struct
from fractions import Fraction

import numpy as np
import pytest

import oracle as O
from polytope_amd import pbrt as P
from polytope_amd import scenes as S

needs_ref = pytest.mark.skipif(not O.reference_available(), reason="reference build (oracle/_ref) not present")


def _bits(a):
    return np.asarray(a, np.float32).view(np.uint32).ravel().tolist()


def _describe(path, tmp_path):
    out = tmp_path / "describe.json"
    O.run_reference(["--pbrt", str(path), "--describe", str(out)], cwd=str(tmp_path))
    return json.loads(out.read_text())


def _assert_same(ref, sf):
    assert (ref["width"], ref["height"], ref["spp"], ref["max_depth"]) == \
        (sf.camera.width, sf.camera.height, sf.spp, sf.max_depth)
    assert ref["camera"]["c2w"] == _bits(sf.camera.camera_to_world)
    assert ref["camera"]["fov"] == _bits([sf.camera.fov])[0]
    assert ref["sky"] == (None if sf.skybox is None else _bits(sf.skybox))
    assert len(ref["meshes"]) == len(sf.meshes)
    unrenderable = dict(sf.unrenderable)
    for i, (rm, m) in enumerate(zip(ref["meshes"], sf.meshes)):
        assert rm["o2w"] == _bits(m.object_to_world), i
        assert rm["w2o"] == _bits(m.world_to_object()), i
        assert rm["light"] == m.is_light, i
        if m.is_light:
            assert rm["spd"] == _bits(m.emission), i
        elif i not in unrenderable:
            assert rm["brdf"] == {S.LAMBERT: "lambert", S.MIRROR: "mirror", S.GLOSSY: "glossy"}[m.brdf], i
            if m.brdf != S.GLOSSY and m.texture is None:
                assert rm["refl"] == _bits(m.refl), i
        for j, axis in enumerate("xyz"):
            assert rm[axis] == _bits(m.verts[:, j]), (i, axis)
        faces = np.asarray(m.faces, np.uint32)
        for j in range(3):
            assert rm[f"fv{j}"] == faces[:, j].tolist(), (i, j)
        if m.normals is not None:
            for j, axis in enumerate(("nx", "ny", "nz")):
                assert rm[axis] == _bits(m.normals[:, j]), (i, axis)
        else:
            assert not rm["has_normals"] or not rm["nx"], i
        if m.texture is not None:           # textured Lambert: texture size and coordinates
            assert (rm["texture"]["width"], rm["texture"]["height"]) == (m.texture.width, m.texture.height), i
            assert rm["u"] == _bits(m.uv[:, 0]) and rm["v"] == _bits(m.uv[:, 1]), i
        elif i not in unrenderable and not m.is_light and m.brdf == S.LAMBERT:
            assert "texture" not in rm, i


# ----------------------------------------------------------------------------- crafted scenes

def _write_ply_ascii(path, v, f, normals=None):
    with open(path, "w") as fh:
        fh.write("ply\nformat ascii 1.0\ncomment crafted\nelement vertex %d\n" % len(v))
        fh.write("property float x\nproperty float32 y\nproperty float z\n")
        if normals is not None:
            fh.write("property float nx\nproperty float ny\nproperty float nz\n")
        fh.write("element face %d\nproperty list uchar int vertex_indices\nend_header\n" % len(f))
        for k, p in enumerate(v):
            row = list(p) + (list(normals[k]) if normals is not None else [])
            fh.write(" ".join("%.9g" % x for x in row) + "\n")
        for a in f:
            fh.write("3 %d %d %d\n" % tuple(a))


def _write_ply_binary(path, v, f, normals=None, big=False):
    e = ">" if big else "<"
    head = "ply\nformat %s 1.0\nelement vertex %d\nproperty float x\nproperty float y\nproperty float z\n" % (
        "binary_big_endian" if big else "binary_little_endian", len(v))
    if normals is not None:
        head += "property float nx\nproperty float ny\nproperty float nz\n"
    head += "element face %d\nproperty list uchar int vertex_indices\nend_header\n" % len(f)
    with open(path, "wb") as fh:
        fh.write(head.encode())
        for k, p in enumerate(v):
            row = list(p) + (list(normals[k]) if normals is not None else [])
            fh.write(struct.pack(e + "%df" % len(row), *row))
        for a in f:
            fh.write(struct.pack(e + "B3i", 3, *[int(x) for x in a]))


def _crafted_meshes(d):
    v, f = S.icosphere(1, radius=0.7, noise=0.05, seed=41)
    n = (v / np.linalg.norm(v, axis=1, keepdims=True)).astype(np.float32)
    _write_ply_ascii(d / "ico.ply", v, f)
    _write_ply_ascii(d / "ico_n.ply", v, f, n)
    _write_ply_binary(d / "ico_le.ply", v, f)                  # doubled vertices (Q6)
    _write_ply_binary(d / "ico_n_be.ply", v, f, n, big=True)
    gv, gf = S.quad(3.0, -1.0)
    with open(d / "quad.obj", "w") as fh:
        fh.write("# crafted\n")
        for p in gv:
            fh.write("v %.9g %.9g %.9g\n" % tuple(p))
        for a in gf:
            fh.write("f %d %d %d\n" % tuple(int(x) + 1 for x in a))


CRAFTED = {
    "transforms_and_lights": """
Scale -1 1 1
LookAt 1 2 3   0 0.5 0
       0 1 0   # up
Camera "perspective" "float fov" 47.5
Film "image" "integer xresolution" [32] "integer yresolution" [ 24 ] "string filename" [ "out.png" ]
Sampler "halton" "integer pixelsamples" [ 3 ]
Integrator "path" "integer maxdepth" [ 6 ]
WorldBegin
LightSource "infinite" "rgb L" [.1 .2 .3]
AttributeBegin
  Rotate 30 1 2 3
  Material "matte" "rgb Kd" [ .25 .5 .75 ]
  Shape "plymesh" "string filename" [ "ico.ply" ]
  Translate 0.5 0 0
  Shape "plymesh" "string filename" "ico_n.ply"
AttributeEnd
AttributeBegin
  AreaLightSource "diffuse" "rgb L" [ 1 1 1 ]
  Shape "plymesh" "string filename" [ "ico.ply" ]
  AreaLightSource "diffuse" "rgb L" [ 2 .5 .25 ]
  Scale 2 3 4
  Shape "plymesh" "string filename" [ "ico_le.ply" ]
AttributeEnd
TransformBegin
  Translate 1 2 3
  Rotate 77 0.3 0 1
  MakeNamedMaterial "m" "string type" [ "matte" ] "rgb Kd" [ .9 .1 .1 ]
  NamedMaterial "m"
  Shape "sphere" "float radius" [ 25.7 ]
TransformEnd
Material "mirror" "rgb Kr" [ .8 .7 .6 ]
Shape "objmesh" "string filename" [ "quad.obj" ]
ObjectBegin "o"
  Shape "plymesh" "string filename" [ "ico_n_be.ply" ]
ObjectEnd
Translate -1 0 0
ObjectInstance "o"
Rotate -15 0 1 0
ObjectInstance "o"
WorldEnd
""",
    "point_light": """
LookAt 0 2 -6  0 0 0  0 1 0
Camera "perspective" "float fov" [ 40 ]
Film "image" "integer xresolution" [ 48 ] "integer yresolution" [ 32 ]
Integrator "path" "integer maxdepth" [ 3 ]
WorldBegin
LightSource "infinite" "rgb L" [ .05 .05 .05 ]
AttributeBegin
  Translate 0 0.5 0
  LightSource "point" "rgb I" [ 8 7 6 ] "point from" [ 0 3 -1 ]
AttributeEnd
Material "matte" "rgb Kd" [ .7 .7 .7 ]
Shape "plymesh" "string filename" [ "ico.ply" ]
Shape "objmesh" "string filename" [ "quad.obj" ]
WorldEnd
""",
    "defaults_and_stacks": """
LookAt 0 0 -5 0 0 0 0 1 0
WorldBegin
LightSource "infinite" "rgb L" [ .5 .6 .7 ]
Material "matte" "rgb Kd" [ 0.3 0.3 0.3 ]
AttributeBegin
  Material "mirror" "rgb Kr" [ .6 .7 .8 ]
  AttributeBegin
    Translate 0 1 0
    Shape "plymesh" "string filename" [ "ico.ply" ]
  AttributeEnd
  Shape "plymesh" "string filename" [ "ico.ply" ]
AttributeEnd
Shape "plymesh" "string filename" [ "ico.ply" ]
Scale 0.5 0.5 0.5
Material "plastic" "rgb Kd" [ .2 .4 .6 ] "rgb Ks" [ .7 .7 .7 ] "float roughness" [ .35 ]
Shape "sphere" "float radius" [ 3 ]
WorldEnd
""",
}


@needs_ref
@pytest.mark.parametrize("name", sorted(CRAFTED))
def test_crafted_scene_matches_reference_parser(name, tmp_path):
    _crafted_meshes(tmp_path)
    path = tmp_path / (name + ".pbrt")
    path.write_text(CRAFTED[name])
    sf = P.read_pbrt(str(path))
    _assert_same(_describe(path, tmp_path), sf)


@needs_ref
def test_crafted_scene_bvh_matches_reference(tmp_path):
    from polytope_amd import build_bvh
    _crafted_meshes(tmp_path)
    path = tmp_path / "s.pbrt"
    path.write_text(CRAFTED["transforms_and_lights"])
    sf = P.read_pbrt(str(path))
    O.run_reference(["--pbrt", str(path), "--nodes", "n.bin", "--pairs", "p.bin"], cwd=str(tmp_path))
    nodes, pairs = build_bvh(sf.meshes)
    assert nodes.tobytes() == (tmp_path / "n.bin").read_bytes()
    assert pairs.tobytes() == (tmp_path / "p.bin").read_bytes()


def test_crafted_scene_quirks(tmp_path):
    """What the reference's rules imply, independent of running it."""
    _crafted_meshes(tmp_path)
    path = tmp_path / "s.pbrt"
    path.write_text(CRAFTED["transforms_and_lights"])
    sf = P.read_pbrt(str(path))
    m = sf.meshes
    assert len(m) == 8 and (sf.camera.width, sf.camera.height, sf.spp, sf.max_depth) == (32, 24, 3, 6)
    assert sf.film_filename == "out.png" and sf.camera.fov == np.float32(47.5)
    # the Translate after the first shape moved it too (shared transform) ...
    assert np.array_equal(m[0].object_to_world, m[1].object_to_world)
    # ... but not its inverse, taken when the shape was created
    assert not np.array_equal(m[0].world_to_object(), m[1].world_to_object())
    # the second AreaLightSource rewrote the first light
    assert m[2].is_light and m[3].is_light and m[2].emission == m[3].emission
    assert m[2].emission == tuple(float(np.float32(np.float32(x) * np.float32(255))) for x in (2, .5, .25))
    # binary PLY without normals: every vertex twice; with normals: once
    assert len(m[3].verts) == 2 * len(m[0].verts) and np.array_equal(m[3].verts[::2], m[0].verts)
    assert len(m[6].verts) == len(m[0].verts) and m[6].normals is not None
    # named matte material: albedo 0.5 whatever Kd says; sphere of radius 25.7: 25 x 25
    assert m[4].brdf == S.LAMBERT and m[4].refl == (0.5, 0.5, 0.5)
    assert len(m[4].verts) == 25 * 25 + 3 and len(m[4].faces) == 2 * 25 * 25
    # the OBJ mesh takes the mirror material
    assert m[5].brdf == S.MIRROR and len(m[5].faces) == 2
    # the two instances share the geometry, and the transform of their scope too:
    # the Rotate after the first one turned both, but only the second's inverse
    assert np.array_equal(m[6].verts, m[7].verts)
    assert np.array_equal(m[6].object_to_world, m[7].object_to_world)
    assert not np.array_equal(m[6].world_to_object(), m[7].world_to_object())


# ----------------------------------------------------------------------------- reference scenes

REF_SCENES = ["bvh-test/bvh-test.pbrt", "lighttest/lighttest.pbrt", "minimum.pbrt", "minimum-absolute.pbrt",
              "minimum-emptyworld.pbrt", "minimum-worldwithlight.pbrt", "minimum-emptyworld-missingcamera.pbrt",
              "two-balls.pbrt", "teapot/teapot.pbrt", "bunny/bunny.pbrt", "test/cuda/bvh-traversal.pbrt",
              "triangle-one/triangle-one.pbrt", "test/instancing/teapot-instancing1.pbrt",
              "test/instancing/teapot-instancing2.pbrt", "test/instancing/teapot-instancing3.pbrt",
              "test/instancing/teapot-instancing4.pbrt", "test/instancing/teapot-instancing5.pbrt",
              "test/texture/rainbow-texture.pbrt", "test/texture/rainbow-texture-2.pbrt"]


@needs_ref
@pytest.mark.parametrize("rel", REF_SCENES)
def test_reference_scene_files(rel, tmp_path):
    path = O.reference_scene(rel)
    sf = P.read_pbrt(path, strict=False)
    _assert_same(_describe(path, tmp_path), sf)


@needs_ref
@pytest.mark.parametrize("rel,what", [("three-balls.pbrt", "LookAt requires exactly 9"),
                                      ("example.pbrt", "blackbody"),
                                      ("dragon/dragon.pbrt", "cannot read")])
def test_reference_scene_files_the_reference_rejects(rel, what):
    with pytest.raises(P.PbrtError, match=what):
        P.read_pbrt(O.reference_scene(rel))


def test_point_lights_are_read_for_next_event_estimation(tmp_path):
    _crafted_meshes(tmp_path)
    path = tmp_path / "s.pbrt"
    path.write_text(CRAFTED["point_light"])
    sf = P.read_pbrt(str(path))
    i255 = tuple(float(np.float32(np.float32(x) * np.float32(255))) for x in (8, 7, 6))
    assert sf.lights == [((0.0, 3.5, -1.0), i255)]
    assert len(sf.meshes) == 2                     # the point light makes no mesh


def test_unrenderable_meshes_are_refused_when_strict(tmp_path):
    _crafted_meshes(tmp_path)
    path = tmp_path / "s.pbrt"
    from polytope_amd.png import encode_png
    (tmp_path / "checks.png").write_bytes(encode_png(S.noise_texture(8, 8, 1).rgba))
    path.write_text('WorldBegin\nShape "plymesh" "string filename" [ "ico.ply" ]\n'
                    'Material "plastic" "rgb Kd" [ .5 .5 .5 ]\nShape "plymesh" "string filename" [ "ico.ply" ]\n'
                    'Texture "checks" "color" "imagemap" "string filename" "checks.png"\n'
                    'Material "matte" "texture Kd" "checks"\nShape "plymesh" "string filename" [ "ico.ply" ]\n'
                    'WorldEnd\n')
    with pytest.raises(P.PbrtError, match="no material"):
        P.read_pbrt(str(path))
    sf = P.read_pbrt(str(path), strict=False)
    assert [i for i, _ in sf.unrenderable] == [0, 2] and "texture coordinates" in sf.unrenderable[1][1]
    g = sf.meshes[1]                     # plastic: GlossyBRDF, Ks and roughness at their defaults
    assert g.brdf == S.GLOSSY and g.refl == (0.5, 0.5, 0.5)
    assert g.specular == tuple(float(np.float32(0.25)) for _ in range(3)) and g.roughness == float(np.float32(0.1))


# ----------------------------------------------------------------------------- writer round trip

def test_writer_output_reads_back_to_the_same_scene(tmp_path):
    v, f = S.icosphere(2, radius=1.0, noise=0.05, seed=31)
    gv, gf = S.quad(6.0, 0.0)
    lv, lf = S.quad(1.5, 4.0)
    src = [S.MeshData(v, f, object_to_world=S.translate((0.25, 1.0, 0.5)) @ S.scale((1.5, 1.0, 1.25)),
                      brdf=S.MIRROR, refl=(0.9, 0.8, 0.7)),
           S.MeshData(gv, gf, brdf=S.LAMBERT, refl=(0.5, 0.6, 0.7)),
           S.MeshData(lv, lf, object_to_world=S.translate((0.0, 0.5, 0.0)), is_light=True,
                      emission=(255.0, 240.0, 200.0))]
    path, meshes, cam, sky = S.write_pbrt(str(tmp_path), src, (0, 2, -5), (0, 0.8, 0), (0, 1, 0), 60.0, 64, 48,
                                          spp=4, max_depth=7, skybox=(40.0, 60.0, 90.0))
    sf = P.read_pbrt(path)
    assert (sf.spp, sf.max_depth, sf.camera.width, sf.camera.height) == (4, 7, 64, 48)
    assert np.array_equal(sf.camera.camera_to_world, cam.camera_to_world) and sf.camera.fov == cam.fov
    assert sf.skybox == sky
    for a, b in zip(sf.meshes, meshes):
        assert np.array_equal(a.verts, b.verts) and np.array_equal(a.faces, b.faces)
        assert np.array_equal(a.object_to_world, b.object_to_world)
        assert (a.brdf, a.refl, a.is_light, a.emission) == (b.brdf, b.refl, b.is_light, b.emission)


# ----------------------------------------------------------------------------- numbers and tokens

def test_stof_rounds_decimals_correctly():
    for s in ["0.1", "1e-3", "3.4028234e38", "-2.5", "7", "1.17549435e-38", "0.5abc"]:
        assert P.stof(s) == np.float32(float(s.rstrip("abc")))
    # decimals just above and below a float32 midpoint that float64 rounds onto it:
    # converting through float64 would round to even, strtof rounds by the exact value
    lo = np.float32(1.0)
    hi = np.nextafter(lo, np.float32(2))
    mid = (Fraction(float(lo)) + Fraction(float(hi))) / 2
    for delta, want in ((Fraction(1, 10 ** 25), hi), (-Fraction(1, 10 ** 25), lo)):
        x = mid + delta
        s = "%d.%s" % (x.numerator // x.denominator,
                       str((x - x.numerator // x.denominator) * 10 ** 40 // 1).rjust(40, "0"))
        assert float(s) == float(mid)                       # float64 lands on the midpoint
        assert P.stof(s) == want
    with pytest.raises(P.PbrtError):
        P.stof("abc")
    assert P.stoi("12x") == 12 and P.stoi(" -3") == -3


def test_scanner_bracket_and_comment_forms():
    t = P.scan('Camera "perspective" "float fov" [ 85 ]   # c\n'
               'LookAt 0 3 4 # eye\n 0 0 0 # at\n 0 1 0\n'
               'Film "image" "integer xresolution" [1920] "integer yresolution" [ 10 20 ]\n'
               'Shape "plymesh" "string filename" [ "a.ply" ]\n')
    assert t[0] == ['Camera', '"perspective"', '"float', 'fov"', '85']
    assert t[1] == ['LookAt'] + "0 3 4 0 0 0 0 1 0".split()
    assert t[2] == ['Film', '"image"', '"integer', 'xresolution"', '1920', '"integer', 'yresolution"', '[', '10',
                    '20', ']']
    assert t[3] == ['Shape', '"plymesh"', '"string', 'filename"', '"a.ply"']
    d = P.lex(t[2])
    assert [(a.name, a.ints) for a in d.args] == [("xresolution", [1920]), ("yresolution", [10, 20])]
    with pytest.raises(P.PbrtError, match="Q5"):
        P.read_pbrt_string('Transform [ 1 0 0 0 0 1 0 0 0 0 1 0 0 0 0 1 ]\nWorldBegin\nWorldEnd\n')
