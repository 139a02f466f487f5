"""Textured Lambert (the reference's LambertBRDF with an image texture) on the
CPU: the PNG decode against lodepng inside the reference build, the texel
lookup against closed forms, and the reference's own lookups and films against
the oracle.

Pins, in order:
* decode: every PNG form lodepng reads (grey 1/2/4/8/16 bits, RGB 8/16,
  palette 1-8 bits with tRNS, grey+alpha, RGBA 8/16, colour keys, all five
  filters, Adam7) decodes here to the bytes the reference parser stores
  (FNV-1a of texture::data from ref_driver --describe); the same forms against
  values the PNG spec fixes, where no reference build is present;
* lookup: texel_index against the index the reference computes (column
  (unsigned)(u * W), row (unsigned)((1 - v) * H), flat row * W + column) on
  hand-picked coordinates, including u = 1 (next row's first texel), v = 0
  (outside: clamped) and negative coordinates;
* the reference's own path: ref_driver --tex-probe runs Scene::intersect and
  LambertBRDF::sample on camera rays of the reference's texture scenes and of
  crafted ones; the oracle's hit, interpolated coordinates, texel index and
  reflectance must match bit for bit;
* whole films: the reference's rainbow-texture scenes are a textured plane
  under a constant sky, so every bounce escapes and the film does not depend
  on the random directions; the reference's film equals the oracle's.
"""
import json
import struct
import zlib

import numpy as np
import pytest

import oracle as O
from polytope_amd import pbrt as P
from polytope_amd import png as G
from polytope_amd import scenes as S

needs_ref = pytest.mark.skipif(not O.reference_available(), reason="reference build (oracle/_ref) not present")


def _fnv(b: bytes) -> str:
    h = 1469598103934665603
    a = np.frombuffer(b, np.uint8)
    for c in a.tolist():
        h = ((h ^ c) * 1099511628211) & 0xFFFFFFFFFFFFFFFF
    return "%016x" % h


def _bits(a):
    return np.asarray(a, np.float32).view(np.uint32).ravel().tolist()


# ----------------------------------------------------------------------------- a PNG writer for every form

def _pack(samples: np.ndarray, bits: int) -> bytes:
    """One scanline of integer samples at `bits` per sample, MSB first."""
    if bits == 8:
        return samples.astype(np.uint8).tobytes()
    if bits == 16:
        return samples.astype(">u2").tobytes()
    bitarr = ((samples[:, None] >> np.arange(bits - 1, -1, -1)) & 1).astype(np.uint8).ravel()
    return np.packbits(bitarr).tobytes()


def _filter_rows(rows, bpp: int, filters):
    out = []
    prior = bytes(len(rows[0])) if rows else b""
    for y, line in enumerate(rows):
        ft = filters[y % len(filters)]
        cur = np.frombuffer(line, np.uint8).astype(np.int32)
        up = np.frombuffer(prior, np.uint8).astype(np.int32)
        left = np.concatenate([np.zeros(bpp, np.int32), cur[:-bpp]]) if len(cur) else cur
        ul = np.concatenate([np.zeros(bpp, np.int32), up[:-bpp]]) if len(up) else up
        if ft == 0:
            f = cur
        elif ft == 1:
            f = cur - left
        elif ft == 2:
            f = cur - up
        elif ft == 3:
            f = cur - ((left + up) >> 1)
        else:
            p = left + up - ul
            pa, pb, pc = np.abs(p - left), np.abs(p - up), np.abs(p - ul)
            f = cur - np.where((pa <= pb) & (pa <= pc), left, np.where(pb <= pc, up, ul))
        out.append(bytes([ft]) + (f & 255).astype(np.uint8).tobytes())
        prior = line
    return b"".join(out)


def make_png(samples: np.ndarray, color: int, bits: int, interlace: int = 0, filters=(0,), plte=None,
             trns=None) -> bytes:
    """samples: (H, W, channels) integers of the colour type's channels."""
    h, w, ch = samples.shape
    bpp = max(1, ch * bits // 8)

    def image(sub):
        rows = [_pack(sub[y].reshape(-1), bits) for y in range(sub.shape[0])]
        return _filter_rows(rows, bpp, filters)

    if interlace == 0:
        raw = image(samples)
    else:
        parts = []
        for x0, y0, dx, dy in ((0, 0, 8, 8), (4, 0, 8, 8), (0, 4, 4, 8), (2, 0, 4, 4), (0, 2, 2, 4),
                               (1, 0, 2, 2), (0, 1, 1, 2)):
            sub = samples[y0::dy, x0::dx]
            if sub.shape[0] and sub.shape[1]:
                parts.append(image(sub))
        raw = b"".join(parts)

    def chunk(t, body):
        return struct.pack(">I", len(body)) + t + body + struct.pack(">I", zlib.crc32(t + body) & 0xFFFFFFFF)

    out = b"\x89PNG\r\n\x1a\n" + chunk(b"IHDR", struct.pack(">IIBBBBB", w, h, bits, color, 0, 0, interlace))
    if plte is not None:
        out += chunk(b"PLTE", np.asarray(plte, np.uint8).tobytes())
    if trns is not None:
        out += chunk(b"tRNS", trns)
    out += chunk(b"tEXt", b"Comment\x00crafted")
    return out + chunk(b"IDAT", zlib.compress(raw, 9)) + chunk(b"IEND", b"")


def _forms():
    rng = np.random.default_rng(7)
    w, h = 13, 11
    f = {}
    for bits in (1, 2, 4, 8, 16):
        f[f"grey{bits}"] = (rng.integers(0, 1 << bits, (h, w, 1)), 0, bits, {})
    f["grey8_key"] = (np.where(rng.random((h, w, 1)) < 0.3, 77, rng.integers(0, 256, (h, w, 1))), 0, 8,
                      {"trns": struct.pack(">H", 77)})
    f["rgb8"] = (rng.integers(0, 256, (h, w, 3)), 2, 8, {"filters": (0, 1, 2, 3, 4)})
    f["rgb16"] = (rng.integers(0, 65536, (h, w, 3)), 2, 16, {"filters": (4, 3, 2, 1)})
    key = rng.integers(0, 256, (h, w, 3))
    key[::3, ::2] = (10, 20, 30)
    f["rgb8_key"] = (key, 2, 8, {"trns": struct.pack(">HHH", 10, 20, 30)})
    for bits in (1, 2, 4, 8):
        n = min(1 << bits, 19)
        pal = rng.integers(0, 256, (n, 3))
        idx = rng.integers(0, min(1 << bits, n + 2), (h, w, 1))      # a few indices past the palette
        f[f"palette{bits}"] = (idx, 3, bits, {"plte": pal, "trns": bytes(rng.integers(0, 256, n // 2).tolist())})
    f["grey_alpha8"] = (rng.integers(0, 256, (h, w, 2)), 4, 8, {"filters": (1, 4)})
    f["grey_alpha16"] = (rng.integers(0, 65536, (h, w, 2)), 4, 16, {})
    f["rgba8"] = (rng.integers(0, 256, (h, w, 4)), 6, 8, {"filters": (0, 1, 2, 3, 4)})
    f["rgba16"] = (rng.integers(0, 65536, (h, w, 4)), 6, 16, {})
    f["rgb8_adam7"] = (rng.integers(0, 256, (h, w, 3)), 2, 8, {"interlace": 1, "filters": (0, 4, 1)})
    f["palette2_adam7"] = (rng.integers(0, 4, (h, w, 1)), 3, 2, {"interlace": 1, "plte": rng.integers(0, 256, (4, 3))})
    f["grey4_adam7"] = (rng.integers(0, 16, (h, w, 1)), 0, 4, {"interlace": 1})
    return f


FORMS = _forms()


def _expected_rgba(samples, color, bits, plte=None, trns=None, **_):
    """RGBA8 that the PNG spec and lodepng's 8-bit conversion give (closed form)."""
    s = np.asarray(samples, np.int64)
    h, w = s.shape[:2]
    out = np.zeros((h, w, 4), np.uint8)
    hi = (lambda v: v >> 8) if bits == 16 else (lambda v: v)
    if color == 0:
        g = (s[..., 0] * 255) // ((1 << bits) - 1) if bits < 8 else hi(s[..., 0])
        out[..., 0] = out[..., 1] = out[..., 2] = g
        out[..., 3] = 255
        if trns is not None:
            out[..., 3][s[..., 0] == struct.unpack(">H", trns)[0]] = 0
    elif color == 2:
        out[..., :3] = hi(s)
        out[..., 3] = 255
        if trns is not None:
            k = struct.unpack(">HHH", trns)
            out[..., 3][(s[..., 0] == k[0]) & (s[..., 1] == k[1]) & (s[..., 2] == k[2])] = 0
    elif color == 3:
        pal = np.zeros((256, 4), np.int64)
        pal[:, 3] = 255
        pal[: len(plte), :3] = plte
        if trns is not None:
            pal[: len(trns), 3] = list(trns)
        out[:] = pal[s[..., 0]]
    elif color == 4:
        out[..., 0] = out[..., 1] = out[..., 2] = hi(s[..., 0])
        out[..., 3] = hi(s[..., 1])
    else:
        out[:] = hi(s)
    return out


@pytest.mark.parametrize("name", sorted(FORMS))
def test_png_forms_decode_to_the_closed_form(name):
    samples, color, bits, kw = FORMS[name]
    blob = make_png(samples, color, bits, **kw)
    got = G.decode_png(blob)
    assert np.array_equal(got, _expected_rgba(samples, color, bits, **kw))


def test_png_round_trip_and_errors():
    img = S.noise_texture(37, 29, 3).rgba
    for filters in ((0,), (1,), (2,), (3,), (4,), (0, 1, 2, 3, 4)):
        assert np.array_equal(G.decode_png(G.encode_png(img, filters)), img)
    blob = bytearray(G.encode_png(img))
    blob[40] ^= 1                                     # inside IDAT: CRC mismatch
    with pytest.raises(G.PngError, match="CRC"):
        G.decode_png(bytes(blob))
    with pytest.raises(G.PngError, match="signature"):
        G.decode_png(b"GIF89a" + bytes(40))
    with pytest.raises(G.PngError, match="colour type"):
        G.decode_png(make_png(np.zeros((2, 2, 1), np.int64), 0, 3))          # grey at 3 bits


# ----------------------------------------------------------------------------- crafted textured scenes

def _ply_uv(path, v, f, uv, names=("u", "v"), normals=None, binary=False):
    props = ["x", "y", "z"] + (["nx", "ny", "nz"] if normals is not None else []) + list(names)
    cols = [v[:, 0], v[:, 1], v[:, 2]] + ([normals[:, k] for k in range(3)] if normals is not None else []) + \
           [uv[:, 0], uv[:, 1]]
    vals = np.stack(cols, 1).astype(np.float32)
    head = "ply\nformat %s 1.0\nelement vertex %d\n" % ("binary_little_endian" if binary else "ascii", len(v))
    head += "".join("property float %s\n" % p for p in props)
    head += "element face %d\nproperty list uchar int vertex_indices\nend_header\n" % len(f)
    with open(path, "wb") as fh:
        fh.write(head.encode())
        if binary:
            fh.write(vals.astype("<f4").tobytes())
            rec = np.zeros(len(f), np.dtype([("n", "u1"), ("i", "<i4", 3)]))
            rec["n"] = 3
            rec["i"] = f
            fh.write(rec.tobytes())
        else:
            fh.write("".join(" ".join("%.9g" % x for x in row) + "\n" for row in vals).encode())
            fh.write("".join("3 %d %d %d\n" % tuple(a) for a in f).encode())


def _crafted(d):
    """A textured scene: a floor with texture coordinates past [0, 1], a textured
    sphere with normals (s/t property names), a binary PLY quad (stored twice,
    Q6), a Kd sphere, a float texture and a PLY without coordinates."""
    (d / "tex_a.png").write_bytes(G.encode_png(S.noise_texture(61, 37, 21).rgba, (0, 1, 2, 3, 4)))
    (d / "tex_b.png").write_bytes(make_png(FORMS["palette4"][0], 3, 4, plte=FORMS["palette4"][3]["plte"]))
    (d / "tex_c.png").write_bytes(make_png(FORMS["rgb8_adam7"][0], 2, 8, interlace=1, filters=(4, 0, 1)))
    gv, gf = S.quad(6.0, 0.0)
    _ply_uv(d / "floor.ply", gv, gf, np.array([[-0.25, -0.25], [2.25, -0.25], [2.25, 2.25], [-0.25, 2.25]]))
    sv, sf = S.latlong_sphere(12, 24, 1.0)
    _ply_uv(d / "ball.ply", sv, sf, S.latlong_uv(12, 24), names=("s", "t"), normals=sv)
    qv, qf = S.quad(1.0, 0.0)
    _ply_uv(d / "card.ply", qv, qf, np.array([[0.1, 0.1], [0.9, 0.1], [0.9, 0.9], [0.1, 0.9]]), binary=True)
    with open(d / "plain.ply", "w") as fh:
        fh.write("ply\nformat ascii 1.0\nelement vertex 4\nproperty float x\nproperty float y\nproperty float z\n"
                 "element face 2\nproperty list uchar int vertex_indices\nend_header\n"
                 "-1 0 -1\n1 0 -1\n1 0 1\n-1 0 1\n3 0 1 2\n3 0 2 3\n")
    return d


TEXTURED = """LookAt 0 3 -7  0 0.8 0  0 1 0
Camera "perspective" "float fov" [ 50 ]
Film "image" "integer xresolution" [ 64 ] "integer yresolution" [ 48 ]
Sampler "halton" "integer pixelsamples" [ 2 ]
Integrator "path" "integer maxdepth" [ 4 ]
WorldBegin
LightSource "infinite" "rgb L" [ 0.7 0.75 0.8 ]
Texture "a" "color" "imagemap" "string filename" [ "tex_a.png" ]
Texture "b" "spectrum" "imagemap" "string filename" "tex_b.png"
Texture "c" "color" "imagemap" "string filename" [ "tex_c.png" ]
Material "matte" "texture Kd" [ "a" ]
Shape "plymesh" "string filename" [ "floor.ply" ]
AttributeBegin
  Translate -1.5 1.2 0.5
  Material "matte" "texture Kd" "b"
  Shape "plymesh" "string filename" [ "ball.ply" ]
AttributeEnd
AttributeBegin
  Translate 1.6 1.0 -0.5
  Rotate -70 1 0 0
  Material "matte" "texture Kd" [ "c" ]
  Shape "plymesh" "string filename" [ "card.ply" ]
AttributeEnd
AttributeBegin
  Translate 1.5 0.8 1.5
  Scale 0.8 0.8 0.8
  Material "matte" "rgb Kd" [ 0.6 0.5 0.4 ]
  Shape "plymesh" "string filename" [ "ball.ply" ]
AttributeEnd
WorldEnd
"""


def _describe(path, tmp_path):
    out = tmp_path / "describe.json"
    O.run_reference(["--pbrt", str(path), "--describe", str(out)], cwd=str(tmp_path))
    return json.loads(out.read_text())


def test_reader_builds_textured_meshes(tmp_path):
    _crafted(tmp_path)
    (tmp_path / "s.pbrt").write_text(TEXTURED)
    sf = P.read_pbrt(str(tmp_path / "s.pbrt"))
    floor, ball, card, kd = sf.meshes
    assert floor.texture is not None and floor.texture.width == 61 and floor.uv[1].tolist() == [2.25, -0.25]
    assert ball.texture is not None and np.array_equal(ball.uv, S.latlong_uv(12, 24))
    assert card.texture is not None and len(card.verts) == 8 and len(card.uv) == 4     # Q6 doubles positions only
    assert kd.texture is None and kd.uv is None and kd.refl == tuple(float(np.float32(x)) for x in (0.6, 0.5, 0.4))
    assert np.array_equal(card.texture.rgba, _expected_rgba(*FORMS["rgb8_adam7"][:3]))
    assert ball.texture is not floor.texture


@needs_ref
def test_textured_scene_matches_reference_parser(tmp_path):
    _crafted(tmp_path)
    (tmp_path / "s.pbrt").write_text(TEXTURED)
    sf = P.read_pbrt(str(tmp_path / "s.pbrt"))
    ref = _describe(tmp_path / "s.pbrt", tmp_path)
    for rm, m in zip(ref["meshes"], sf.meshes):
        assert rm["has_uv"]                       # every PLY mesh (mesh_parsers.cpp:320)
        if m.texture is None:
            assert "texture" not in rm
            continue
        t = rm["texture"]
        assert (t["width"], t["height"], t["format"]) == (m.texture.width, m.texture.height, "rgba")
        assert t["fnv1a"] == _fnv(m.texture.rgba.tobytes())
        assert rm["u"] == _bits(m.uv[:, 0]) and rm["v"] == _bits(m.uv[:, 1])


@needs_ref
def test_png_forms_decode_like_the_reference(tmp_path):
    """Each crafted PNG as a texture of a one-triangle scene: the reference
    parser's decoded bytes equal this decoder's."""
    v = np.array([[-1, -1, 0], [-1, 1, 0], [1, -1, 0]], np.float32)
    _ply_uv(tmp_path / "tri.ply", v, np.array([[0, 1, 2]]), np.array([[0, 0], [0, 1], [1, 0]]))
    lines = ['LookAt 0 0 -10 0 0 0 0 1 0', 'Camera "perspective" "float fov" [ 10 ]', "WorldBegin"]
    names = sorted(FORMS)
    for name in names:
        samples, color, bits, kw = FORMS[name]
        (tmp_path / f"{name}.png").write_bytes(make_png(samples, color, bits, **kw))
        lines += [f'Texture "{name}" "color" "imagemap" "string filename" [ "{name}.png" ]',
                  f'Material "matte" "texture Kd" [ "{name}" ]', 'Shape "plymesh" "string filename" [ "tri.ply" ]']
    (tmp_path / "forms.pbrt").write_text("\n".join(lines + ["WorldEnd"]) + "\n")
    ref = _describe(tmp_path / "forms.pbrt", tmp_path)
    for name, rm in zip(names, ref["meshes"]):
        samples, color, bits, kw = FORMS[name]
        mine = G.decode_png((tmp_path / f"{name}.png").read_bytes())
        assert rm["texture"]["fnv1a"] == _fnv(mine.tobytes()), name
        assert (rm["texture"]["width"], rm["texture"]["height"]) == (mine.shape[1], mine.shape[0])


@needs_ref
def test_reference_texture_file_decodes_identically(tmp_path):
    path = O.reference_scene("test/texture/rainbow-texture.pbrt")
    ref = _describe(path, tmp_path)
    sf = P.read_pbrt(path)
    t = ref["meshes"][0]["texture"]
    assert (t["width"], t["height"]) == (1197, 873)
    assert t["fnv1a"] == _fnv(sf.meshes[0].texture.rgba.tobytes())


def test_reader_texture_errors_and_unrenderable_meshes(tmp_path):
    _crafted(tmp_path)
    head = "WorldBegin\n"
    cases = {
        'Texture "a" "color" "imagemap" "string filename" "tex_a.png"\n'
        'Texture "a" "color" "imagemap" "string filename" "tex_a.png"\n': "already previously been defined",
        'Texture "a" "rgb" "imagemap" "string filename" "tex_a.png"\n': "texture type",
        'Texture "a" "color" "checkerboard" "string filename" "tex_a.png"\n': "texture class",
        'Texture "a" "color" "imagemap" "float scale" [ 2 ]\n': "insufficient",
        'Texture "a" "color" "imagemap" "string filename" "missing.png"\n': "Failed to read",
        'Texture "a" "color" "imagemap" "string filename" "floor.ply"\n': "Failed to read",
        'Material "matte" "texture Kd" "nope"\n': "hasn't been defined yet",
    }
    for body, what in cases.items():
        (tmp_path / "e.pbrt").write_text(head + body + "WorldEnd\n")
        with pytest.raises(P.PbrtError, match=what):
            P.read_pbrt(str(tmp_path / "e.pbrt"))
    (tmp_path / "u.pbrt").write_text(
        head + 'Texture "g" "float" "imagemap" "string filename" "tex_a.png"\n'
        'Texture "a" "color" "imagemap" "string filename" "tex_a.png"\n'
        'Material "matte" "texture Kd" "g"\nShape "plymesh" "string filename" "floor.ply"\n'
        'Material "matte" "texture Kd" "a"\nShape "plymesh" "string filename" "plain.ply"\n'
        'Shape "sphere" "float radius" 1\nShape "plymesh" "string filename" "floor.ply"\nWorldEnd\n')
    with pytest.raises(P.PbrtError, match="float texture"):
        P.read_pbrt(str(tmp_path / "u.pbrt"))
    sf = P.read_pbrt(str(tmp_path / "u.pbrt"), strict=False)
    assert [i for i, _ in sf.unrenderable] == [0, 1, 2]
    assert "texture coordinates" in sf.unrenderable[1][1] and "texture coordinates" in sf.unrenderable[2][1]
    assert sf.meshes[3].texture is not None


def test_writer_round_trip_keeps_textures(tmp_path):
    meshes, cam, sky = S.textured_render_scene(1, 32, 24)
    meshes = [m for m in meshes if not np.any(m.object_to_world[:3, :3] != np.diag(np.diag(m.object_to_world[:3, :3])))]
    path, out, cam2, sky2 = S.write_pbrt(str(tmp_path), meshes, (0, 2.2, -6), (0, 0.8, 1.5), (0, 1, 0), 60.0, 32, 24,
                                         spp=1, max_depth=3, skybox=sky)
    sf = P.read_pbrt(path)
    for a, b in zip(out, sf.meshes):
        assert (a.texture is None) == (b.texture is None)
        if a.texture is not None:
            assert np.array_equal(a.texture.rgba, b.texture.rgba) and np.array_equal(a.uv, b.uv)


# ----------------------------------------------------------------------------- texel lookup

def test_texel_index_closed_forms():
    W, H = 10, 8
    assert O.texel_index(0.0, 1.0, W, H) == 0                    # top-left texel
    assert O.texel_index(0.55, 0.5, W, H) == 4 * W + 5
    assert O.texel_index(0.999, 0.001, W, H) == (H - 1) * W + 9   # bottom-right
    assert O.texel_index(1.0, 0.5, W, H) == 5 * W                 # column W: the next row's first texel
    assert O.texel_index(1.0, 0.125 + 1e-6, W, H) == 7 * W        # still inside the image
    assert O.texel_index(0.3, 0.0, W, H) == (H - 1) * W + 3       # row H: outside, clamped
    assert O.texel_index(1.0, 0.0, W, H) == (H - 1) * W + 9       # past the last texel: clamped
    assert O.texel_index(-0.5, 1.5, W, H) == 0                    # negative column and row
    assert O.texel_index(2.5, 0.5, W, H) == 4 * W + 25            # column 25: two rows on, column 5
    assert O.texel_index(float("nan"), 0.5, W, H) == 4 * W
    assert O.texel_index(1e30, -1e30, W, H) == (H - 1) * W + 9
    # the interpolation is (u0 a + u1 b) + u2 c in float32, unfused
    a, b, c = np.float32(0.1), np.float32(0.7), np.float32(0.2)
    uv = np.float32([0.3, 0.9, 0.6, 0.2, 0.15, 0.55])
    tu, tv = O.tex_lerp(uv, float(a), float(b), float(c))
    assert tu == float(np.float32(np.float32(uv[0] * a) + np.float32(uv[2] * b)) + np.float32(uv[4] * c))
    assert tv == float(np.float32(np.float32(uv[1] * a) + np.float32(uv[3] * b)) + np.float32(uv[5] * c))


TEX_PROBE = np.dtype([("flags", "<u4"), ("mesh", "<u4"), ("face", "<u4"), ("tu", "<f4"), ("tv", "<f4"),
                      ("index", "<u4"), ("r", "<f4"), ("g", "<f4"), ("b", "<f4")])


def _camera_rays(cam, step):
    from polytope_amd import camera_desc
    import ctypes as C
    cd = camera_desc(cam)
    rays = []
    for y in range(0, cam.height, step):
        for x in range(0, cam.width, step):
            out = np.zeros(6, np.float32)
            O.lib().orc_camera_ray(C.byref(cd), C.c_float(x + 0.5), C.c_float(y + 0.5), out.ctypes.data)
            rays.append(out)
    return np.array(rays, np.float32)


def _probe_matches(path, tmp_path, step):
    sf = P.read_pbrt(str(path))
    rays = _camera_rays(sf.camera, step)
    S.write_rays(str(tmp_path / "rays.bin"), rays)
    O.run_reference(["--pbrt", str(path), "--rays", "rays.bin", "--tex-probe", "probe.bin"], cwd=str(tmp_path))
    ref = np.fromfile(tmp_path / "probe.bin", TEX_PROBE)
    orc = O.OracleScene(sf.meshes, sf.skybox)
    hits = orc.intersect(rays)
    assert np.array_equal(ref["flags"] & 1, hits["hit"])
    seen = {"textured": 0, "inside": 0, "outside": 0}
    for i in np.nonzero(ref["flags"] & 4)[0]:
        h, m = hits[i], sf.meshes[int(ref["mesh"][i])]
        assert (h["mesh"], h["face"]) == (ref["mesh"][i], ref["face"][i])
        f = m.faces[h["face"]]
        uv6 = np.concatenate([m.uv[f[0]], m.uv[f[1]], m.uv[f[2]]])
        tu, tv = O.tex_lerp(uv6, float(h["u"]), float(h["v"]), float(h["w"]))
        assert _bits([tu, tv]) == _bits([ref["tu"][i], ref["tv"][i]]), i
        seen["textured"] += 1
        if ref["flags"][i] & 8:
            k = O.texel_index(tu, tv, m.texture.width, m.texture.height)
            assert k == ref["index"][i], i
            px = m.texture.rgba.reshape(-1, 4)[k, :3].astype(np.float32) * np.float32(1.0 / 255.0)
            assert _bits(px) == _bits([ref["r"][i], ref["g"][i], ref["b"][i]]), i
            seen["inside"] += 1
        else:
            seen["outside"] += 1
    return seen


def test_texture_lookups_match_reference_golden(tmp_path):
    """The reference's lookups on the crafted scene (tests/golden/tex_probe_crafted.npz,
    written by tests/golden/make_golden.py) against the oracle's, without the
    reference build: the scene's files are regenerated here from the seeds."""
    from pathlib import Path
    with np.load(Path(__file__).parent / "golden" / "tex_probe_crafted.npz", allow_pickle=False) as z:
        rays, ref = z["rays"], z["probe"].reshape(-1).view(TEX_PROBE)
    _crafted(tmp_path)
    (tmp_path / "s.pbrt").write_text(TEXTURED)
    sf = P.read_pbrt(str(tmp_path / "s.pbrt"))
    assert np.array_equal(rays, _camera_rays(sf.camera, 1))
    hits = O.OracleScene(sf.meshes, sf.skybox).intersect(rays)
    assert np.array_equal(ref["flags"] & 1, hits["hit"])
    inside = 0
    for i in np.nonzero(ref["flags"] & 4)[0]:
        h, m = hits[i], sf.meshes[int(ref["mesh"][i])]
        assert (h["mesh"], h["face"]) == (ref["mesh"][i], ref["face"][i])
        f = m.faces[h["face"]]
        tu, tv = O.tex_lerp(np.concatenate([m.uv[f[0]], m.uv[f[1]], m.uv[f[2]]]), float(h["u"]), float(h["v"]),
                            float(h["w"]))
        assert _bits([tu, tv]) == _bits([ref["tu"][i], ref["tv"][i]]), i
        if ref["flags"][i] & 8:
            k = O.texel_index(tu, tv, m.texture.width, m.texture.height)
            assert k == ref["index"][i], i
            px = m.texture.rgba.reshape(-1, 4)[k, :3].astype(np.float32) * np.float32(1.0 / 255.0)
            assert _bits(px) == _bits([ref["r"][i], ref["g"][i], ref["b"][i]]), i
            inside += 1
    assert inside > 1000


@needs_ref
def test_texture_lookups_match_reference_on_crafted_scene(tmp_path):
    _crafted(tmp_path)
    (tmp_path / "s.pbrt").write_text(TEXTURED.replace("[ 64 ]", "[ 160 ]").replace("[ 48 ]", "[ 120 ]"))
    seen = _probe_matches(tmp_path / "s.pbrt", tmp_path, 1)
    assert seen["inside"] > 5000


@needs_ref
@pytest.mark.parametrize("rel", ["test/texture/rainbow-texture.pbrt", "test/texture/rainbow-texture-2.pbrt"])
def test_texture_lookups_match_reference_on_its_scenes(rel, tmp_path):
    seen = _probe_matches(O.reference_scene(rel), tmp_path, 3)
    assert seen["inside"] > 1000


@needs_ref
@pytest.mark.parametrize("rel", ["test/texture/rainbow-texture.pbrt", "test/texture/rainbow-texture-2.pbrt"])
def test_reference_film_of_its_texture_scenes_equals_the_oracle(rel, tmp_path):
    """A textured plane under a constant sky: every bounce escapes, so the film
    is sky x texel at every covered pixel whatever the random directions are,
    and the reference's film is reproducible."""
    path = O.reference_scene(rel)
    sf = P.read_pbrt(path)
    O.run_reference(["--pbrt", path, "--render", "film.bin", "--threads", "8"], cwd=str(tmp_path), timeout=900)
    ref = np.fromfile(tmp_path / "film.bin", np.float32).reshape(sf.camera.height, sf.camera.width, 3)
    film = O.OracleScene(sf.meshes, sf.skybox).render(sf.camera, sf.spp, sf.max_depth, seed=5)
    same = (film.view(np.uint32) == ref.view(np.uint32)).all(2)
    covered = (ref != ref[0, 0]).any(2)
    assert covered.mean() > 0.05
    assert same.all(), np.argwhere(~same)[:10]
