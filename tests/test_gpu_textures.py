"""Textured Lambert on the GPU, in every render form, against the oracle: films
bit for bit.  Scenes: the seeded textured scene (a floor whose texture
coordinates run past the image, so lookups take the next row's texel or clamp;
a textured lat-long sphere; a Kd sphere; a mirror) and a scene file written
here with PNG textures and PLY u/v properties (nothing of the reference is
read).  Batches mix textured, glossy and plain scenes; next-event estimation
uses the texel as Kd."""
from __future__ import annotations

import numpy as np
import pytest

import oracle as O
from polytope_amd import Device, Scene, Server, _abi, render, render_batch
from polytope_amd import pbrt as P
from polytope_amd import scenes as S
from tests.test_textures import TEXTURED, _crafted

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")


@pytest.fixture(scope="module")
def dev():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    d = Device(0)
    yield d
    d.close()


def _same(a, b):
    return np.array_equal(a.view(np.uint32), b.view(np.uint32))


@pytest.mark.parametrize("spp,depth", [(3, 5), (2, 1)])
def test_textured_all_forms_match_oracle(dev, spp, depth):
    print(f"case textured scene spp={spp} depth={depth}", flush=True)
    meshes, cam, sky = S.textured_render_scene(2, 80, 60)
    ref = O.OracleScene(meshes, sky).render(cam, spp, depth, seed=11)
    assert len(np.unique(ref.reshape(-1, 3), axis=0)) > 500          # textures visible
    sc = Scene(dev, meshes, sky)
    assert _same(render(sc, cam, spp, depth, 11), ref)
    for mode in (_abi.RENDER_REGEN, _abi.RENDER_WAVEFRONT):
        dev.set_tuning(render_mode=mode)
        try:
            (b,) = render_batch([(sc, cam, spp, depth, 11)])
        finally:
            dev.set_tuning()
        assert _same(b, ref), mode
    sc.close()


def test_textured_scene_file_matches_oracle(dev, tmp_path):
    print("case textured scene file", flush=True)
    _crafted(tmp_path)
    (tmp_path / "s.pbrt").write_text(TEXTURED)
    sf = P.read_pbrt(str(tmp_path / "s.pbrt"))
    ref = O.OracleScene(sf.meshes, sf.skybox).render(sf.camera, 2, 4, 5)
    sc = Scene(dev, sf.meshes, sf.skybox)
    assert _same(render(sc, sf.camera, 2, 4, 5), ref)
    assert _same(render_batch([(sc, sf.camera, 2, 4, 5)])[0], ref)
    sc.close()


def test_textured_mixed_batch_nee_and_server(dev):
    print("case textured mixed batch + NEE + server", flush=True)
    tm, tcam, tsky = S.textured_render_scene(3, 48, 36)
    pm, pcam, psky = S.plain_render_scene(1, 40, 30)
    gv, gf = S.icosphere(3, radius=0.7, center=(0.0, 0.7, 0.0), noise=0.03, seed=8)
    fv, ff = S.quad(4.0, 0.0)
    gm = [S.MeshData(gv, gf, brdf=S.GLOSSY, refl=(0.3, 0.5, 0.7), specular=(0.8, 0.8, 0.7), roughness=0.3),
          S.MeshData(fv, ff, refl=(0.5, 0.5, 0.5))]
    gcam = S.Camera(S.look_at((0.0, 1.5, -3.5), (0.0, 0.6, 0.0)), 55.0, 44, 33)
    lights = [((0.5, 4.0, -1.0), (900.0, 880.0, 860.0))]
    t = Scene(dev, tm, tsky, lights=lights)
    p = Scene(dev, pm, psky)
    g = Scene(dev, gm, (70.0, 90.0, 120.0))
    jobs = [(t, tcam, 2, 4, 3), (p, pcam, 2, 4, 4), (g, gcam, 2, 4, 5), (t, tcam, 2, 4, 6, True)]
    refs = [O.OracleScene(tm, tsky, lights=lights).render(tcam, 2, 4, 3),
            O.OracleScene(pm, psky).render(pcam, 2, 4, 4),
            O.OracleScene(gm, (70.0, 90.0, 120.0)).render(gcam, 2, 4, 5),
            O.OracleScene(tm, tsky, lights=lights).render(tcam, 2, 4, 6, nee=True)]
    for f, r in zip(render_batch(jobs), refs):
        assert _same(f, r)
    assert _same(render(t, tcam, 2, 4, 6, nee=True), refs[3])
    dev.set_tuning(render_mode=_abi.RENDER_WAVEFRONT)
    try:
        wf = render_batch(jobs)
    finally:
        dev.set_tuning()
    for f, r in zip(wf, refs):
        assert _same(f, r)
    srv = Server(dev, hbm_budget_bytes=2 << 30, paths_per_wave=1 << 15)
    try:
        ids = [srv.submit(*j) for j in jobs]
        for i, r in zip(ids, refs):
            assert _same(srv.wait(i)[0], r)
    finally:
        srv.close()
    for s in (t, p, g):
        s.close()


def test_textured_scene_rejects_bad_descriptions(dev):
    print("case textured scene validation", flush=True)
    meshes, cam, sky = S.textured_render_scene(1, 16, 12)
    bad = [S.MeshData(m.verts, m.faces, object_to_world=m.object_to_world, brdf=m.brdf, refl=m.refl,
                      texture=m.texture, uv=m.uv) for m in meshes]
    bad[0].uv = bad[0].uv[:2]                             # faces index past the coordinates
    with pytest.raises(Exception, match="texture coordinates"):
        Scene(dev, bad, sky)
    bad[0].uv = meshes[0].uv
    bad[0].brdf = S.MIRROR                                # only Lambert meshes take a texture
    with pytest.raises(Exception, match="Lambert"):
        Scene(dev, bad, sky)


def test_texture_size_limits_and_far_coordinates(dev):
    """Textures at the size limits (1x1, 32767x1, 1x32767) on quads whose
    coordinates run far outside [0, 1] and below zero: every lookup that the
    reference would make outside the image clamps, on the GPU as in the oracle."""
    print("case texture limits", flush=True)
    rng = np.random.default_rng(3)
    texs = [S.Texture(rng.integers(0, 256, (h, w, 4), dtype=np.uint8)) for w, h in ((1, 1), (32767, 1), (1, 32767))]
    meshes = []
    for k, t in enumerate(texs):
        v, f = S.quad(1.0, 0.0)
        uv = np.array([[-3.0, -2.5], [4.5, -3.0], [5.0, 6.0], [-2.0, 4.0]], np.float32) * (k + 1)
        o2w = (S.translate((-2.3 + 2.3 * k, 0.0, 0.0))).astype(np.float32)
        meshes.append(S.MeshData(v, f, object_to_world=o2w, texture=t, uv=uv))
    cam = S.Camera(S.look_at((0.0, 4.0, -3.0), (0.0, 0.0, 0.0)), 60.0, 96, 64)
    sky = (200.0, 210.0, 220.0)
    ref = O.OracleScene(meshes, sky).render(cam, 2, 3, seed=4)
    sc = Scene(dev, meshes, sky)
    assert _same(render(sc, cam, 2, 3, 4), ref)
    assert _same(render_batch([(sc, cam, 2, 3, 4)])[0], ref)
    assert len(np.unique(ref.reshape(-1, 3), axis=0)) > 50
    sc.close()
