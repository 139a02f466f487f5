"""The glossy ("plastic") lobe on the GPU, in every render form, against the
oracle: films bit for bit.  Scenes mix glossy meshes of several roughnesses
with Lambert, mirror and emitting ones under a skybox; batches mix glossy and
non-glossy scenes (the round then runs the kernel variant with the lobe, and
each film still equals the film of its job rendered alone), and next-event
estimation stays at Lambert vertices."""
from __future__ import annotations

import numpy as np
import pytest

import oracle as O
from polytope_amd import Device, Scene, _abi, render, render_batch
from polytope_amd import scenes as S

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")


@pytest.fixture(scope="module")
def dev():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    d = Device(0)
    yield d
    d.close()


def _glossy_scene(w=72, h=54):
    meshes = []
    for k, rough in enumerate((0.0, 0.2, 0.6, 1.0)):
        v, f = S.icosphere(3, radius=0.6, center=(-2.1 + 1.4 * k, 0.6, 0.0), noise=0.04, seed=60 + k)
        meshes.append(S.MeshData(v, f, brdf=S.GLOSSY, refl=(0.2, 0.5, 0.8), specular=(0.9, 0.8, 0.6),
                                 roughness=rough))
    gv, gf = S.quad(5.0, 0.0)
    meshes.append(S.MeshData(gv, gf, brdf=S.LAMBERT, refl=(0.6, 0.6, 0.55)))
    mv, mf = S.icosphere(2, radius=0.5, center=(0.0, 1.8, 1.2), noise=0.0, seed=70)
    meshes.append(S.MeshData(mv, mf, brdf=S.MIRROR, refl=(0.9, 0.9, 0.9)))
    lv, lf = S.quad(0.8, 3.5)
    meshes.append(S.MeshData(lv, lf, is_light=True, emission=(255.0, 230.0, 200.0)))
    cam = S.Camera(S.look_at((0.0, 2.2, -5.0), (0.0, 0.6, 0.0)), 55.0, w, h)
    return meshes, cam, (60.0, 80.0, 110.0)


def _same(a, b):
    return np.array_equal(a.view(np.uint32), b.view(np.uint32))


@pytest.mark.parametrize("spp,depth", [(3, 4), (2, 1)])
def test_glossy_all_forms_match_oracle(dev, spp, depth):
    print(f"case glossy spp={spp} depth={depth}", flush=True)
    meshes, cam, sky = _glossy_scene()
    ref = O.OracleScene(meshes, sky).render(cam, spp, depth, seed=9)
    assert (ref != ref[0, 0]).any(2).mean() > 0.3
    sc = Scene(dev, meshes, sky)
    assert _same(render(sc, cam, spp, depth, 9), ref)
    for mode in (_abi.RENDER_REGEN, _abi.RENDER_WAVEFRONT):
        dev.set_tuning(render_mode=mode)
        try:
            (b,) = render_batch([(sc, cam, spp, depth, 9)])
        finally:
            dev.set_tuning()
        assert _same(b, ref), mode
    sc.close()


def test_glossy_mixed_batch_and_next_event_estimation(dev):
    print("case glossy mixed batch + NEE", flush=True)
    meshes, cam, sky = _glossy_scene(48, 36)
    lights = [((1.0, 3.0, -1.5), (900.0, 850.0, 800.0))]
    plain_meshes, pcam, psky = S.plain_render_scene(1, 40, 30)
    g = Scene(dev, meshes, sky, lights=lights)
    p = Scene(dev, plain_meshes, psky)
    jobs = [(g, cam, 2, 4, 3), (p, pcam, 2, 4, 4), (g, cam, 2, 4, 5, True)]
    films = render_batch(jobs)
    refs = [O.OracleScene(meshes, sky, lights=lights).render(cam, 2, 4, 3),
            O.OracleScene(plain_meshes, psky).render(pcam, 2, 4, 4),
            O.OracleScene(meshes, sky, lights=lights).render(cam, 2, 4, 5, nee=True)]
    for f, r in zip(films, refs):
        assert _same(f, r)
    alone = render_batch([jobs[1]])[0]                  # the plain scene alone uses the lobe-free kernel
    assert _same(alone, films[1])
    assert _same(render(g, cam, 2, 4, 5, nee=True), refs[2])
    g.close()
    p.close()
