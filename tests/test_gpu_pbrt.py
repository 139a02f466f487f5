"""A scene file read by the .pbrt reader renders on the GPU exactly as the
oracle renders the same meshes: the crafted scene with rotated, shared and
instanced transforms, vertex normals, a tessellated sphere, emitters, a mirror
OBJ mesh and a skybox (files generated here; nothing of the reference is read).
Bar: films bit for bit (plain and batched)."""
from __future__ import annotations

import numpy as np
import pytest

import oracle as O
from polytope_amd import Device, Scene, render, render_batch
from polytope_amd import pbrt as P
from tests.test_pbrt_reader import CRAFTED, _crafted_meshes

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")


@pytest.fixture(scope="module")
def dev():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    d = Device(0)
    yield d
    d.close()


@pytest.mark.parametrize("name", sorted(CRAFTED))
def test_gpu_renders_scene_file_like_oracle(dev, name, tmp_path):
    print(f"case pbrt scene {name}", flush=True)
    _crafted_meshes(tmp_path)
    path = tmp_path / "s.pbrt"
    path.write_text(CRAFTED[name])
    sf = P.read_pbrt(str(path))
    cam = sf.camera
    spp, depth = 2, min(sf.max_depth, 5)
    sc = Scene(dev, sf.meshes, sf.skybox)
    film = render(sc, cam, spp, depth, 7)
    batched = render_batch([(sc, cam, spp, depth, 7)])[0]
    ref = O.OracleScene(sf.meshes, sf.skybox).render(cam, spp, depth, 7)
    assert np.array_equal(film.view(np.uint32), ref.view(np.uint32))
    assert np.array_equal(batched.view(np.uint32), ref.view(np.uint32))
    assert (ref != ref[0, 0]).any(2).mean() > 0.01         # the scene is visible
    sc.close()


def test_gpu_renders_scene_file_point_light_like_oracle(dev, tmp_path):
    print("case pbrt scene point light, next-event estimation", flush=True)
    _crafted_meshes(tmp_path)
    path = tmp_path / "s.pbrt"
    path.write_text(CRAFTED["point_light"])
    sf = P.read_pbrt(str(path))
    sc = Scene(dev, sf.meshes, sf.skybox, lights=sf.lights)
    film = render(sc, sf.camera, 4, sf.max_depth, 3, nee=True)
    ref = O.OracleScene(sf.meshes, sf.skybox, lights=sf.lights).render(sf.camera, 4, sf.max_depth, 3, nee=True)
    plain = O.OracleScene(sf.meshes, sf.skybox).render(sf.camera, 4, sf.max_depth, 3)
    assert np.array_equal(film.view(np.uint32), ref.view(np.uint32))
    assert (ref > plain).any()                             # the light adds to the film
    sc.close()
