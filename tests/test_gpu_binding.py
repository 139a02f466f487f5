"""The drop-in boundary at the reference's own runner API: poly::GpuRunner
(polytope_amd/binding/GpuRunner.cpp), an AbstractRunner compiled against the
reference's headers and sources (oracle/_ref/gpu_runner_check, built by
oracle/ref/Makefile where the reference tree exists), loads a scene into the
reference's Scene / PerspectiveCamera / PathTraceIntegrator / BoxFilter and
renders it through libptx_hip.so.

Bar: the film the reference's BoxFilter returns is bit-identical to the
reference's own CPU film on the deterministic mirror scene, with the library's
BVH or with the tree the reference built, and to the oracle's film on a
Lambert scene (same RNG stream).
"""
from __future__ import annotations

import subprocess

import numpy as np
import pytest

import oracle as O
from polytope_amd import scenes as S
from tests import _fixtures as F

pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(not O.GPU_RUNNER_CHECK.exists(),
                                 reason="oracle/_ref/gpu_runner_check not built (needs the reference tree)")]
torch = pytest.importorskip("torch")


@pytest.fixture(scope="module", autouse=True)
def gpu():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")


def run_check(tmp_path, meshes, cam, spp, depth, sky, seed, scene_bvh=False):
    scene = tmp_path / "scene.bin"
    film = tmp_path / "film.bin"
    S.write_ptxm(str(scene), meshes, cam, spp, depth, sky)
    args = [str(O.GPU_RUNNER_CHECK), "--ptxm", str(scene), "--out", str(film), "--seed", str(seed)]
    if scene_bvh:
        args.append("--scene-bvh")
    print("case", " ".join(args[1:]), flush=True)
    subprocess.run(args, check=True, timeout=300)
    return np.fromfile(film, np.float32).reshape(cam.height, cam.width, 3)


@pytest.mark.parametrize("scene_bvh", [False, True])
def test_gpu_runner_mirror_equals_reference_cpu_film(tmp_path, scene_bvh):
    d = F.load("mirror_render")
    cam, spp, depth = F.camera_of(d)
    film = run_check(tmp_path, F.meshes_of(d), cam, spp, depth, d["sky"], 1, scene_bvh)
    assert np.array_equal(film.view(np.uint32), d["ref_film"].view(np.uint32))


def test_gpu_runner_lambert_equals_oracle(tmp_path):
    meshes, cam, sky = S.plain_render_scene(1, 96, 72)
    film = run_check(tmp_path, meshes, cam, 3, 5, sky, 7)
    ref = O.OracleScene(meshes, sky).render(cam, 3, 5, seed=7)
    assert np.array_equal(film.view(np.uint32), ref.view(np.uint32))


def test_gpu_runner_textured_scene_file_equals_oracle(tmp_path):
    """A scene file with PNG textures and PLY u/v read by the reference's own
    parser (lodepng decode, texture_directive, PLY u/v) and rendered through
    GpuRunner: the oracle's film of the same file read by polytope_amd.pbrt."""
    from polytope_amd import pbrt as P
    from tests.test_textures import TEXTURED, _crafted
    _crafted(tmp_path)
    (tmp_path / "s.pbrt").write_text(TEXTURED)
    film_path = tmp_path / "film.bin"
    args = [str(O.GPU_RUNNER_CHECK), "--pbrt", str(tmp_path / "s.pbrt"), "--out", str(film_path), "--seed", "4"]
    print("case", " ".join(args[1:]), flush=True)
    subprocess.run(args, check=True, timeout=300)
    sf = P.read_pbrt(str(tmp_path / "s.pbrt"))
    film = np.fromfile(film_path, np.float32).reshape(sf.camera.height, sf.camera.width, 3)
    ref = O.OracleScene(sf.meshes, sf.skybox).render(sf.camera, sf.spp, sf.max_depth, seed=4)
    assert np.array_equal(film.view(np.uint32), ref.view(np.uint32))
