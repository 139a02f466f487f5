"""Full-size render parity: the batched render workload (a displaced icosphere
of 1,310,720 faces, 32 rigid instances and a ground quad: 43,253,762 faces in
one flat BVH; 70 % Lambert / 30 % mirror), rendered in one batch of jobs of
uneven resolution; the oracle recomputes every 61st row of each job with the
tree the library built (the builder's parity is checked on its own in
test_abi_and_builder).  Bar: those rows bit-identical; every film finite and
within [0, 255]."""
from __future__ import annotations

import numpy as np
import pytest

import oracle as O
from polytope_amd import Device, Scene, render_batch
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


@pytest.mark.timeout(900)
def test_gpu_full_size_batched_render_rows_match_oracle(dev):
    print("case full-size render: building 43M-face scene", flush=True)
    meshes, sky, jobs = S.batched_render_workload(8, 32, 16, 8, 8)
    jobs = [jobs[0], jobs[7], jobs[15]]                 # 256x256, mid size, 1920x1080; 8 spp, depth 8
    sc = Scene(dev, meshes, sky)
    print(f"case full-size render: scene built ({sc.info()['build_seconds']:.1f} s)", flush=True)
    films = render_batch([(sc, cam, spp, depth, seed) for cam, spp, depth, seed in jobs])
    for f in films:
        assert np.isfinite(f).all() and f.min() >= 0 and f.max() <= 255
    oracle = O.OracleScene(meshes, sky, bvh=sc.bvh())
    for (cam, spp, depth, seed), film in zip(jobs, films):
        print(f"case full-size render: oracle rows of {cam.width}x{cam.height}", flush=True)
        ref = oracle.render(cam, spp, depth, seed, row_step=61)
        rows = np.arange(0, cam.height, 61)
        assert np.array_equal(film[rows].view(np.uint32), ref[rows].view(np.uint32))
    sc.close()
