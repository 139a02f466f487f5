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
    meshes, sky, jobs = S.batched_render_workload(8, 32, 16, 64, 8)
    # 256x256 at the workload's 64 spp; a mid-size job and 1920x1080 at 8 spp; depth 8
    jobs = [jobs[0], (jobs[7][0], 8, jobs[7][2], jobs[7][3]), (jobs[15][0], 8, jobs[15][2], jobs[15][3])]
    assert jobs[0][1] == 64
    sc = Scene(dev, meshes, sky)
    print(f"case full-size render: scene built ({sc.info()['build_seconds']:.1f} s)", flush=True)
    films = render_batch([(sc, cam, spp, depth, seed) for cam, spp, depth, seed in jobs])
    for f in films:
        assert np.isfinite(f).all() and f.min() >= 0 and f.max() <= 255
    # the oracle builds its own tree (the restated bvh::bound + compact); it is
    # the library's byte for byte
    print("case full-size render: oracle tree", flush=True)
    oracle = O.OracleScene(meshes, sky)
    on, op = oracle.bvh()
    ln, lp = sc.bvh()
    assert np.array_equal(on.view(np.uint8), ln.view(np.uint8)) and np.array_equal(op, lp)
    del on, op, ln, lp
    for (cam, spp, depth, seed), film in zip(jobs, films):
        step = 16 if spp == 64 else 61
        print(f"case full-size render: oracle rows of {cam.width}x{cam.height} at {spp} spp", flush=True)
        ref = oracle.render(cam, spp, depth, seed, row_step=step)
        rows = np.arange(0, cam.height, step)
        assert np.array_equal(film[rows].view(np.uint32), ref[rows].view(np.uint32))
    sc.close()
