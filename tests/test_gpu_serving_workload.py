"""GPU test of the render service on its serving workload: a Poisson stream of
32 render jobs at 512 per second over displaced-icosphere scenes of 10,242 to
1,966,082 faces, 256x256 to 1920x1080, spp log-uniform in 4..256, depth 4..12
(polytope_amd.scenes.serving_requests), through the continuous-batching server
with jobs arriving while others run.

Every job's film equals the same job rendered alone (ptx_render_batch) bit for
bit, and two rows of every film equal the oracle's restatement of the
reference integrator (PathTraceIntegrator.cpp:10-135 with BoxFilter.cpp:14-71)
on the same tree bit for bit."""
from __future__ import annotations

import time

import numpy as np
import pytest

import oracle as O
from polytope_amd import Device, Scene, Server, render_batch
from polytope_amd import _abi
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


@pytest.mark.timeout(600)
def test_serving_stream_equals_alone_and_oracle(dev):
    print("case serving workload: building 6 scenes (10k..2M faces)", flush=True)
    scene_defs, reqs = S.serving_requests(32, 512.0)     # arrivals compressed so that jobs overlap
    tris = [sum(m.num_faces for m in d[0]) for d in scene_defs]
    assert min(tris) <= 10_500 and max(tris) >= 1_900_000
    assert min(r[3] for r in reqs) >= 4 and max(r[3] for r in reqs) <= 256
    assert {r[4] for r in reqs} <= set(range(4, 13))
    scenes = [Scene(dev, m, sky) for m, sky in scene_defs]
    srv = Server(dev, hbm_budget_bytes=64 << 30)
    try:
        print("case serving workload: 32 requests at 512/s", flush=True)
        ids = []
        t0 = time.perf_counter()
        for arrival, si, cam, spp, depth, seed in reqs:
            delay = arrival - (time.perf_counter() - t0)
            if delay > 0:
                time.sleep(delay)
            ids.append(srv.submit(scenes[si], cam, spp, depth, seed))
        films = []
        overlapped = 0
        ends = []
        for jid in ids:
            film, st = srv.wait(jid)
            assert st["state"] == _abi.JOB_DONE
            assert st["samples_done"] == st["samples_total"]
            films.append(film)
            ends.append((st["start_time"], st["end_time"]))
        # continuous batching: jobs ran at the same time as others
        for k, (s0, e0) in enumerate(ends):
            overlapped += any(s1 < e0 and s0 < e1 for j, (s1, e1) in enumerate(ends) if j != k)
        assert overlapped >= len(ends) // 2, overlapped
    finally:
        srv.close()
    print("case serving workload: each job alone", flush=True)
    for k, ((_, si, cam, spp, depth, seed), film) in enumerate(zip(reqs, films)):
        (alone,) = render_batch([(scenes[si], cam, spp, depth, seed)])
        assert np.array_equal(alone.view(np.uint32), film.view(np.uint32)), k
    print("case serving workload: oracle rows", flush=True)
    oracles = {}
    for k, ((_, si, cam, spp, depth, seed), film) in enumerate(zip(reqs, films)):
        if si not in oracles:
            oracles[si] = O.OracleScene(scene_defs[si][0], scene_defs[si][1])      # the oracle's own tree
        step = max(1, cam.height // 2)
        ref = oracles[si].render(cam, spp, depth, seed, row_step=step)
        rows = np.arange(0, cam.height, step)
        assert np.array_equal(film[rows].view(np.uint32), ref[rows].view(np.uint32)), k
    for sc in scenes:
        sc.close()
