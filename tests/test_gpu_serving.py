"""GPU tests of the continuous-batching server (ptx_server_*) and the
autotuner: every job's film is identical to the same job rendered alone,
whatever else is in flight; admission respects the HBM budget; cancellation
and budget rejection report errors without disturbing other jobs."""
from __future__ import annotations

import time

import numpy as np
import pytest

import oracle as O
from polytope_amd import Device, PtxError, Scene, Server, autotune, render_batch
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


def _scenes(dev):
    meshes_a, cam_a, sky_a = S.plain_render_scene(1, 64, 48)
    v, f = S.icosphere(4, noise=0.05, seed=8)
    meshes_b = [S.MeshData(v, f, refl=(0.7, 0.5, 0.3)),
                S.MeshData(*S.quad(4.0, -1.1), brdf=S.MIRROR, refl=(0.8, 0.8, 0.8))]
    return Scene(dev, meshes_a, sky_a), cam_a, Scene(dev, meshes_b, (100.0, 120.0, 150.0)), meshes_b


@pytest.mark.parametrize("mode", [_abi.RENDER_REGEN, _abi.RENDER_WAVEFRONT])
def test_server_results_equal_alone(dev, mode):
    sa, cam_a, sb, _ = _scenes(dev)
    print(f"case server alone mode={mode}", flush=True)
    cam_b = S.Camera(S.look_at((0.3, 0.5, -3.0), (0.0, 0.0, 0.0)), 50.0, 40, 30)
    cam_c = S.Camera(S.look_at((0.3, 0.5, -3.0), (0.0, 0.0, 0.0)), 70.0, 96, 40)
    jobs = [(sa, cam_a, 6, 4, 21), (sb, cam_b, 9, 7, 22), (sb, cam_c, 3, 3, 23), (sa, cam_a, 2, 2, 24),
            (sb, cam_b, 17, 5, 25)]
    alone = [render_batch([j])[0] for j in jobs]
    dev.set_tuning(render_mode=mode)
    srv = Server(dev, hbm_budget_bytes=8 << 30, paths_per_wave=20000)
    try:
        ids = []
        for i, j in enumerate(jobs):        # staggered arrivals: later jobs join rounds in progress
            ids.append(srv.submit(*j))
            if i == 1:
                time.sleep(0.01)
        for jid, a in zip(ids, alone):
            film, st = srv.wait(jid)
            assert st["state"] == _abi.JOB_DONE
            assert st["end_time"] >= st["start_time"] >= st["submit_time"]
            assert st["samples_done"] == st["samples_total"]
            assert np.array_equal(film.view(np.uint32), a.view(np.uint32))
    finally:
        srv.close()
        dev.set_tuning()


def test_server_more_jobs_than_a_round_holds(dev):
    # 70 jobs in flight: rounds take 64 of them in rotation; every film equals the
    # same job rendered alone
    sa, cam_a, sb, _ = _scenes(dev)
    small = S.Camera(cam_a.camera_to_world, 60.0, 16, 12)
    jobs = [(sa, small, 2 + i % 3, 1 + i % 4, 300 + i) for i in range(70)]
    srv = Server(dev, hbm_budget_bytes=8 << 30, max_jobs_in_flight=128, paths_per_wave=20000)
    try:
        ids = [srv.submit(*j) for j in jobs]
        films = [srv.wait(i)[0] for i in ids]
    finally:
        srv.close()
    for i in (0, 5, 63, 64, 69):
        (alone,) = render_batch([jobs[i]])
        assert np.array_equal(alone.view(np.uint32), films[i].view(np.uint32))


def test_server_budget_admission_and_rejection(dev):
    sa, cam_a, sb, _ = _scenes(dev)
    # a budget smaller than the path pool itself
    with pytest.raises(PtxError) as e:
        Server(dev, hbm_budget_bytes=1 << 20, paths_per_wave=1 << 16)
    assert e.value.code == _abi.PTX_ERR_BUDGET
    srv = Server(dev, hbm_budget_bytes=2 << 30, paths_per_wave=1 << 16)
    try:
        big = S.Camera(cam_a.camera_to_world, 60.0, 15000, 15000)   # film + records ~17 GB > budget
        with pytest.raises(PtxError) as e:
            srv.submit(sa, big, 1, 1, 1)
        assert e.value.code == _abi.PTX_ERR_BUDGET
        ids = [srv.submit(sa, cam_a, 3, 3, 30 + i) for i in range(6)]
        films = [srv.wait(i)[0] for i in ids]
        ref = render_batch([(sa, cam_a, 3, 3, 30)])[0]
        assert np.array_equal(films[0].view(np.uint32), ref.view(np.uint32))
    finally:
        srv.close()


def test_server_admits_one_job_at_a_time_when_limited(dev):
    sa, cam_a, sb, _ = _scenes(dev)
    srv = Server(dev, hbm_budget_bytes=4 << 30, max_jobs_in_flight=1, paths_per_wave=1 << 16)
    try:
        ids = [srv.submit(sa, cam_a, 2, 2, 50 + i) for i in range(4)]
        sts = [srv.wait(i)[1] for i in ids]
        for a, b in zip(sts, sts[1:]):
            assert b["start_time"] >= a["end_time"]
    finally:
        srv.close()


def test_server_cancel(dev):
    sa, cam_a, sb, _ = _scenes(dev)
    srv = Server(dev, hbm_budget_bytes=4 << 30, paths_per_wave=4096)
    try:
        keep = srv.submit(sa, cam_a, 4, 3, 41)
        drop = srv.submit(sa, cam_a, 200, 5, 42)
        srv.cancel(drop)
        with pytest.raises(PtxError) as e:
            srv.wait(drop)
        assert e.value.code == _abi.PTX_ERR_CANCELLED
        film, st = srv.wait(keep)
        assert np.array_equal(film.view(np.uint32), render_batch([(sa, cam_a, 4, 3, 41)])[0].view(np.uint32))
        with pytest.raises(PtxError) as e:
            srv.status(12345678)
        assert e.value.code == _abi.PTX_ERR_NOT_FOUND
    finally:
        srv.close()


def test_autotune_keeps_results(dev):
    sa, cam_a, sb, meshes_b = _scenes(dev)
    before = render_batch([(sb, S.Camera(S.look_at((0.3, 0.5, -3.0), (0, 0, 0)), 50.0, 40, 30), 4, 4, 3)])[0]
    t = autotune(dev, sb)
    assert t["lds_stack"] in (8, 12, 16) and t["paths_per_wave"] >= (1 << 20)
    after = render_batch([(sb, S.Camera(S.look_at((0.3, 0.5, -3.0), (0, 0, 0)), 50.0, 40, 30), 4, 4, 3)])[0]
    assert np.array_equal(before.view(np.uint32), after.view(np.uint32))
    dev.set_tuning()


def test_server_job_keeps_its_scene(dev):
    # the caller closes its scene handle right after submitting: the jobs hold
    # their own reference, and their films equal the same jobs rendered alone
    _, cam_a, _, meshes_b = _scenes(dev)
    cam_b = S.Camera(S.look_at((0.3, 0.5, -3.0), (0.0, 0.0, 0.0)), 50.0, 48, 36)
    keep = Scene(dev, meshes_b, (100.0, 120.0, 150.0))
    alone = [render_batch([(keep, cam_b, 5 + i, 4, 70 + i)])[0] for i in range(3)]
    keep.close()
    srv = Server(dev, hbm_budget_bytes=4 << 30, paths_per_wave=20000)
    try:
        tmp = Scene(dev, meshes_b, (100.0, 120.0, 150.0))
        ids = [srv.submit(tmp, cam_b, 5 + i, 4, 70 + i) for i in range(3)]
        tmp.close()
        del tmp
        for jid, a in zip(ids, alone):
            film, st = srv.wait(jid)
            assert st["state"] == _abi.JOB_DONE
            assert np.array_equal(film.view(np.uint32), a.view(np.uint32))
    finally:
        srv.close()


def test_server_serves_every_job_when_the_pool_is_full(dev):
    # ten 64x48 jobs on a 20,000-slot pool: a round holds six of them, so the
    # rest must be first in line for the next round; by the time the first job
    # has its eight samples every job has been served
    sa, cam_a, _, _ = _scenes(dev)
    srv = Server(dev, hbm_budget_bytes=4 << 30, paths_per_wave=20000)
    try:
        ids = [srv.submit(sa, cam_a, 8, 3, 90 + i) for i in range(10)]
        first, _ = srv.wait(ids[0])
        for jid in ids[1:]:
            st = srv.status(jid)
            assert st["samples_done"] > 0, st
        films = [first] + [srv.wait(jid)[0] for jid in ids[1:]]
    finally:
        srv.close()
    for i in (0, 6, 9):
        (alone,) = render_batch([(sa, cam_a, 8, 3, 90 + i)])
        assert np.array_equal(alone.view(np.uint32), films[i].view(np.uint32))


def test_batch_recovers_after_an_allocation_failure(dev):
    # a batch whose path pool cannot be allocated fails with PTX_ERR_OOM; once
    # memory is back, the engine's tables are allocated again and the next batch
    # renders bit-identically
    sa, cam_a, _, _ = _scenes(dev)
    job = (sa, cam_a, 3, 3, 61)
    dev.set_tuning(paths_per_wave=1 << 16)
    (before,) = render_batch([job])
    free, _ = torch.cuda.mem_get_info()
    print(f"case allocation failure: filling {free / 2**30:.1f} GiB but about 1 GiB", flush=True)
    # the filler lives only inside `hog` (a list), so no frame or traceback
    # snapshot keeps the memory once the list is cleared
    hog = []
    for spare in (1 << 30, 3 << 29, 1 << 31, 5 << 29):       # leave less than the 3.2 GB pool
        try:
            hog.append(torch.empty(max(0, free - spare), dtype=torch.uint8, device="cuda"))
            break
        except RuntimeError:
            continue
    if not hog:
        pytest.skip("could not fill device memory")
    code = None
    try:
        dev.set_tuning(paths_per_wave=1 << 28)          # 3.2 GB of path results
        big = S.Camera(cam_a.camera_to_world, 60.0, 1024, 1024)
        try:
            render_batch([(sa, big, 256, 2, 1)])
        except PtxError as err:
            code = err.code
        # the library cleared the failure from the thread's HIP error state:
        # the caller's own next GPU call does not report it
        assert float(torch.ones(16, device="cuda").sum().item()) == 16.0
    finally:
        hog.clear()
        torch.cuda.synchronize()
        torch.cuda.empty_cache()
        dev.set_tuning(paths_per_wave=1 << 16)
    free_after, _ = torch.cuda.mem_get_info()
    assert free_after > free - (1 << 30), (free_after, free)
    assert code == _abi.PTX_ERR_OOM
    (after,) = render_batch([job])
    dev.set_tuning()
    assert np.array_equal(before.view(np.uint32), after.view(np.uint32))
