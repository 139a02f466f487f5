"""GPU parity of the render paths (ptx_render: plain; ptx_render_batch: path
regeneration and wavefront rounds) against the reference's golden film and the
oracle.

Bar: per-pixel relative error <= 1e-4 against the oracle under the same RNG
stream (the paths share the oracle's operation order, so the films are checked
bit for bit as well); the deterministic mirror scene matches the reference's
own film bit for bit; every job renders the same alone and batched with others.
"""
from __future__ import annotations

import numpy as np
import pytest

import oracle as O
from polytope_amd import Device, Scene, render, render_batch
from polytope_amd import _abi
from polytope_amd import scenes as S
from tests import _fixtures as F

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")

REL_TOL = 1e-4


@pytest.fixture(scope="module")
def dev():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    d = Device(0)
    yield d
    d.close()


def rel_err(a, b):
    return float((np.abs(a - b) / np.maximum(np.abs(b), 1e-3)).max())


MODES = [_abi.RENDER_REGEN, _abi.RENDER_WAVEFRONT]


def batch(dev, jobs, mode, paths_per_wave=0):
    dev.set_tuning(paths_per_wave=paths_per_wave, render_mode=mode)
    try:
        return render_batch(jobs)
    finally:
        dev.set_tuning()


def test_render_mirror_matches_reference_golden(dev):
    d = F.load("mirror_render")
    cam, spp, depth = F.camera_of(d)
    print("case mirror golden", flush=True)
    sc = Scene(dev, F.meshes_of(d), d["sky"])
    plain = render(sc, cam, spp, depth, seed=1)
    assert np.array_equal(plain.view(np.uint32), d["ref_film"].view(np.uint32))
    for mode in MODES:
        (batched,) = batch(dev, [(sc, cam, spp, depth, 1)], mode)
        assert np.array_equal(batched.view(np.uint32), d["ref_film"].view(np.uint32)), mode


def _lambert_scene(w=80, h=60):
    meshes, cam, sky = S.plain_render_scene(1, w, h)
    return meshes, cam, sky


@pytest.mark.parametrize("spp,depth", [(4, 5), (3, 1), (2, 0), (2, 8)])
def test_render_plain_and_batched_match_oracle(dev, spp, depth):
    meshes, cam, sky = _lambert_scene()
    print(f"case lambert spp={spp} depth={depth}", flush=True)
    ref = O.OracleScene(meshes, sky).render(cam, spp, depth, seed=5)
    sc = Scene(dev, meshes, sky)
    plain = render(sc, cam, spp, depth, seed=5)
    assert rel_err(plain, ref) <= REL_TOL
    assert np.array_equal(plain.view(np.uint32), ref.view(np.uint32))
    for mode in MODES:
        (batched,) = batch(dev, [(sc, cam, spp, depth, 5)], mode)
        assert rel_err(batched, ref) <= REL_TOL, mode
        assert np.array_equal(batched.view(np.uint32), ref.view(np.uint32)), mode


def test_render_mixed_materials_emitter_and_debug_pixel(dev):
    # mirror + Lambert + emitter, a film wider than 128 so pixel (128, 128) exists
    v, f = S.icosphere(3, radius=1.0, center=(0.0, 1.0, 0.0), noise=0.05, seed=3)
    gv, gf = S.quad(6.0, 0.0)
    lv, lf = S.quad(1.5, 4.0)
    meshes = [S.MeshData(v, f, brdf=S.MIRROR, refl=(0.9, 0.8, 0.7)),
              S.MeshData(gv, gf, refl=(0.5, 0.6, 0.7)),
              S.MeshData(lv, lf, is_light=True, emission=(255.0, 240.0, 200.0))]
    cam = S.Camera(S.look_at((0.0, 2.0, -5.0), (0.0, 0.8, 0.0)), 60.0, 160, 136)
    sky = (40.0, 60.0, 90.0)
    ref = O.OracleScene(meshes, sky).render(cam, 4, 6, seed=9)
    sc = Scene(dev, meshes, sky)
    plain = render(sc, cam, 4, 6, seed=9)
    assert np.array_equal(plain.view(np.uint32), ref.view(np.uint32))
    for mode in MODES:
        (batched,) = batch(dev, [(sc, cam, 4, 6, 9)], mode)
        assert np.array_equal(batched.view(np.uint32), ref.view(np.uint32)), mode


@pytest.mark.parametrize("mode", MODES)
def test_render_batch_alone_equals_mixed(dev, mode):
    # jobs of uneven resolution, spp and depth over two scenes: each film is the same
    # rendered alone, in the mix, and with a path capacity small enough to force many rounds
    meshes_a, cam_a, sky_a = _lambert_scene(64, 48)
    v, f = S.icosphere(4, noise=0.05, seed=8)
    meshes_b = [S.MeshData(v, f, refl=(0.7, 0.5, 0.3)),
                S.MeshData(*S.quad(4.0, -1.1), brdf=S.MIRROR, refl=(0.8, 0.8, 0.8))]
    sa = Scene(dev, meshes_a, sky_a)
    sb = Scene(dev, meshes_b, (100.0, 120.0, 150.0))
    cam_b = S.Camera(S.look_at((0.3, 0.5, -3.0), (0.0, 0.0, 0.0)), 50.0, 40, 30)
    cam_c = S.Camera(S.look_at((0.3, 0.5, -3.0), (0.0, 0.0, 0.0)), 70.0, 96, 40)
    jobs = [(sa, cam_a, 5, 4, 11), (sb, cam_b, 9, 7, 12), (sb, cam_c, 2, 3, 13), (sa, cam_a, 1, 2, 14)]
    print(f"case alone/mixed mode={mode}", flush=True)
    alone = [batch(dev, [j], mode)[0] for j in jobs]
    mixed = batch(dev, jobs, mode)
    small = batch(dev, jobs, mode, paths_per_wave=5000)
    for a, m, s, j in zip(alone, mixed, small, jobs):
        assert np.array_equal(a.view(np.uint32), m.view(np.uint32))
        assert np.array_equal(a.view(np.uint32), s.view(np.uint32))
        plain = render(*j)
        assert np.array_equal(a.view(np.uint32), plain.view(np.uint32))
    ref = O.OracleScene(meshes_b, (100.0, 120.0, 150.0)).render(cam_b, 9, 7, seed=12)
    assert np.array_equal(mixed[1].view(np.uint32), ref.view(np.uint32))


@pytest.mark.parametrize("mode", MODES)
def test_render_batch_more_jobs_than_lds_table(dev, mode):
    # 70 jobs: more than the regeneration kernel's 64-chunk LDS table, so its rounds
    # take at most 64 jobs each (the wavefront form takes all 70 in one round);
    # films are unchanged
    meshes, cam, sky = _lambert_scene(16, 12)
    sc = Scene(dev, meshes, sky)
    jobs = [(sc, cam, 1 + i % 3, 1 + i % 4, 100 + i) for i in range(70)]
    print(f"case 70 jobs mode={mode}", flush=True)
    mixed = batch(dev, jobs, mode)
    for i in (0, 1, 33, 64, 69):
        (alone,) = render_batch([jobs[i]])
        assert np.array_equal(alone.view(np.uint32), mixed[i].view(np.uint32))


def _nee_scene():
    v, f = S.icosphere(3, radius=1.0, center=(0.0, 1.0, 0.0), noise=0.05, seed=3)
    gv, gf = S.quad(6.0, 0.0)
    lv, lf = S.quad(1.0, 4.5)
    meshes = [S.MeshData(v, f, brdf=S.MIRROR, refl=(0.9, 0.8, 0.7)),
              S.MeshData(gv, gf, refl=(0.5, 0.6, 0.7)),
              S.MeshData(*S.icosphere(2, radius=0.6, center=(1.8, 0.6, -0.5), noise=0.0), refl=(0.7, 0.3, 0.3)),
              S.MeshData(lv, lf, is_light=True, emission=(255.0, 240.0, 200.0))]
    lights = [((0.0, 3.0, -2.5), (3000.0, 2800.0, 2500.0)), ((-2.5, 1.5, 0.5), (600.0, 900.0, 1200.0)),
              ((0.0, -1.0, 0.0), (5000.0, 5000.0, 5000.0))]      # the last one is under the floor
    cam = S.Camera(S.look_at((0.0, 2.0, -5.0), (0.0, 0.8, 0.0)), 60.0, 144, 96)
    return meshes, lights, cam, (40.0, 60.0, 90.0)


@pytest.mark.parametrize("depth", [0, 1, 5])
def test_next_event_estimation_matches_oracle(dev, depth):
    meshes, lights, cam, sky = _nee_scene()
    print(f"case nee depth={depth}", flush=True)
    ref = O.OracleScene(meshes, sky, lights=lights).render(cam, 3, depth, seed=21, nee=True)
    sc = Scene(dev, meshes, sky, lights=lights)
    plain = render(sc, cam, 3, depth, 21, nee=True)
    assert np.array_equal(plain.view(np.uint32), ref.view(np.uint32))
    (batched,) = batch(dev, [(sc, cam, 3, depth, 21, True)], _abi.RENDER_REGEN)
    assert np.array_equal(batched.view(np.uint32), ref.view(np.uint32))
    (wave,) = batch(dev, [(sc, cam, 3, depth, 21, True)], _abi.RENDER_WAVEFRONT)   # shadow-connect stage
    assert np.array_equal(wave.view(np.uint32), ref.view(np.uint32))
    (wave_rounds,) = batch(dev, [(sc, cam, 3, depth, 21, True)], _abi.RENDER_WAVEFRONT, paths_per_wave=4096)
    assert np.array_equal(wave_rounds.view(np.uint32), ref.view(np.uint32))          # one sample per round
    if depth:
        off = O.OracleScene(meshes, sky, lights=lights).render(cam, 3, depth, seed=21)
        assert (ref >= off).all() and (ref > off).any()


@pytest.mark.parametrize("nee", [False, True])
def test_plain_path_retracing_primaries_equals_cached(dev, nee):
    # ptx_render traces each pixel's primary ray once by default; tracing it
    # again for every sample (the reference CUDA path's shape) gives the same film
    meshes, lights, cam, sky = _nee_scene()
    sc = Scene(dev, meshes, sky, lights=lights)
    print(f"case plain retrace nee={nee}", flush=True)
    ref = O.OracleScene(meshes, sky, lights=lights).render(cam, 3, 5, seed=31, nee=nee)
    cached = render(sc, cam, 3, 5, 31, nee=nee)
    dev.set_tuning(plain_primary=_abi.PLAIN_RETRACE)
    try:
        retraced = render(sc, cam, 3, 5, 31, nee=nee)
    finally:
        dev.set_tuning()
    assert np.array_equal(cached.view(np.uint32), ref.view(np.uint32))
    assert np.array_equal(retraced.view(np.uint32), ref.view(np.uint32))


def test_next_event_jobs_mixed_with_plain_jobs(dev):
    # one round holding a next-event job and an ordinary job of the same scene:
    # the ordinary job's film is the one it has without next-event jobs around
    meshes, lights, cam, sky = _nee_scene()
    sc = Scene(dev, meshes, sky, lights=lights)
    small = S.Camera(cam.camera_to_world, 60.0, 64, 40)
    mixed = render_batch([(sc, cam, 2, 4, 5, True), (sc, small, 3, 4, 6)])
    alone = render_batch([(sc, small, 3, 4, 6)])[0]
    assert np.array_equal(mixed[1].view(np.uint32), alone.view(np.uint32))
    ref = O.OracleScene(meshes, sky, lights=lights).render(cam, 2, 4, seed=5, nee=True)
    assert np.array_equal(mixed[0].view(np.uint32), ref.view(np.uint32))
    # the wavefront form: the shadow-connect stage serves the next-event job only
    wave = batch(dev, [(sc, cam, 2, 4, 5, True), (sc, small, 3, 4, 6)], _abi.RENDER_WAVEFRONT)
    assert np.array_equal(wave[0].view(np.uint32), ref.view(np.uint32))
    assert np.array_equal(wave[1].view(np.uint32), alone.view(np.uint32))


def test_scene_and_server_outlive_their_device_handle():
    """Closing a device handle while scenes and a server made on it live: the
    scene refuses new work with an error and is destroyed cleanly afterwards,
    the server still finishes its job, and no HIP error is left for the
    caller's own GPU work."""
    from polytope_amd import PtxError, Server
    print("case device closed before its scene and server", flush=True)
    d = Device(0)
    meshes, cam, sky = S.plain_render_scene(1, 24, 18)
    sc = Scene(d, meshes, sky)
    srv = Server(d, hbm_budget_bytes=1 << 30, paths_per_wave=1 << 14)
    jid = srv.submit(sc, cam, 2, 3, 9)
    d.close()
    with pytest.raises(PtxError, match="closed"):
        render(sc, cam, 1, 1, 1)
    film, st = srv.wait(jid)
    assert np.array_equal(film.view(np.uint32), O.OracleScene(meshes, sky).render(cam, 2, 3, 9).view(np.uint32))
    srv.close()
    sc.close()
    x = torch.ones(4, device="cuda") * 2
    torch.cuda.synchronize()
    assert float(x.sum()) == 8.0


def test_render_batch_more_scenes_than_lds_table(dev):
    # 70 scenes: more than the regeneration kernel's 64-entry scene table, so the
    # batch runs as consecutive groups of at most 64 scenes in the default form;
    # every film equals the job rendered alone and the wavefront form's
    meshes, cam, sky = _lambert_scene(16, 12)
    scenes = [Scene(dev, meshes, (sky[0] + i, sky[1], sky[2])) for i in range(70)]
    jobs = [(scenes[i], cam, 1 + i % 2, 1 + i % 3, 300 + i) for i in range(70)]
    print("case 70 scenes", flush=True)
    grouped = render_batch(jobs)
    queued = batch(dev, jobs, _abi.RENDER_WAVEFRONT)
    for i in range(70):
        assert np.array_equal(grouped[i].view(np.uint32), queued[i].view(np.uint32)), i
    for i in (0, 63, 64, 69):
        (alone,) = render_batch([jobs[i]])
        assert np.array_equal(alone.view(np.uint32), grouped[i].view(np.uint32)), i
    for s in scenes:
        s.close()


def test_job_buffers_reused_across_calls(dev):
    # A finished job's film and primary records stay with the engine as spares
    # and a later job of at most that size takes them (the film is zeroed, the
    # primary records rewritten): the same job gives the same film first, after
    # a larger and a smaller job left their buffers, and in the plain path.
    meshes, cam_s, sky = _lambert_scene(40, 30)
    sc = Scene(dev, meshes, sky)
    cam_l = S.Camera(cam_s.camera_to_world, 60.0, 96, 72)
    cam_m = S.Camera(cam_s.camera_to_world, 60.0, 80, 60)     # fits the larger job's buffers (at most 2x)
    print("case job buffers reused", flush=True)
    (first,) = render_batch([(sc, cam_s, 3, 4, 17)])
    render_batch([(sc, cam_l, 2, 3, 18), (sc, cam_s, 1, 2, 19)])      # larger buffers, then a spare of this size
    (again,) = render_batch([(sc, cam_s, 3, 4, 17)])
    (mid,) = render_batch([(sc, cam_m, 2, 5, 20)])                     # takes the larger job's spares
    plain = render(sc, cam_s, 3, 4, 17)
    (last,) = render_batch([(sc, cam_m, 2, 5, 20), (sc, cam_s, 3, 4, 17)])[1:]
    for f in (again, plain, last):
        assert np.array_equal(f.view(np.uint32), first.view(np.uint32))
    ref = O.OracleScene(meshes, sky).render(cam_m, 2, 5, 20)
    assert np.array_equal(mid.view(np.uint32), ref.view(np.uint32))
    sc.close()
