"""GPU parity of the closest-hit path (ptx_intersect_closest*, HIP kernel
k_closest_hit) against the reference's golden fixtures and the oracle.

Bar: hit/miss, (mesh, face) and the reference's per-ray counters bit-exact;
t, u, v, w within 2 ulp (this path reproduces them bit-exactly, and the tests
check bits); the shading frame (Location with its repeated offset, normal,
tangents, error bound) bit-exact where the reference is defined.
"""
from __future__ import annotations

import numpy as np
import pytest

import oracle as O
from polytope_amd import COUNTER_DTYPE, FRAME_DTYPE, HIT_DTYPE, Device, Scene
from polytope_amd import scenes as S
from tests import _fixtures as F

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


@pytest.fixture(scope="module")
def dev():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    d = Device(0)
    yield d
    d.close()


def gpu_query(scene: Scene, rays: np.ndarray):
    """Runs the device entry point on torch-allocated HBM buffers."""
    n = rays.shape[0]
    r = torch.from_numpy(np.ascontiguousarray(rays, np.float32)).cuda()
    hits = torch.zeros(max(n, 1) * 6, dtype=torch.int32, device="cuda")
    cnt = torch.zeros(max(n, 1) * 3, dtype=torch.int32, device="cuda")
    frm = torch.zeros(max(n, 1) * 15, dtype=torch.float32, device="cuda")
    stream = torch.cuda.current_stream().cuda_stream
    scene.intersect_device(r.data_ptr(), n, hits.data_ptr(), cnt.data_ptr(), frm.data_ptr(), stream)
    torch.cuda.synchronize()
    h = hits.cpu().numpy().view(HIT_DTYPE)[:n]
    c = cnt.cpu().numpy().view(COUNTER_DTYPE)[:n]
    f = frm.cpu().numpy().view(FRAME_DTYPE)[:n]
    return h, c, f


def bits(a):
    return np.ascontiguousarray(a).view(np.uint32)


def assert_same_hits(h, c, f, ref, frames=True, frame_mask=None):
    hit = ref["hit"] == 1
    assert np.array_equal(h["face"] != 0xFFFFFFFF, hit)
    assert np.array_equal(h["mesh"][hit], ref["mesh"][hit])
    assert np.array_equal(h["face"][hit], ref["face"][hit])
    for k in ("t", "u", "v", "w"):
        assert np.array_equal(bits(h[k][hit]), bits(ref[k][hit])), k
    assert np.array_equal(c["bb_hits"], ref["bb_hits"])
    assert np.array_equal(c["bb_misses"], ref["bb_misses"])
    assert np.array_equal(c["tri_isects"], ref["tri_isects"])
    if frames:
        m = hit if frame_mask is None else hit & frame_mask
        for k in ("location", "normal", "tangent1", "tangent2", "error"):
            assert np.array_equal(bits(f[k][m]), bits(ref[k][m])), k


@pytest.mark.parametrize("name", F.NAMES)
def test_gpu_matches_reference_golden(dev, name):
    d = F.load(name)
    print(f"case golden:{name}", flush=True)
    sc = Scene(dev, F.meshes_of(d), d.get("sky"))
    nodes, pairs = sc.bvh()
    assert nodes.tobytes() == d["ref_nodes"].tobytes()
    h, c, f = gpu_query(sc, d["rays"])
    ref = d["ref_hits"].view(O.HIT_RECORD)
    mask = None
    if not F.single_mesh(d):
        mask = np.isfinite(ref["location"]).all(1) & np.isfinite(ref["tangent2"]).all(1)
    assert_same_hits(h, c, f, ref, frame_mask=mask)


def _scene_cases():
    v, f = S.icosphere(6, noise=0.05, seed=2)
    m = [S.MeshData(v, f)]
    lo, hi = S.scene_bounds(m)
    yield "ico6", m, None, np.concatenate([S.rays_incoherent(lo, hi, 30000, 5), S.rays_at_features(v, f, 30000, 6),
                                           S.closest_hit_workload(6, 8192)[1]])
    v4, f4 = S.icosphere(4, noise=0.05, seed=3)
    meshes = [S.MeshData(v4, f4, object_to_world=S.rigid(100 + i, 3.0)) for i in range(8)]
    gv, gf = S.quad(8.0, -4.0)
    meshes.append(S.MeshData(gv, gf))
    lo, hi = S.scene_bounds(meshes)
    yield "instances", meshes, None, S.rays_incoherent(lo, hi, 40000, 7)
    sv, sf = S.sliver_grid(160, 6.0, aspect=1000.0, seed=5)
    m = [S.MeshData(sv, sf)]
    yield "sliver160", m, None, S.rays_at_features(sv, sf, 40000, seed=8, dist=4.0, center=(0.0, 0.0, 0.0))
    meshes, cam, sky = S.plain_render_scene(1, 64, 48)
    lo, hi = S.scene_bounds(meshes)
    yield "plain", meshes, sky, S.rays_incoherent(lo, hi, 40000, 9)


@pytest.mark.parametrize("case", list(_scene_cases()), ids=lambda c: c[0])
def test_gpu_matches_oracle(dev, case):
    name, meshes, sky, rays = case
    print(f"case oracle:{name} rays={len(rays)}", flush=True)
    ref = O.OracleScene(meshes, sky).intersect(rays)
    sc = Scene(dev, meshes, sky)
    h, c, f = gpu_query(sc, rays)
    # frames: the oracle defines them for every hit (hit-mesh semantics)
    assert_same_hits(h, c, f, ref)


@pytest.mark.parametrize("lds", [8, 12, 16, 24])
def test_gpu_lds_depths_agree(dev, lds):
    # lds = 8 forces deep trees to spill the deferred stack to global memory
    v, f = S.icosphere(6, noise=0.3, seed=4)
    meshes = [S.MeshData(v, f)]
    lo, hi = S.scene_bounds(meshes)
    rays = S.rays_incoherent(lo, hi, 20000, 3)
    print(f"case lds={lds}", flush=True)
    ref = O.OracleScene(meshes).intersect(rays)
    dev.set_tuning(lds_stack=lds)
    try:
        sc = Scene(dev, meshes)
        assert sc.info()["max_depth"] > 8
        h, c, fr = gpu_query(sc, rays)
    finally:
        dev.set_tuning()
    assert_same_hits(h, c, fr, ref)


def test_gpu_reference_cuda_unit_tree(dev):
    # tests/cuda/test_bvh_traversal.cpp:16-203: hand-built 3-node tree over the two
    # triangles of scenes/test/cuda/cuda-bvh-traversal1.ply (root, low child x<-1,
    # high child x>1), given to the library as a prebuilt reference BVH.
    v = np.array([[-3, -1, 1], [-3, 2, 1], [-1, -1, -1], [1, -1, -1], [3, 2, 1], [3, -1, 1]], np.float32)
    f = np.array([[0, 1, 2], [3, 4, 5]], np.uint32)
    nodes = np.zeros(3, O.NODE_DTYPE)
    nodes[0] = ((-3, -1, -1), (3, 2, 1), 2, 0, 0)
    nodes[1] = ((-3, -1, -1), (-1, 2, 1), 0, 1, 0)
    nodes[2] = ((1, -1, -1), (3, 2, 1), 1, 1, 0)
    pairs = np.array([[0, 0], [0, 1]], np.uint32)
    sc = Scene(dev, [S.MeshData(v, f)], bvh=(nodes, pairs))
    rays = np.array([[2, 0, 10, 0, 0, -1],                       # test2: over the high child
                     [2, 0, 0, -1, 0, 0],                        # test3: inside high, toward low
                     [-1.701634, -0.397628, -0.295537, 0.985882, -0.038071, 0.163058]], np.float32)
    h, c, _ = gpu_query(sc, rays)
    # test2: root and high child hit, low child missed; the high leaf's face is hit
    assert (c["bb_hits"][0], c["bb_misses"][0]) == (2, 1)
    assert h["face"][0] == 1
    # test3/test4: all three boxes hit
    assert c["bb_hits"][1] == 3 and c["bb_hits"][2] == 3
    o = O.OracleScene([S.MeshData(v, f)])      # oracle builds its own tree: compare ids only
    ref = o.intersect(rays)
    assert np.array_equal(h["face"], np.where(ref["hit"] == 1, ref["face"], 0xFFFFFFFF))


def test_gpu_edge_inputs(dev):
    v, f = S.icosphere(3, noise=0.05, seed=2)
    meshes = [S.MeshData(v, f)]
    sc = Scene(dev, meshes)
    # n = 0 is a no-op
    h, c, fr = gpu_query(sc, np.zeros((0, 6), np.float32))
    assert h.shape == (0,)
    # axis-aligned directions (infinite inverse components, -0.0), origins inside the
    # mesh and on its box faces, one ray
    rays = np.array([[0, 0, 0, 1, 0, 0], [0, 0, 0, -0.0, 0, 1], [0, 0, 0, 0, -1, -0.0],
                     [5, 0, 0, -1, 0, 0], [0, 5, 0, 0, 1, 0], [1.2, 1.2, 1.2, -1, -1, -1],
                     [0.3, 0.1, 0.2, 0.6, 0.8, 0.0]], np.float32)
    rays[5, 3:] /= np.float32(np.sqrt(3))
    ref = O.OracleScene(meshes).intersect(rays)
    h, c, fr = gpu_query(sc, rays)
    assert_same_hits(h, c, fr, ref)
    one = gpu_query(sc, rays[:1])
    assert one[0]["face"][0] == h["face"][0]


def test_gpu_host_and_device_entry_points_agree(dev):
    meshes, rays, _ = S.closest_hit_workload(5, 20000, seed=2)
    sc = Scene(dev, meshes)
    h_dev, c_dev, _ = gpu_query(sc, rays)
    h_host, c_host = sc.intersect(rays, counters=True)
    assert h_host.tobytes() == h_dev.tobytes()
    assert c_host.tobytes() == c_dev.tobytes()


def test_gpu_host_entry_point_pipelined_copies(dev):
    # The host-buffer entry point splits the arrays over up to eight host threads,
    # each copying 2 MiB chunks through its own pinned slots: sizes with one
    # chunk, exactly a chunk of rays (2 MiB / 24 B = 87,381.33), ragged shares of
    # several chunks per thread, with and without counters, a smaller call after
    # a larger one (staging reused), and caller-supplied output arrays.
    meshes, rays, _ = S.closest_hit_workload(5, (1 << 21) + 4099, seed=3)
    sc = Scene(dev, meshes)
    h_all, c_all, _ = gpu_query(sc, rays)
    for n, counters in ((len(rays), True), (87_381, False), (87_382, True), (700_001, False),
                        (1, True), (len(rays) - 1, False), (174_763, True)):
        print(f"case host pipeline n={n} counters={counters}", flush=True)
        got = sc.intersect(rays[:n], counters=counters)
        h = got[0] if counters else got
        assert h.tobytes() == h_all[:n].tobytes(), n
        if counters:
            assert got[1].tobytes() == c_all[:n].tobytes(), n
    # into the caller's arrays, twice (second call: pages already mapped)
    n = 1_000_003
    with pytest.raises(ValueError):
        sc.intersect(rays[:n], out=(np.zeros(n * 24, np.uint8), None))     # wrong dtype
    with pytest.raises(ValueError):
        sc.intersect(rays[:n], counters=True, out=(np.zeros(n, HIT_DTYPE), None))
    hits = np.zeros(n, HIT_DTYPE)
    cnt = np.zeros(n, COUNTER_DTYPE)
    for _ in range(2):
        h, c = sc.intersect(rays[:n], counters=True, out=(hits, cnt))
        assert h is hits and c is cnt
        assert hits.tobytes() == h_all[:n].tobytes()
        assert cnt.tobytes() == c_all[:n].tobytes()


def test_gpu_full_size_closest_hit_sample(dev):
    # 327,680 triangles, 2^22 rays (half coherent, half incoherent): the whole batch
    # runs on the GPU; every 61st ray is checked against the oracle, and every ray
    # with the size-independent properties (hit <=> t finite, barycentrics sum to ~1,
    # counters consistent: bb_hits + bb_misses is odd = 1 + 2 * interior visits).
    meshes, rays, n_coh = S.closest_hit_workload(7, 1 << 22, seed=2)
    print(f"case full-size rays={len(rays)}", flush=True)
    sc = Scene(dev, meshes)
    h, c, _ = gpu_query(sc, rays)
    hit = h["face"] != 0xFFFFFFFF
    assert np.isfinite(h["t"][hit]).all() and (h["t"][hit] > 0).all()
    s = h["u"][hit].astype(np.float64) + h["v"][hit] + h["w"][hit]
    assert np.abs(s - 1).max() < 1e-5
    assert ((c["bb_hits"] + c["bb_misses"]) % 2 == 1).all()
    assert hit[:n_coh].mean() > 0.35       # the coherent grid is framed on the sphere
    idx = np.arange(0, len(rays), 61)
    ref = O.OracleScene(meshes).intersect(rays[idx])
    assert_same_hits(h[idx], c[idx], None, ref, frames=False)


def _query_entry(entry, scene, rays):
    """hits, counters and frames from one C-ABI closest-hit entry point."""
    import ctypes as C  # noqa: F401  (the library's argtypes come from _abi)
    from polytope_amd import PtxError, _abi
    n = rays.shape[0]
    lib = _abi.load()
    r = torch.from_numpy(np.ascontiguousarray(rays, np.float32)).cuda()
    hits = torch.full((max(n, 1) * 6,), -7, dtype=torch.int32, device="cuda")
    cnt = torch.full((max(n, 1) * 3,), -7, dtype=torch.int32, device="cuda")
    frm = torch.full((max(n, 1) * 15,), -7.0, dtype=torch.float32, device="cuda")
    rc = getattr(lib, entry)(scene.handle, r.data_ptr(), n, hits.data_ptr(), cnt.data_ptr(), frm.data_ptr(),
                             torch.cuda.current_stream().cuda_stream)
    if rc != _abi.PTX_OK:
        raise PtxError(rc, lib.ptx_last_error().decode())
    torch.cuda.synchronize()
    return (hits.cpu().numpy().view(HIT_DTYPE)[:n].copy(), cnt.cpu().numpy().view(COUNTER_DTYPE)[:n].copy(),
            frm.cpu().numpy().view(FRAME_DTYPE)[:n].copy())


def _window_cases():
    yield "workload_subdiv6", S.closest_hit_workload(6, 1 << 17, seed=7)[:2]
    v, f = S.icosphere(6, noise=0.3, seed=4)          # deep tree: the stack spills at LDS depth 8
    m = [S.MeshData(v, f)]
    lo, hi = S.scene_bounds(m)
    yield "deep_noise", (m, S.rays_incoherent(lo, hi, 30000, 3))
    for name, meshes, _sky, rays in _scene_cases():
        if name in ("instances", "sliver160"):
            yield name, (meshes, rays)


@pytest.mark.parametrize("lds", [8, 12])
@pytest.mark.parametrize("case", list(_window_cases()), ids=lambda c: c[0])
def test_gpu_window_and_node_record_forms_agree(dev, case, lds):
    # The window form (ray_windows=1: once the ray queue is drained, steps over
    # a node's record and its first-visited child's in one 128-byte block, up
    # to three faces per leaf step) against the one-record form and the
    # oracle: hits, counters and frames bit for bit.  The first query on a fresh scene through the no-allocation entry
    # point runs the one-record form (it never builds the windows); the device
    # entry point builds them, and the no-allocation entry point then uses them.
    name, (meshes, rays) = case
    print(f"case windows:{name} lds={lds} rays={len(rays)}", flush=True)
    ref = O.OracleScene(meshes).intersect(rays)
    dev.set_tuning(lds_stack=lds, ray_windows=1)
    try:
        sc = Scene(dev, meshes)
        node_form = _query_entry("ptx_intersect_closest_device_noalloc", sc, rays)
        win_form = _query_entry("ptx_intersect_closest_device", sc, rays)
        win_noalloc = _query_entry("ptx_intersect_closest_device_noalloc", sc, rays)
        dev.set_tuning(lds_stack=lds, ray_windows=2)
        off_form = _query_entry("ptx_intersect_closest_device", sc, rays)
        sc.close()
    finally:
        dev.set_tuning()
    for a, b in ((node_form, win_form), (off_form, win_form), (win_noalloc, win_form)):
        for x, y in zip(a, b):
            assert x.tobytes() == y.tobytes()
    assert_same_hits(*win_form, ref)


def test_gpu_results_independent_of_grid_and_thresholds(dev):
    # Work ids are dealt to waves by the grid's wave count (three quarters of
    # the 64-ray batches statically, the rest from one counter), so the grid
    # size, the refill and leaf thresholds and the LDS depth change which
    # wave traces which ray and when; hits and the reference's counters must
    # not change.  Ray counts around the batch and grid sizes included.
    meshes, rays, _ = S.closest_hit_workload(6, 1 << 16, seed=11)
    sc = Scene(dev, meshes)
    base = gpu_query(sc, rays)
    settings = [dict(blocks_per_cu=1), dict(blocks_per_cu=2), dict(refill_lanes=16), dict(leaf_lanes=32),
                dict(lds_stack=8, blocks_per_cu=3)]
    try:
        for kw in settings:
            dev.set_tuning(**kw)
            print(f"case grid/thresholds {kw}", flush=True)
            h, c, f = gpu_query(sc, rays)
            assert h.tobytes() == base[0].tobytes(), kw
            assert c.tobytes() == base[1].tobytes(), kw
            assert f.tobytes() == base[2].tobytes(), kw
            for n in (1, 63, 64, 65, 4097, 40000):
                hs, cs, _ = gpu_query(sc, rays[:n])
                assert hs.tobytes() == base[0][:n].tobytes(), (kw, n)
                assert cs.tobytes() == base[1][:n].tobytes(), (kw, n)
    finally:
        dev.set_tuning()
        sc.close()
