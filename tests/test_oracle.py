"""Pins the oracle (oracle/oracle.cpp, the CPU restatement of the reference hot
path) before it is trusted as the checker of the GPU path.

* Golden fixtures produced by the compiled reference (tests/golden): BVH bytes,
  leaf pairs and every per-ray hit field, bit for bit; the mirror-scene film.
* The reference's own unit-test cases: BoundingBox Hits1 (tests/TestBoundingBox.cpp:10-30),
  the four CUDA slab cases with CPU semantics (tests/cuda/test_bvh_traversal.cpp:16-203),
  one-triangle hits in both windings (tests/shapes/test_mesh.cpp:35-75) and
  intersect_stay_above_surface (test_mesh.cpp:77-117).
* Mathematics: Philox4x32-10 known-answer vectors (Salmon et al. 2011), sin/cos
  against libm, t/u/v/w against exact rational arithmetic on integer-coordinate
  triangles, watertightness on closed meshes and on a sliver grid, closest
  hits against a float64 brute-force scan of every face.
"""
from __future__ import annotations

import math
from fractions import Fraction

import numpy as np
import pytest

import oracle as O
from polytope_amd import scenes as S
from tests import _fixtures as F

BIT_FIELDS = ["hit", "mesh", "face", "t", "u", "v", "w", "bb_hits", "bb_misses", "tri_isects"]
FRAME_FIELDS = ["location", "normal", "tangent1", "tangent2", "error"]


def _bits(a):
    return np.ascontiguousarray(a).view(np.uint32)


@pytest.mark.parametrize("name", F.NAMES)
def test_oracle_matches_reference_golden(name):
    d = F.load(name)
    meshes = F.meshes_of(d)
    sc = O.OracleScene(meshes, d.get("sky"))
    nodes, pairs = sc.bvh()
    assert nodes.tobytes() == d["ref_nodes"].tobytes(), "compact BVH differs from bvh::bound + compact"
    assert np.array_equal(pairs, d["ref_pairs"])
    ref = d["ref_hits"].view(O.HIT_RECORD)
    got = sc.intersect(d["rays"])
    for f in BIT_FIELDS:
        assert np.array_equal(_bits(got[f]), _bits(ref[f])), f
    if F.single_mesh(d):
        for f in FRAME_FIELDS:
            assert np.array_equal(_bits(got[f]), _bits(ref[f])), f
    else:
        # Multi-mesh: where the faces tested after the hit belong to another mesh the
        # reference builds the frame from that mesh (mesh.cpp:369-378) and may read out
        # of range; compare the rays where its frame is finite.
        ok = np.isfinite(ref["location"]).all(1) & np.isfinite(ref["tangent2"]).all(1)
        assert ok.mean() > 0.98
        for f in FRAME_FIELDS:
            assert np.array_equal(_bits(got[f][ok]), _bits(ref[f][ok])), f


def test_oracle_mirror_render_matches_reference_golden():
    d = F.load("mirror_render")
    meshes = F.meshes_of(d)
    cam, spp, depth = F.camera_of(d)
    sc = O.OracleScene(meshes, d["sky"])
    film = sc.render(cam, spp, depth, seed=1)
    ref = d["ref_film"]
    assert film.shape == ref.shape
    assert np.array_equal(film.view(np.uint32), ref.view(np.uint32))
    assert (ref != ref[0, 0]).any(2).mean() > 0.3      # most pixels see the mirror geometry


def test_oracle_mirror_multimesh_render_close_to_reference_golden():
    # Three meshes + emitter: pixels whose paths end a traversal on a face of another
    # mesh than the hit one differ (the reference builds the hit frame from that mesh).
    d = F.load("mirror_multimesh")
    cam, spp, depth = F.camera_of(d)
    film = O.OracleScene(F.meshes_of(d), d["sky"]).render(cam, spp, depth, seed=1)
    same = (film == d["ref_film"]).all(2)
    assert same.mean() > 0.97


@pytest.mark.skipif(not O.reference_available(), reason="reference build (oracle/_ref) not present")
@pytest.mark.parametrize("seed", [21, 22, 23])
def test_oracle_matches_live_reference(seed, tmp_path):
    v, f = S.icosphere(3, noise=0.08, seed=seed)
    gv, gf = S.quad(3.0, -1.2)
    meshes = [S.MeshData(v, f)]
    lo, hi = S.scene_bounds(meshes)
    rays = np.concatenate([S.rays_incoherent(lo, hi, 1024, seed), S.rays_at_features(v, f, 1024, seed + 1)])
    S.write_ptxm(str(tmp_path / "s.bin"), meshes)
    S.write_rays(str(tmp_path / "r.bin"), rays)
    O.run_reference(["--ptxm", "s.bin", "--nodes", "n.bin", "--pairs", "p.bin", "--rays", "r.bin", "--hits", "h.bin"],
                    cwd=str(tmp_path))
    sc = O.OracleScene(meshes)
    nodes, pairs = sc.bvh()
    assert nodes.tobytes() == (tmp_path / "n.bin").read_bytes()
    got = sc.intersect(rays)
    assert got.tobytes() == (tmp_path / "h.bin").read_bytes()


# ----------------------------------------------------------------------------- RNG and sin/cos

def test_philox_known_answers():
    # Random123 known-answer vectors for philox4x32, 10 rounds.
    assert list(O.philox([0, 0, 0, 0], [0, 0])) == [0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8]
    assert list(O.philox([0xFFFFFFFF] * 4, [0xFFFFFFFF] * 2)) == [0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD]
    assert list(O.philox([0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344], [0xA4093822, 0x299F31D0])) == \
        [0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1]


def test_sincos_accuracy():
    two_pi = np.float32(2.0) * np.float32(math.pi)
    worst = 0.0
    for u in np.linspace(0.0, 1.0, 20001, dtype=np.float32):
        th = float(two_pi * u)
        s, c = O.sincos(th)
        es, ec = math.sin(th), math.cos(th)
        worst = max(worst, abs(s - es), abs(c - ec))
        for got, exp in ((s, es), (c, ec)):
            if abs(exp) > 1e-2:
                ulp = np.spacing(np.float32(abs(exp)))
                assert abs(got - exp) <= 2 * ulp, (th, got, exp)
    assert worst < 2.5e-7


# ----------------------------------------------------------------------------- reference unit-test cases

def test_box_hits_reference_cases():
    inf = float("inf")
    # TestBoundingBox.cpp:10-30
    assert O.box_hits([0, 0, 0, 1, 1, 1], [0.5, 0.5, 10, 0, 0, -1], inf)
    root, low, high = [-3, -1, -1, 3, 2, 1], [-3, -1, -1, -1, 2, 1], [1, -1, -1, 3, 2, 1]
    # test_bvh_traversal.cpp test1/test2: ray over the high child, straight down
    r = [2, 0, 10, 0, 0, -1]
    assert O.box_hits(root, r, inf) and O.box_hits(high, r, inf) and not O.box_hits(low, r, inf)
    # test3: origin inside the high child, pointed at the low child
    r = [2, 0, 0, -1, 0, 0]
    assert O.box_hits(root, r, inf) and O.box_hits(low, r, inf) and O.box_hits(high, r, inf)
    # test4: origin inside the low child, pointed at the high child
    r = [-1.701634, -0.397628, -0.295537, 0.985882, -0.038071, 0.163058]
    assert O.box_hits(root, r, inf) and O.box_hits(low, r, inf) and O.box_hits(high, r, inf)
    # MinT prunes: the high child starts at t = 1 along +x from x = 0
    assert not O.box_hits(high, [0, 0, 0, 1, 0, 0], 0.5)
    assert O.box_hits(high, [0, 0, 0, 1, 0, 0], 1.5)


def _one_triangle(order):
    v = np.array([[0, 0, 0], [0, 1, 0], [1, 0, 0]], np.float32)
    return [S.MeshData(v, np.array([order], np.uint32))]


@pytest.mark.parametrize("order,oz,dz,nz", [((0, 1, 2), -10, 1, -1), ((0, 2, 1), -10, 1, -1),
                                            ((0, 1, 2), 10, -1, 1), ((0, 2, 1), 10, -1, 1)])
def test_one_triangle_both_windings(order, oz, dz, nz):
    # test_mesh.cpp:35-75: hits from both sides, normal flipped toward the ray.
    sc = O.OracleScene(_one_triangle(order))
    h = sc.intersect(np.array([[0.2, 0.2, oz, 0, 0, dz]], np.float32))[0]
    assert h["hit"] == 1 and h["face"] == 0
    assert h["normal"][0] == 0 and h["normal"][1] == 0 and h["normal"][2] == nz
    for got, exp in zip(h["location"], (0.2, 0.2, 0.0)):
        assert abs(got - exp) <= 4 * np.spacing(np.float32(max(abs(exp), 1e-30))) + 1e-6


def test_stay_above_surface():
    # test_mesh.cpp:77-117: rays from ever farther origins toward (0.2, 0.2, z) hit the
    # face with the normal flipped toward them.  Directions are normalised as
    # Vector::Normalize does (Vectors.cpp:94-99).  Distances stop at 3e5: farther, the
    # conservative t bound (mesh.cpp:312-322) grows past t itself (its deltaX term scales
    # with the sheared z distance) and the reference rejects the hit, so its own test
    # loop to 1e7 cannot pass there.
    v = np.array([[0, 0, 0.2], [0, 1, 0.2], [1, 0, 0.2]], np.float32)
    sc = O.OracleScene([S.MeshData(v, np.array([[0, 1, 2]], np.uint32))])
    f32 = np.float32
    ks = np.unique(np.geomspace(100, 300000, 400).astype(np.int64))
    o = np.stack([np.full(ks.size, 1.464552, f32), np.full(ks.size, 3.635245, f32),
                  (f32(-1.0364) * ks.astype(f32)).astype(f32)], 1)
    d = (np.array([0.2, 0.2, 0.2], f32) - o).astype(f32)
    length = np.sqrt(d[:, 0] * d[:, 0] + d[:, 1] * d[:, 1] + d[:, 2] * d[:, 2]).astype(f32)
    d = (d * (f32(1.0) / length)[:, None]).astype(f32)
    hits = sc.intersect(np.concatenate([o, d], 1))
    assert (hits["hit"] == 1).all()
    assert (hits["normal"][:, 2] == -1).all()
    assert np.all(hits["location"][:, 2] <= f32(0.2))   # the origin offset moves toward the ray's side


# ----------------------------------------------------------------------------- exact arithmetic

def test_t_and_barycentrics_against_exact_rationals():
    rng = np.random.default_rng(5)
    checked = 0
    for _ in range(300):
        tri = rng.integers(-8, 9, size=(3, 3)).astype(np.float32)
        if np.linalg.norm(np.cross(tri[1] - tri[0], tri[2] - tri[0])) < 1:
            continue
        bary = rng.integers(0, 5, size=3).astype(np.float64)
        if bary.sum() == 0:
            continue
        target = (bary[:, None] * tri).sum(0) / bary.sum()
        o = (target + rng.integers(-20, 21, size=3)).astype(np.float32)
        d = (target - o.astype(np.float64))
        if np.abs(d).max() == 0:
            continue
        d = (d / np.abs(d).max()).astype(np.float32)
        sc = O.OracleScene([S.MeshData(tri, np.array([[0, 1, 2]], np.uint32))])
        h = sc.intersect(np.array([[*o, *d]], np.float32))[0]
        # exact: solve o + t d = a u + b v + c w with u + v + w = 1 over rationals
        A = [[Fraction(float(tri[k][i])) for k in range(3)] + [Fraction(-float(d[i]))] for i in range(3)]
        A.append([Fraction(1)] * 3 + [Fraction(0)])
        rhs = [Fraction(float(o[i])) for i in range(3)] + [Fraction(1)]
        M = [row + [r] for row, r in zip(A, rhs)]
        n = 4
        for c in range(n):
            p = next((r for r in range(c, n) if M[r][c] != 0), None)
            if p is None:
                break
            M[c], M[p] = M[p], M[c]
            for r in range(n):
                if r != c and M[r][c] != 0:
                    fct = M[r][c] / M[c][c]
                    M[r] = [x - fct * y for x, y in zip(M[r], M[c])]
        else:
            sol = [M[i][n] / M[i][i] for i in range(n)]
            u, v, w, t = sol
            if min(u, v, w) < 0 or t <= 0:
                continue
            if h["hit"] != 1:
                # the reference rejects hits inside its conservative t error bound only
                continue
            # t: several roundings (shear, products, quotient) -> 16 ulp of t; barycentrics sum to 1, their
            # error is absolute -> 16 ulp of 1.0 (a dropped term or wrong sign errs by far more)
            et = float(t)
            assert abs(float(h["t"]) - et) <= 16 * float(np.spacing(np.float32(et))), (h["t"], et)
            for got, exp in ((h["u"], u), (h["v"], v), (h["w"], w)):
                assert abs(float(got) - float(exp)) <= 16 * float(np.spacing(np.float32(1.0))), (got, float(exp))
            checked += 1
    assert checked > 100


# ----------------------------------------------------------------------------- invariants

def test_watertight_closed_icosphere():
    v, f = S.icosphere(4, noise=0.0)
    sc = O.OracleScene([S.MeshData(v, f)])
    rng = S.Rng(3)
    d = S.unit_vectors(20000, rng).astype(np.float32)
    o = np.zeros_like(d)
    # plus rays aimed exactly at vertices and edge midpoints from the centre
    feat = S.rays_at_features(v, f, 20000, seed=4, dist=0.0, center=(0.0, 0.0, 0.0))
    rays = np.concatenate([np.concatenate([o, d], 1), feat])
    rays = rays[np.isfinite(rays).all(1)]
    hits = sc.intersect(rays)
    assert (hits["hit"] == 1).all(), "a ray from inside a closed mesh leaked"


def test_watertight_sliver_grid():
    # Rays aimed exactly at shared vertices and edge midpoints of a grid of long thin
    # triangles (aspect up to 1000:1) must hit.  Grazing rays (|d_y| < 0.25) are left
    # out: for them the reference's conservative t bound (mesh.cpp:312-322) may reject
    # a hit by design, which is not a leak.
    v, f = S.sliver_grid(64, 4.0, aspect=1000.0, seed=5)
    sc = O.OracleScene([S.MeshData(v, f)])
    rays = S.rays_at_features(v, f, 20000, seed=6, dist=2.0, center=(0.0, 0.0, 0.0))
    hits = sc.intersect(rays)
    o = rays[:, :3].astype(np.float64)
    d = rays[:, 3:].astype(np.float64)
    t_plane = -o[:, 1] / d[:, 1]
    p = o + t_plane[:, None] * d
    aimed = (np.abs(p[:, 0]) < 1.9) & (np.abs(p[:, 2]) < 1.9) & (t_plane > 0) & (np.abs(d[:, 1]) > 0.25)
    assert aimed.sum() > 1000
    assert (hits["hit"][aimed] == 1).all(), "a ray aimed at a shared edge or vertex leaked"


def test_closest_hit_against_brute_force():
    v, f = S.icosphere(2, noise=0.1, seed=9)
    gv, gf = S.quad(2.0, -1.0)
    meshes = [S.MeshData(v, f), S.MeshData(gv, gf)]
    sc = O.OracleScene(meshes)
    lo, hi = S.scene_bounds(meshes)
    rays = S.rays_incoherent(lo, hi, 3000, 10)
    got = sc.intersect(rays)
    tris = np.concatenate([v[f].astype(np.float64), gv[gf].astype(np.float64)])
    ids = np.concatenate([np.stack([np.zeros(len(f)), np.arange(len(f))], 1),
                          np.stack([np.ones(len(gf)), np.arange(len(gf))], 1)]).astype(np.int64)
    checked = 0
    for i, r in enumerate(rays.astype(np.float64)):
        o, d = r[:3], r[3:]
        e1 = tris[:, 1] - tris[:, 0]
        e2 = tris[:, 2] - tris[:, 0]
        p = np.cross(d, e2)
        det = (e1 * p).sum(1)
        with np.errstate(divide="ignore", invalid="ignore"):
            s = o - tris[:, 0]
            uu = (s * p).sum(1) / det
            q = np.cross(s, e1)
            vv = (d * q).sum(1) / det
            t = (e2 * q).sum(1) / det
        ok = (np.abs(det) > 1e-12) & (uu >= 0) & (vv >= 0) & (uu + vv <= 1) & (t > 1e-9)
        tt = np.where(ok, t, np.inf)
        order = np.argsort(tt)
        if not np.isfinite(tt[order[0]]):
            assert got["hit"][i] == 0
            checked += 1
            continue
        if np.isfinite(tt[order[1]]) and tt[order[1]] - tt[order[0]] < 1e-5:
            continue          # near-tie: the winner depends on test order
        margins = (uu[order[0]], vv[order[0]], 1 - uu[order[0]] - vv[order[0]])
        if min(margins) < 1e-5:
            continue          # on an edge: the winner depends on test order
        assert got["hit"][i] == 1
        assert (got["mesh"][i], got["face"][i]) == tuple(ids[order[0]])
        assert abs(got["t"][i] - tt[order[0]]) <= 1e-5 * max(1.0, tt[order[0]])
        checked += 1
    assert checked > 2500


def test_stress_feature_rays_do_not_leak():
    # the stress workload at reduced size (4 displaced spheres of 4,096 faces and a
    # 64x64 sliver grid): every ray aimed exactly at a shared vertex or edge
    # midpoint, arriving non-grazing, hits at or before its target
    meshes, rays, n_feat, t_target = S.stress_workload(1 << 18, spheres=4, n_lat=32, n_lon=64, grid_n=64)
    ref = O.OracleScene(meshes).intersect(rays[:n_feat])
    hit = ref["hit"] == 1
    assert hit.all()
    assert (ref["t"] <= t_target * (1 + 1e-5) + 1e-6).all()
    # a BVH handed in from outside gives the same hits as the restated build
    sc = O.OracleScene(meshes)
    again = O.OracleScene(meshes, bvh=sc.bvh()).intersect(rays[:4096])
    assert np.array_equal(again.view(np.uint8), ref[:4096].view(np.uint8))


def test_next_event_estimation_closed_form():
    # one Lambert plane y = 0 (Kd), one point light straight above at height h, a
    # camera looking down, depth 1 and no sky: a pixel's radiance is exactly the
    # direct term Kd * I * cos / (pi * d^2) at its hit point (the bounce ray then
    # escapes to black); a blocker between plane and light makes its shadow black
    kd = (0.5, 0.6, 0.7)
    intensity = (1000.0, 2000.0, 3000.0)
    h = 3.0
    gv, gf = S.quad(10.0, 0.0)
    plane = S.MeshData(gv, gf, refl=kd)
    cam = S.Camera(S.look_at((0.0, 6.0, 0.0), (0.0, 0.0, 0.0), up=(0.0, 0.0, 1.0)), 60.0, 24, 24)
    lights = [((0.0, h, 0.0), intensity)]
    film = O.OracleScene([plane], None, lights=lights).render(cam, 1, 1, 3, nee=True)
    # hit points of the pixel-centre rays on y = 0
    rays = S.rays_coherent(cam, cam.width, cam.height).astype(np.float64)
    t = -rays[:, 1] / rays[:, 4]
    p = rays[:, :3] + t[:, None] * rays[:, 3:]
    w = np.array([0.0, h, 0.0]) - p
    d2 = (w * w).sum(1)
    cosv = w[:, 1] / np.sqrt(d2)
    expect = np.array(kd)[None, :] * np.array(intensity)[None, :] * (cosv / (np.pi * d2))[:, None]
    got = film.reshape(-1, 3).astype(np.float64)
    assert np.abs(got - expect).max() / expect.max() < 1e-4
    # depth 0 stops before any vertex, so there is no direct term
    assert not O.OracleScene([plane], None, lights=lights).render(cam, 1, 0, 3, nee=True).any()
    # a small blocker at y = 1.5 under the light: plane points whose segment to the
    # light crosses it (and that the camera still sees) are black, the others keep
    # their value; points within 2 % of the shadow's edge are not checked
    bv, bf = S.quad(0.5, 1.5)
    film2 = O.OracleScene([plane, S.MeshData(bv, bf, refl=(0.1, 0.1, 0.1))], None,
                          lights=lights).render(cam, 1, 1, 3, nee=True).reshape(-1, 3)
    cam_o = rays[:, :3]
    seen = np.abs(cam_o + (p - cam_o) * ((6.0 - 1.5) / 6.0))[:, [0, 2]].max(1) > 0.52    # camera ray misses the blocker
    at = p + (np.array([0.0, h, 0.0]) - p) * (1.5 / h)                                 # light segment at y = 1.5
    r = np.abs(at[:, [0, 2]]).max(1)
    shadow, lit = seen & (r < 0.49), seen & (r > 0.51)
    assert shadow.sum() >= 4 and lit.sum() >= 100
    assert (film2[shadow] == 0).all()
    assert np.array_equal(film2[lit], film.reshape(-1, 3)[lit])
    # off by default: the same scene without the flag renders black
    assert not O.OracleScene([plane], None, lights=lights).render(cam, 1, 1, 3).any()


def test_render_ray_accounting():
    # the counted render traces one ray per sample per bounce and sums the
    # per-ray counters of the reference's intersect (checked against the
    # counters orc_intersect returns for the same primary rays at depth 0)
    import ctypes as C
    from polytope_amd import camera_desc
    meshes, cam, sky = S.plain_render_scene(1, 40, 30)
    sc = O.OracleScene(meshes, sky)
    film, st = sc.render(cam, 2, 0, 3, stats=True)
    assert st["samples"] == 2 * 40 * 30 and st["rays"] == st["samples"]
    cd = camera_desc(cam)
    rays = np.zeros((40 * 30, 6), np.float32)
    for y in range(30):
        for x in range(40):
            O.lib().orc_camera_ray(C.byref(cd), C.c_float(x + 0.5), C.c_float(y + 0.5), rays[y * 40 + x].ctypes.data)
    ref = sc.intersect(rays)
    assert st["boxes"] == 2 * int(ref["bb_hits"].sum() + ref["bb_misses"].sum())
    assert st["tris"] == 2 * int(ref["tri_isects"].sum())
    assert np.array_equal(film, sc.render(cam, 2, 0, 3))
    # deeper paths trace more rays, never more than depth + 1 per sample
    _, st5 = sc.render(cam, 2, 5, 3, stats=True)
    assert st["rays"] < st5["rays"] <= 6 * st5["samples"]
