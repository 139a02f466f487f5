"""Writes the golden fixtures in tests/golden/ from the reference polytope CPU
path, compiled from /root/reference by oracle/ref/Makefile (oracle/_ref/ref_driver).

Each fixture (.npz, no pickles) holds the inputs — meshes (packed vertices,
faces, object-to-world matrices), rays — and the reference's outputs for them:
the compact BVH (32-byte nodes), the leaf (mesh, face) pairs and the 100-byte
per-ray hit records of bvh::intersect_compact + Mesh::intersect (hit flag,
mesh, face, t, u, v, w, Location, normal, tangents, error bound and the
per-ray counters).  Render fixtures hold the reference film of a
deterministic (mirror + emitter + skybox) scene.  The texture fixture holds
camera rays of the textured scene that tests/test_textures.py writes and the
reference's Scene::intersect + LambertBRDF::sample texture lookups for them.

Run in a container where /root/reference exists:  python tests/golden/make_golden.py
"""
from __future__ import annotations

import sys
import tempfile
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[2]
sys.path.insert(0, str(ROOT))

from polytope_amd import scenes as S  # noqa: E402
import oracle as O  # noqa: E402

OUT = Path(__file__).resolve().parent


def pack_meshes(meshes):
    d = {"n_meshes": np.array(len(meshes), np.uint32)}
    for i, m in enumerate(meshes):
        d[f"m{i}_verts"] = m.verts.astype(np.float32)
        d[f"m{i}_faces"] = m.faces.astype(np.uint32)
        d[f"m{i}_o2w"] = np.asarray(m.object_to_world, np.float32)
        d[f"m{i}_mat"] = np.array([m.brdf, *m.refl, 1.0 if m.is_light else 0.0, *m.emission], np.float32)
    return d


def reference_outputs(meshes, rays, sky=None, camera=None, spp=1, max_depth=5):
    with tempfile.TemporaryDirectory() as d:
        S.write_ptxm(f"{d}/s.bin", meshes, camera=camera, spp=spp, max_depth=max_depth, skybox=sky)
        args = ["--ptxm", f"{d}/s.bin", "--nodes", f"{d}/n.bin", "--pairs", f"{d}/p.bin"]
        if rays is not None:
            S.write_rays(f"{d}/r.bin", rays)
            args += ["--rays", f"{d}/r.bin", "--hits", f"{d}/h.bin"]
        if camera is not None:
            args += ["--render", f"{d}/f.bin", "--threads", "8"]
        O.run_reference(args, cwd=d)
        out = {"nodes": np.fromfile(f"{d}/n.bin", np.uint8),
               "pairs": np.fromfile(f"{d}/p.bin", np.uint32).reshape(-1, 2)}
        if rays is not None:
            out["hits"] = np.fromfile(f"{d}/h.bin", np.uint8)
        if camera is not None:
            out["film"] = np.fromfile(f"{d}/f.bin", np.float32).reshape(camera.height, camera.width, 3)
        return out


def save(name, meshes, rays, sky=None, camera=None, spp=1, max_depth=5):
    ref = reference_outputs(meshes, rays, sky, camera, spp, max_depth)
    d = pack_meshes(meshes)
    if rays is not None:
        d["rays"] = rays.astype(np.float32)
    if sky is not None:
        d["sky"] = np.asarray(sky, np.float32)
    if camera is not None:
        d["camera"] = np.concatenate([np.asarray(camera.camera_to_world, np.float32).ravel(),
                                      np.array([camera.fov, camera.width, camera.height, spp, max_depth],
                                               np.float32)])
    d.update({f"ref_{k}": v for k, v in ref.items()})
    np.savez_compressed(OUT / f"{name}.npz", **d)
    print(f"wrote {name}.npz: {sum(v.nbytes for v in d.values()) / 1e3:.0f} kB raw")


def main():
    if O.build_reference() is None:
        raise SystemExit("reference tree not found; fixtures are only regenerated where /root/reference exists")
    # 1. displaced icosphere, 3 subdivisions: mixed + feature-aimed rays
    v, f = S.icosphere(3, noise=0.05, seed=2)
    meshes = [S.MeshData(v, f)]
    _, rays, _ = S.closest_hit_workload(3, 2048, seed=2)
    edge = S.rays_at_features(v, f, 1024, seed=7)
    save("ico3", meshes, np.concatenate([rays, edge]))
    # 2. displaced icosphere, 5 subdivisions, incoherent + feature rays
    v, f = S.icosphere(5, noise=0.05, seed=11)
    meshes = [S.MeshData(v, f)]
    lo, hi = S.scene_bounds(meshes)
    rays = np.concatenate([S.rays_incoherent(lo, hi, 1536, 12), S.rays_at_features(v, f, 1536, seed=13)])
    save("ico5", meshes, rays)
    # 3. the plain-render scene (ground quad + 3 lat-long spheres, 4 meshes)
    meshes, cam, sky = S.plain_render_scene(1, 64, 48)
    lo, hi = S.scene_bounds(meshes)
    save("plain_multimesh", meshes, S.rays_incoherent(lo, hi, 2048, 9), sky=sky)
    # 4. sliver grid (aspect up to 1000:1), rays aimed at its shared edges and vertices
    v, f = S.sliver_grid(48, 4.0, aspect=1000.0, seed=5)
    meshes = [S.MeshData(v, f)]
    rays = S.rays_at_features(v, f, 2048, seed=14, dist=3.0, center=(0.0, 0.0, 0.0))
    save("sliver48", meshes, rays)
    # 5. deterministic render (mirror only, no RNG draws): one mirror mesh holding a
    #    displaced icosphere and a ground quad, constant skybox
    v, f = S.icosphere(3, radius=1.0, center=(0.0, 1.0, 0.0), noise=0.05, seed=3)
    gv, gf = S.quad(6.0, 0.0)
    mv = np.concatenate([v, gv])
    mf = np.concatenate([f, gf + len(v)]).astype(np.uint32)
    meshes = [S.MeshData(mv, mf, brdf=S.MIRROR, refl=(0.9, 0.8, 0.7))]
    cam = S.Camera(S.look_at((0.0, 2.0, -5.0), (0.0, 0.8, 0.0)), 60.0, 64, 48)
    save("mirror_render", meshes, None, sky=(40.0, 60.0, 90.0), camera=cam, spp=2, max_depth=5)
    # 6. the same with three meshes and an emitter (the reference's multi-mesh hit frame applies)
    lv, lf = S.quad(1.5, 4.0)
    meshes = [S.MeshData(v, f, brdf=S.MIRROR, refl=(0.9, 0.8, 0.7)),
              S.MeshData(gv, gf, brdf=S.MIRROR, refl=(0.5, 0.6, 0.7)),
              S.MeshData(lv, lf, is_light=True, emission=(255.0, 240.0, 200.0))]
    save("mirror_multimesh", meshes, None, sky=(40.0, 60.0, 90.0), camera=cam, spp=2, max_depth=5)
    # 7. textured scene file (PNG textures, PLY u/v, coordinates past the image):
    #    the reference's texture lookups on one camera ray per pixel
    save_tex_probe("tex_probe_crafted")


def save_tex_probe(name):
    from polytope_amd import pbrt as P
    from tests.test_textures import TEX_PROBE, TEXTURED, _camera_rays, _crafted
    with tempfile.TemporaryDirectory() as d:
        d = Path(d)
        _crafted(d)
        (d / "s.pbrt").write_text(TEXTURED)
        sf = P.read_pbrt(str(d / "s.pbrt"))
        rays = _camera_rays(sf.camera, 1)
        S.write_rays(str(d / "rays.bin"), rays)
        O.run_reference(["--pbrt", str(d / "s.pbrt"), "--rays", "rays.bin", "--tex-probe", "probe.bin"], cwd=str(d))
        probe = np.fromfile(d / "probe.bin", TEX_PROBE)
    np.savez_compressed(OUT / f"{name}.npz", rays=rays, probe=probe.view(np.uint32).reshape(len(rays), -1))
    print(f"wrote {name}.npz: {len(rays)} rays, {int((probe['flags'] & 8).astype(bool).sum())} texel lookups")

def save_glossy_draws(name="glossy_draws"):
    """The reference's GlossyBRDF::sample on 300 incoming directions and
    roughnesses, with the uniforms its generator handed out (the reference
    driver's --glossy-cases): the inputs and the reference's rows."""
    import struct
    from tests.test_glossy import _cases
    cases = _cases(300, 11)
    with tempfile.TemporaryDirectory() as d:
        with open(f"{d}/in.bin", "wb") as fh:
            fh.write(struct.pack("<I", len(cases)))
            for v, r, sd in cases:
                fh.write(struct.pack("<4fI", float(v[0]), float(v[1]), float(v[2]), float(r), sd))
        O.run_reference(["--glossy-cases", "in.bin", "out.bin"], cwd=d)
        rows = np.fromfile(f"{d}/out.bin", np.float32).reshape(len(cases), 132)
    np.savez_compressed(OUT / f"{name}.npz", incoming=np.array([c[0] for c in cases], np.float32),
                        roughness=np.array([c[1] for c in cases], np.float32),
                        seed=np.array([c[2] for c in cases], np.uint32), rows=rows)
    print(f"wrote {name}.npz: {len(cases)} cases")


# Scenes whose reference BVH is pinned by hash (tests/test_abi_and_builder.py):
# name -> generator of the meshes.
LARGE_BVH_SCENES = {
    "closest_hit_subdiv8": lambda: S.closest_hit_workload(8, 1024, seed=2)[0],
    "stress": lambda: S.stress_workload(1 << 10)[0],
    "batched_render": lambda: S.batched_render_workload(8, 32, 16, 64, 8)[0],
}


def save_bvh_hashes(names=tuple(LARGE_BVH_SCENES), name="bvh_hashes"):
    """SHA-256 of the reference's own node and leaf-pair arrays (bvh::bound +
    bvh::compact, 32-byte nodes and (mesh, face) pairs as the driver writes
    them) for the large workload scenes: 1.31M, 10.5M and 43.25M faces."""
    import hashlib
    import json
    import time
    path = OUT / f"{name}.json"
    table = json.loads(path.read_text()) if path.exists() else {}
    for key in names:
        meshes = LARGE_BVH_SCENES[key]()
        with tempfile.TemporaryDirectory() as d:
            S.write_ptxm(f"{d}/s.bin", meshes)
            t0 = time.time()
            O.run_reference(["--ptxm", f"{d}/s.bin", "--nodes", f"{d}/n.bin", "--pairs", f"{d}/p.bin"], cwd=d,
                            timeout=3600)
            dt = time.time() - t0
            nodes = np.fromfile(f"{d}/n.bin", np.uint8)
            pairs = np.fromfile(f"{d}/p.bin", np.uint8)
        table[key] = {"faces": int(sum(m.num_faces for m in meshes)), "nodes": int(nodes.size // 32),
                      "pairs": int(pairs.size // 8),
                      "sha256_nodes": hashlib.sha256(nodes.tobytes()).hexdigest(),
                      "sha256_pairs": hashlib.sha256(pairs.tobytes()).hexdigest(),
                      "reference_seconds": round(dt, 1)}
        del meshes, nodes, pairs
        path.write_text(json.dumps(table, indent=1) + "\n")
        print(f"{key}: {table[key]}", flush=True)


if __name__ == "__main__":
    if O.build_reference() is None:
        raise SystemExit("reference tree not found; fixtures are only regenerated where /root/reference exists")
    what = sys.argv[1:] or ["all"]
    if "all" in what:
        main()
        save_glossy_draws()
        save_bvh_hashes()
    if "glossy" in what:
        save_glossy_draws()
    if "bvh_hashes" in what:
        save_bvh_hashes()
    for key in what:
        if key in LARGE_BVH_SCENES:
            save_bvh_hashes((key,))
