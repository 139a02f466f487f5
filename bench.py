#!/usr/bin/env python3
"""Benchmark: incoherent closest-hit throughput on MI355X.

Workload (one GPU): a displaced icosphere subdivided 8 times (1,310,720 FP32
triangles, radial value-noise displacement of +/-5 % of the radius) queried with
2^22 rays — 2^21 coherent pinhole rays on a 2048x1024 grid aimed at the sphere
and 2^21 incoherent rays (origins uniform in the bounding box, directions
uniform on the sphere).  One step = one closest-hit launch over the 2^21
incoherent rays, inputs resident in HBM; outputs are hit flag, (mesh, face),
t and barycentrics per ray.

Prints one JSON line: metric/value (incoherent Mrays/s), the roofline of the
traversal kernel (algorithmic bytes per ray from the reference's own per-ray
counters, which the kernel reproduces exactly), the coherent rate, and the CPU
baseline (the reference CPU intersector when its build is present, otherwise
the oracle restatement) timed on this host on a bounded sample.
"""
from __future__ import annotations

import argparse
import json
import os
import subprocess
import sys
import tempfile
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

HBM_PEAK_GBS = 8000.0


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--subdiv", type=int, default=8)
    ap.add_argument("--rays-log2", type=int, default=22)
    ap.add_argument("--cpu-sample-log2", type=int, default=21)
    ap.add_argument("--cpu-repeat", type=int, default=36, help="passes over the CPU sample (about 10 s)")
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--skip-cpu", action="store_true")
    ap.add_argument("--lds", type=int, default=0, help="LDS stack entries per lane (0 = library default)")
    ap.add_argument("--refill", type=int, default=0, help="idle lanes that trigger a refill (0 = default)")
    ap.add_argument("--blocks-per-cu", type=int, default=0)
    ap.add_argument("--leaf-lanes", type=int, default=0,
                    help="lanes a traversal wave gathers at leaves before a leaf step (0 = default)")
    ap.add_argument("--ray-windows", type=int, default=0,
                    help="closest-hit record form: 0 default, 1 window records, 2 one node record per step")
    ap.add_argument("--no-coherent", action="store_true", help="skip the coherent-rate measurement")
    ap.add_argument("--skip-render", action="store_true", help="skip the plain vs batched render comparison")
    ap.add_argument("--skip-serving", action="store_true", help="skip the Poisson serving run")
    ap.add_argument("--render-subdiv", type=int, default=8)
    ap.add_argument("--render-instances", type=int, default=32)
    ap.add_argument("--render-spp", type=int, default=64)
    ap.add_argument("--shade-lanes", type=int, default=0,
                    help="lanes a path-regeneration wave gathers before it shades (0 = default)")
    ap.add_argument("--render-paths", type=int, default=0, help="path slots per round (0 = default)")
    ap.add_argument("--serving-requests", type=int, default=32)
    ap.add_argument("--serving-rate", type=float, default=8.0)
    ap.add_argument("--serving-loaded-requests", type=int, default=64,
                    help="second stream at a rate near the GPU's capacity")
    ap.add_argument("--serving-loaded-rate", type=float, default=96.0,
                    help="offers about 2.8 G samples/s, about 90 %% of what the server delivers on a burst")
    ap.add_argument("--render-batched-only", action="store_true",
                    help="render section: the batched form and its accounting only (profiling runs)")
    ap.add_argument("--skip-accounting", action="store_true",
                    help="skip the render workload's ray and byte accounting (oracle on a row sample)")
    ap.add_argument("--calibrate-copy-mb", type=int, default=0,
                    help="after the run, copy this many MiB device-to-device (profiler calibration)")
    return ap.parse_args()


def algorithmic_bytes(counters: np.ndarray) -> np.ndarray:
    """B_ray = 24 (ray) + 20 (hit) + 32 * N_box + 36 * N_tri, with N_box the nodes the
    reference pops and box-tests (num_bb_hits + num_bb_misses) and N_tri its face tests."""
    n_box = counters["bb_hits"].astype(np.float64) + counters["bb_misses"]
    return 24.0 + 20.0 + 32.0 * n_box + 36.0 * counters["tri_isects"].astype(np.float64)


def host_threads(args) -> int:
    """Threads for the CPU baselines: the cores this process may run on (nproc),
    or fewer when OMP_NUM_THREADS declares a smaller share of the host."""
    if args.cpu_threads:
        return args.cpu_threads
    n = len(os.sched_getaffinity(0))
    env = os.environ.get("OMP_NUM_THREADS", "")
    if env.isdigit() and int(env) > 0:
        n = min(n, int(env))
    return max(1, n)


def cpu_model() -> str:
    try:
        for ln in open("/proc/cpuinfo"):
            if ln.startswith("model name"):
                return ln.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(meshes, rays: np.ndarray, threads: int, repeat: int):
    import oracle as O
    from polytope_amd import scenes as S
    if O.reference_available():
        with tempfile.TemporaryDirectory() as d:
            S.write_ptxm(f"{d}/s.bin", meshes)
            S.write_rays(f"{d}/r.bin", rays)
            S.write_rays(f"{d}/r1.bin", rays[: 1 << 16])

            def timed(path, nthreads, reps):
                p = subprocess.run([str(O.REF_DRIVER), "--ptxm", f"{d}/s.bin", "--time-rays", path,
                                    "--threads", str(nthreads), "--repeat", str(reps)],
                                   capture_output=True, text=True, timeout=900, cwd=d)
                line = [ln[ln.index("{\"rays\""):] for ln in p.stdout.splitlines() if "{\"rays\"" in ln]
                return json.loads(line[-1]) if p.returncode == 0 and line else None

            r = timed(f"{d}/r.bin", threads, repeat)
            r1 = timed(f"{d}/r1.bin", 1, 1)
            if r:
                return {"value": round(r["mrays_per_s"], 4), "unit": "Mrays/s", "cores": threads,
                        "kind": "reference",
                        "sample": f"the workload's {len(rays)} incoherent rays x {repeat} repeats "
                                  f"({r['seconds']:.1f} s), reference bvh::intersect_compact built from "
                                  f"its own sources, {threads} threads over contiguous ray chunks",
                        "one_thread_mrays_per_s": round(r1["mrays_per_s"], 4) if r1 else None,
                        "cpu_model": cpu_model()}
    sc = O.OracleScene(meshes)
    sc.intersect_ids(rays[: min(len(rays), 4096)], threads)
    t0 = time.perf_counter()
    sc.intersect_ids(rays, threads)
    dt = time.perf_counter() - t0
    return {"value": round(len(rays) / dt / 1e6, 4), "unit": "Mrays/s", "cores": threads, "kind": "port",
            "sample": f"the workload's first {len(rays)} incoherent rays, oracle restatement, {threads} threads",
            "cpu_model": cpu_model()}


def render_accounting(orc, jobs, t_batched: float, threads: int, row_step: int = 64):
    """Rays and algorithmic bytes of the batched render: the oracle's
    restatement of the integrator on every row_step-th row of every job counts
    the rays a sample traces and the reference's per-ray counters over them,
    scaled to the whole films.  The batched path traces a pixel's primary ray
    once (not once per sample), so it traces (spp - 1) * pixels fewer rays than
    the integrator.  Bytes: per traced ray B_ray = 24 + 20 + 32 * N_box + 36 *
    N_tri, per sample 64 B of primary record and 12 B of result."""
    tot = {"rays": 0.0, "boxes": 0.0, "tris": 0.0, "samples": 0.0}
    gpu_rays = 0.0
    t0 = time.perf_counter()
    for cam, spp, depth, seed in jobs:
        _, st = orc.render(cam, spp, depth, seed, threads=threads, row_step=row_step, stats=True)
        rows = len(range(0, cam.height, row_step))
        scale = cam.height / rows
        for k in tot:
            tot[k] += st[k] * scale
        gpu_rays += (st["rays"] - (spp - 1) * rows * cam.width) * scale
    dt = time.perf_counter() - t0
    rays = tot["rays"]
    mean_box, mean_tri = tot["boxes"] / rays, tot["tris"] / rays
    b_ray = 24 + 20 + 32 * mean_box + 36 * mean_tri
    total_bytes = gpu_rays * b_ray + tot["samples"] * (64 + 12)
    return {"method": f"oracle integrator on every {row_step}th row of every job ({dt:.1f} s), scaled to the films",
            "samples": round(tot["samples"]), "integrator_rays": round(rays),
            "rays_per_sample": round(rays / tot["samples"], 4), "extend_rays": round(gpu_rays),
            "extend_mrays_per_s": round(gpu_rays / t_batched / 1e6, 1),
            "mean_box_tests_per_ray": round(mean_box, 2), "mean_face_tests_per_ray": round(mean_tri, 2),
            "bytes_per_ray": round(b_ray, 1), "algorithmic_bytes": round(total_bytes),
            "achieved_gbs": round(total_bytes / t_batched / 1e9, 1),
            "frac_of_hbm_peak": round(total_bytes / t_batched / 1e9 / HBM_PEAK_GBS, 4)}


def render_section(dev, args):
    """Plain (one job per launch series) vs batched (all jobs' samples in shared
    rounds: path regeneration, and the wavefront form for comparison) samples/s
    on the 16-job instanced-icosphere workload."""
    from polytope_amd import Scene, render, render_batch
    from polytope_amd import _abi
    from polytope_amd import scenes as S
    t0 = time.perf_counter()
    meshes, sky, jobs = S.batched_render_workload(args.render_subdiv, args.render_instances, 16, args.render_spp, 8)
    scene = Scene(dev, meshes, sky)
    info = scene.info()
    print(f"[bench] render scene: {sum(m.num_faces for m in meshes)} triangles, BVH build "
          f"{info['build_seconds']:.1f} s, {info['device_bytes'] / 1e9:.2f} GB on device "
          f"({time.perf_counter() - t0:.1f} s)", file=sys.stderr, flush=True)
    samples = sum(c.width * c.height * spp for c, spp, _, _ in jobs)
    batch = [(scene, c, spp, d, seed) for c, spp, d, seed in jobs]
    dev.set_tuning(shade_lanes=args.shade_lanes, paths_per_wave=args.render_paths, leaf_lanes=args.leaf_lanes)
    # warm-up on the same batch: the engine allocates its path slots (up to
    # 2^30, 12 B each) on the first batch that needs them, as a long-running
    # process does once
    render_batch(batch)
    tb = time.perf_counter()
    films_b = render_batch(batch)
    t_batched = time.perf_counter() - tb
    print(f"[bench] batched: {samples / t_batched / 1e6:.1f} Msamples/s", file=sys.stderr, flush=True)
    if args.render_batched_only:
        accounting = None
        if not args.skip_accounting:
            import oracle as O
            orc = O.OracleScene(meshes, sky, bvh=scene.bvh())
            accounting = render_accounting(orc, jobs, t_batched, host_threads(args))
            del orc
        scene.close()
        return {"workload": f"icosphere subdiv {args.render_subdiv} x {args.render_instances + 1} instances + ground "
                            f"({sum(m.num_faces for m in meshes)} tris), 16 jobs, {args.render_spp} spp, depth 8",
                "samples": samples, "batched_s": round(t_batched, 3),
                "batched_samples_per_s": round(samples / t_batched, 1), "accounting": accounting}
    dev.set_tuning(render_mode=_abi.RENDER_WAVEFRONT, paths_per_wave=args.render_paths, leaf_lanes=args.leaf_lanes)
    render_batch([(scene, jobs[0][0], 1, 8, 1)])
    tw = time.perf_counter()
    films_w = render_batch(batch)
    t_wave = time.perf_counter() - tw
    dev.set_tuning()
    print(f"[bench] batched wavefront: {samples / t_wave / 1e6:.1f} Msamples/s", file=sys.stderr, flush=True)
    # plain: one job per call, one launch per sample; each pixel's primary ray
    # traced once (the default), and again for every sample (the reference
    # CUDA path's shape)
    tp = time.perf_counter()
    films_p = [render(scene, c, spp, d, seed) for c, spp, d, seed in jobs]
    t_plain = time.perf_counter() - tp
    print(f"[bench] plain: {samples / t_plain / 1e6:.1f} Msamples/s", file=sys.stderr, flush=True)
    dev.set_tuning(plain_primary=_abi.PLAIN_RETRACE)
    tu = time.perf_counter()
    films_u = [render(scene, c, spp, d, seed) for c, spp, d, seed in jobs]
    t_uncached = time.perf_counter() - tu
    dev.set_tuning()
    print(f"[bench] plain, primary rays traced every sample: {samples / t_uncached / 1e6:.1f} Msamples/s",
          file=sys.stderr, flush=True)
    same = all(np.array_equal(a.view(np.uint32), b.view(np.uint32)) and
               np.array_equal(a.view(np.uint32), w.view(np.uint32)) and
               np.array_equal(a.view(np.uint32), u.view(np.uint32))
               for a, b, w, u in zip(films_b, films_p, films_w, films_u))
    del films_u
    # the same geometry and BVH with textured Lambert and glossy meshes
    bvh = scene.bvh()
    mat_scene = Scene(dev, S.with_materials(meshes), sky, bvh=bvh)
    scene.close()
    mat_batch = [(mat_scene, c, spp, d, seed) for c, spp, d, seed in jobs]
    render_batch([(mat_scene, jobs[0][0], 1, 8, 1)])
    tm = time.perf_counter()
    films_m = render_batch(mat_batch)
    t_mat = time.perf_counter() - tm
    tmp = time.perf_counter()
    films_mp = [render(mat_scene, c, spp, d, seed) for c, spp, d, seed in jobs]
    t_mat_plain = time.perf_counter() - tmp
    same_m = all(np.array_equal(a.view(np.uint32), b.view(np.uint32)) for a, b in zip(films_m, films_mp))
    print(f"[bench] textured + glossy: batched {samples / t_mat / 1e6:.1f}, plain {samples / t_mat_plain / 1e6:.1f} "
          f"Msamples/s", file=sys.stderr, flush=True)
    mat_scene.close()
    cpu = None
    accounting = None
    orc = None
    if not (args.skip_cpu and args.skip_accounting):
        import oracle as O
        orc = O.OracleScene(meshes, sky, bvh=bvh)
    del bvh
    if not args.skip_accounting:
        accounting = render_accounting(orc, jobs, t_batched, host_threads(args))
        print(f"[bench] render accounting: {accounting['extend_rays'] / 1e9:.3f} G extend rays, "
              f"{accounting['extend_mrays_per_s']:.0f} Mrays/s, {accounting['achieved_gbs']:.0f} GB/s of algorithmic "
              f"bytes", file=sys.stderr, flush=True)
    if not args.skip_cpu:
        # CPU baseline: the oracle (port of the reference's integrator) on the same
        # scene and tree, the whole first job, on the host's cores
        threads = host_threads(args)
        cam0, spp0, d0, seed0 = jobs[0]
        tc = time.perf_counter()
        film0 = orc.render(cam0, spp0, d0, seed0, threads=threads)
        t_cpu = time.perf_counter() - tc
        n_cpu = cam0.height * cam0.width * spp0
        cpu = {"value": round(n_cpu / t_cpu, 1), "unit": "samples/s", "cores": threads, "kind": "port",
               "sample": f"job 0 ({cam0.width}x{cam0.height}, {spp0} spp, depth {d0}): {n_cpu} samples in "
                         f"{t_cpu:.2f} s, oracle restatement of PathTraceIntegrator on the same tree",
               "film_equals_gpu": bool(np.array_equal(film0.view(np.uint32), films_b[0].view(np.uint32)))}
        print(f"[bench] render CPU baseline: {n_cpu / t_cpu / 1e6:.3f} Msamples/s", file=sys.stderr, flush=True)
    del orc
    return {"workload": f"icosphere subdiv {args.render_subdiv} x {args.render_instances + 1} instances + ground "
                        f"({sum(m.num_faces for m in meshes)} tris), 16 jobs 256x256..1920x1080, "
                        f"{args.render_spp} spp, depth 8, 70% Lambert / 30% mirror",
            "samples": samples, "batched_s": round(t_batched, 3), "plain_s": round(t_plain, 3),
            "batched_samples_per_s": round(samples / t_batched, 1), "plain_samples_per_s": round(samples / t_plain, 1),
            "batched_vs_plain": round(t_plain / t_batched, 3),
            "plain_uncached_samples_per_s": round(samples / t_uncached, 1),
            "batched_vs_plain_uncached": round(t_uncached / t_batched, 3),
            "plain_note": "plain = one job per call, one launch per sample, each pixel's primary ray traced once "
                          "(as the batched path does); plain_uncached traces it again for every sample, as the "
                          "reference CUDA path does",
            "accounting": accounting,
            "wavefront_samples_per_s": round(samples / t_wave, 1), "films_identical": same,
            "bvh_build_s": round(info["build_seconds"], 2),
            "materials": {"workload": "the same meshes and BVH, Lambert meshes textured (512x512 RGBA8), mirrors "
                                      "glossy (roughness 0.2)",
                          "batched_samples_per_s": round(samples / t_mat, 1),
                          "plain_samples_per_s": round(samples / t_mat_plain, 1), "films_identical": same_m},
            "cpu_baseline": cpu}


def plain_cpu_render_section(dev, args, threads: int):
    """The reference's CPU-runnable render (ground quad + three 1,600-face spheres,
    640x480, 16 spp, depth 5): the reference's own TileRunner on the host's cores
    beside the GPU's plain and batched paths on the same scene."""
    import oracle as O
    from polytope_amd import Scene, render, render_batch
    from polytope_amd import scenes as S
    meshes, cam, sky = S.plain_render_scene(1, 640, 480)
    spp, depth = 16, 5
    samples = cam.width * cam.height * spp
    out = {"workload": "ground quad + 3 lat-long spheres (4,802 faces), 640x480, 16 spp, depth 5",
           "samples": samples}
    scene = Scene(dev, meshes, sky)
    render(scene, cam, 1, depth, 1)
    t = time.perf_counter()
    render(scene, cam, spp, depth, 1)
    out["gpu_plain_samples_per_s"] = round(samples / (time.perf_counter() - t), 1)
    render_batch([(scene, cam, 1, depth, 1)])
    t = time.perf_counter()
    render_batch([(scene, cam, spp, depth, 1)])
    out["gpu_batched_samples_per_s"] = round(samples / (time.perf_counter() - t), 1)
    scene.close()
    if O.reference_available() and not args.skip_cpu:
        with tempfile.TemporaryDirectory() as d:
            S.write_ptxm(f"{d}/c1.bin", meshes, cam, spp, depth, sky)
            p = subprocess.run([str(O.REF_DRIVER), "--ptxm", f"{d}/c1.bin", "--time-render", "--threads",
                                str(threads)], capture_output=True, text=True, timeout=900, cwd=d)
            line = [ln[ln.index("{\"samples\""):] for ln in p.stdout.splitlines() if "{\"samples\"" in ln]
            if p.returncode == 0 and line:
                r = json.loads(line[-1])
                out["cpu_baseline"] = {"value": round(r["msamples_per_s"] * 1e6, 1), "unit": "samples/s",
                                       "cores": threads, "kind": "reference",
                                       "sample": f"the whole frame ({r['seconds']:.2f} s): the reference's "
                                                 f"TileRunner over PathTraceIntegrator, {threads} threads"}
                out["gpu_batched_vs_cpu"] = round(out["gpu_batched_samples_per_s"] / out["cpu_baseline"]["value"], 1)
    print(f"[bench] plain CPU-runnable render: GPU plain {out['gpu_plain_samples_per_s'] / 1e6:.1f}, "
          f"batched {out['gpu_batched_samples_per_s'] / 1e6:.1f} Msamples/s", file=sys.stderr, flush=True)
    return out


def _latency_stats(lat, samples: float, span: float) -> dict:
    lat = np.array(lat)
    return {"throughput_samples_per_s": round(samples / span, 1),
            "latency_p50_s": round(float(np.percentile(lat, 50)), 4),
            "latency_p95_s": round(float(np.percentile(lat, 95)), 4),
            "latency_p99_s": round(float(np.percentile(lat, 99)), 4)}


def serving_section(dev, args, n_requests: int, rate: float, burst: bool = False):
    """Poisson stream of render jobs (or, with burst, all of them at once) through
    the continuous-batching server, and the same stream through the plain path:
    one job per render() call, in arrival order, as a caller without the server
    would serve it.  Latency = completion - arrival."""
    from polytope_amd import Scene, Server, render
    from polytope_amd import scenes as S
    scene_defs, reqs = S.serving_requests(n_requests, rate)
    if burst:
        reqs = [(0.0,) + r[1:] for r in reqs]
    scenes_ = [Scene(dev, m, sky) for m, sky in scene_defs]
    total = sum(cam.width * cam.height * spp for _, _, cam, spp, _, _ in reqs)
    # warm both paths (kernel loading, engine buffers) outside the timed streams
    render(scenes_[0], reqs[0][2], 1, 4, 1)
    srv = Server(dev, hbm_budget_bytes=64 << 30)
    srv.wait(srv.submit(scenes_[0], reqs[0][2], 1, 4, 1))
    ids = []
    t0 = time.perf_counter()
    for arr, si, cam, spp, depth, seed in reqs:
        delay = arr - (time.perf_counter() - t0)
        if delay > 0:
            time.sleep(delay)
        ids.append(srv.submit(scenes_[si], cam, spp, depth, seed))
    # the server's times are seconds on its own clock; a job is submitted at its arrival
    lat, first, last = [], None, None
    for jid in ids:
        _, st = srv.wait(jid)
        lat.append(st["end_time"] - st["submit_time"])
        first = st["submit_time"] if first is None else min(first, st["submit_time"])
        last = st["end_time"] if last is None else max(last, st["end_time"])
    srv.close()
    out = {"requests": len(reqs), "rate_hz": None if burst else rate, "arrivals": "all at t=0" if burst else "Poisson",
           "scenes_tris": [sum(m.num_faces for m in d[0]) for d in scene_defs], "samples": total}
    if not burst:
        out["offered_samples_per_s"] = round(total / max(1e-9, reqs[-1][0]), 1)
    out.update(_latency_stats(lat, total, last - first))
    # the plain path over the same stream
    lat_p = []
    t0 = time.perf_counter()
    for arr, si, cam, spp, depth, seed in reqs:
        delay = arr - (time.perf_counter() - t0)
        if delay > 0:
            time.sleep(delay)
        render(scenes_[si], cam, spp, depth, seed)
        lat_p.append(time.perf_counter() - (t0 + arr))
    span = time.perf_counter() - t0
    out["plain"] = _latency_stats(lat_p, total, span)
    out["server_vs_plain_throughput"] = round(out["throughput_samples_per_s"] / out["plain"]["throughput_samples_per_s"], 3)
    out["server_vs_plain_p50_latency"] = round(out["plain"]["latency_p50_s"] / out["latency_p50_s"], 3)
    for s_ in scenes_:
        s_.close()
    print(f"[bench] serving {out['arrivals']} ({len(reqs)} jobs): server {out['throughput_samples_per_s'] / 1e6:.0f} "
          f"Msamples/s p50 {out['latency_p50_s']:.3f} s p99 {out['latency_p99_s']:.3f} s; plain "
          f"{out['plain']['throughput_samples_per_s'] / 1e6:.0f} Msamples/s p50 {out['plain']['latency_p50_s']:.3f} s "
          f"p99 {out['plain']['latency_p99_s']:.3f} s", file=sys.stderr, flush=True)
    return out


def main():
    args = parse()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus must be 1",
                          "n_gpus_requested": args.gpus}))
        sys.exit(2)
    import torch
    from polytope_amd import COUNTER_DTYPE, HIT_DTYPE, Device, Scene
    from polytope_amd import scenes as S

    n_rays = 1 << args.rays_log2
    t0 = time.perf_counter()
    meshes, rays, n_coh = S.closest_hit_workload(args.subdiv, n_rays, seed=2)
    print(f"[bench] workload built: {meshes[0].num_faces} triangles, {n_rays} rays "
          f"({time.perf_counter() - t0:.1f} s)", file=sys.stderr, flush=True)
    dev = Device(0)
    if args.lds or args.refill or args.blocks_per_cu or args.leaf_lanes or args.ray_windows:
        dev.set_tuning(lds_stack=args.lds, refill_lanes=args.refill, blocks_per_cu=args.blocks_per_cu,
                       leaf_lanes=args.leaf_lanes, ray_windows=args.ray_windows)
    sc = Scene(dev, meshes)
    info = sc.info()
    print(f"[bench] scene: {info['num_nodes']} nodes, build {info['build_seconds']:.2f} s, "
          f"{info['device_bytes'] / 1e6:.1f} MB on device", file=sys.stderr, flush=True)

    d_rays = torch.from_numpy(rays).cuda()
    n_inc = n_rays - n_coh
    inc_ptr = d_rays.data_ptr() + n_coh * 24
    hits = torch.empty(n_rays * 6, dtype=torch.int32, device="cuda")
    cnt = torch.empty(n_rays * 3, dtype=torch.int32, device="cuda")
    stream = torch.cuda.current_stream()

    # algorithmic bytes of the incoherent half (untimed run with counters)
    sc.intersect_device(inc_ptr, n_inc, hits.data_ptr(), cnt.data_ptr(), 0, stream.cuda_stream)
    torch.cuda.synchronize()
    c = cnt[: n_inc * 3].cpu().numpy().view(COUNTER_DTYPE)
    b_ray = algorithmic_bytes(c)
    mean_b = float(b_ray.mean())
    coherent_mrays = None
    if not args.no_coherent:                      # coherent rate (outside the step loop)
        for _ in range(2):
            sc.intersect_device(d_rays.data_ptr(), n_coh, hits.data_ptr(), 0, 0, stream.cuda_stream)
        torch.cuda.synchronize()
        tc = time.perf_counter()
        for _ in range(5):
            sc.intersect_device(d_rays.data_ptr(), n_coh, hits.data_ptr(), 0, 0, stream.cuda_stream)
        torch.cuda.synchronize()
        coherent_mrays = round(5 * n_coh / (time.perf_counter() - tc) / 1e6, 2)

    for _ in range(args.warmup):
        sc.intersect_device(inc_ptr, n_inc, hits.data_ptr(), 0, 0, stream.cuda_stream)
    torch.cuda.synchronize()
    ev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in range(args.steps)]
    t_start = time.perf_counter()
    for s, e in ev:
        s.record(stream)
        sc.intersect_device(inc_ptr, n_inc, hits.data_ptr(), 0, 0, stream.cuda_stream)
        e.record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t_start
    launch_ms = float(np.mean([s.elapsed_time(e) for s, e in ev]))
    ms_per_step = wall * 1e3 / args.steps
    mrays = n_inc / (ms_per_step / 1e3) / 1e6
    achieved_gbs = mean_b * n_inc / (launch_ms / 1e3) / 1e9

    # PCIe-inclusive rate, not the metric: the host-buffer entry point on the
    # same rays (24 B of ray in, 24 B of hit out, between pageable numpy arrays,
    # the hits array reused between calls)
    host_rays = np.ascontiguousarray(rays[n_coh:])
    host_hits = np.zeros(n_inc, HIT_DTYPE)
    sc.intersect(host_rays, out=(host_hits, None))
    th = time.perf_counter()
    for _ in range(5):
        sc.intersect(host_rays, out=(host_hits, None))
    host_mrays = round(5 * n_inc / (time.perf_counter() - th) / 1e6, 2)
    del host_rays, host_hits

    traffic = None
    tf = ROOT / "profiles" / "closest_hit_traffic.json"
    if tf.exists():
        try:
            traffic = json.loads(tf.read_text()).get("hbm_bytes_per_launch")
        except Exception:
            traffic = None

    cpu = None
    if not args.skip_cpu:
        threads = host_threads(args)
        sample = rays[n_coh: n_coh + (1 << args.cpu_sample_log2)]
        print(f"[bench] cpu baseline on {len(sample)} rays, {threads} threads", file=sys.stderr, flush=True)
        cpu = cpu_baseline(meshes, sample, threads, args.cpu_repeat)

    out = {
        "metric": "incoherent closest-hit Mrays/s @1.3M tris",
        "value": round(mrays, 2),
        "unit": "Mrays/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded displaced icosphere, seeded rays)",
        "config": {"workload": "closest-hit microbench: icosphere subdiv 8 (1,310,720 tris), 2^21 incoherent "
                               "rays per step (of 2^22: half coherent 2048x1024, half incoherent)",
                   "triangles": meshes[0].num_faces, "rays_per_step": n_inc, "parallelism": "single GPU"},
        "roofline": {"bound": "hbm", "achieved": round(achieved_gbs, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved_gbs / HBM_PEAK_GBS, 4), "traffic": traffic,
                     "bytes_per_ray": round(mean_b, 1), "kernel": "k_closest_hit",
                     "launch_ms": round(launch_ms, 4)},
        "coherent_mrays_per_s": coherent_mrays,
        "host_buffers_mrays_per_s": host_mrays,
        "tuning": dev.get_tuning(),
        "cpu_baseline": cpu,
    }
    sc.close()
    del d_rays, hits, cnt
    torch.cuda.empty_cache()
    if not args.skip_render:
        out["render"] = render_section(dev, args)
        if not args.render_batched_only:
            out["render_cpu_runnable"] = plain_cpu_render_section(dev, args, host_threads(args))
    if not args.skip_serving:
        out["serving"] = serving_section(dev, args, args.serving_requests, args.serving_rate)
        out["serving_loaded"] = serving_section(dev, args, args.serving_loaded_requests, args.serving_loaded_rate)
        out["serving_burst"] = serving_section(dev, args, args.serving_loaded_requests, 1.0, burst=True)
    print(json.dumps(out), flush=True)
    if args.calibrate_copy_mb:
        src = torch.empty(args.calibrate_copy_mb << 20, dtype=torch.uint8, device="cuda")
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        print(f"[bench] calibration copy of {args.calibrate_copy_mb} MiB done", file=sys.stderr, flush=True)
    dev.close()


if __name__ == "__main__":
    main()
