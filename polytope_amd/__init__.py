"""polytope_amd — MI355X-native path-tracing hot path for polytope scenes.

Python face of libptx_hip.so (C ABI in include/ptx.h).  The host-side objects
mirror the reference's scene API: a :class:`Scene` is built from meshes
(``mesh_geometry`` + ``Mesh``: packed vertices, faces, object-to-world
transform, Lambert (constant or textured), mirror or glossy material,
optional emission) and an optional
constant skybox; its BVH is the reference's binned-SAH tree
(src/cpu/acceleration/bvh.cpp), built on the host.  Queries and renders run
on the GPU only: there is no CPU fallback, and a missing library or device
raises.
"""
from __future__ import annotations

import ctypes as C
from typing import Iterable, List, Optional, Sequence, Tuple

import numpy as np

from . import _abi
from ._abi import (BVH_NODE_DTYPE, COUNTER_DTYPE, FRAME_DTYPE, HIT_DTYPE, PtxError, check, load)
from .scenes import Camera, MeshData, Texture
from .pbrt import PbrtError, SceneFile, read_pbrt, read_pbrt_string

__all__ = ["Device", "Scene", "PtxError", "render", "render_batch", "Server", "autotune", "HIT_DTYPE", "COUNTER_DTYPE", "FRAME_DTYPE", "BVH_NODE_DTYPE",
           "mesh_descs", "texture_descs", "build_bvh", "camera_desc", "job_desc", "MeshData", "Camera", "Texture"]


def _ptr(a: Optional[np.ndarray], ctype):
    if a is None:
        return C.cast(None, C.POINTER(ctype))
    return a.ctypes.data_as(C.POINTER(ctype))


def texture_table(meshes: Sequence[MeshData]) -> List[Texture]:
    """The distinct textures of ``meshes`` in order of first use: the scene's
    texture table: ptx_mesh_desc.texture = k selects entry k - 1."""
    out: List[Texture] = []
    seen = {}
    for m in meshes:
        t = getattr(m, "texture", None)
        if t is not None and id(t) not in seen:
            seen[id(t)] = len(out)
            out.append(t)
    return out


def texture_descs(meshes: Sequence[MeshData]) -> Tuple[C.Array, list]:
    """The ptx_texture_desc array of ``texture_table(meshes)`` with the buffers it points into."""
    table = texture_table(meshes)
    arr = (_abi.TextureDesc * max(1, len(table)))()
    keep = []
    for i, t in enumerate(table):
        a = np.ascontiguousarray(t.rgba, np.uint8)
        keep.append(a)
        arr[i].width, arr[i].height = a.shape[1], a.shape[0]
        arr[i].rgba = a.ctypes.data_as(C.POINTER(C.c_uint8))
    return arr, keep


def mesh_descs(meshes: Sequence[MeshData]) -> Tuple[C.Array, list]:
    """Builds the ptx_mesh_desc array; returns it with the numpy buffers it points
    into.  A textured mesh's ``texture`` field indexes ``texture_table(meshes)``."""
    keep = []
    arr = (_abi.MeshDesc * len(meshes))()
    tex_index = {id(t): i for i, t in enumerate(texture_table(meshes))}
    for i, m in enumerate(meshes):
        v = np.ascontiguousarray(m.verts, np.float32)
        f = np.ascontiguousarray(m.faces, np.uint32)
        xs = [np.ascontiguousarray(v[:, k]) for k in range(3)]
        fs = [np.ascontiguousarray(f[:, k]) for k in range(3)]
        ns = [None, None, None]
        if m.normals is not None:
            nn = np.ascontiguousarray(m.normals, np.float32)
            ns = [np.ascontiguousarray(nn[:, k]) for k in range(3)]
        keep += xs + fs + [n for n in ns if n is not None]
        d = arr[i]
        d.num_vertices = v.shape[0]
        d.num_faces = f.shape[0]
        d.x, d.y, d.z = (_ptr(x, C.c_float) for x in xs)
        d.nx, d.ny, d.nz = (_ptr(x, C.c_float) for x in ns)
        d.fv0, d.fv1, d.fv2 = (_ptr(x, C.c_uint32) for x in fs)
        d.object_to_world[:] = [float(x) for x in np.asarray(m.object_to_world, np.float32).ravel()]
        d.world_to_object[:] = [float(x) for x in m.world_to_object().ravel()]
        d.brdf = int(m.brdf)
        d.refl[:] = [float(x) for x in m.refl]
        d.is_light = 1 if m.is_light else 0
        d.emission[:] = [float(x) for x in m.emission]
        d.specular[:] = [float(x) for x in m.specular]
        d.roughness = float(m.roughness)
        if getattr(m, "texture", None) is not None:
            d.texture = tex_index[id(m.texture)] + 1           # 0: untextured
        if getattr(m, "uv", None) is not None:
            uv = np.ascontiguousarray(m.uv, np.float32).reshape(-1, 2)
            us, vs = np.ascontiguousarray(uv[:, 0]), np.ascontiguousarray(uv[:, 1])
            keep += [us, vs]
            d.u, d.v = _ptr(us, C.c_float), _ptr(vs, C.c_float)
            d.num_uv = uv.shape[0]
    return arr, keep


def camera_desc(cam: Camera) -> _abi.CameraDesc:
    c = _abi.CameraDesc()
    c.camera_to_world[:] = [float(x) for x in np.asarray(cam.camera_to_world, np.float32).ravel()]
    c.fov = cam.fov
    c.width = cam.width
    c.height = cam.height
    c.one_over_width, c.one_over_height, c.aspect_ratio, c.tan_fov_over_2 = cam.derived()
    return c


def job_desc(cam: Camera, spp: int, max_depth: int, seed: int, nee: bool = False) -> _abi.JobDesc:
    """nee: next-event estimation to the scene's point lights (PTX_JOB_NEE)."""
    j = _abi.JobDesc()
    j.flags = _abi.JOB_NEE if nee else 0
    j.camera = camera_desc(cam)
    j.spp = spp
    j.max_depth = max_depth
    j.seed = seed
    return j


def build_bvh(meshes: Sequence[MeshData], threads: int = 0) -> Tuple[np.ndarray, np.ndarray]:
    """The reference BVH of ``meshes`` (compact nodes, (P, 2) leaf pairs), built on
    the host by the library's builder (ptx_bvh_build); needs no GPU."""
    lib = load()
    arr, keep = mesh_descs(meshes)
    faces = sum(int(m.faces.shape[0]) for m in meshes)
    nodes = np.zeros(max(1, 2 * faces - 1), BVH_NODE_DTYPE)      # a binary tree over F faces has <= 2F-1 nodes
    pairs = np.zeros((max(1, faces), 2), np.uint32)
    nn, np_ = C.c_uint32(nodes.shape[0]), C.c_uint32(pairs.shape[0])
    marr = C.cast(arr, C.POINTER(_abi.MeshDesc))
    check(lib.ptx_bvh_build(marr, len(meshes), threads, nodes.ctypes.data, C.byref(nn), pairs.ctypes.data,
                            C.byref(np_)))
    nodes, pairs = nodes[:nn.value].copy(), pairs[:np_.value].copy()
    return nodes, pairs


class Device:
    """An opened gfx950 device (ptx_device_open)."""

    def __init__(self, ordinal: int = 0):
        lib = load()
        h = C.c_void_p()
        check(lib.ptx_device_open(ordinal, C.byref(h)))
        self.handle = h
        self.ordinal = ordinal

    def close(self) -> None:
        if self.handle:
            load().ptx_device_close(self.handle)
            self.handle = None

    def set_tuning(self, paths_per_wave: int = 0, block_size: int = 0, lds_stack: int = 0, refill_lanes: int = 0,
                   blocks_per_cu: int = 0, render_mode: int = 0, shade_lanes: int = 0, leaf_lanes: int = 0,
                   plain_primary: int = 0, ray_windows: int = 0) -> None:
        """render_mode: 0 auto, 1 path regeneration, 2 wavefront queues;
        shade_lanes: lanes a regeneration wave gathers before it shades;
        leaf_lanes: lanes a traversal wave gathers at leaves before a leaf step;
        plain_primary: render(), 0/1 primary rays traced once per pixel, 2 traced again every sample;
        ray_windows: closest-hit queries, 0/2 one node record per step, 1 window records (node + first-visited
        child in one 128-byte block) once the ray queue is drained; same results, slower on MI355X."""
        t = _abi.Tuning(paths_per_wave, block_size, lds_stack, refill_lanes, blocks_per_cu, render_mode, shade_lanes,
                        leaf_lanes, plain_primary, ray_windows)
        check(load().ptx_set_tuning(self.handle, C.byref(t)))

    def get_tuning(self) -> dict:
        t = _abi.Tuning()
        check(load().ptx_get_tuning(self.handle, C.byref(t)))
        return {k: getattr(t, k) for k, _ in _abi.Tuning._fields_}

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Scene:
    """A scene resident on the device (ptx_scene_create).

    ``bvh``: optional prebuilt reference BVH as (nodes structured array of
    BVH_NODE_DTYPE, pairs uint32 array of shape (P, 2)); by default the library
    builds it on the host."""

    def __init__(self, device: Device, meshes: Sequence[MeshData], skybox: Optional[Sequence[float]] = None,
                 bvh: Optional[Tuple[np.ndarray, np.ndarray]] = None, build_threads: int = 0,
                 lights: Optional[Sequence[Tuple[Sequence[float], Sequence[float]]]] = None):
        """lights: point lights [(position xyz, intensity rgb), ...] for jobs
        rendered with next-event estimation."""
        lib = load()
        self.device = device
        self.meshes = list(meshes)
        arr, keep = mesh_descs(self.meshes)
        tex, tex_keep = texture_descs(self.meshes)
        keep += tex_keep
        desc = _abi.SceneDesc()
        desc.meshes = C.cast(arr, C.POINTER(_abi.MeshDesc))
        desc.num_meshes = len(self.meshes)
        if tex_keep:
            desc.textures = C.cast(tex, C.POINTER(_abi.TextureDesc))
            desc.num_textures = len(tex_keep)
        if bvh is not None:
            nodes = np.ascontiguousarray(bvh[0]).view(BVH_NODE_DTYPE)
            pairs = np.ascontiguousarray(bvh[1], np.uint32).reshape(-1)
            keep += [nodes, pairs]
            desc.nodes = C.cast(nodes.ctypes.data, C.POINTER(_abi.BvhNode))
            desc.num_nodes = nodes.shape[0]
            desc.leaf_pairs = pairs.ctypes.data_as(C.POINTER(C.c_uint32))
            desc.num_pairs = pairs.shape[0] // 2
        desc.has_skybox = 1 if skybox is not None else 0
        if skybox is not None:
            desc.skybox[:] = [float(x) for x in skybox]
        desc.build_threads = build_threads
        if lights:
            flat = np.ascontiguousarray([[*p, *i] for p, i in lights], np.float32).reshape(-1)
            keep.append(flat)
            desc.lights = flat.ctypes.data_as(C.POINTER(C.c_float))
            desc.num_lights = len(lights)
        h = C.c_void_p()
        check(lib.ptx_scene_create(device.handle, C.byref(desc), C.byref(h)))
        self.handle = h
        self.skybox = skybox

    def info(self) -> dict:
        i = _abi.SceneInfo()
        check(load().ptx_scene_get_info(self.handle, C.byref(i)))
        return {k: getattr(i, k) for k, _ in _abi.SceneInfo._fields_}

    def bvh(self) -> Tuple[np.ndarray, np.ndarray]:
        """The reference-layout BVH (compact nodes, (P, 2) leaf pairs) the scene uses."""
        inf = self.info()
        nodes = np.zeros(inf["num_nodes"], BVH_NODE_DTYPE)
        pairs = np.zeros((inf["num_pairs"], 2), np.uint32)
        check(load().ptx_scene_copy_bvh(self.handle, nodes.ctypes.data, pairs.ctypes.data))
        return nodes, pairs

    def intersect(self, rays: np.ndarray, counters: bool = False, out: Optional[tuple] = None):
        """Closest hit of host rays (n, 6) float32 -> structured hits (and counters).

        ``out``: ``(hits, counters)`` arrays to write into instead of new ones
        (``counters`` None when not asked for), e.g. reused between calls so
        that their pages are already mapped."""
        rays = np.ascontiguousarray(rays, np.float32)
        n = rays.shape[0]
        if out is None:
            hits = np.zeros(n, HIT_DTYPE)
            cnt = np.zeros(n, COUNTER_DTYPE) if counters else None
        else:
            hits, cnt = out
            for a, dt in ((hits, HIT_DTYPE), (cnt, COUNTER_DTYPE)):
                if a is None:
                    continue
                if a.dtype != dt or a.shape != (n,) or not a.flags.c_contiguous or not a.flags.writeable:
                    raise ValueError(f"out arrays must be writable contiguous ({n},) arrays of {dt}")
            if hits is None or (cnt is None) == counters:
                raise ValueError("out must hold a hits array, and a counters array exactly when counters=True")
        check(load().ptx_intersect_closest(self.handle, rays.ctypes.data, n, hits.ctypes.data,
                                           cnt.ctypes.data if cnt is not None else None))
        return (hits, cnt) if counters else hits

    def intersect_device(self, rays_ptr: int, n: int, hits_ptr: int, counters_ptr: int = 0,
                         frames_ptr: int = 0, stream: int = 0) -> None:
        """Closest hit on device buffers (raw device pointers, e.g. torch ``data_ptr()``)."""
        check(load().ptx_intersect_closest_device(self.handle, rays_ptr or None, n, hits_ptr or None,
                                                  counters_ptr or None, frames_ptr or None, stream or None))

    def close(self) -> None:
        if getattr(self, "handle", None):
            load().ptx_scene_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def render(scene: Scene, cam: Camera, spp: int, max_depth: int, seed: int, nee: bool = False) -> np.ndarray:
    """Plain GPU render of one job (ptx_render): one launch per sample, one thread
    per pixel path.  Returns the (H, W, 3) BoxFilter film (0-255 units)."""
    film = np.zeros((cam.height, cam.width, 3), np.float32)
    job = job_desc(cam, spp, max_depth, seed, nee)
    check(load().ptx_render(scene.handle, C.byref(job), film.ctypes.data))
    return film


def render_batch(jobs: Sequence[tuple]) -> List[np.ndarray]:
    """Batched render (ptx_render_batch) of (scene, camera, spp, max_depth, seed[,
    nee]) jobs in one call; returns one film per job."""
    n = len(jobs)
    handles = (C.c_void_p * n)(*[j[0].handle.value for j in jobs])
    descs = (_abi.JobDesc * n)(*[job_desc(j[1], j[2], j[3], j[4], len(j) > 5 and bool(j[5])) for j in jobs])
    films = [np.zeros((j[1].height, j[1].width, 3), np.float32) for j in jobs]
    ptrs = (C.c_void_p * n)(*[f.ctypes.data for f in films])
    check(load().ptx_render_batch(handles, descs, n, ptrs))
    return films


def autotune(device: Device, scene: Scene) -> dict:
    """Times candidate tunings on ``scene`` and keeps the fastest on ``device``:
    traversal LDS stack depth x refill threshold (ray queries), then round form x
    path slots per round, shading and leaf gathering thresholds (rendering)."""
    t = _abi.Tuning()
    check(load().ptx_autotune(device.handle, scene.handle, C.byref(t)))
    return {k: getattr(t, k) for k, _ in _abi.Tuning._fields_}


def _status_dict(st: _abi.JobStatus) -> dict:
    return {k: getattr(st, k) for k, _ in _abi.JobStatus._fields_}


def film_to_rgba8(film: np.ndarray) -> np.ndarray:
    """(H, W, 3) float film -> (H, W, 4) uint8 as PNGFilm writes it (clamp at 255, truncate)."""
    film = np.ascontiguousarray(film, np.float32)
    h, w = film.shape[:2]
    out = np.zeros((h, w, 4), np.uint8)
    check(load().ptx_film_to_rgba8(film.ctypes.data, w, h, out.ctypes.data))
    return out


def write_png(film: np.ndarray, path: str) -> None:
    film = np.ascontiguousarray(film, np.float32)
    check(load().ptx_film_write_png(film.ctypes.data, film.shape[1], film.shape[0], str(path).encode()))


def write_pfm(film: np.ndarray, path: str) -> None:
    film = np.ascontiguousarray(film, np.float32)
    check(load().ptx_film_write_pfm(film.ctypes.data, film.shape[1], film.shape[0], str(path).encode()))


class Server:
    """Continuous-batching render service (ptx_server_*): jobs are admitted as
    HBM budget and job slots free up and share the GPU round by round."""

    def __init__(self, device: Device, hbm_budget_bytes: int, max_jobs_in_flight: int = 0,
                 paths_per_wave: int = 0):
        cfg = _abi.ServerConfig(hbm_budget_bytes, max_jobs_in_flight, paths_per_wave)
        h = C.c_void_p()
        check(load().ptx_server_create(device.handle, C.byref(cfg), C.byref(h)))
        self.handle = h
        self.device = device
        self._dims = {}
        self._scenes = {}       # job id -> Scene, kept until the job is waited for

    def submit(self, scene: Scene, cam: Camera, spp: int, max_depth: int, seed: int, nee: bool = False) -> int:
        job = job_desc(cam, spp, max_depth, seed, nee)
        jid = C.c_uint64()
        check(load().ptx_server_submit(self.handle, scene.handle, C.byref(job), C.byref(jid)))
        self._dims[jid.value] = (cam.height, cam.width)
        self._scenes[jid.value] = scene
        return jid.value

    def status(self, job_id: int) -> dict:
        st = _abi.JobStatus()
        check(load().ptx_server_status(self.handle, job_id, C.byref(st)))
        return _status_dict(st)

    def wait(self, job_id: int):
        """Blocks until the job ends; returns (film, status).  Raises PtxError for
        failed or cancelled jobs (status attached as .status)."""
        h, w = self._dims.pop(job_id)
        film = np.zeros((h, w, 3), np.float32)
        st = _abi.JobStatus()
        rc = load().ptx_server_wait(self.handle, job_id, film.ctypes.data, C.byref(st))
        self._scenes.pop(job_id, None)
        if rc != _abi.PTX_OK:
            err = PtxError(rc, load().ptx_last_error().decode(errors="replace"))
            err.status = _status_dict(st)
            raise err
        return film, _status_dict(st)

    def cancel(self, job_id: int) -> None:
        check(load().ptx_server_cancel(self.handle, job_id))

    def close(self) -> None:
        if getattr(self, "handle", None):
            load().ptx_server_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
