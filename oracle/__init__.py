"""Oracle: CPU restatement of polytope's hot path (oracle/oracle.cpp) and the
reference build driver (oracle/ref) — TEST INFRASTRUCTURE ONLY.

Only tests/, __graft_entry__.smoke() and the cpu_baseline leg of bench.py may
import this package, and only as the checker.  The product (polytope_amd) never
imports it.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess
from pathlib import Path
from typing import Optional, Sequence, Tuple

import numpy as np

HERE = Path(__file__).resolve().parent
LIB = HERE / "_build" / "liboracle.so"
REF_DRIVER = HERE / "_ref" / "ref_driver"
# the reference's runner API driving the GPU library through poly::GpuRunner
GPU_RUNNER_CHECK = HERE / "_ref" / "gpu_runner_check"
REFERENCE_ROOT = Path("/root/reference")

HIT_RECORD = np.dtype([("hit", "<u4"), ("mesh", "<u4"), ("face", "<u4"), ("t", "<f4"), ("u", "<f4"),
                       ("v", "<f4"), ("w", "<f4"), ("location", "<f4", 3), ("normal", "<f4", 3),
                       ("tangent1", "<f4", 3), ("tangent2", "<f4", 3), ("error", "<f4", 3),
                       ("bb_hits", "<u4"), ("bb_misses", "<u4"), ("tri_isects", "<u4")])
assert HIT_RECORD.itemsize == 100

NODE_DTYPE = np.dtype([("p0", "<f4", 3), ("p1", "<f4", 3), ("offset", "<u4"), ("num_faces", "<u2"),
                       ("flags", "<u2")])


def build_oracle() -> Path:
    subprocess.run(["make", "-s", "-C", str(HERE)], check=True)
    return LIB


def build_reference() -> Optional[Path]:
    """Compiles the reference CPU path from /root/reference into oracle/_ref (only
    where the reference tree exists)."""
    if not (REFERENCE_ROOT / "src" / "cpu" / "acceleration" / "bvh.cpp").exists():
        return None
    subprocess.run(["make", "-s", "-j8", "-C", str(HERE / "ref"), f"REF={REFERENCE_ROOT}"], check=True)
    return REF_DRIVER


_lib = None


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        if not LIB.exists():
            build_oracle()
        L = C.CDLL(str(LIB))
        L.orc_scene_create.restype = C.c_void_p
        L.orc_scene_create.argtypes = [C.c_void_p, C.c_uint32, C.c_int, C.c_void_p]
        L.orc_scene_set_lights.restype = None
        L.orc_scene_set_lights.argtypes = [C.c_void_p, C.c_void_p, C.c_uint32]
        L.orc_scene_set_textures.restype = None
        L.orc_scene_set_textures.argtypes = [C.c_void_p, C.c_void_p, C.c_uint32]
        L.orc_tex_lerp.restype = None
        L.orc_tex_lerp.argtypes = [C.c_void_p, C.c_float, C.c_float, C.c_float, C.c_void_p]
        L.orc_texel_index.restype = C.c_uint32
        L.orc_texel_index.argtypes = [C.c_float, C.c_float, C.c_uint32, C.c_uint32]
        L.orc_scene_create_prebuilt.restype = C.c_void_p
        L.orc_scene_create_prebuilt.argtypes = [C.c_void_p, C.c_uint32, C.c_int, C.c_void_p, C.c_void_p, C.c_uint32,
                                                C.c_void_p, C.c_uint32]
        L.orc_scene_destroy.argtypes = [C.c_void_p]
        L.orc_num_nodes.restype = C.c_uint32
        L.orc_num_nodes.argtypes = [C.c_void_p]
        L.orc_num_pairs.restype = C.c_uint32
        L.orc_num_pairs.argtypes = [C.c_void_p]
        L.orc_copy_bvh.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p]
        L.orc_intersect.argtypes = [C.c_void_p, C.c_void_p, C.c_uint32, C.c_void_p, C.c_int]
        L.orc_intersect_ids.restype = C.c_uint64
        L.orc_intersect_ids.argtypes = [C.c_void_p, C.c_void_p, C.c_uint32, C.c_void_p, C.c_int]
        L.orc_render.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_int, C.c_uint32]
        L.orc_render_counted.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_int, C.c_uint32,
                                         C.c_void_p]
        L.orc_philox4x32_10.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p]
        L.orc_sincos.argtypes = [C.c_float, C.POINTER(C.c_float), C.POINTER(C.c_float)]
        L.orc_acos.restype = C.c_float
        L.orc_acos.argtypes = [C.c_float]
        L.orc_glossy_sample.restype = C.c_int
        L.orc_glossy_sample.argtypes = [C.c_void_p, C.c_float, C.c_void_p, C.c_uint32, C.c_int, C.c_void_p,
                                        C.POINTER(C.c_uint32)]
        L.orc_box_hits.argtypes = [C.c_void_p, C.c_void_p, C.c_float, C.POINTER(C.c_int)]
        L.orc_camera_ray.argtypes = [C.c_void_p, C.c_float, C.c_float, C.c_void_p]
        _lib = L
    return _lib


class OrcJob(C.Structure):
    _fields_ = [("width", C.c_uint32), ("height", C.c_uint32), ("spp", C.c_uint32),
                ("max_depth", C.c_uint32), ("seed", C.c_uint32), ("flags", C.c_uint32)]


class OracleScene:
    """The oracle's scene: world vertices, the restated reference BVH."""

    def __init__(self, meshes, skybox: Optional[Sequence[float]] = None, bvh=None, lights=None):
        """bvh: optional (nodes, pairs) in the reference compact layout to use
        instead of building the restated tree; lights: optional point lights
        [(position xyz, intensity rgb), ...] for next-event estimation."""
        from polytope_amd import mesh_descs, texture_descs
        arr, keep = mesh_descs(meshes)
        tex, tex_keep = texture_descs(meshes)
        sky = (C.c_float * 3)(*(skybox if skybox is not None else (0.0, 0.0, 0.0)))
        if bvh is None:
            self.h = lib().orc_scene_create(C.cast(arr, C.c_void_p), len(meshes), 1 if skybox is not None else 0,
                                            C.cast(sky, C.c_void_p))
        else:
            nodes = np.ascontiguousarray(bvh[0], NODE_DTYPE)
            pairs = np.ascontiguousarray(bvh[1], np.uint32).reshape(-1, 2)
            self.h = lib().orc_scene_create_prebuilt(C.cast(arr, C.c_void_p), len(meshes),
                                                     1 if skybox is not None else 0, C.cast(sky, C.c_void_p),
                                                     nodes.ctypes.data, len(nodes), pairs.ctypes.data, len(pairs))
        self._keep = (arr, keep)
        if tex_keep:
            lib().orc_scene_set_textures(self.h, C.cast(tex, C.c_void_p), len(tex_keep))
        if lights:
            flat = np.ascontiguousarray([[*p, *i] for p, i in lights], np.float32)
            lib().orc_scene_set_lights(self.h, flat.ctypes.data, len(flat))

    def bvh(self) -> Tuple[np.ndarray, np.ndarray]:
        n = lib().orc_num_nodes(self.h)
        p = lib().orc_num_pairs(self.h)
        nodes = np.zeros(n, NODE_DTYPE)
        pairs = np.zeros((p, 2), np.uint32)
        lib().orc_copy_bvh(self.h, nodes.ctypes.data, pairs.ctypes.data)
        return nodes, pairs

    def intersect(self, rays: np.ndarray, threads: int = 0) -> np.ndarray:
        rays = np.ascontiguousarray(rays, np.float32)
        out = np.zeros(rays.shape[0], HIT_RECORD)
        lib().orc_intersect(self.h, rays.ctypes.data, rays.shape[0], out.ctypes.data, threads or os.cpu_count())
        return out

    def intersect_ids(self, rays: np.ndarray, threads: int) -> Tuple[np.ndarray, int]:
        rays = np.ascontiguousarray(rays, np.float32)
        ids = np.zeros(rays.shape[0], np.uint32)
        s = lib().orc_intersect_ids(self.h, rays.ctypes.data, rays.shape[0], ids.ctypes.data, threads)
        return ids, int(s)

    def render(self, cam, spp: int, max_depth: int, seed: int, threads: int = 0, row_step: int = 1,
               nee: bool = False, stats: bool = False):
        """The film (rows y % row_step == 0); with stats, also {rays, boxes,
        tris, samples}: rays traced and the reference's per-ray counters summed
        over them."""
        from polytope_amd import camera_desc
        cd = camera_desc(cam)
        job = OrcJob(cam.width, cam.height, spp, max_depth, seed, 1 if nee else 0)
        film = np.zeros((cam.height, cam.width, 3), np.float32)
        counts = np.zeros(4, np.uint64)
        lib().orc_render_counted(self.h, C.byref(cd), C.byref(job), film.ctypes.data, threads or os.cpu_count(),
                                 row_step, counts.ctypes.data if stats else None)
        if stats:
            return film, dict(zip(("rays", "boxes", "tris", "samples"), (int(x) for x in counts)))
        return film

    def __del__(self):
        try:
            if self.h:
                lib().orc_scene_destroy(self.h)
                self.h = None
        except Exception:
            pass


def philox(ctr, key) -> np.ndarray:
    c = (C.c_uint32 * 4)(*ctr)
    k = (C.c_uint32 * 2)(*key)
    o = (C.c_uint32 * 4)()
    lib().orc_philox4x32_10(C.cast(c, C.c_void_p), C.cast(k, C.c_void_p), C.cast(o, C.c_void_p))
    return np.array(list(o), np.uint32)


def sincos(theta: float) -> Tuple[float, float]:
    s, c = C.c_float(), C.c_float()
    lib().orc_sincos(C.c_float(theta), C.byref(s), C.byref(c))
    return s.value, c.value


def acos(x: float) -> float:
    """The project-owned acos the glossy lobe uses (CPU and GPU alike)."""
    return lib().orc_acos(C.c_float(x))


def glossy_sample(incoming, roughness: float, uniforms, libm: bool):
    """GlossyBRDF::sample on given uniforms (u0, u1 per attempt); libm selects
    the C library's acos/sin/cos over the project-owned ones.  Returns
    (direction or None, attempts drawn)."""
    v = np.ascontiguousarray(incoming, np.float32)
    u = np.ascontiguousarray(uniforms, np.float32)
    out = np.zeros(3, np.float32)
    used = C.c_uint32()
    ok = lib().orc_glossy_sample(v.ctypes.data, C.c_float(roughness), u.ctypes.data, len(u) // 2, 1 if libm else 0,
                                 out.ctypes.data, C.byref(used))
    return (out if ok else None), used.value


def tex_lerp(uv, u: float, v: float, w: float) -> Tuple[float, float]:
    """Texture coordinates of a hit with barycentric weights u, v, w on a face
    whose corners have coordinates uv = (u0, v0, u1, v1, u2, v2)."""
    c = np.ascontiguousarray(uv, np.float32)
    out = np.zeros(2, np.float32)
    lib().orc_tex_lerp(c.ctypes.data, C.c_float(u), C.c_float(v), C.c_float(w), out.ctypes.data)
    return float(out[0]), float(out[1])


def texel_index(tu: float, tv: float, width: int, height: int) -> int:
    """Flat texel index (row * width + column) the texture lookup reads."""
    return int(lib().orc_texel_index(C.c_float(tu), C.c_float(tv), width, height))


def box_hits(box, ray, min_t: float) -> bool:
    b = (C.c_float * 6)(*box)
    r = (C.c_float * 6)(*ray)
    h = C.c_int()
    lib().orc_box_hits(C.cast(b, C.c_void_p), C.cast(r, C.c_void_p), C.c_float(min_t), C.byref(h))
    return bool(h.value)


def reference_available() -> bool:
    return REF_DRIVER.exists()


def reference_scene(rel: str) -> str:
    """Path of one of the reference's own scene files (scenes/<rel>), read in
    place by the scene-file tests where the reference tree is present."""
    return str(REFERENCE_ROOT / "scenes" / rel)


def run_reference(args: Sequence[str], cwd: Optional[str] = None, timeout: int = 600) -> subprocess.CompletedProcess:
    return subprocess.run([str(REF_DRIVER), *args], cwd=cwd, capture_output=True, text=True, timeout=timeout,
                          check=True)
