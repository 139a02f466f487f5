"""ctypes mirror of include/ptx.h and the loader for libptx_hip.so.

The library is loaded from this directory only (built in-tree by
polytope_amd.build); a missing library raises instead of falling back to any
other implementation.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

import numpy as np

LIB_PATH = Path(__file__).resolve().parent / "libptx_hip.so"

PTX_OK = 0
PTX_ERR_INVALID = 1
PTX_ERR_HIP = 2
PTX_ERR_OOM = 3
PTX_ERR_BUDGET = 4
PTX_ERR_NOT_FOUND = 5
PTX_ERR_CANCELLED = 6
PTX_ERR_NO_DEVICE = 7

JOB_QUEUED, JOB_RUNNING, JOB_DONE, JOB_FAILED, JOB_CANCELLED = range(5)

_fp = C.POINTER(C.c_float)
_up = C.POINTER(C.c_uint32)


class MeshDesc(C.Structure):
    _fields_ = [("num_vertices", C.c_uint32), ("num_faces", C.c_uint32),
                ("x", _fp), ("y", _fp), ("z", _fp),
                ("nx", _fp), ("ny", _fp), ("nz", _fp),
                ("fv0", _up), ("fv1", _up), ("fv2", _up),
                ("object_to_world", C.c_float * 16), ("world_to_object", C.c_float * 16),
                ("brdf", C.c_int32), ("refl", C.c_float * 3),
                ("is_light", C.c_int32), ("emission", C.c_float * 3),
                ("specular", C.c_float * 3), ("roughness", C.c_float),
                ("u", _fp), ("v", _fp), ("num_uv", C.c_uint32), ("texture", C.c_uint32)]


class TextureDesc(C.Structure):
    _fields_ = [("width", C.c_uint32), ("height", C.c_uint32), ("rgba", C.POINTER(C.c_uint8))]


class BvhNode(C.Structure):
    _fields_ = [("p0", C.c_float * 3), ("p1", C.c_float * 3), ("offset", C.c_uint32),
                ("num_faces", C.c_uint16), ("flags", C.c_uint16)]


BVH_NODE_DTYPE = np.dtype([("p0", "<f4", 3), ("p1", "<f4", 3), ("offset", "<u4"),
                           ("num_faces", "<u2"), ("flags", "<u2")])


class SceneDesc(C.Structure):
    _fields_ = [("meshes", C.POINTER(MeshDesc)), ("num_meshes", C.c_uint32),
                ("nodes", C.POINTER(BvhNode)), ("num_nodes", C.c_uint32),
                ("leaf_pairs", _up), ("num_pairs", C.c_uint32),
                ("has_skybox", C.c_int32), ("skybox", C.c_float * 3),
                ("build_threads", C.c_uint32), ("lights", C.POINTER(C.c_float)), ("num_lights", C.c_uint32),
                ("textures", C.POINTER(TextureDesc)), ("num_textures", C.c_uint32)]


class SceneInfo(C.Structure):
    _fields_ = [("num_nodes", C.c_uint32), ("num_pairs", C.c_uint32), ("num_interior", C.c_uint32),
                ("num_leaves", C.c_uint32), ("max_depth", C.c_uint32), ("device_bytes", C.c_uint64),
                ("build_seconds", C.c_double), ("upload_seconds", C.c_double)]


class CameraDesc(C.Structure):
    _fields_ = [("camera_to_world", C.c_float * 16), ("fov", C.c_float),
                ("width", C.c_uint32), ("height", C.c_uint32),
                ("one_over_width", C.c_float), ("one_over_height", C.c_float),
                ("aspect_ratio", C.c_float), ("tan_fov_over_2", C.c_float)]


class JobDesc(C.Structure):
    _fields_ = [("camera", CameraDesc), ("spp", C.c_uint32), ("max_depth", C.c_uint32), ("seed", C.c_uint32),
                ("flags", C.c_uint32)]


JOB_NEE = 1


class Tuning(C.Structure):
    _fields_ = [("paths_per_wave", C.c_uint32), ("block_size", C.c_uint32), ("lds_stack", C.c_uint32),
                ("refill_lanes", C.c_uint32), ("blocks_per_cu", C.c_uint32), ("render_mode", C.c_uint32),
                ("shade_lanes", C.c_uint32), ("leaf_lanes", C.c_uint32), ("plain_primary", C.c_uint32),
                ("ray_windows", C.c_uint32)]


RENDER_AUTO, RENDER_REGEN, RENDER_WAVEFRONT = 0, 1, 2
PLAIN_DEFAULT, PLAIN_CACHED, PLAIN_RETRACE = 0, 1, 2
WINDOWS_DEFAULT, WINDOWS_ON, WINDOWS_OFF = 0, 1, 2


class ServerConfig(C.Structure):
    _fields_ = [("hbm_budget_bytes", C.c_uint64), ("max_jobs_in_flight", C.c_uint32),
                ("paths_per_wave", C.c_uint32)]


class JobStatus(C.Structure):
    _fields_ = [("state", C.c_int32), ("error", C.c_int32), ("submit_time", C.c_double),
                ("start_time", C.c_double), ("end_time", C.c_double),
                ("samples_done", C.c_uint64), ("samples_total", C.c_uint64)]


HIT_DTYPE = np.dtype([("mesh", "<u4"), ("face", "<u4"), ("t", "<f4"), ("u", "<f4"), ("v", "<f4"), ("w", "<f4")])
COUNTER_DTYPE = np.dtype([("bb_hits", "<u4"), ("bb_misses", "<u4"), ("tri_isects", "<u4")])
FRAME_DTYPE = np.dtype([("location", "<f4", 3), ("normal", "<f4", 3), ("tangent1", "<f4", 3),
                        ("tangent2", "<f4", 3), ("error", "<f4", 3)])

# name -> (restype, argtypes)
SIGNATURES = {
    "ptx_last_error": (C.c_char_p, []),
    "ptx_version": (C.c_char_p, []),
    "ptx_device_open": (C.c_int, [C.c_int, C.POINTER(C.c_void_p)]),
    "ptx_device_close": (None, [C.c_void_p]),
    "ptx_scene_create": (C.c_int, [C.c_void_p, C.POINTER(SceneDesc), C.POINTER(C.c_void_p)]),
    "ptx_scene_destroy": (None, [C.c_void_p]),
    "ptx_scene_get_info": (C.c_int, [C.c_void_p, C.POINTER(SceneInfo)]),
    "ptx_scene_copy_bvh": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p]),
    "ptx_bvh_build": (C.c_int, [C.POINTER(MeshDesc), C.c_uint32, C.c_uint32, C.c_void_p, C.POINTER(C.c_uint32),
                                C.c_void_p, C.POINTER(C.c_uint32)]),
    "ptx_intersect_closest": (C.c_int, [C.c_void_p, C.c_void_p, C.c_uint32, C.c_void_p, C.c_void_p]),
    "ptx_intersect_closest_device": (C.c_int, [C.c_void_p, C.c_void_p, C.c_uint32, C.c_void_p, C.c_void_p,
                                               C.c_void_p, C.c_void_p]),
    "ptx_intersect_closest_device_noalloc": (C.c_int, [C.c_void_p, C.c_void_p, C.c_uint32, C.c_void_p, C.c_void_p,
                                                       C.c_void_p, C.c_void_p]),
    "ptx_render": (C.c_int, [C.c_void_p, C.POINTER(JobDesc), C.c_void_p]),
    "ptx_render_batch": (C.c_int, [C.POINTER(C.c_void_p), C.POINTER(JobDesc), C.c_uint32,
                                   C.POINTER(C.c_void_p)]),
    "ptx_set_tuning": (C.c_int, [C.c_void_p, C.POINTER(Tuning)]),
    "ptx_get_tuning": (C.c_int, [C.c_void_p, C.POINTER(Tuning)]),
    "ptx_autotune": (C.c_int, [C.c_void_p, C.c_void_p, C.POINTER(Tuning)]),
    "ptx_film_to_rgba8": (C.c_int, [C.c_void_p, C.c_uint32, C.c_uint32, C.c_void_p]),
    "ptx_film_write_png": (C.c_int, [C.c_void_p, C.c_uint32, C.c_uint32, C.c_char_p]),
    "ptx_film_write_pfm": (C.c_int, [C.c_void_p, C.c_uint32, C.c_uint32, C.c_char_p]),
    "ptx_server_create": (C.c_int, [C.c_void_p, C.POINTER(ServerConfig), C.POINTER(C.c_void_p)]),
    "ptx_server_destroy": (None, [C.c_void_p]),
    "ptx_server_submit": (C.c_int, [C.c_void_p, C.c_void_p, C.POINTER(JobDesc), C.POINTER(C.c_uint64)]),
    "ptx_server_status": (C.c_int, [C.c_void_p, C.c_uint64, C.POINTER(JobStatus)]),
    "ptx_server_wait": (C.c_int, [C.c_void_p, C.c_uint64, C.c_void_p, C.POINTER(JobStatus)]),
    "ptx_server_cancel": (C.c_int, [C.c_void_p, C.c_uint64]),
}

_lib = None


def load() -> C.CDLL:
    """Loads libptx_hip.so from the package directory (raises if it is not built)."""
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise RuntimeError(f"{LIB_PATH} is not built; run polytope_amd.build.build() first")
        lib = C.CDLL(str(LIB_PATH))
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(lib, name)
            fn.restype = res
            fn.argtypes = args
        _lib = lib
    return _lib


class PtxError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"ptx error {code}: {msg}")
        self.code = code


def check(rc: int) -> None:
    if rc != PTX_OK:
        raise PtxError(rc, load().ptx_last_error().decode(errors="replace"))
