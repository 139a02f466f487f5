"""Seeded synthetic scenes and ray sets for the closest-hit and render workloads.

Every generator is deterministic from its seed and uses only IEEE basic
operations (add, mul, div, sqrt) and integer hashing, so the same call yields
the same bits on any host: sines and cosines come from ``_sincos64`` below, a
fixed-term Taylor evaluation in float64, not from libm.

Scenes are lists of :class:`MeshData`, the host-side analogue of the
reference's ``mesh_geometry`` + ``Mesh`` pair (src/cpu/shapes/mesh.h:17-113):
packed vertices, three face-index arrays, a 4x4 object-to-world matrix and a
material (Lambert ``refl`` or an image texture with per-vertex texture
coordinates, perfect mirror, or the glossy lobe), optionally an emitter.
"""
from __future__ import annotations

import math
import struct
from dataclasses import dataclass, field
from typing import List, Optional, Sequence, Tuple

import numpy as np

LAMBERT = 1
MIRROR = 2
GLOSSY = 3        # the reference's "plastic" material (GlossyBRDF); refl holds Kd

IDENTITY = np.eye(4, dtype=np.float32)


class Texture:
    """An RGBA8 image texture (poly::texture, src/cpu/shading/texture.h:12-42):
    ``rgba`` is (H, W, 4) uint8, rows top to bottom.  Meshes that share one
    Texture object share one device copy."""

    def __init__(self, rgba: np.ndarray, name: str = ""):
        a = np.ascontiguousarray(rgba, np.uint8)
        if a.ndim != 3 or a.shape[2] != 4 or a.shape[0] < 1 or a.shape[1] < 1:
            raise ValueError("texture must be (H, W, 4) uint8")
        self.rgba = a
        self.name = name

    @property
    def width(self) -> int:
        return int(self.rgba.shape[1])

    @property
    def height(self) -> int:
        return int(self.rgba.shape[0])


@dataclass
class MeshData:
    verts: np.ndarray                      # (V, 3) float32, packed vertices
    faces: np.ndarray                      # (F, 3) uint32
    object_to_world: np.ndarray = field(default_factory=lambda: IDENTITY.copy())
    normals: Optional[np.ndarray] = None   # (V, 3) float32 vertex normals
    brdf: int = LAMBERT
    refl: Tuple[float, float, float] = (0.5, 0.5, 0.5)
    is_light: bool = False
    emission: Tuple[float, float, float] = (0.0, 0.0, 0.0)
    specular: Tuple[float, float, float] = (0.25, 0.25, 0.25)   # glossy: Ks
    roughness: float = 0.1                                        # glossy roughness
    w2o: Optional[np.ndarray] = None       # world-to-object if given (the inverse of object_to_world otherwise)
    texture: Optional[Texture] = None      # Lambert: the reflectance comes from this texture instead of refl
    uv: Optional[np.ndarray] = None        # (N, 2) float32 per-vertex texture coordinates, indexed like verts

    @property
    def num_faces(self) -> int:
        return int(self.faces.shape[0])

    def world_to_object(self) -> np.ndarray:
        if self.w2o is not None:
            return np.asarray(self.w2o, np.float32)
        return np.linalg.inv(self.object_to_world.astype(np.float64)).astype(np.float32)


@dataclass
class Camera:
    """Pinhole camera: camera-to-world matrix plus the values
    ``PerspectiveCamera``'s constructor derives (src/cpu/cameras/PerspectiveCamera.cpp:17-25)."""
    camera_to_world: np.ndarray
    fov: float
    width: int
    height: int

    def derived(self) -> Tuple[float, float, float, float]:
        f32 = np.float32
        one_over_w = f32(1.0) / f32(self.width)
        one_over_h = f32(1.0) / f32(self.height)
        aspect = f32(self.width) * one_over_h
        pi_over_180 = f32(math.pi / 180.0)
        half = f32(self.fov) * pi_over_180 * f32(0.5)
        tan_f2 = f32(math.tan(float(half)))
        return float(one_over_w), float(one_over_h), float(aspect), float(tan_f2)


# --------------------------------------------------------------------------- maths helpers

def _sincos64(x: np.ndarray) -> Tuple[np.ndarray, np.ndarray]:
    """sin and cos of float64 angles with basic operations only: reduce by
    k*pi/2 (k = round(x*2/pi)), then Taylor series to degree 21 on [-pi/4, pi/4]
    (truncation error below 1e-17)."""
    x = np.asarray(x, dtype=np.float64)
    k = np.floor(x * (2.0 / math.pi) + 0.5)
    r = x - k * (math.pi / 2.0)
    r2 = r * r
    s = np.zeros_like(r)
    c = np.zeros_like(r)
    for n in range(21, 0, -2):          # sin: r - r^3/3! + ...
        s = s * r2 + ((-1.0) ** ((n - 1) // 2)) / math.factorial(n)
    s = s * r
    for n in range(20, -1, -2):         # cos: 1 - r^2/2! + ...
        c = c * r2 + ((-1.0) ** (n // 2)) / math.factorial(n)
    q = (k.astype(np.int64)) & 3
    sin = np.where(q == 0, s, np.where(q == 1, c, np.where(q == 2, -s, -c)))
    cos = np.where(q == 0, c, np.where(q == 1, -s, np.where(q == 2, -c, s)))
    return sin, cos


def _hash_u32(*keys: np.ndarray) -> np.ndarray:
    """Integer hash (xorshift-multiply finaliser, as in MurmurHash3 fmix32) of
    several uint32 keys combined by a multiply-xor chain."""
    h = np.full(np.broadcast(*keys).shape, 0x9E3779B9, dtype=np.uint64)
    for k in keys:
        h = (h ^ np.asarray(k, dtype=np.uint64)) * np.uint64(0x85EBCA6B) & np.uint64(0xFFFFFFFF)
        h ^= h >> np.uint64(13)
    h = (h * np.uint64(0xC2B2AE35)) & np.uint64(0xFFFFFFFF)
    h ^= h >> np.uint64(16)
    return h.astype(np.uint32)


class Rng:
    """Counter-based uniform stream: value i of stream s is hash(seed, s, i)."""

    def __init__(self, seed: int):
        self.seed = np.uint32(seed)
        self.stream = 0

    def uniform(self, n: int) -> np.ndarray:
        self.stream += 1
        i = np.arange(n, dtype=np.uint64)
        lo = (i & np.uint64(0xFFFFFFFF)).astype(np.uint32)
        hi = (i >> np.uint64(32)).astype(np.uint32)
        h = _hash_u32(np.uint32(self.seed), np.uint32(self.stream), lo, hi)
        return (h >> np.uint32(8)).astype(np.float64) * (1.0 / (1 << 24))


def value_noise(p: np.ndarray, seed: int, freq: float) -> np.ndarray:
    """Trilinear value noise in [-1, 1] on an integer lattice scaled by ``freq``,
    with smoothstep weights; lattice values from an integer hash."""
    q = p * freq
    i0 = np.floor(q)
    f = q - i0
    w = f * f * (3.0 - 2.0 * f)
    ii = i0.astype(np.int64)
    out = np.zeros(p.shape[0], dtype=np.float64)
    for dx in (0, 1):
        for dy in (0, 1):
            for dz in (0, 1):
                h = _hash_u32(np.uint32(seed),
                              (ii[:, 0] + dx).astype(np.uint32),
                              (ii[:, 1] + dy).astype(np.uint32),
                              (ii[:, 2] + dz).astype(np.uint32))
                val = (h >> np.uint32(8)).astype(np.float64) * (2.0 / (1 << 24)) - 1.0
                wx = w[:, 0] if dx else 1.0 - w[:, 0]
                wy = w[:, 1] if dy else 1.0 - w[:, 1]
                wz = w[:, 2] if dz else 1.0 - w[:, 2]
                out += val * wx * wy * wz
    return out


# --------------------------------------------------------------------------- meshes

def icosphere(subdiv: int, radius: float = 1.0, center=(0.0, 0.0, 0.0), noise: float = 0.05,
              seed: int = 2, noise_freq: float = 3.0) -> Tuple[np.ndarray, np.ndarray]:
    """Icosahedron subdivided ``subdiv`` times (20 * 4**subdiv faces), vertices
    projected to the unit sphere, then displaced radially by value noise of
    amplitude ``noise`` (fraction of the radius)."""
    t = (1.0 + math.sqrt(5.0)) / 2.0
    v = np.array([[-1, t, 0], [1, t, 0], [-1, -t, 0], [1, -t, 0],
                  [0, -1, t], [0, 1, t], [0, -1, -t], [0, 1, -t],
                  [t, 0, -1], [t, 0, 1], [-t, 0, -1], [-t, 0, 1]], dtype=np.float64)
    f = np.array([[0, 11, 5], [0, 5, 1], [0, 1, 7], [0, 7, 10], [0, 10, 11],
                  [1, 5, 9], [5, 11, 4], [11, 10, 2], [10, 7, 6], [7, 1, 8],
                  [3, 9, 4], [3, 4, 2], [3, 2, 6], [3, 6, 8], [3, 8, 9],
                  [4, 9, 5], [2, 4, 11], [6, 2, 10], [8, 6, 7], [9, 8, 1]], dtype=np.int64)
    v /= np.sqrt((v * v).sum(1))[:, None]
    for _ in range(subdiv):
        nv = v.shape[0]
        a, b, c = f[:, 0], f[:, 1], f[:, 2]
        edges = np.stack([np.stack([a, b], 1), np.stack([b, c], 1), np.stack([c, a], 1)], 1).reshape(-1, 2)
        lo = edges.min(1)
        hi = edges.max(1)
        key = lo * nv + hi
        uniq, inv = np.unique(key, return_inverse=True)
        mid = (v[uniq // nv] + v[uniq % nv]) * 0.5
        mid /= np.sqrt((mid * mid).sum(1))[:, None]
        v = np.concatenate([v, mid], 0)
        m = (inv + nv).reshape(-1, 3)
        mab, mbc, mca = m[:, 0], m[:, 1], m[:, 2]
        f = np.concatenate([np.stack([a, mab, mca], 1), np.stack([b, mbc, mab], 1),
                            np.stack([c, mca, mbc], 1), np.stack([mab, mbc, mca], 1)], 0)
    if noise:
        v = v * (1.0 + noise * value_noise(v, seed, noise_freq))[:, None]
    v = v * radius + np.asarray(center, dtype=np.float64)
    return v.astype(np.float32), f.astype(np.uint32)


def latlong_sphere(n_lat: int = 20, n_lon: int = 40, radius: float = 1.0,
                   center=(0.0, 0.0, 0.0)) -> Tuple[np.ndarray, np.ndarray]:
    """Latitude-longitude sphere: (n_lat+1) rings of n_lon vertices (the pole rings
    collapse to a point), 2*n_lat*n_lon faces including zero-area pole faces."""
    th = np.arange(n_lat + 1, dtype=np.float64) * (math.pi / n_lat)
    ph = np.arange(n_lon, dtype=np.float64) * (2.0 * math.pi / n_lon)
    st, ct = _sincos64(th)
    sp, cp = _sincos64(ph)
    x = st[:, None] * cp[None, :]
    y = np.repeat(ct[:, None], n_lon, 1)
    z = st[:, None] * sp[None, :]
    v = np.stack([x, y, z], -1).reshape(-1, 3) * radius + np.asarray(center, dtype=np.float64)
    i = np.arange(n_lat)[:, None]
    j = np.arange(n_lon)[None, :]
    a = i * n_lon + j
    b = i * n_lon + (j + 1) % n_lon
    c = (i + 1) * n_lon + j
    d = (i + 1) * n_lon + (j + 1) % n_lon
    f = np.concatenate([np.stack([a, c, b], -1).reshape(-1, 3), np.stack([b, c, d], -1).reshape(-1, 3)], 0)
    return v.astype(np.float32), f.astype(np.uint32)


def quad(half: float, y: float = 0.0) -> Tuple[np.ndarray, np.ndarray]:
    """Axis-aligned square in the plane y = const, two triangles."""
    v = np.array([[-half, y, -half], [half, y, -half], [half, y, half], [-half, y, half]], dtype=np.float32)
    f = np.array([[0, 1, 2], [0, 2, 3]], dtype=np.uint32)
    return v, f


def sliver_grid(n: int, size: float, aspect: float = 1000.0, y: float = 0.0,
                seed: int = 5) -> Tuple[np.ndarray, np.ndarray]:
    """Regular (n+1) x (n+1) grid mesh in the plane y=const whose column widths
    vary so cells reach ``aspect``:1; 2*n*n long thin triangles sharing edges."""
    rng = Rng(seed)
    w = 1.0 + (aspect - 1.0) * rng.uniform(n) ** 4
    xs = np.concatenate([[0.0], np.cumsum(1.0 / w)])
    xs = xs / xs[-1] * size - size / 2
    zs = np.linspace(-size / 2, size / 2, n + 1)
    X, Z = np.meshgrid(xs, zs, indexing="xy")
    v = np.stack([X.ravel(), np.full(X.size, y), Z.ravel()], 1)
    i = np.arange(n)[:, None]
    j = np.arange(n)[None, :]
    a = i * (n + 1) + j
    b = a + 1
    c = a + (n + 1)
    d = c + 1
    f = np.concatenate([np.stack([a, c, b], -1).reshape(-1, 3), np.stack([b, c, d], -1).reshape(-1, 3)], 0)
    return v.astype(np.float32), f.astype(np.uint32)


def translate(t) -> np.ndarray:
    m = np.eye(4, dtype=np.float32)
    m[:3, 3] = np.asarray(t, dtype=np.float32)
    return m


def scale(s) -> np.ndarray:
    s = np.broadcast_to(np.asarray(s, dtype=np.float32), (3,))
    return np.diag([s[0], s[1], s[2], 1.0]).astype(np.float32)


def rigid(seed: int, extent: float) -> np.ndarray:
    """Seeded rotation (unit quaternion from hashed uniforms) plus translation
    uniform in [-extent, extent]^3."""
    u = Rng(seed).uniform(7)
    s1, c1 = _sincos64(np.array([2 * math.pi * u[1]]))
    s2, c2 = _sincos64(np.array([2 * math.pi * u[2]]))
    a, b = math.sqrt(1 - u[0]), math.sqrt(u[0])
    qx, qy, qz, qw = a * s1[0], a * c1[0], b * s2[0], b * c2[0]
    r = np.array([[1 - 2 * (qy * qy + qz * qz), 2 * (qx * qy - qz * qw), 2 * (qx * qz + qy * qw)],
                  [2 * (qx * qy + qz * qw), 1 - 2 * (qx * qx + qz * qz), 2 * (qy * qz - qx * qw)],
                  [2 * (qx * qz - qy * qw), 2 * (qy * qz + qx * qw), 1 - 2 * (qx * qx + qy * qy)]])
    m = np.eye(4)
    m[:3, :3] = r
    m[:3, 3] = (u[3:6] * 2 - 1) * extent
    return m.astype(np.float32)


def look_at(eye, target, up=(0.0, 1.0, 0.0)) -> np.ndarray:
    """Camera-to-world matrix for a left-handed camera looking down +z."""
    eye = np.asarray(eye, dtype=np.float64)
    d = np.asarray(target, dtype=np.float64) - eye
    d /= np.linalg.norm(d)
    right = np.cross(np.asarray(up, dtype=np.float64), d)
    right /= np.linalg.norm(right)
    new_up = np.cross(d, right)
    m = np.eye(4)
    m[:3, 0] = right
    m[:3, 1] = new_up
    m[:3, 2] = d
    m[:3, 3] = eye
    return m.astype(np.float32)


# --------------------------------------------------------------------------- rays

def rays_coherent(cam: Camera, nx: int, ny: int) -> np.ndarray:
    """Primary pinhole rays through an nx x ny grid of pixel centres of ``cam``
    (camera rays as the reference camera computes them, in float32)."""
    f32 = np.float32
    oow, ooh, aspect, tanf2 = (f32(x) for x in Camera(cam.camera_to_world, cam.fov, nx, ny).derived())
    xs = (np.arange(nx, dtype=np.float32) + f32(0.5)) * oow
    ys = (np.arange(ny, dtype=np.float32) + f32(0.5)) * ooh
    cx = (f32(2) * xs - f32(1)) * aspect * tanf2
    cy = (f32(1) - f32(2) * ys) * tanf2
    CX, CY = np.meshgrid(cx, cy, indexing="xy")
    d = np.stack([CX.ravel(), CY.ravel(), np.ones(CX.size, np.float32)], 1).astype(np.float32)
    d = d / np.sqrt((d * d).sum(1, dtype=np.float32))[:, None]
    m = cam.camera_to_world.astype(np.float32)
    w = (d @ m[:3, :3].T).astype(np.float32)
    w = w / np.sqrt((w * w).sum(1, dtype=np.float32))[:, None]
    o = np.broadcast_to(m[:3, 3], w.shape)
    return np.concatenate([o, w], 1).astype(np.float32)


def unit_vectors(n: int, rng: Rng) -> np.ndarray:
    z = rng.uniform(n) * 2.0 - 1.0
    phi = rng.uniform(n) * (2.0 * math.pi)
    s, c = _sincos64(phi)
    r = np.sqrt(np.maximum(0.0, 1.0 - z * z))
    return np.stack([r * c, r * s, z], 1)


def rays_incoherent(bmin, bmax, n: int, seed: int) -> np.ndarray:
    """Origins uniform in the box [bmin, bmax], directions uniform on the sphere."""
    rng = Rng(seed)
    bmin = np.asarray(bmin, np.float64)
    bmax = np.asarray(bmax, np.float64)
    o = bmin + (bmax - bmin) * np.stack([rng.uniform(n), rng.uniform(n), rng.uniform(n)], 1)
    d = unit_vectors(n, rng)
    return np.concatenate([o, d], 1).astype(np.float32)


def rays_at_features(verts: np.ndarray, faces: np.ndarray, n: int, seed: int,
                     dist: float = 3.0, center=None) -> np.ndarray:
    """Rays aimed exactly at shared mesh features: half at vertices, half at edge
    midpoints (computed in float32), from origins on a sphere of radius ``dist``
    around ``center``.  Direction = normalize(target - origin)."""
    rng = Rng(seed)
    center = verts.mean(0) if center is None else np.asarray(center)
    fi = (rng.uniform(n) * faces.shape[0]).astype(np.int64) % faces.shape[0]
    ci = (rng.uniform(n) * 3).astype(np.int64) % 3
    a = verts[faces[fi, ci]].astype(np.float32)
    b = verts[faces[fi, (ci + 1) % 3]].astype(np.float32)
    mid = ((a + b) * np.float32(0.5)).astype(np.float32)
    target = np.where((np.arange(n) % 2 == 0)[:, None], a, mid)
    o = (center + unit_vectors(n, rng) * dist).astype(np.float32)
    d = (target.astype(np.float64) - o.astype(np.float64))
    d /= np.linalg.norm(d, axis=1)[:, None]
    return np.concatenate([o, d.astype(np.float32)], 1).astype(np.float32)


# --------------------------------------------------------------------------- files

def write_ptxm(path: str, meshes: Sequence[MeshData], camera: Optional[Camera] = None,
               spp: int = 1, max_depth: int = 5, skybox: Optional[Sequence[float]] = None) -> None:
    """Scene container read by the reference test driver (oracle/ref/ref_driver.cpp)."""
    out = bytearray()
    out += struct.pack("<III", 0x4D585450, 1, len(meshes))
    for m in meshes:
        v = np.ascontiguousarray(m.verts, np.float32)
        f = np.ascontiguousarray(m.faces, np.uint32)
        out += struct.pack("<III", v.shape[0], f.shape[0], 1 if m.normals is not None else 0)
        out += np.ascontiguousarray(m.object_to_world, np.float32).tobytes()
        out += struct.pack("<I3fI3f", m.brdf, *m.refl, 1 if m.is_light else 0, *m.emission)
        for k in range(3):
            out += np.ascontiguousarray(v[:, k]).tobytes()
        if m.normals is not None:
            nn = np.ascontiguousarray(m.normals, np.float32)
            for k in range(3):
                out += np.ascontiguousarray(nn[:, k]).tobytes()
        for k in range(3):
            out += np.ascontiguousarray(f[:, k]).tobytes()
    if camera is not None:
        out += struct.pack("<I", 1) + np.ascontiguousarray(camera.camera_to_world, np.float32).tobytes()
        out += struct.pack("<fIIII", camera.fov, camera.width, camera.height, spp, max_depth)
    else:
        out += struct.pack("<I", 0) + np.eye(4, dtype=np.float32).tobytes() + struct.pack("<fIIII", 90.0, 1, 1, 1, 1)
    if skybox is not None:
        out += struct.pack("<I3f", 1, *skybox)
    else:
        out += struct.pack("<I3f", 0, 0.0, 0.0, 0.0)
    with open(path, "wb") as fh:
        fh.write(bytes(out))


def look_at_lh(eye, target, up) -> np.ndarray:
    """Camera-to-world matrix exactly as the reference's pbrt scene block makes
    it from ``LookAt``: Transform::LookAt(eye, lookAt, up, false)
    (src/cpu/structures/Transform.cpp:322-365) composed onto the identity with
    Matrix4x4::operator*= (src/cpu/structures/Matrix4x4.cpp:244-283), all in
    float32 with the same operation order (so signed zeros match too)."""
    f = np.float32
    ex, ey, ez = (f(v) for v in eye)
    tx, ty, tz = (f(v) for v in target)
    ux, uy, uz = (f(v) for v in up)

    def normalize(x, y, z):                                   # Vector::Normalize
        ld = f(1.0) / f(np.sqrt(x * x + y * y + z * z))
        return x * ld, y * ld, z * ld

    def cross(a, b):                                          # Vector::Cross
        return (a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2], a[0] * b[1] - a[1] * b[0])

    with np.errstate(all="ignore"):
        d = normalize(tx - ex, ty - ey, tz - ez)
        u = normalize(ux, uy, uz)
        left = cross(u, d)
        if f(np.sqrt(left[0] * left[0] + left[1] * left[1] + left[2] * left[2])) == 0:
            raise ValueError("LookAt: up is parallel to the view direction")
        left = normalize(*left)
        new_up = cross(d, left)
        m = np.zeros((4, 4), np.float32)
        m[0, 0], m[1, 0], m[2, 0] = left
        m[0, 1], m[1, 1], m[2, 1] = new_up
        m[0, 2], m[1, 2], m[2, 2] = d
        m[0, 3], m[1, 3], m[2, 3], m[3, 3] = ex, ey, ez, f(1)
        ident = np.eye(4, dtype=np.float32)
        out = np.zeros((4, 4), np.float32)
        for i in range(4):
            for j in range(4):
                acc = ident[i, 0] * m[0, j]
                for k in range(1, 4):
                    acc = f(acc + ident[i, k] * m[k, j])
                out[i, j] = acc
    return out


def _fmt(x) -> str:
    return "%.9g" % float(np.float32(x))


def write_pbrt(directory: str, meshes: Sequence[MeshData], eye, target, up, fov: float, width: int, height: int,
               spp: int, max_depth: int, skybox: Optional[Sequence[float]] = None, name: str = "scene"):
    """Writes ``name``.pbrt and one ASCII PLY per mesh into ``directory``, in the
    subset of the format the reference's parser reads (pbrt_parser.cpp,
    mesh_parsers.cpp): LookAt, Camera "perspective", Film, Sampler, Integrator
    "path", LightSource "infinite", per mesh AttributeBegin / Translate / Scale /
    Material "matte" | "mirror" | "plastic" / AreaLightSource "diffuse" /
    Shape "plymesh".  A textured mesh's texture is written as an RGBA PNG with a
    Texture "color" "imagemap" directive (once per texture) and its PLY carries
    u, v vertex properties; the mesh must have per-vertex ``uv``.
    Meshes must have a translate-and-scale object-to-world (no rotation), which
    the reference composes exactly.  Floats are written with 9 significant
    digits (exact for float32).  Emission and skybox are written as value / 255
    because the reference multiplies by 255 (pbrt_parser.cpp:1281-1283, 1340).

    Returns (path, meshes, camera, skybox) as the reference will see them:
    emission and sky after the / 255 * 255 round trip and the camera built by
    look_at_lh, so that rendering them here and there gives the same film."""
    import os
    f = np.float32
    os.makedirs(directory, exist_ok=True)
    lines = ["LookAt " + " ".join(_fmt(v) for v in (*eye, *target, *up)),
             'Camera "perspective" "float fov" [ %s ]' % _fmt(fov),
             'Film "image" "integer xresolution" [ %d ] "integer yresolution" [ %d ] "string filename" [ "%s.png" ]'
             % (width, height, name),
             'Sampler "halton" "integer pixelsamples" [ %d ]' % spp,
             'Integrator "path" "integer maxdepth" [ %d ]' % max_depth,
             "WorldBegin"]
    sky_out = None
    if skybox is not None:
        l = [f(f(v) / f(255)) for v in skybox]
        lines.append('LightSource "infinite" "rgb L" [ %s ]' % " ".join(_fmt(v) for v in l))
        sky_out = tuple(float(f(v) * f(255)) for v in l)
    out_meshes = []
    tex_names = {}
    from .png import encode_png
    for i, m in enumerate(meshes):
        if m.texture is not None and id(m.texture) not in tex_names:
            tname = f"{name}_tex{len(tex_names)}"
            with open(os.path.join(directory, tname + ".png"), "wb") as fh:
                fh.write(encode_png(m.texture.rgba, filters=(0, 1, 2, 3, 4)))
            tex_names[id(m.texture)] = tname
            lines.append('Texture "%s" "color" "imagemap" "string filename" [ "%s.png" ]' % (tname, tname))
    for i, m in enumerate(meshes):
        o2w = np.asarray(m.object_to_world, np.float32)
        lin = o2w[:3, :3]
        if np.any(lin != np.diag(np.diag(lin))) or np.any(o2w[3] != np.array([0, 0, 0, 1], np.float32)):
            raise ValueError(f"mesh {i}: write_pbrt supports translate-and-scale transforms only")
        ply = f"{name}_mesh{i}.ply"
        v = np.asarray(m.verts, np.float32)
        fc = np.asarray(m.faces, np.uint32)
        uv = None
        if m.texture is not None:
            if m.uv is None or len(m.uv) != len(v):
                raise ValueError(f"mesh {i}: a textured mesh needs one (u, v) per vertex")
            uv = np.asarray(m.uv, np.float32).reshape(-1, 2)
        with open(os.path.join(directory, ply), "w") as fh:
            fh.write("ply\nformat ascii 1.0\nelement vertex %d\nproperty float x\nproperty float y\n"
                     "property float z\n%selement face %d\nproperty list uchar int vertex_indices\nend_header\n"
                     % (len(v), "property float u\nproperty float v\n" if uv is not None else "", len(fc)))
            if uv is None:
                fh.write("".join("%s %s %s\n" % (_fmt(a), _fmt(b), _fmt(c)) for a, b, c in v))
            else:
                fh.write("".join("%s %s %s %s %s\n" % (_fmt(a), _fmt(b), _fmt(c), _fmt(tu), _fmt(tv))
                                 for (a, b, c), (tu, tv) in zip(v, uv)))
            fh.write("".join("3 %d %d %d\n" % (a, b, c) for a, b, c in fc))
        lines.append("AttributeBegin")
        t = o2w[:3, 3]
        s = np.diag(lin)
        if np.any(t != 0):
            lines.append("  Translate " + " ".join(_fmt(x) for x in t))
        if np.any(s != 1):
            lines.append("  Scale " + " ".join(_fmt(x) for x in s))
        if m.brdf == MIRROR:
            lines.append('  Material "mirror" "rgb Kr" [ %s ]' % " ".join(_fmt(x) for x in m.refl))
        elif m.brdf == GLOSSY:
            lines.append('  Material "plastic" "rgb Kd" [ %s ] "rgb Ks" [ %s ] "float roughness" [ %s ]'
                         % (" ".join(_fmt(x) for x in m.refl), " ".join(_fmt(x) for x in m.specular),
                            _fmt(m.roughness)))
        elif m.texture is not None:
            lines.append('  Material "matte" "texture Kd" [ "%s" ]' % tex_names[id(m.texture)])
        else:
            lines.append('  Material "matte" "rgb Kd" [ %s ]' % " ".join(_fmt(x) for x in m.refl))
        emission = m.emission
        if m.is_light:
            l = [f(f(x) / f(255)) for x in m.emission]
            lines.append('  AreaLightSource "diffuse" "rgb L" [ %s ]' % " ".join(_fmt(x) for x in l))
            emission = tuple(float(f(x) * f(255)) for x in l)
        lines.append('  Shape "plymesh" "string filename" [ "%s" ]' % ply)
        lines.append("AttributeEnd")
        textured = m.texture is not None and not m.is_light and m.brdf == LAMBERT
        out_meshes.append(MeshData(v, fc, object_to_world=o2w, brdf=m.brdf,
                                   refl=(1.0, 1.0, 1.0) if textured else tuple(float(f(x)) for x in m.refl),
                                   is_light=m.is_light, emission=emission,
                                   specular=tuple(float(f(x)) for x in m.specular), roughness=float(f(m.roughness)),
                                   texture=m.texture if textured else None, uv=uv if textured else None))
    lines.append("WorldEnd")
    path = os.path.join(directory, name + ".pbrt")
    with open(path, "w") as fh:
        fh.write("\n".join(lines) + "\n")
    cam = Camera(look_at_lh(eye, target, up), float(f(fov)), width, height)
    return path, out_meshes, cam, sky_out


def write_rays(path: str, rays: np.ndarray) -> None:
    rays = np.ascontiguousarray(rays, np.float32)
    with open(path, "wb") as fh:
        fh.write(struct.pack("<I", rays.shape[0]))
        fh.write(rays.tobytes())


# --------------------------------------------------------------------------- workloads

def scene_bounds(meshes: Sequence[MeshData]) -> Tuple[np.ndarray, np.ndarray]:
    lo = np.full(3, np.inf)
    hi = np.full(3, -np.inf)
    for m in meshes:
        w = (np.c_[m.verts.astype(np.float64), np.ones(len(m.verts))] @ m.object_to_world.astype(np.float64).T)[:, :3]
        lo = np.minimum(lo, w.min(0))
        hi = np.maximum(hi, w.max(0))
    return lo, hi


def closest_hit_workload(subdiv: int, n_rays: int, seed: int = 2):
    """Closest-hit microbench: displaced icosphere (radius 1 at the origin) and
    n_rays rays, half coherent pinhole rays on a 2:1 grid aimed at the sphere,
    half incoherent (origin uniform in the bounding box, direction uniform)."""
    v, f = icosphere(subdiv, noise=0.05, seed=seed)
    meshes = [MeshData(v, f)]
    n_coh = n_rays // 2
    nx = int(round(math.sqrt(n_coh * 2)))
    ny = max(1, n_coh // nx)
    cam = Camera(look_at((0.0, 0.3, -2.5), (0.0, 0.0, 0.0)), 44.0, nx, ny)
    coh = rays_coherent(cam, nx, ny)
    lo, hi = scene_bounds(meshes)
    inc = rays_incoherent(lo, hi, n_rays - coh.shape[0], seed + 100)
    return meshes, np.concatenate([coh, inc], 0), coh.shape[0]


def plain_render_scene(seed: int = 1, width: int = 640, height: int = 480):
    """Ground quad plus three 20x40 lat-long spheres (1,600 faces each), radius
    U[0.5,1.5], Lambert albedo U[0.2,0.8]^3, constant skybox, 60-degree camera."""
    rng = Rng(seed)
    u = rng.uniform(32)
    meshes = []
    gv, gf = quad(12.0, 0.0)
    meshes.append(MeshData(gv, gf, refl=(0.6, 0.6, 0.6)))
    xs = (-2.6, 0.0, 2.6)
    for i in range(3):
        r = 0.5 + u[i]
        sv, sf = latlong_sphere(20, 40, 1.0)
        o2w = translate((xs[i], r, 1.0 + 2.0 * u[3 + i])) @ scale(r)
        refl = tuple(float(0.2 + 0.6 * u[6 + 3 * i + k]) for k in range(3))
        meshes.append(MeshData(sv, sf, object_to_world=o2w.astype(np.float32), refl=refl))
    cam = Camera(look_at((0.0, 2.2, -6.0), (0.0, 0.8, 1.5)), 60.0, width, height)
    return meshes, cam, (180.0, 200.0, 230.0)


def noise_texture(width: int, height: int, seed: int) -> Texture:
    """Seeded RGBA8 texture: 8x8-texel blocks of one colour each plus a per-texel
    jitter of up to 31, alpha 255."""
    rng = Rng(seed)
    bw, bh = (width + 7) // 8, (height + 7) // 8
    base = (rng.uniform(bw * bh * 3) * 224.0).astype(np.uint32).reshape(bh, bw, 3)
    jit = (rng.uniform(width * height * 3) * 32.0).astype(np.uint32).reshape(height, width, 3)
    rgb = np.repeat(np.repeat(base, 8, 0), 8, 1)[:height, :width] + jit
    rgba = np.concatenate([rgb, np.full((height, width, 1), 255, np.uint32)], 2)
    return Texture(rgba.astype(np.uint8), f"noise{seed}")


def latlong_uv(n_lat: int, n_lon: int) -> np.ndarray:
    """Texture coordinates for latlong_sphere's vertices: u = j / n_lon around,
    v = 1 - i / n_lat from the north ring down."""
    i = np.repeat(np.arange(n_lat + 1, dtype=np.float64), n_lon)
    j = np.tile(np.arange(n_lon, dtype=np.float64), n_lat + 1)
    return np.stack([j / n_lon, 1.0 - i / n_lat], 1).astype(np.float32)


def textured_render_scene(seed: int = 1, width: int = 640, height: int = 480):
    """plain_render_scene's layout with textures: a ground quad whose texture
    coordinates run from -0.25 to 2.25 (texels past the image's edge: the next
    row's first texel where that index stays inside the image, else the edge),
    a textured lat-long sphere, a constant-Kd sphere and a mirror sphere; two
    seeded noise textures (61 x 37 and 256 x 128 texels)."""
    rng = Rng(seed)
    u = rng.uniform(16)
    floor_tex = noise_texture(61, 37, seed + 10)
    ball_tex = noise_texture(256, 128, seed + 11)
    gv, gf = quad(12.0, 0.0)
    guv = np.array([[-0.25, -0.25], [2.25, -0.25], [2.25, 2.25], [-0.25, 2.25]], np.float32)
    meshes = [MeshData(gv, gf, texture=floor_tex, uv=guv)]
    sv, sf = latlong_sphere(20, 40, 1.0)
    suv = latlong_uv(20, 40)
    xs = (-2.6, 0.0, 2.6)
    for i in range(3):
        r = 0.5 + u[i]
        o2w = (translate((xs[i], r, 1.0 + 2.0 * u[3 + i])) @ scale(r)).astype(np.float32)
        if i == 0:
            meshes.append(MeshData(sv, sf, object_to_world=o2w, texture=ball_tex, uv=suv))
        elif i == 1:
            meshes.append(MeshData(sv, sf, object_to_world=o2w, brdf=MIRROR, refl=(0.9, 0.85, 0.8)))
        else:
            meshes.append(MeshData(sv, sf, object_to_world=o2w, refl=tuple(float(0.2 + 0.6 * u[6 + k])
                                                                                for k in range(3))))
    cam = Camera(look_at((0.0, 2.2, -6.0), (0.0, 0.8, 1.5)), 60.0, width, height)
    return meshes, cam, (180.0, 200.0, 230.0)


def planar_uv(verts: np.ndarray, repeat: float = 1.0) -> np.ndarray:
    """Texture coordinates by projection onto the x-z plane of the vertices'
    bounding box, scaled by ``repeat``."""
    v = np.asarray(verts, np.float64)
    lo, hi = v.min(0), v.max(0)
    span = np.where(hi - lo > 0, hi - lo, 1.0)
    return np.stack([(v[:, 0] - lo[0]) / span[0] * repeat, (v[:, 2] - lo[2]) / span[2] * repeat],
                    1).astype(np.float32)


def with_materials(meshes: Sequence[MeshData], seed: int = 5, texture_size: int = 512) -> List[MeshData]:
    """The same geometry with the f4 materials: every Lambert mesh textured with
    one seeded ``texture_size``^2 noise texture (planar coordinates, repeated 8
    times on large meshes), every mirror turned into the glossy lobe (Kd its
    reflectance, Ks 0.9, roughness 0.2)."""
    tex = noise_texture(texture_size, texture_size, seed)
    out = []
    uv_cache = {}
    for m in meshes:
        if m.is_light:
            out.append(m)
        elif m.brdf == LAMBERT:
            key = id(m.verts)
            if key not in uv_cache:
                uv_cache[key] = planar_uv(m.verts, 8.0 if m.num_faces <= 2 else 1.0)
            out.append(MeshData(m.verts, m.faces, object_to_world=m.object_to_world, normals=m.normals,
                                texture=tex, uv=uv_cache[key]))
        else:
            out.append(MeshData(m.verts, m.faces, object_to_world=m.object_to_world, normals=m.normals,
                                brdf=GLOSSY, refl=m.refl, specular=(0.9, 0.9, 0.9), roughness=0.2))
    return out


def batched_render_workload(subdiv: int = 8, instances: int = 32, jobs: int = 16, spp: int = 64,
                            max_depth: int = 8, seed: int = 3):
    """Batched wavefront workload: a displaced icosphere (the shared geometry of
    ``instances`` + 1 meshes: the original plus instanced copies under seeded rigid
    transforms) over a 2-triangle ground plane; 70 % of the meshes Lambertian
    (albedo U[0.2, 0.8]^3), 30 % perfect mirrors; ``jobs`` render jobs whose
    resolutions step from 256x256 to 1920x1080, ``spp`` samples, depth ``max_depth``,
    seeds seed..seed+jobs-1.  Returns (meshes, skybox, [(camera, spp, depth, seed)])."""
    v, f = icosphere(subdiv, radius=1.0, noise=0.05, seed=seed)
    rng = Rng(seed)
    u = rng.uniform(4 * (instances + 1))
    meshes = []
    for i in range(instances + 1):
        o2w = IDENTITY.copy() if i == 0 else rigid(1000 + i, 9.0)
        o2w = o2w.copy()
        o2w[1, 3] = abs(float(o2w[1, 3])) + 1.0          # keep the copies above the ground
        mirror = u[4 * i] < 0.3
        refl = tuple(float(0.2 + 0.6 * u[4 * i + 1 + k]) for k in range(3))
        meshes.append(MeshData(v, f, object_to_world=o2w.astype(np.float32),
                               brdf=MIRROR if mirror else LAMBERT, refl=(0.9, 0.9, 0.9) if mirror else refl))
    gv, gf = quad(40.0, 0.0)
    meshes.append(MeshData(gv, gf, refl=(0.5, 0.5, 0.5)))
    job_list = []
    for j in range(jobs):
        t = j / max(1, jobs - 1)
        w = int(round(256 + (1920 - 256) * t))
        h = int(round(256 + (1080 - 256) * t))
        ang = 2 * math.pi * j / jobs
        s_, c_ = _sincos64(np.array([ang]))
        eye = (float(22.0 * c_[0]), 9.0, float(22.0 * s_[0]))
        cam = Camera(look_at(eye, (0.0, 2.0, 0.0)), 55.0, w, h)
        job_list.append((cam, spp, max_depth, seed + j))
    return meshes, (150.0, 170.0, 200.0), job_list


def stress_workload(n_rays: int = 1 << 24, spheres: int = 32, n_lat: int = 256, n_lon: int = 512,
                    grid_n: int = 1024, seed: int = 5):
    """Stress and watertightness workload: ``spheres`` displaced lat-long spheres of
    2*n_lat*n_lon faces (262,144 at the defaults; radius 1, +/-5 % radial value
    noise) on an 8-wide layout above a grid_n x grid_n sliver grid (2*grid_n^2
    faces, cells up to 1000:1), every mesh in world space with an identity
    transform so feature positions are exact floats.  Meshes alternate Lambert and
    mirror.  Of the n_rays rays the first quarter is aimed exactly at shared
    features (vertices and float32 edge midpoints; half on the spheres, from
    origins 2.5 radii around the sphere's centre, half on the grid, from origins
    at distance 4 in random directions); the rest are incoherent (origin uniform in
    the scene box, direction uniform).  Feature rays are never grazing: they
    arrive within 53 degrees of the radial direction (spheres) or of the grid's
    normal (grid targets off its border), so each one enters the surface at its
    target and a watertight intersector hits at t <= |target - origin| (up to
    rounding).  Returns (meshes, rays, n_feature, t_target)."""
    rng = Rng(seed)
    sv, sf = latlong_sphere(n_lat, n_lon, 1.0)
    unit = sv.astype(np.float64)
    meshes = []
    centers = []
    cols = 8
    for i in range(spheres):
        cx = (i % cols - (cols - 1) / 2.0) * 3.0
        cz = (i // cols - (max(1, (spheres + cols - 1) // cols) - 1) / 2.0) * 3.0
        c = np.array([cx, 1.2, cz])
        d = unit * (1.0 + 0.05 * value_noise(unit, seed + i, 3.0))[:, None] + c
        centers.append(c)
        u = rng.uniform(3)
        meshes.append(MeshData(d.astype(np.float32), sf, brdf=MIRROR if i % 3 == 2 else LAMBERT,
                               refl=tuple(float(0.2 + 0.6 * x) for x in u)))
    size = 3.0 * cols + 6.0
    gv, gf = sliver_grid(grid_n, size, 1000.0, 0.0, seed)
    meshes.append(MeshData(gv, gf, refl=(0.5, 0.5, 0.5)))
    n_feat = n_rays // 4
    n_sph = n_feat // 2
    n_grid = n_feat - n_sph

    def targets(verts, faces, n):
        fi = (rng.uniform(n) * faces.shape[0]).astype(np.int64) % faces.shape[0]
        ci = (rng.uniform(n) * 3).astype(np.int64) % 3
        a = verts[faces[fi, ci]].astype(np.float32)
        b = verts[faces[fi, (ci + 1) % 3]].astype(np.float32)
        mid = ((a + b) * np.float32(0.5)).astype(np.float32)
        return np.where((np.arange(n) % 2 == 0)[:, None], a, mid)

    def cone(axis, n):
        """Unit vectors within asin(0.8) (53 degrees) of the unit rows of ``axis``."""
        d = axis + 0.8 * unit_vectors(n, rng)
        return d / np.linalg.norm(d, axis=1)[:, None]

    which = (rng.uniform(n_sph) * spheres).astype(np.int64) % spheres
    tgt_s = np.zeros((n_sph, 3), np.float32)
    org_s = np.zeros((n_sph, 3), np.float64)
    for i in range(spheres):
        sel = np.nonzero(which == i)[0]
        tgt_s[sel] = targets(meshes[i].verts, sf, len(sel))
        radial = tgt_s[sel].astype(np.float64) - centers[i]
        radial /= np.linalg.norm(radial, axis=1)[:, None]
        org_s[sel] = tgt_s[sel] + cone(radial, len(sel)) * 2.5
    # grid targets: features of cells off the border (a border edge has one face)
    cell = np.arange(gf.shape[0]) % (grid_n * grid_n)
    ci_, cj_ = cell // grid_n, cell % grid_n
    inner = np.nonzero((ci_ >= 1) & (ci_ <= grid_n - 2) & (cj_ >= 1) & (cj_ <= grid_n - 2))[0]
    tgt_g = targets(gv, gf[inner], n_grid)
    up = np.zeros((n_grid, 3))
    up[:, 1] = np.where(rng.uniform(n_grid) < 0.5, -1.0, 1.0)     # from above or below the grid
    org_g = tgt_g.astype(np.float64) + cone(up, n_grid) * 4.0
    tgt = np.concatenate([tgt_s, tgt_g], 0).astype(np.float64)
    org = np.concatenate([org_s, org_g], 0).astype(np.float32).astype(np.float64)
    d = tgt - org
    t_target = np.linalg.norm(d, axis=1)
    d /= t_target[:, None]
    feat = np.concatenate([org, d], 1).astype(np.float32)
    lo, hi = scene_bounds(meshes)
    inc = rays_incoherent(lo, hi, n_rays - n_feat, seed + 100)
    return meshes, np.concatenate([feat, inc], 0), n_feat, t_target.astype(np.float32)


# Scenes of the serving stream: the icosphere subdivisions of each scene's
# spheres, from 10,242 faces (two spheres of 5,120 and the ground) to
# 1,966,082 (one sphere of 1,310,720 and two of 327,680).
SERVING_SCENES = ((4, 4), (5,), (6,), (7,), (8,), (8, 7, 7))


def serving_scene(spec, index: int, seed: int = 4):
    """One scene of the serving stream: displaced icospheres of subdivisions
    ``spec`` side by side (radius 1 for the first, 0.5 for the others) above a
    ground quad (mirror for odd ``index``), with a constant sky."""
    meshes = []
    for k, sd in enumerate(spec):
        center = (0.0, 0.0, 0.0) if k == 0 else (-1.6 + 3.2 * (k - 1), -0.5, 0.8)
        v, f = icosphere(sd, radius=1.0 if k == 0 else 0.5, center=center, noise=0.05, seed=seed + 7 * index + k)
        meshes.append(MeshData(v, f, refl=(0.7, 0.6, 0.5) if k == 0 else (0.5, 0.6, 0.7),
                               brdf=LAMBERT if k != 1 else MIRROR))
    gv, gf = quad(6.0, -1.05)
    meshes.append(MeshData(gv, gf, brdf=MIRROR if index % 2 else LAMBERT, refl=(0.6, 0.6, 0.6)))
    return meshes, (140.0, 160.0, 190.0)


def serving_requests(n: int, rate_hz: float, seed: int = 4, scenes=SERVING_SCENES):
    """Poisson request stream for the render service: ``n`` jobs at rate ``rate_hz``
    over the displaced-icosphere scenes ``scenes`` (10k to 2M faces by default),
    resolution between 256x256 and 1920x1080 (log-uniform area), spp log-uniform in
    [4, 256], depth uniform in 4..12.  Returns (scenes, [(arrival_s, scene_index,
    camera, spp, depth, seed)])."""
    scenes_ = [serving_scene(spec, i, seed) for i, spec in enumerate(scenes)]
    rng = Rng(seed)
    u = [rng.uniform(n) for _ in range(6)]
    gaps = -np.log(np.maximum(u[0], 1e-12)) / rate_hz      # exponential inter-arrival times
    arrivals = np.cumsum(gaps) - gaps[0]
    reqs = []
    for i in range(n):
        area = math.exp(math.log(256 * 256) + (math.log(1920 * 1080) - math.log(256 * 256)) * u[1][i])
        w = int(min(1920, max(256, round(math.sqrt(area * 16 / 9)))))
        h = int(min(1080, max(256, round(area / w))))
        spp = int(round(math.exp(math.log(4) + (math.log(256) - math.log(4)) * u[2][i])))
        depth = 4 + int(u[3][i] * 9) % 9
        sc = int(u[4][i] * len(scenes)) % len(scenes)
        cam = Camera(look_at((0.4, 0.6, -3.2), (0.0, 0.0, 0.0)), 50.0, w, h)
        reqs.append((float(arrivals[i]), sc, cam, spp, depth, seed + 100 + i))
    return scenes_, reqs
