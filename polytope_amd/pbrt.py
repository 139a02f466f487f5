"""Reader for the pbrt-v3 scene subset and PLY / OBJ meshes that the reference
understands, with the reference's own semantics, so that a scene file read
here describes exactly the scene the reference would render from it.

What is restated (file:line in the reference tree):

* the scanner that splits the file into one token list per directive,
  including its bracket and comment handling (pbrt_parser.cpp:299-392);
* the lexer of directive arguments (pbrt_parser.cpp:394-665): typed
  ``"float fov" [ 85 ]`` arguments, the fixed-arity LookAt / Translate /
  Scale / Rotate / Transform forms, ``std::stof`` / ``std::stoi`` parsing;
* the scene block (pbrt_parser.cpp:925-1242): the first Sampler's first
  integer argument, the first Film, every transform directive composed into
  the camera transform, the first perspective Camera, the first Integrator;
* the world block (pbrt_parser.cpp:1245-1706): attribute / transform stacks,
  Material, MakeNamedMaterial / NamedMaterial, AreaLightSource, LightSource
  "infinite", ObjectBegin / ObjectEnd / ObjectInstance, and Shape "plymesh",
  "objmesh" and "sphere";
* the PLY reader (mesh_parsers.cpp:14-417), the OBJ reader
  (mesh_parsers.cpp:530-581) and the sphere tessellator (tesselators.cpp:10-84);
* transforms as float32 4x4 matrices composed with Matrix4x4::operator*=
  (Matrix4x4.cpp:244-283), each carrying the inverse the reference tracks
  alongside (Transform.cpp:61-67, 230-300).

The reference's quirks are kept, because they decide which scene it renders:

* meshes keep a reference to the transform that was active when they were
  created, so a later Translate / Scale / Rotate in the same attribute scope
  moves them too (pbrt_parser.cpp:1462, 1475-1479, 1697-1698); their inverse
  is the one current at creation (pbrt_parser.cpp:1461);
* a second AreaLightSource in a scope rewrites the light that earlier meshes
  of that scope already hold (pbrt_parser.cpp:1276-1284);
* Q5: ``Transform [16 values]`` cannot be read (the lexer keeps 15 values and
  the directive then reads a 16th, pbrt_parser.cpp:414-416, 713): an error;
* Q6: binary PLY files without normals store every vertex twice
  (mesh_parsers.cpp:474-479), so their faces index the wrong vertices;
* Q7: MakeNamedMaterial "matte" ignores Kd and uses 0.5 (pbrt_parser.cpp:1352-1374);
* Q9: Shape "sphere" ignores its radius for the geometry's transform and uses
  max((int) radius, 20) meridians and parallels (pbrt_parser.cpp:1624-1643),
  with the tessellator's duplicated pole (tesselators.cpp:15-16);
* LightSource "point" is not read by the reference (pbrt_parser.cpp:1331-1350).
  Here such lights are collected separately for the library's opt-in
  next-event estimation; they do not change anything else.

The "plastic" material is the reference's GlossyBRDF (Kd, Ks, roughness).
``Texture "name" "color" "imagemap"`` directives load PNG files as the
reference does (pbrt_parser.cpp:730-776, decoded to RGBA8 as lodepng does,
see png.py) and ``Material "matte" "texture Kd"`` makes a textured Lambert
material (pbrt_parser.cpp:866-875); its meshes take their texture coordinates
from their PLY files' u/v (or s/t) vertex properties.  The reference marks
every PLY mesh as having texture coordinates, uninitialised when the file has
none (mesh_parsers.cpp:320, 340); a textured mesh without them, or one whose
texture is a "float" (one byte per texel, which the reference indexes four
bytes at a time), cannot be shaded.

Not supported, with an error: shapes without a material, and anything the
reference itself would stop on.
"""
from __future__ import annotations

import ctypes
import ctypes.util
import math
import os
import re
from dataclasses import dataclass, field
from fractions import Fraction
from typing import Dict, List, Optional, Tuple

import numpy as np

from .png import PngError, read_png
from .scenes import GLOSSY, LAMBERT, MIRROR, Camera, MeshData, Texture

f32 = np.float32

SCENE_DIRECTIVES = {"Camera", "Film", "Integrator", "LookAt", "PixelFilter", "Rotate", "Sampler", "Scale",
                    "Transform", "Translate"}
WORLD_DIRECTIVES = {"AreaLightSource", "AttributeBegin", "AttributeEnd", "LightSource", "MakeNamedMaterial",
                    "Material", "NamedMaterial", "ObjectBegin", "ObjectEnd", "ObjectInstance", "Rotate", "Scale",
                    "Shape", "Texture", "Transform", "TransformBegin", "TransformEnd", "Translate", "WorldBegin",
                    "WorldEnd"}
DIRECTIVES = SCENE_DIRECTIVES | WORLD_DIRECTIVES

# pbrt_parser.cpp:269-272 and constants.h:33-43
DEFAULT_FOV = f32(90.0)
DEFAULT_WIDTH, DEFAULT_HEIGHT = 640, 360
DEFAULT_SAMPLES = 8
PI = f32(math.pi)
TWO_PI = f32(2.0) * PI
PI_OVER_2 = PI / f32(2.0)
PI_OVER_180 = f32(math.pi / 180.0)


class PbrtError(ValueError):
    """A scene the reference would reject, or one this reader does not support."""


# ----------------------------------------------------------------------------- numbers

_FLOAT_PREFIX = re.compile(r"\s*([+-]?(?:(?:\d+\.?\d*|\.\d+)(?:[eE][+-]?\d+)?|inf(?:inity)?|nan))", re.I)
_INT_PREFIX = re.compile(r"\s*([+-]?\d+)")


def _round_f32(text: str) -> np.float32:
    """The float32 nearest to the decimal `text`, ties to even: what strtof
    returns.  Rounding through float64 is exact unless the float64 lands on a
    float32 midpoint; then the exact rational decides."""
    d = float(text)
    r = f32(d)
    if not np.isfinite(r) or float(r) == d:
        return r
    lo, hi = (r, np.nextafter(r, f32(np.inf))) if float(r) < d else (np.nextafter(r, f32(-np.inf)), r)
    mid = (Fraction(float(lo)) + Fraction(float(hi))) / 2
    if Fraction(d) != mid:
        return r
    exact = Fraction(text.strip())
    if exact < mid:
        return lo
    if exact > mid:
        return hi
    return lo if (int(lo.view(np.uint32)) & 1) == 0 else hi


def stof(token: str) -> np.float32:
    """std::stof: the longest leading decimal number, correctly rounded
    (trailing characters are ignored); no number is an error."""
    m = _FLOAT_PREFIX.match(token)
    if not m:
        raise PbrtError(f"failed to parse [{token}] as a float")
    s = m.group(1)
    low = s.lower().lstrip("+-")
    if low.startswith("inf"):
        return f32(-np.inf) if s.startswith("-") else f32(np.inf)
    if low == "nan":
        return f32(np.nan)
    v = _round_f32(s)
    if not np.isfinite(v):
        raise PbrtError(f"[{token}] is out of range for a float")
    return v


def stoi(token: str) -> int:
    """std::stoi: the leading decimal integer; out of the int range is an error."""
    m = _INT_PREFIX.match(token)
    if not m:
        raise PbrtError(f"failed to parse [{token}] as an int")
    v = int(m.group(1))
    if not -(1 << 31) <= v < (1 << 31):
        raise PbrtError(f"[{token}] is out of range for an int")
    return v


_libm = ctypes.CDLL(ctypes.util.find_library("m") or "libm.so.6")
for _fn in ("sinf", "cosf"):
    getattr(_libm, _fn).restype = ctypes.c_float
    getattr(_libm, _fn).argtypes = [ctypes.c_float]


def sinf(x) -> np.float32:
    """The C library's sinf, as std::sin(float) calls it."""
    return f32(_libm.sinf(float(x)))


def cosf(x) -> np.float32:
    return f32(_libm.cosf(float(x)))


# ----------------------------------------------------------------------------- transforms

def mat_mul(a: np.ndarray, b: np.ndarray) -> np.ndarray:
    """a * b in float32 with each entry summed left to right
    (Matrix4x4.cpp:173-283)."""
    out = np.empty((4, 4), np.float32)
    for i in range(4):
        for j in range(4):
            acc = f32(a[i, 0] * b[0, j])
            for k in range(1, 4):
                acc = f32(acc + f32(a[i, k] * b[k, j]))
            out[i, j] = acc
    return out


def mat_inverse(m: np.ndarray) -> np.ndarray:
    """Matrix4x4::Inverse (Matrix4x4.cpp:112-170): Gauss-Jordan with full
    pivoting in float32.  The reference calls an unqualified abs() on the
    floats, which here resolves to the integer abs, so pivots are compared by
    their truncated magnitudes (int(|x|) >= big, big starting at 0): the last
    candidate of equal truncated size wins."""
    minv = np.array(m, np.float32)
    indxc, indxr, ipiv = [0] * 4, [0] * 4, [0] * 4
    for i in range(4):
        irow = icol = -1
        big = 0
        for j in range(4):
            if ipiv[j] != 1:
                for k in range(4):
                    if ipiv[k] == 0:
                        a = abs(int(minv[j, k])) if np.isfinite(minv[j, k]) else 0
                        if a >= big:
                            big, irow, icol = a, j, k
                    elif ipiv[k] > 1:
                        raise PbrtError("Singular matrix in MatrixInvert")
        ipiv[icol] += 1
        if irow != icol:
            minv[[irow, icol]] = minv[[icol, irow]]
        indxr[i], indxc[i] = irow, icol
        if minv[icol, icol] == 0:
            raise PbrtError("Singular matrix in MatrixInvert")
        pivinv = f32(1.0) / minv[icol, icol]
        minv[icol, icol] = f32(1.0)
        for j in range(4):
            minv[icol, j] = f32(minv[icol, j] * pivinv)
        for j in range(4):
            if j != icol:
                save = minv[j, icol]
                minv[j, icol] = f32(0.0)
                for k in range(4):
                    minv[j, k] = f32(minv[j, k] - f32(minv[icol, k] * save))
    for j in range(3, -1, -1):
        if indxr[j] != indxc[j]:
            minv[:, [indxr[j], indxc[j]]] = minv[:, [indxc[j], indxr[j]]]
    return minv


@dataclass(eq=False)
class Xf:
    """A reference Transform: matrix and the inverse it tracks.  Mutable and
    shared by reference, as the reference shares std::shared_ptr<Transform>."""
    m: np.ndarray = field(default_factory=lambda: np.eye(4, dtype=np.float32))
    inv: np.ndarray = field(default_factory=lambda: np.eye(4, dtype=np.float32))

    def copy(self) -> "Xf":
        return Xf(self.m.copy(), self.inv.copy())

    def compose(self, other: "Xf") -> None:
        """Transform::operator*= (Transform.cpp:61-67): both matrices right-multiplied."""
        self.m = mat_mul(self.m, other.m)
        self.inv = mat_mul(self.inv, other.inv)


def translate(x, y, z) -> Xf:                                   # Transform.cpp:248-264
    m = np.eye(4, dtype=np.float32)
    m[:3, 3] = [x, y, z]
    inv = np.eye(4, dtype=np.float32)
    inv[:3, 3] = [-f32(x), -f32(y), -f32(z)]
    return Xf(m, inv)


def scale(x, y, z) -> Xf:                                       # Transform.cpp:284-300
    m = np.diag([x, y, z, 1.0]).astype(np.float32)
    inv = np.diag([f32(1.0) / f32(x), f32(1.0) / f32(y), f32(1.0) / f32(z), f32(1.0)]).astype(np.float32)
    return Xf(m, inv)


def rotate(angle_deg, x, y, z) -> Xf:
    """rotate_directive (pbrt_parser.cpp:667-682) and Transform::Rotate
    (Transform.cpp:214-228), whose inverse is Matrix4x4::Inverse."""
    angle = f32(f32(angle_deg) * PI_OVER_180)
    x, y, z = f32(x), f32(y), f32(z)
    li = f32(1.0) / f32(np.sqrt(f32(f32(f32(x * x) + f32(y * y)) + f32(z * z))))
    x, y, z = f32(x * li), f32(y * li), f32(z * li)
    s, c = sinf(angle), cosf(angle)
    omc = f32(f32(1.0) - c)
    m = np.array([
        [c + x * x * omc, x * y * omc - z * s, x * z * omc + y * s, 0],
        [y * x * omc + z * s, c + y * y * omc, y * z * omc - x * s, 0],
        [z * x * omc - y * s, z * y * omc + x * s, c + z * z * omc, 0],
        [0, 0, 0, 1]], dtype=np.float32)
    # the expressions above in float32, left to right
    m[0, 0] = f32(c + f32(f32(x * x) * omc))
    m[0, 1] = f32(f32(f32(x * y) * omc) - f32(z * s))
    m[0, 2] = f32(f32(f32(x * z) * omc) + f32(y * s))
    m[1, 0] = f32(f32(f32(y * x) * omc) + f32(z * s))
    m[1, 1] = f32(c + f32(f32(y * y) * omc))
    m[1, 2] = f32(f32(f32(y * z) * omc) - f32(x * s))
    m[2, 0] = f32(f32(f32(z * x) * omc) - f32(y * s))
    m[2, 1] = f32(f32(f32(z * y) * omc) + f32(x * s))
    m[2, 2] = f32(c + f32(f32(z * z) * omc))
    return Xf(m, mat_inverse(m))


def look_at(eye, target, up) -> Xf:
    """Transform::LookAt(eye, lookAt, up, false) (Transform.cpp:322-365), whose
    inverse is Matrix4x4::Inverse."""
    from .scenes import look_at_lh
    m = look_at_lh(eye, target, up)          # identity * LookAt, which is LookAt itself
    return Xf(m, mat_inverse(m))


# ----------------------------------------------------------------------------- scanner and lexer

def scan(text: str) -> List[List[str]]:
    """One token list per directive (pbrt_parser.cpp:299-392): words split on
    whitespace, a word starting with '#' ends the line, a directive name starts
    a new list, other words go to the current one, and brackets are split off
    words (a single value in brackets loses its brackets)."""
    tokens: List[List[str]] = []
    target = -1
    for line in text.split("\n"):
        if line == "":
            continue
        tokens.append([])
        for word in line.split():
            if word.startswith("#"):
                break
            if word in DIRECTIVES:
                target += 1
            if target < 0 or target >= len(tokens):
                raise PbrtError(f"token [{word}] before the first directive, or two directives on one line")
            cur = tokens[target]
            first_b, last_b = word[0] == "[", word[-1] == "]"
            ti = len(cur) - 1
            if len(word) == 1:
                if last_b:
                    if ti >= 1 and cur[ti - 1] == "[":
                        del cur[-2]
                    else:
                        cur.append("]")
                else:
                    cur.append(word)
            elif first_b:
                if last_b:
                    cur.append(word[1:-1])
                else:
                    cur.append("[")
                    cur.append(word[1:])
            elif last_b:
                if ti >= 0 and cur[ti] == "[":
                    del cur[-1]
                    cur.append(word[:-1])
                else:
                    cur.append(word[:-1])
                    cur.append("]")
            else:
                cur.append(word)
        if not tokens[-1]:
            tokens.pop()
    return tokens


_ARG_TYPES = {"bool": "bool", "float": "float", "integer": "int", "normal": "normal", "normal3": "normal",
              "point2": "point2", "point": "point3", "point3": "point3", "rgb": "rgb", "color": "rgb",
              "spectrum": "spectrum", "string": "string", "vector2": "vector2", "vector": "vector3",
              "vector3": "vector3", "texture": "texture"}


@dataclass
class Arg:
    type: str
    name: str
    floats: Optional[List[np.float32]] = None
    ints: Optional[List[int]] = None
    string: Optional[str] = None
    boolean: Optional[bool] = None
    other: Optional[List[np.float32]] = None    # values of point / normal / vector arguments, which the
                                                # reference lexer drops; only point lights read them


@dataclass
class Directive:
    identifier: str
    type: str = ""
    name: str = ""
    class_name: str = ""
    args: List[Arg] = field(default_factory=list)


def _quoted(t: str) -> bool:
    return t[0] == '"' and t[-1] == '"'


def _start_quoted(t: str) -> bool:
    return t[0] == '"' and t[-1] != '"'


def _end_quoted(t: str) -> bool:
    return t[0] != '"' and t[-1] == '"'


def _float_array(line: List[str], n: int, brackets: bool) -> Arg:
    """LexFloatArrayArgument (pbrt_parser.cpp:394-428): exactly n values (the
    bracketed form reads one fewer, Q5)."""
    if len(line) != n + 1 + (2 if brackets else 0):
        raise PbrtError(f"{line[0]} requires exactly {n} arguments, but found {len(line)}")
    if brackets and (line[1] != "[" or line[-1] != "]"):
        raise PbrtError(f"{line[0]}: expected brackets around its values")
    start = 2 if brackets else 1
    return Arg("float", "", floats=[stof(line[i]) for i in range(start, n + 1)])


def lex(line: List[str]) -> Directive:
    """pbrt_parser::lex (pbrt_parser.cpp:430-665)."""
    d = Directive(line[0])
    if len(line) == 1:
        return d
    start = 2
    fixed = {"LookAt": (9, False), "Rotate": (4, False), "Scale": (3, False), "Transform": (16, True),
             "Translate": (3, False)}
    if d.identifier in fixed:
        n, br = fixed[d.identifier]
        d.args = [_float_array(line, n, br)]
        return d
    if d.identifier == "Texture":
        d.name, d.type, d.class_name = line[1][1:-1], line[2][1:-1], line[3][1:-1]
        if not d.name or not d.type or not d.class_name:
            raise PbrtError("Texture needs a name, a type and a class")
        start = 4
    elif _quoted(line[1]):
        d.type = line[1][1:-1]
    else:
        raise PbrtError(f"second token (identifier) {line[0]} isn't quoted, but should be")
    if len(line) == 2:
        return d
    in_value = False
    cur: Optional[Arg] = None
    i = start
    while i < len(line):
        tok = line[i]
        if _start_quoted(tok) and i + 1 < len(line) and _end_quoted(line[i + 1]):
            kind = tok[1:]
            if kind not in _ARG_TYPES:
                raise PbrtError(f"Given argument type [{kind}] is not a known (or implemented, yet) PBRT type")
            cur = Arg(_ARG_TYPES[kind], line[i + 1][:-1])
            d.args.append(cur)
            in_value = True
            i += 2
            continue
        if tok == "[":
            in_value = True
        elif tok == "]":
            in_value = False
        elif in_value:
            if cur is None:
                raise PbrtError(f"{line[0]}: a value before any argument")
            if _quoted(tok):
                value = tok[1:-1]
                if cur.type in ("texture", "string"):
                    cur.string = value
                elif cur.type == "bool":
                    if value not in ("true", "false"):
                        raise PbrtError(f"{line[0]}: failed to parse [{value}] as a bool")
                    cur.boolean = value == "true"
                else:
                    raise PbrtError(f"{line[0]}: a quoted value for a {cur.type} argument")
            elif cur.type in ("rgb", "float"):
                cur.floats = (cur.floats or []) + [stof(tok)]
            elif cur.type == "int":
                cur.ints = (cur.ints or []) + [stoi(tok)]
            elif cur.type in ("point2", "point3", "normal", "vector2", "vector3"):
                # the reference keeps no values for these types (pbrt_parser.cpp:651-653)
                try:
                    cur.other = (cur.other or []) + [stof(tok)]
                except PbrtError:
                    pass
            # unquoted tokens of string, texture, bool and spectrum arguments are ignored
        else:
            raise PbrtError(f"lex(): current line has invalid token [{tok}]")
        i += 1
    return d


# ----------------------------------------------------------------------------- meshes

@dataclass(eq=False)
class Geometry:
    """mesh_geometry (mesh.h:17-72): packed vertices, optional normals, faces."""
    x: List[np.ndarray] = field(default_factory=list)
    y: List[np.ndarray] = field(default_factory=list)
    z: List[np.ndarray] = field(default_factory=list)
    nx: List[np.ndarray] = field(default_factory=list)
    ny: List[np.ndarray] = field(default_factory=list)
    nz: List[np.ndarray] = field(default_factory=list)
    faces: List[np.ndarray] = field(default_factory=list)
    has_vertex_normals: bool = False
    # u_packed / v_packed: one entry per vertex a PLY file declares (not per
    # stored vertex: Q6 stores binary vertices twice but their u, v once), and
    # whether the file gave them
    u: List[np.ndarray] = field(default_factory=list)
    v: List[np.ndarray] = field(default_factory=list)
    uv_given: List[bool] = field(default_factory=list)
    has_vertex_uvs: bool = False

    def add_vertices(self, x, y, z, nx=None, ny=None, nz=None) -> None:
        self.x.append(np.asarray(x, np.float32))
        self.y.append(np.asarray(y, np.float32))
        self.z.append(np.asarray(z, np.float32))
        if nx is not None:
            self.nx.append(np.asarray(nx, np.float32))
            self.ny.append(np.asarray(ny, np.float32))
            self.nz.append(np.asarray(nz, np.float32))

    def add_faces(self, f: np.ndarray) -> None:
        self.faces.append(np.asarray(f, np.uint32).reshape(-1, 3))

    def add_uv(self, u, v, given: bool) -> None:
        self.u.append(np.asarray(u, np.float32))
        self.v.append(np.asarray(v, np.float32))
        self.uv_given.append(given)

    def uv(self) -> Optional[np.ndarray]:
        """(N, 2) texture coordinates when every vertex got them from its file."""
        if not self.has_vertex_uvs or not self.uv_given or not all(self.uv_given):
            return None
        return np.stack([np.concatenate(self.u), np.concatenate(self.v)], 1).astype(np.float32)

    def arrays(self):
        cat = lambda parts: np.concatenate(parts) if parts else np.zeros(0, np.float32)
        v = np.stack([cat(self.x), cat(self.y), cat(self.z)], 1).astype(np.float32)
        f = np.concatenate(self.faces) if self.faces else np.zeros((0, 3), np.uint32)
        n = None
        if self.has_vertex_normals and self.nx and sum(len(a) for a in self.nx) == len(v):
            n = np.stack([cat(self.nx), cat(self.ny), cat(self.nz)], 1).astype(np.float32)
        return v, f, n


_PLY_TYPES = {}
for _names, _t in ((("char", "int8", "int_8"), "char"), (("uchar", "u_char", "uint8", "u_int8", "uint_8", "u_int_8"), "uchar"),
                   (("short", "int16", "int_16"), "short"),
                   (("ushort", "u_short", "uint16", "u_int16", "uint_16", "u_int_16"), "ushort"),
                   (("int", "int32", "int_32"), "int"), (("uint", "u_int", "uint32", "u_int32", "uint_32", "u_int_32"), "uint"),
                   (("float", "float32", "float_32"), "float"), (("double", "float64", "float_64"), "double"),
                   (("list",), "list")):
    for _n in _names:
        _PLY_TYPES[_n] = _t

_VERTEX_PROPS = {"x": "x", "y": "y", "z": "z", "nx": "nx", "ny": "ny", "nz": "nz", "u": "u", "s": "u", "v": "v",
                 "t": "v"}


def read_ply(path: str, geom: Geometry) -> None:
    """ply_parser::parse_file (mesh_parsers.cpp:14-417) into geom (appending,
    as the reference appends to an object's geometry).  Every vertex property
    is read as a 4-byte float in binary files and parsed with stof in ASCII
    ones, whatever its declared type; faces must have 3 indices."""
    with open(path, "rb") as fh:
        data = fh.read()
    pos = 0

    def line() -> Optional[str]:
        nonlocal pos
        if pos >= len(data):
            return None
        e = data.find(b"\n", pos)
        e = len(data) if e < 0 else e
        s = data[pos:e].decode("latin-1")
        pos = e + 1
        return s

    first = line()
    if first is None or (first.split()[:1] != ["ply"]):
        raise PbrtError(f"{path}: Missing PLY header")
    fmt = (line() or "").split()
    if len(fmt) < 3 or fmt[0] != "format":
        raise PbrtError(f"{path}: Missing \"format\" header")
    if fmt[1] not in ("ascii", "binary_little_endian", "binary_big_endian"):
        raise PbrtError(f"{path}: Unsupported format [{fmt[1]}]")
    if fmt[2] != "1.0":
        raise PbrtError(f"{path}: Unsupported format version [{fmt[2]}]")
    mode = fmt[1]
    elements: List[dict] = []
    seen = set()
    while True:
        ln = line()
        if ln is None:
            raise PbrtError(f"{path}: no end_header")
        w = ln.split()
        word = w[0] if w else ""
        if word == "element":
            name = w[1] if len(w) > 1 else ""
            if not name:
                raise PbrtError(f"{path}: Element name may not be empty")
            count = stoi(w[2] if len(w) > 2 else "")
            if count < 0:
                raise PbrtError(f"{path}: Element [{name}] must declare a non-negative number of instances")
            elements.append({"kind": name if name in ("vertex", "face") else "other", "count": count, "props": []})
        elif word == "property":
            if not elements:
                raise PbrtError(f"{path}: Found property with no valid preceding element")
            el = elements[-1]
            ptype = _PLY_TYPES.get(w[1] if len(w) > 1 else "")
            if ptype is None:
                raise PbrtError(f"{path}: Unrecognized property type")
            if el["kind"] == "vertex":
                pname = _VERTEX_PROPS.get(w[2] if len(w) > 2 else "", "unknown")
                seen.add(pname)
                el["props"].append(pname)
            elif el["kind"] == "face":
                if ptype != "list":
                    raise PbrtError(f"{path}: Face property must have type list (for now)")
                if len(w) < 5 or _PLY_TYPES.get(w[2]) is None or _PLY_TYPES.get(w[3]) is None:
                    raise PbrtError(f"{path}: Face property list has an unknown type")
                if w[4] not in ("vertex_indices", "vertex_index", "vertices", "vertex"):
                    raise PbrtError(f"{path}: Face property list has unknown name [{w[4]}]")
                el["props"].append("list")
            else:
                raise PbrtError(f"{path}: Property outside of element")
        elif word == "end_header":
            break
    for axis in ("x", "y", "z"):
        if axis not in seen:
            raise PbrtError(f"{path}: Header missing element vertex {axis} property")
    nn = sum(a in seen for a in ("nx", "ny", "nz"))
    if nn not in (0, 3):
        raise PbrtError(f"{path}: Header has {nn} normal directions per vertex; must have all (3) or none (0)")
    if sum(a in seen for a in ("u", "v")) not in (0, 2):
        raise PbrtError(f"{path}: Header has one of u, v")
    has_normals = nn == 3
    has_uv = "u" in seen and "v" in seen
    geom.has_vertex_normals = has_normals
    geom.has_vertex_uvs = True        # set for every PLY file (mesh_parsers.cpp:320, 340)
    endian = ">" if mode == "binary_big_endian" else "<"
    for el in elements:
        if el["kind"] == "vertex":
            n, props = el["count"], el["props"]
            if mode == "ascii":
                vals = np.zeros((n, len(props)), np.float32)
                for i in range(n):
                    ln = line()
                    if ln is None:
                        raise PbrtError(f"{path}: Error reading vertex {i}")
                    words = ln.split()
                    for k in range(len(props)):
                        if k >= len(words):
                            raise PbrtError(f"{path}: vertex {i} has too few values")
                        vals[i, k] = stof(words[k])
            else:
                size = n * len(props) * 4
                if pos + size > len(data):
                    raise PbrtError(f"{path}: vertex data truncated")
                vals = np.frombuffer(data, endian + "f4", n * len(props), pos).reshape(n, len(props))
                pos += size
            col = {p: vals[:, k] for k, p in enumerate(props) if p != "unknown"}
            x, y, z = col["x"], col["y"], col["z"]
            if has_uv:
                geom.add_uv(col["u"], col["v"], True)
            else:
                geom.add_uv(np.zeros(n, np.float32), np.zeros(n, np.float32), False)
            if has_normals:
                geom.add_vertices(x, y, z, col["nx"], col["ny"], col["nz"])
            elif mode == "ascii":
                geom.add_vertices(x, y, z)
            else:
                # Q6: each vertex stored twice, in place (vertex i at 2i and 2i + 1)
                geom.add_vertices(np.repeat(x, 2), np.repeat(y, 2), np.repeat(z, 2))
        elif el["kind"] == "face":
            n = el["count"]
            if mode == "ascii":
                f = np.zeros((n, 3), np.uint32)
                for i in range(n):
                    ln = line()
                    if ln is None:
                        raise PbrtError(f"{path}: Failed to read face line")
                    words = ln.split()
                    if stoi(words[0] if words else "") != 3:
                        raise PbrtError(f"{path}: Face instance has wrong number of vertex indices (expected 3)")
                    f[i] = [stoi(words[k]) & 0xFFFFFFFF for k in (1, 2, 3)]
            else:
                rec = np.dtype([("n", "u1"), ("i", endian + "i4", 3)])
                size = n * 13
                if pos + size > len(data):
                    raise PbrtError(f"{path}: face data truncated")
                a = np.frombuffer(data, rec, n, pos)
                pos += size
                if n and (a["n"] != 3).any():
                    raise PbrtError(f"{path}: Face has wrong number of vertex indices (expected 3)")
                f = a["i"].astype(np.int64).astype(np.uint32)
            geom.add_faces(f)
        else:
            raise PbrtError(f"{path}: Unknown element type")


def read_obj(path: str, geom: Geometry) -> None:
    """obj_parser::parse_file (mesh_parsers.cpp:530-581): 'v x y z' and
    'f a b c' lines (1-based), others ignored."""
    xs, ys, zs, faces = [], [], [], []
    with open(path, "r", encoding="latin-1") as fh:
        for ln in fh:
            w = ln.split()
            if not w or w[0][0] == "#":
                continue
            if w[0][0] == "v":
                xs.append(stof(w[1])); ys.append(stof(w[2])); zs.append(stof(w[3]))
            elif w[0][0] == "f":
                faces.append([(stoi(w[k]) - 1) & 0xFFFFFFFF for k in (1, 2, 3)])
    geom.add_vertices(np.array(xs, np.float32), np.array(ys, np.float32), np.array(zs, np.float32))
    geom.add_faces(np.array(faces, np.uint32).reshape(-1, 3))


def sphere_geometry(meridians: int, parallels: int) -> Geometry:
    """SphereTesselator::Create (tesselators.cpp:10-84), with the duplicated
    north pole and the face indices that follow from it."""
    g = Geometry()
    xs, ys, zs = [f32(0), f32(0)], [f32(1), f32(1)], [f32(0), f32(0)]
    starts = [0]
    m_step = f32(TWO_PI / f32(meridians))
    p_step = f32(PI / f32(parallels + 1))
    for i in range(parallels):
        vert = f32(PI_OVER_2 - f32(p_step * f32(i + 1)))
        y, r = sinf(vert), cosf(vert)
        starts.append(i * meridians + 1)
        for j in range(meridians):
            h = f32(m_step * f32(j))
            xs.append(f32(cosf(h) * r)); ys.append(y); zs.append(f32(sinf(h) * r))
    xs.append(f32(0)); ys.append(f32(-1)); zs.append(f32(0))
    nv = len(xs)
    faces = [(v + 1, v, 0) for v in range(1, meridians)]
    faces.append((1, meridians, 0))
    for p in range(1, parallels):
        top, bot = starts[p], starts[p + 1]
        for m in range(1, meridians):
            tr, tl, bl, br = top + m - 1, top + m, bot + m - 1, bot + m
            faces.append((br, bl, tr))
            faces.append((tr, tl, br))
        tl, tr, bl, br = top, top + meridians - 1, bot, bot + meridians - 1
        faces.append((bl, br, tr))
        faces.append((tr, tl, bl))
    bot = starts[parallels]
    for v in range(meridians - 1):
        faces.append((v + bot, v + 1 + bot, nv - 1))
    faces.append((nv - 2, bot, nv - 1))
    g.add_vertices(np.array(xs, np.float32), np.array(ys, np.float32), np.array(zs, np.float32))
    g.add_faces(np.array(faces, np.uint32))
    return g


# ----------------------------------------------------------------------------- scene

@dataclass(eq=False)
class _Material:
    brdf: int                     # LAMBERT, MIRROR, GLOSSY, or -1: none the GPU path can render
    refl: Tuple[float, float, float]
    name: str = ""
    why: str = ""
    specular: Tuple[float, float, float] = (0.25, 0.25, 0.25)
    roughness: float = 0.1
    texture: Optional[Texture] = None


@dataclass(eq=False)
class _Texture:
    texture: Texture
    grey: bool                    # a "float" texture: one byte per texel


@dataclass(eq=False)
class _Mesh:
    o2w: Xf                       # shared with the scope that created the mesh
    w2o: np.ndarray               # snapshot at creation
    geom: Geometry
    material: Optional[_Material]
    light: Optional[list]         # shared [r, g, b]


@dataclass
class PointLight:
    """A LightSource "point" (the reference does not read these): world
    position and intensity in the same 0-255 units as the other lights."""
    position: Tuple[float, float, float]
    intensity: Tuple[float, float, float]


@dataclass
class SceneFile:
    meshes: List[MeshData]
    camera: Camera
    skybox: Optional[Tuple[float, float, float]]
    spp: int
    max_depth: int
    film_filename: str
    point_lights: List[PointLight]
    # meshes the GPU path can trace but not shade, with the reason (no
    # material, a texture without texture coordinates, ...); they carry the
    # default Lambert 0.5 material
    unrenderable: List[Tuple[int, str]] = field(default_factory=list)

    @property
    def lights(self) -> List[Tuple[Tuple[float, float, float], Tuple[float, float, float]]]:
        """The point lights in the form Scene(..., lights=) takes."""
        return [(l.position, l.intensity) for l in self.point_lights]


def _first_int(d: Directive) -> Optional[Arg]:
    for a in d.args:
        if a.type == "int":
            return a
    return None


def read_pbrt(path: str, strict: bool = True) -> SceneFile:
    """pbrt_parser::parse_file (pbrt_parser.cpp:275-292, 907-1721).  With
    strict, a mesh the GPU path cannot shade (see SceneFile.unrenderable) is
    an error; without, such meshes are kept for closest-hit queries."""
    try:
        with open(path, "r", encoding="latin-1") as fh:
            text = fh.read()
    except OSError as e:
        raise PbrtError(f"cannot read [{path}]: {e}") from e
    cut = path.rfind("/")
    base, fname = (path[: cut + 1], path[cut + 1:]) if cut >= 0 else ("", path)
    return _parse([lex(t) for t in scan(text)], base, fname, strict)


def read_pbrt_string(text: str, base: str = "", strict: bool = True) -> SceneFile:
    """pbrt_parser::parse_string (pbrt_parser.cpp:294-297); mesh files are
    looked up under `base` (a directory prefix ending in '/')."""
    return _parse([lex(t) for t in scan(text)], base, "", strict)


def _parse(directives: List[Directive], base: str, fname: str, strict: bool) -> SceneFile:
    scene_d, world_d = [], []
    cur = scene_d
    for d in directives:
        if d.identifier == "WorldBegin":
            cur = world_d
        cur.append(d)

    # sampler: the first Sampler directive's first integer argument (pbrt_parser.cpp:925-966)
    spp = DEFAULT_SAMPLES
    for d in scene_d:
        if d.identifier == "Sampler":
            a = _first_int(d)
            if a is not None and a.name == "pixelsamples":
                spp = a.ints[0]
            break

    # film: the first Film directive (pbrt_parser.cpp:968-1034)
    width, height = DEFAULT_WIDTH, DEFAULT_HEIGHT
    film_filename = "polytope.png"
    for d in scene_d:
        if d.identifier == "Film":
            if d.type == "image":
                dot = fname.rfind(".")
                film_filename = fname if dot < 0 else fname[:dot] + ".png"
                fx = fy = None
                for a in d.args:
                    if a.type == "int" and a.name == "xresolution":
                        fx = a.ints[0]
                    elif a.type == "int" and a.name == "yresolution":
                        fy = a.ints[0]
                    elif a.type == "string" and a.name == "filename":
                        film_filename = a.string
                width = fx if fx is not None else DEFAULT_WIDTH
                height = fy if fy is not None else DEFAULT_HEIGHT
            break

    # camera: every scene-block transform composed in order, then the first
    # perspective Camera (pbrt_parser.cpp:1082-1186)
    cam_xf = Xf()
    for d in scene_d:
        v = d.args[0].floats if d.args and d.args[0].floats is not None else None
        if d.identifier == "LookAt":
            cam_xf.compose(look_at(v[0:3], v[3:6], v[6:9]))
        elif d.identifier == "Rotate":
            cam_xf.compose(rotate(*v))
        elif d.identifier == "Scale":
            cam_xf.compose(scale(*v))
        elif d.identifier == "Translate":
            cam_xf.compose(translate(*v))
        elif d.identifier == "Transform":
            raise PbrtError("Transform: the reference's lexer keeps 15 of its 16 values (Q5)")
        elif d.identifier not in SCENE_DIRECTIVES and d.identifier != "WorldBegin":
            raise PbrtError(f"Indentifier [{d.identifier}] is not valid in the scene block")
    fov = DEFAULT_FOV
    for d in scene_d:
        if d.identifier == "Camera" and d.type == "perspective":
            for a in d.args:
                if a.type == "float" and a.name == "fov":
                    fov = a.floats[0]
            break

    # integrator: the first Integrator directive (pbrt_parser.cpp:1188-1233)
    max_depth = 5
    for d in scene_d:
        if d.identifier == "Integrator":
            if d.type == "path":
                for a in d.args:
                    if a.type == "int" and a.name == "maxdepth":
                        max_depth = a.ints[0]
                        break
            break

    # world (pbrt_parser.cpp:1237-1706)
    named: List[_Material] = []
    mat_stack: List[_Material] = []
    light_stack: List[list] = []
    xf_stack: List[Xf] = []
    material: Optional[_Material] = None
    light: Optional[list] = None
    xf = Xf()
    objects: Dict[str, Geometry] = {}
    cur_geom: Optional[Geometry] = None
    in_object = False
    sky = None
    meshes: List[_Mesh] = []
    points: List[PointLight] = []
    textures: Dict[str, _Texture] = {}

    def add_mesh(o2w: Xf, w2o: np.ndarray, geom: Geometry) -> None:
        meshes.append(_Mesh(o2w, w2o, geom, None if light is not None else material, light))

    for d in world_d:
        ident = d.identifier
        if ident not in WORLD_DIRECTIVES:
            continue
        v = d.args[0].floats if d.args and d.args[0].floats is not None else None
        if ident == "AreaLightSource":
            if d.type != "diffuse":
                continue
            for a in d.args:
                if a.name == "L":
                    if light is None:
                        light = [f32(0), f32(0), f32(0)]
                    light[0], light[1], light[2] = (f32(a.floats[k] * f32(255)) for k in range(3))
                    break
        elif ident == "AttributeBegin":
            if material is not None:
                mat_stack.append(material)
            if light is not None:
                light_stack.append(light)
            xf_stack.append(xf)
            xf = xf.copy()
        elif ident == "AttributeEnd":
            if mat_stack:
                material = mat_stack.pop()
            light = light_stack.pop() if light_stack else None
            if xf_stack:
                xf = xf_stack.pop()
        elif ident == "LightSource":
            if d.type == "infinite":
                for a in d.args:
                    if a.name == "L":
                        sky = tuple(float(f32(a.floats[k] * f32(255))) for k in range(3))
                        break
            elif d.type == "point":
                points.append(_point_light(d, xf))
        elif ident == "MakeNamedMaterial":
            brdf = None
            refl = (0.0, 0.0, 0.0)
            for a in d.args:
                if a.type == "string" and a.name == "type" and a.string == "matte":
                    brdf = LAMBERT                      # Q7: albedo 0.5, Kd ignored
                if a.type == "rgb" and a.name == "Kd":
                    refl = tuple(float(x) for x in a.floats[:3])
            if brdf is None:
                named.append(_Material(-1, refl, d.type, f"named material [{d.type}] has no BRDF"))
            else:
                named.append(_Material(LAMBERT, (0.5, 0.5, 0.5), d.type))
        elif ident == "Material":
            material = _material(d, textures)
        elif ident == "NamedMaterial":
            for m in named:
                if m.name == d.type:
                    material = m
                    break
        elif ident == "ObjectBegin":
            if in_object:
                raise PbrtError("ObjectBegin directive cannot be nested.")
            if not d.type:
                raise PbrtError("ObjectBegin directive cannot specify an empty name.")
            if d.type in objects:
                raise PbrtError(f"ObjectBegin: [{d.type}] has already been defined")
            in_object = True
            cur_geom = Geometry()
            objects[d.type] = cur_geom
        elif ident == "ObjectEnd":
            if not in_object:
                raise PbrtError("ObjectEnd directive without previous matching ObjectBegin directive")
            in_object = False
            cur_geom = None
        elif ident == "ObjectInstance":
            if in_object:
                raise PbrtError("ObjectInstance directive cannot appear between ObjectBegin and ObjectEnd")
            if d.type not in objects:
                raise PbrtError(f"ObjectInstance: [{d.type}] hasn't been defined yet")
            add_mesh(xf, xf.inv.copy(), objects[d.type])
        elif ident == "Rotate":
            xf.compose(rotate(*v))
        elif ident == "Scale":
            xf.compose(scale(*v))
        elif ident == "Translate":
            xf.compose(translate(*v))
        elif ident == "Transform":
            raise PbrtError("Transform: the reference's lexer keeps 15 of its 16 values (Q5)")
        elif ident == "TransformBegin":
            xf_stack.append(xf)
            xf = xf.copy()
        elif ident == "TransformEnd":
            if xf_stack:
                xf = xf_stack.pop()
        elif ident == "Texture":
            if d.name in textures:
                raise PbrtError(f"Texture name [{d.name}] has already previously been defined.")
            textures[d.name] = _texture(d, base)
        elif ident == "Shape":
            if d.type not in ("plymesh", "objmesh", "sphere"):
                continue
            w2o = xf.inv.copy()
            if d.type == "sphere":
                for a in d.args:
                    if a.type != "float":
                        continue
                    radius = a.floats[0]
                    n = max(int(radius), 20)
                    g = sphere_geometry(n, n)
                    if in_object:
                        cur_geom = g
                    else:
                        add_mesh(xf, w2o, g)
                continue
            filename = None
            for a in d.args:
                if a.name == "filename":
                    if a.string is None:
                        raise PbrtError(f"Shape {d.type}: filename is not a string")
                    filename = a.string
            if filename is None:
                continue
            g = cur_geom if in_object else Geometry()
            try:
                (read_ply if d.type == "plymesh" else read_obj)(base + filename, g)
            except OSError as e:
                raise PbrtError(f"Shape {d.type}: cannot read [{base + filename}]: {e}") from e
            if in_object:
                cur_geom = g
            else:
                add_mesh(xf, w2o, g)

    out: List[MeshData] = []
    unrenderable: List[Tuple[int, str]] = []
    for i, m in enumerate(meshes):
        v, f, n = m.geom.arrays()
        md = MeshData(v, f, object_to_world=m.o2w.m.copy(), normals=n, w2o=m.w2o.copy())
        if m.light is not None:
            md.is_light = True
            md.emission = tuple(float(c) for c in m.light)
        elif m.material is None:
            unrenderable.append((i, "no material (the reference would dereference a null material)"))
        elif m.material.brdf < 0:
            unrenderable.append((i, m.material.why))
        elif m.material.texture is not None and m.geom.uv() is None:
            unrenderable.append((i, "textured material on a mesh without texture coordinates in its file "
                                    "(the reference reads uninitialised ones)"))
        else:
            md.brdf = m.material.brdf
            md.refl = m.material.refl
            if md.brdf == GLOSSY:
                md.specular = m.material.specular
                md.roughness = m.material.roughness
            if m.material.texture is not None:
                md.texture = m.material.texture
                md.uv = m.geom.uv()
        out.append(md)
    if strict and unrenderable:
        i, why = unrenderable[0]
        raise PbrtError(f"shape {i}: {why} (read_pbrt(..., strict=False) keeps it for closest-hit queries)")
    cam = Camera(cam_xf.m.copy(), float(fov), int(width), int(height))
    return SceneFile(out, cam, sky, int(spp), int(max_depth), film_filename, points, unrenderable)


_TEXTURE_CACHE: Dict[Tuple[str, int, int], np.ndarray] = {}


def _load_png(path: str) -> np.ndarray:
    """read_png with a small cache keyed by the file's identity."""
    st = os.stat(path)
    key = (os.path.realpath(path), st.st_mtime_ns, st.st_size)
    img = _TEXTURE_CACHE.get(key)
    if img is None:
        img = read_png(path)
        if len(_TEXTURE_CACHE) >= 8:
            _TEXTURE_CACHE.pop(next(iter(_TEXTURE_CACHE)))
        _TEXTURE_CACHE[key] = img
    return img


def _texture(d: Directive, base: str) -> _Texture:
    """texture_directive (pbrt_parser.cpp:730-776): an imagemap PNG, decoded to
    RGBA for "color"/"spectrum" textures and to one grey byte per texel for
    "float" ones.  The file is looked up at base + '/' + filename."""
    if d.type in ("color", "spectrum"):
        grey = False
    elif d.type == "float":
        grey = True
    else:
        raise PbrtError(f"Unknown/unimplemented texture type [{d.type}]")
    if d.class_name != "imagemap":
        raise PbrtError(f"Unknown/unimplemented texture class [{d.class_name}]")
    for a in d.args:
        if a.type == "string" and a.name == "filename":
            path = base + "/" + (a.string or "")
            try:
                rgba = _load_png(path)
            except (OSError, PngError) as e:
                raise PbrtError(f"Failed to read/decode png file [{a.string}]: {e}") from e
            return _Texture(Texture(rgba, d.name), grey)
    raise PbrtError("Texture directives / arguments insufficient")


def _material(d: Directive, textures: Dict[str, _Texture]) -> Optional[_Material]:
    """material_directive (pbrt_parser.cpp:778-905): matte (rgb Kd or texture
    Kd), mirror (rgb Kr) and plastic.  Unknown types make no material (None);
    materials the GPU path cannot render are kept with brdf -1 and refused
    when a shape uses them."""
    if d.type == "matte":
        for a in d.args:
            if a.name != "Kd":
                continue
            if a.type == "rgb":
                return _Material(LAMBERT, tuple(float(x) for x in a.floats[:3]))
            if a.type == "texture":
                t = textures.get(a.string or "")
                if t is None:
                    raise PbrtError(f"Material references a texture [{a.string}] that hasn't been defined yet")
                if t.grey:
                    return _Material(-1, (1.0, 1.0, 1.0), why=f"texture [{a.string}] is a float texture (one byte "
                                     "per texel, which the reference reads as four)")
                return _Material(LAMBERT, (1.0, 1.0, 1.0), texture=t.texture)
        return _Material(-1, (0.0, 0.0, 0.0), why="Material matte without Kd makes no material in the reference")
    if d.type == "mirror":
        for a in d.args:
            if a.name == "Kr" and a.type == "rgb":
                return _Material(MIRROR, tuple(float(x) for x in a.floats[:3]))
        return _Material(-1, (0.0, 0.0, 0.0), why="Material mirror without rgb Kr makes no material in the reference")
    if d.type == "plastic":
        # pbrt_parser.cpp:791-842: Kd and Ks default to 0.25, roughness to 0.1;
        # only rgb Kd / Ks and a float roughness are read
        kd, ks, rough = [f32(0.25)] * 3, [f32(0.25)] * 3, f32(0.1)
        for a in d.args:
            if a.name == "Kd" and a.type == "rgb":
                kd = a.floats[:3]
            elif a.name == "Ks" and a.type == "rgb":
                ks = a.floats[:3]
            elif a.name == "roughness" and a.type == "float":
                rough = a.floats[0]
        return _Material(GLOSSY, tuple(float(x) for x in kd), specular=tuple(float(x) for x in ks),
                         roughness=float(rough))
    return None


def _point_light(d: Directive, xf: Xf) -> PointLight:
    """pbrt's LightSource "point" ("point from", "rgb I"), read here although
    the reference drops it: the position goes through the current transform."""
    pos = (0.0, 0.0, 0.0)
    inten = (1.0, 1.0, 1.0)
    raw = d.args
    for a in raw:
        if a.name == "I" and a.type == "rgb" and a.floats:
            inten = tuple(float(f32(a.floats[k] * f32(255))) for k in range(3))
    for a in raw:
        if a.name == "from" and a.other:
            pos = tuple(float(x) for x in a.other[:3])
    p = np.array([*pos, 1.0], np.float64)
    w = xf.m.astype(np.float64) @ p
    return PointLight((float(f32(w[0])), float(f32(w[1])), float(f32(w[2]))), inten)
