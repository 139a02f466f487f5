"""Loading of the golden fixtures written by tests/golden/make_golden.py."""
from __future__ import annotations

from pathlib import Path

import numpy as np

from polytope_amd import scenes as S

GOLDEN = Path(__file__).resolve().parent / "golden"
NAMES = ["ico3", "ico5", "plain_multimesh", "sliver48"]


def load(name: str) -> dict:
    with np.load(GOLDEN / f"{name}.npz", allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


def meshes_of(d: dict):
    out = []
    for i in range(int(d["n_meshes"])):
        mat = d[f"m{i}_mat"]
        out.append(S.MeshData(d[f"m{i}_verts"], d[f"m{i}_faces"], object_to_world=d[f"m{i}_o2w"],
                              brdf=int(mat[0]), refl=tuple(float(x) for x in mat[1:4]),
                              is_light=bool(mat[4] > 0), emission=tuple(float(x) for x in mat[5:8])))
    return out


def camera_of(d: dict):
    c = d["camera"]
    cam = S.Camera(c[:16].reshape(4, 4), float(c[16]), int(c[17]), int(c[18]))
    return cam, int(c[19]), int(c[20])


def single_mesh(d: dict) -> bool:
    return int(d["n_meshes"]) == 1
