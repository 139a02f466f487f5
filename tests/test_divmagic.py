"""The render kernel's division by a chunk's pixel count
(polytope_amd/csrc/divmagic.h, Granlund and Montgomery 1994, Fig. 4.1) equals
integer division: the header is compiled for the host and checked against `/`
on exhaustive small cases, boundaries and random 32-bit pairs."""
from __future__ import annotations

import shutil
import subprocess
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]


def test_division_by_invariant_divisor_is_exact(tmp_path):
    hipcc = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    if not Path(hipcc).exists():
        pytest.skip("hipcc not found")
    exe = tmp_path / "divmagic_check"
    subprocess.run([hipcc, "-O2", "-std=c++17", "-x", "hip", "--offload-arch=gfx950",
                    f"-I{ROOT / 'polytope_amd' / 'csrc'}", str(ROOT / "tests" / "divmagic_check.cpp"),
                    "-o", str(exe)], check=True, capture_output=True, timeout=300)
    out = subprocess.run([str(exe)], capture_output=True, text=True, timeout=300)
    assert out.returncode == 0, out.stdout
    assert " 0 mismatches" in out.stdout
