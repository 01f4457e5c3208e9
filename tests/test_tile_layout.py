"""Tile bookkeeping of the fused stage kernel, checked on the host: the face-run table
and the LDS layout of every tile shape (tests/host/tile_layout_check.hip, built with
hipcc as a host program; no GPU needed)."""
from __future__ import annotations

import shutil
import subprocess
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]


def test_face_runs_and_lds_layout(tmp_path):
    hipcc = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    if not Path(hipcc).exists():
        pytest.skip("hipcc not available")
    exe = tmp_path / "tile_layout_check"
    subprocess.run([hipcc, "-std=c++17", "-O1", "-I", str(ROOT / "chimera_amd" / "csrc"),
                    str(ROOT / "tests" / "host" / "tile_layout_check.hip"), "-o", str(exe)],
                   check=True, capture_output=True, timeout=300)
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=60)
    print(r.stdout)
    lines = r.stdout.strip().splitlines()
    assert r.returncode == 0 and len(lines) == 5
    assert all(line.endswith("errors 0") for line in lines)
    # two 32 x 16 blocks must fit one CU's 160 KiB of LDS
    lds = {l.split(":")[0]: int(l.split()[2]) for l in lines}
    assert 2 * lds["32x16"] <= 160 * 1024
