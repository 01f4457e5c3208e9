"""Shared fixtures.  GPU tests are marked ``gpu``; everything else runs on the CPU."""
from __future__ import annotations

import subprocess
import sys
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parent.parent
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library)")
    config.addinivalue_line("markers", "slow: long-running")


@pytest.fixture(scope="session")
def orc():
    """The CPU oracle (oracle/liborc.so), built on demand with gcc."""
    from oracle.oracle import COracle, LIBORC
    if not LIBORC.exists():
        subprocess.run(["make", "-s", "-C", str(ROOT / "oracle")], check=True)
    return COracle()


@pytest.fixture(scope="session")
def gpu_lib():
    """The HIP library; GPU tests fail loudly when it is missing."""
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU visible")
    from chimera_amd import _lib
    return _lib.lib()


def rel_l1(a: np.ndarray, b: np.ndarray) -> float:
    den = float(np.sum(np.abs(b)))
    num = float(np.sum(np.abs(a - b)))
    return num / den if den > 0 else num


def rel_l1_fields(a: np.ndarray, b: np.ndarray) -> list[float]:
    return [rel_l1(a[q], b[q]) for q in range(a.shape[0])]
