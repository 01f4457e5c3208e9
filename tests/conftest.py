"""Shared fixtures.  GPU tests are marked ``gpu``; everything else runs on the CPU."""
from __future__ import annotations

import os
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


# ---- programs that write the reference's snapshot files ---------------------------------

REF_MAIN = ROOT / "oracle" / "_ref" / "chimera_ref_main"


def _h5_name(problem_type, n, seed, extra) -> str:
    tag = "" if extra is None else "_" + "_".join(f"{x:g}" for x in extra)
    return f"p{problem_type}_N{n}_s{seed}{tag}.h5"


def ref_env() -> dict:
    """Environment for the reference's OpenMP executables: idle threads sleep instead of
    spinning, so that they do not starve other test processes on a shared CPU."""
    env = dict(os.environ)
    env.setdefault("OMP_WAIT_POLICY", "passive")
    return env


@pytest.fixture(scope="session")
def ref_main_run(tmp_path_factory):
    """run(problem_type, N, seed=0, extra=None) -> HDF5 path of the reference's own main
    program (oracle/_ref/chimera_ref_main: its modules on the CPU, writing through
    libchimera_h5); run.stdout[key] keeps what it printed.  Cached per argument set."""
    if not REF_MAIN.exists():
        pytest.skip("oracle/_ref/chimera_ref_main is not built")
    base = tmp_path_factory.mktemp("ref_main")
    cache, stdout = {}, {}

    def run(problem_type, n, seed=0, extra=None):
        key = (problem_type, n, seed, extra)
        if key not in cache:
            name = _h5_name(*key)
            argv = [str(REF_MAIN), str(problem_type), str(n), str(seed), f"{base}/", name]
            argv += [f"{x:g}" for x in extra or ()]
            r = subprocess.run(argv, capture_output=True, text=True, timeout=900, env=ref_env())
            assert r.returncode == 0, r.stderr
            cache[key], stdout[key] = base / name, r.stdout
        return cache[key]

    run.stdout = stdout
    return run


@pytest.fixture(scope="session")
def cli_run(gpu_lib, tmp_path_factory):
    """The same signature for chimera_amd.cli (the resident GPU path); run.stdout[key]
    keeps its progress lines."""
    from chimera_amd import cli
    base = tmp_path_factory.mktemp("cli")
    cache, stdout = {}, {}

    def run(problem_type, n, seed=0, extra=None):
        key = (problem_type, n, seed, extra)
        if key not in cache:
            name = _h5_name(*key)
            print(f"[cli] problem {problem_type} N={n} seed={seed} extra={extra}", flush=True)
            lines = []
            ms, beta = extra or (-1.0, -1.0)
            opts = {"problem_type": problem_type, "N": n, "seed": seed, "out_path": f"{base}/",
                    "h5_filename": name, "target_M_s": ms, "target_beta": beta}
            cache[key] = Path(cli.run(opts, out=lines.append))
            stdout[key] = "\n".join(lines)
        return cache[key]

    run.stdout = stdout
    return run
