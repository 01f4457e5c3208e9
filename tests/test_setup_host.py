"""CPU checks of the device set-ups' building blocks: the exp with the C library's bits
(chimera_amd/csrc/libm_exp.hpp, built for the host from the same source the device
kernels use) against the host's libm, its 2^(k/128) table against the generator, and
the set-up entry points' argument checks (host only, no GPU)."""
from __future__ import annotations

import ctypes as C
import importlib.util
import math
import shutil
import subprocess
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
CSRC = ROOT / "chimera_amd" / "csrc"


def _generator():
    spec = importlib.util.spec_from_file_location("gen_exp_table", CSRC / "gen_exp_table.py")
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


def test_exp_table_is_the_generators():
    assert (CSRC / "exp_table.hpp").read_text() == _generator().render()


def test_exp_table_entries_are_powers_of_two():
    """H_k (1 + T_k) = 2^(k/128) to within the rounding of H_k and T_k (closed form,
    independent of the generator's arithmetic)."""
    import struct
    rows = _generator().table()
    for k in range(128):
        t = struct.unpack("<d", struct.pack("<Q", rows[2 * k]))[0]
        h = struct.unpack("<d", struct.pack("<Q", (rows[2 * k + 1] + ((k << 52) // 128))
                                            & 0xFFFFFFFFFFFFFFFF))[0]
        assert 1.0 <= h < 2.0
        assert abs(h - 2.0 ** (k / 128)) <= 2.0 ** -52            # nearest double
        assert abs(t) <= 2.0 ** -53                               # a sub-ulp correction
        assert math.isclose(h + h * t, 2.0 ** (k / 128), rel_tol=4e-16)


def test_exp_has_the_host_libm_bits(tmp_path):
    """Every exponent argument of the Kelvin-Helmholtz set-up (N = 4 ... 2048) and 4M
    random arguments of the supported range give libm's exp bit for bit.  (A restatement
    without the fused multiply-adds misses on ~0.05% of the arguments.)"""
    hipcc = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    if not Path(hipcc).exists():
        pytest.skip("hipcc not available")
    exe = tmp_path / "libm_exp_check"
    subprocess.run([hipcc, "-std=c++17", "-O1", "-ffp-contract=off", "-I", str(CSRC),
                    str(ROOT / "tests" / "host" / "libm_exp_check.hip"), "-o", str(exe)],
                   check=True, capture_output=True, timeout=300)
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=120)
    print(r.stdout)
    lines = r.stdout.strip().splitlines()
    assert r.returncode == 0 and len(lines) == 8
    for line in lines:
        name, count, _, bad = line.split()
        assert int(count) > 0 and int(bad) == 0, line


@pytest.fixture(scope="module")
def L():
    from chimera_amd import _lib
    return _lib.lib()


def test_problem_info(L):
    from chimera_amd._lib import BC
    dx, gamma, bc = C.c_double(), C.c_double(), BC()
    want = {1: (5.0 / 3.0, 1), 2: (1.4, 1), 3: (5.0 / 3.0, 1), 4: (1.4, 3)}
    for p, (g, flag) in want.items():
        assert L.chimera_problem_info(p, 96, C.byref(dx), C.byref(gamma), C.byref(bc)) == 0
        assert dx.value == 1.0 / 96 and gamma.value == g
        assert (bc.xlo, bc.xhi, bc.ylo, bc.yhi) == (flag,) * 4
    assert L.chimera_problem_info(2, 8, None, None, None) == 0


def test_setup_arguments_rejected_on_host(L):
    fake = (C.c_void_p * 6)(*([16] * 6))        # never dereferenced: rejected first
    for p, n, what in ((5, 64, b"GRF"), (0, 64, b"1-5"), (6, 64, b"1-5"), (1, 3, b"N = 3")):
        assert L.chimera_problem_info(p, n, None, None, None) == -1
        assert L.chimera_setup_problem(p, n, fake, None) == -1
        assert what in L.chimera_last_error()
    assert L.chimera_setup_problem(1, 64, None, None) == -1
    assert L.chimera_ctx_setup_problem(None, 1) == -1
    assert b"NULL" in L.chimera_last_error()
