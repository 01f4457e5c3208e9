"""The stage kernel's short FP64 division and square root (the Newton core of hipcc's
IEEE forms without div_scale / div_fixup and without the sqrt pre-scale, used by Rusanov
and HLLE in tiles whose data is in range) give the bits of a / b and sqrt(x): checked on
the GPU over random operands with edge mantissas, in the general range the kernels
claim and in the ranges the fluxes produce (tests/host/short_forms_check.hip)."""
from __future__ import annotations

import ctypes
from pathlib import Path

import pytest

LIB = Path(__file__).resolve().parent / "host" / "libshort_forms_check.so"
N = 1 << 24

# (name, mode, exponent range of a / x, exponent range of b, random signs)
CASES = [
    ("quotients, |a|, |b|, |a/b| in [2^-900, 2^700]", 0, (-900, 700), (-900, 700), 1),
    ("flux quotients: a in [2^-40, 2^234], b in [2^-27, 2^66]", 0, (-40, 234), (-27, 66), 0),
    ("reciprocals: a in [1, 2), b in [2^-40, 2^68]", 0, (0, 0), (-40, 68), 0),
    ("wave-uniform divisor (P_th / (g - 1))", 2, (-40, 133), (-100, 100), 0),
    ("exact quotients a = b k", 3, (0, 0), (-300, 300), 1),
    ("sqrt, x in [2^-700, 2^700]", 1, (-700, 700), (0, 0), 0),
    ("sqrt, x in [2^-106, 2^261] (wave speeds)", 1, (-106, 261), (0, 0), 0),
]


@pytest.fixture(scope="module")
def check():
    if not LIB.exists():
        pytest.fail(f"{LIB.name} is not built (run __graft_entry__.build())")
    # torch first, so that the library binds to the HIP runtime torch loads
    import torch
    assert torch.cuda.is_available()
    torch.cuda.set_device(0)
    lib = ctypes.CDLL(str(LIB))
    f = lib.short_forms_check
    f.restype = ctypes.c_int
    f.argtypes = [ctypes.c_ulonglong, ctypes.c_longlong, ctypes.c_int, ctypes.c_int,
                  ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                  ctypes.POINTER(ctypes.c_ulonglong), ctypes.POINTER(ctypes.c_ulonglong)]
    return f


@pytest.mark.gpu
@pytest.mark.parametrize("name,mode,ea,eb,sign", CASES, ids=[c[0].split(":")[0] for c in CASES])
def test_short_forms_bit_exact(check, name, mode, ea, eb, sign):
    bad, used = ctypes.c_ulonglong(), ctypes.c_ulonglong()
    seed = 0x243F6A8885A308D3 ^ (mode << 8) ^ (ea[0] & 0xFFFF)
    rc = check(seed, N, mode, ea[0], ea[1], eb[0], eb[1], sign, ctypes.byref(bad),
               ctypes.byref(used))
    print(f"{name}: {used.value} operands, {bad.value} differ")
    assert rc == 0
    assert used.value >= N // 4
    assert bad.value == 0
