"""Snapshot files in the layout of CHIMERA's HDF5 writer (src/mhd_write_h5.f90:12-177),
through libchimera_h5.so (HDF5 C API; no h5py needed).

    write_snapshots(path, fields, times, gamma, M_s, beta)   # fields: name -> (n, ny, nx)
    f = read(path)    # {"rho": [(ny, nx) per snapshot], ..., "time": (n,), "parameters": {...}}
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

import numpy as np

FIELDS = ("rho", "P", "Bx", "By", "Vx", "Vy")
PARAMS = ("gamma", "M_s", "beta")
_LIB = Path(__file__).resolve().parent / "libchimera_h5.so"
_lib = None


class H5Error(RuntimeError):
    pass


def lib():
    global _lib
    if _lib is None:
        if not _LIB.exists():
            raise H5Error(f"{_LIB} is not built (chimera_amd._build.build_h5; needs the HDF5 C library)")
        L = C.CDLL(str(_LIB))
        dp = C.POINTER(C.c_double)
        L.chimera_h5_write_snapshots.argtypes = [C.c_char_p, C.c_int, C.c_int, C.c_int] + [dp] * 7 + \
            [C.c_double] * 3
        L.chimera_h5_group_count.argtypes = [C.c_char_p, C.c_char_p, C.POINTER(C.c_int)]
        L.chimera_h5_dataset_dims.argtypes = [C.c_char_p, C.c_char_p, C.POINTER(C.c_int),
                                              C.POINTER(C.c_longlong)]
        L.chimera_h5_read_dataset.argtypes = [C.c_char_p, C.c_char_p, dp, C.c_longlong]
        L.chimera_h5_last_error.restype = C.c_char_p
        _lib = L
    return _lib


def _check(rc: int, what: str):
    if rc != 0:
        raise H5Error(f"{what}: {lib().chimera_h5_last_error().decode()} (rc {rc})")


def _dp(a: np.ndarray):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def write_snapshots(path, fields: dict, times, gamma: float, M_s: float, beta: float) -> None:
    """fields: the six primitives (P = thermal pressure), each (nsnap, ny, nx) float64."""
    arrs = [np.ascontiguousarray(fields[k], dtype=np.float64) for k in FIELDS]
    if arrs[0].ndim != 3:
        raise ValueError(f"fields: need (nsnap, ny, nx) arrays, got shape {arrs[0].shape}")
    nsnap, ny, nx = arrs[0].shape
    for k, a in zip(FIELDS, arrs):
        if a.shape != (nsnap, ny, nx):
            raise ValueError(f"field {k} of shape {a.shape}: need {(nsnap, ny, nx)}")
    t = np.ascontiguousarray(times, dtype=np.float64)
    if t.shape != (nsnap,):
        raise ValueError(f"times of shape {t.shape}: need ({nsnap},)")
    _check(lib().chimera_h5_write_snapshots(str(path).encode(), nsnap, nx, ny, *map(_dp, arrs), _dp(t),
                                            float(gamma), float(M_s), float(beta)),
           "chimera_h5_write_snapshots")


def group_count(path, group: str) -> int:
    n = C.c_int()
    _check(lib().chimera_h5_group_count(str(path).encode(), group.encode(), C.byref(n)), group)
    return n.value


def dims(path, dataset: str) -> tuple:
    r = C.c_int()
    d = (C.c_longlong * 2)()
    _check(lib().chimera_h5_dataset_dims(str(path).encode(), dataset.encode(), C.byref(r), d), dataset)
    return tuple(int(d[k]) for k in range(r.value))


def read_dataset(path, dataset: str) -> np.ndarray:
    shape = dims(path, dataset)
    out = np.empty(shape, dtype=np.float64)
    _check(lib().chimera_h5_read_dataset(str(path).encode(), dataset.encode(), _dp(out), out.size),
           dataset)
    return out


def read(path) -> dict:
    """Every snapshot (in SNAPSHOT order), the times and the parameters."""
    out = {}
    for f in FIELDS:
        n = group_count(path, f)
        out[f] = [read_dataset(path, f"{f}/SNAPSHOT{k + 1}") for k in range(n)]
    out["time"] = read_dataset(path, "time/sim_time")
    out["parameters"] = {p: float(read_dataset(path, f"parameters/{p}")[0]) for p in PARAMS}
    return out
