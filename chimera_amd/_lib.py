"""ctypes binding of libchimera_hip.so (the C ABI declared in include/chimera_hip.h).

The library is required: importing the GPU entry points without it raises
immediately (there is no CPU fallback in the product path).
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

import numpy as np

PKG = Path(__file__).resolve().parent
LIB_PATH = PKG / "libchimera_hip.so"

c_dp = C.POINTER(C.c_double)
c_u32p = C.POINTER(C.c_uint32)


class BC(C.Structure):
    """chimera_bc_t: per-side flags (1 periodic, 2 fixed, 3 outflow, 4 inflow)."""
    _fields_ = [("xlo", C.c_int), ("xhi", C.c_int), ("ylo", C.c_int), ("yhi", C.c_int)]

    @classmethod
    def of(cls, bc):
        if isinstance(bc, BC):
            return bc
        return cls(*[int(v) for v in bc])


class Scheme(C.Structure):
    """chimera_scheme_t (the compile-time choices of src/mhd_config.f90:27-37)."""
    _fields_ = [("riemann_solver", C.c_int), ("slope_limiter", C.c_int),
                ("use_slope_limiting", C.c_int), ("fallback_2_mood", C.c_int)]

    @classmethod
    def of(cls, s):
        if isinstance(s, Scheme):
            return s
        s = list(s)
        if len(s) == 2:
            s += [1, 1]
        if len(s) != 4:
            raise ValueError("scheme is (solver, limiter[, use_slope_limiting, mood])")
        return cls(*[int(v) for v in s])


class ChimeraError(RuntimeError):
    pass


def state_ptrs(x, shape=None, what: str = "state"):
    """The 6 field pointers of a (6, ny, nx) FP64 state, after checking it: a
    C-contiguous numpy array, or a contiguous CUDA tensor on the current device.  The
    library copies or writes nx * ny doubles per field, so a wrong shape, dtype or
    layout is rejected here (ValueError / TypeError) before any copy."""
    if isinstance(x, np.ndarray):
        if x.dtype != np.float64 or not x.flags.c_contiguous:
            raise ValueError(f"{what}: need a C-contiguous float64 array (got {x.dtype})")
        ptrs = [x[q].ctypes.data for q in range(x.shape[0])] if x.ndim == 3 else None
    else:
        import torch
        if not isinstance(x, torch.Tensor):
            raise TypeError(f"{what}: need a numpy array or a torch tensor, got {type(x).__name__}")
        if not x.is_cuda or x.dtype != torch.float64 or not x.is_contiguous():
            raise ValueError(f"{what}: need a contiguous float64 CUDA tensor "
                             f"(got {x.dtype} on {x.device}, contiguous={x.is_contiguous()})")
        if x.device.index != torch.cuda.current_device():
            raise ValueError(f"{what}: tensor on {x.device}, current device is "
                             f"cuda:{torch.cuda.current_device()}")
        ptrs = [x[q].data_ptr() for q in range(x.shape[0])] if x.ndim == 3 else None
    if x.ndim != 3 or x.shape[0] != 6 or (shape is not None and tuple(x.shape) != tuple(shape)):
        want = tuple(shape) if shape is not None else "(6, ny, nx)"
        raise ValueError(f"{what}: shape {tuple(x.shape)}, need {want}")
    return (C.c_void_p * 6)(*ptrs)


_lib = None


def lib() -> C.CDLL:
    """Load the HIP library (raises when it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not LIB_PATH.exists():
        raise ChimeraError(
            f"{LIB_PATH} is missing: build it with `python -m chimera_amd._build` "
            "(or __graft_entry__.build()); there is no CPU fallback")
    L = C.CDLL(str(LIB_PATH))
    _declare(L)
    _lib = L
    return L


def check(rc: int, what: str = "") -> None:
    if rc != 0:
        msg = lib().chimera_last_error().decode(errors="replace")
        raise ChimeraError(f"{what} failed with status {rc}: {msg}")


def _declare(L: C.CDLL) -> None:
    i, d, vp = C.c_int, C.c_double, C.c_void_p
    pBC, pSch = C.POINTER(BC), C.POINTER(Scheme)
    L.chimera_last_error.restype = C.c_char_p
    L.chimera_last_error.argtypes = []
    sigs = {
        "chimera_device_count": [C.POINTER(i)],
        "chimera_fill_ghost_cells": [c_dp, c_dp, c_dp, i, i, pBC],
        "chimera_compute_gradients": [c_dp, d, i, i, c_dp, c_dp],
        "chimera_apply_slope_limiter": [c_dp, d, i, i, c_dp, c_dp, i],
        "chimera_reconstruction": [c_dp, c_dp, c_dp, d, i, i, c_dp, c_dp, c_dp, c_dp, i, pBC, i],
        "chimera_thermal_pressure_check": [i, i] + [c_dp] * 30,
        "chimera_compute_fluxes": [c_dp] * 12 + [d, i, i] + [c_dp] * 5 + [i],
        "chimera_update_conserved": [c_dp, c_dp, c_dp, d, d, i, i, pBC],
        "chimera_constrained_transport": [c_dp, c_dp, c_dp, c_dp, d, d, i, i, pBC],
        "chimera_compute_curl_2d": [c_dp, d, i, i, c_dp, c_dp, pBC],
        "chimera_compute_divB": [c_dp, c_dp, d, i, i, c_dp, pBC],
        "chimera_average_face_to_cell_B": [c_dp, c_dp, i, i, c_dp, c_dp, pBC],
        "chimera_get_conserved": [c_dp] * 6 + [d, d, i, i] + [c_dp] * 4,
        "chimera_get_primitive": [c_dp] * 6 + [d, d, i, i] + [c_dp] * 4,
        "chimera_apply_conserved_floors": [c_dp] * 6 + [d, d, i, i],
        "chimera_stage_fluxes": [c_dp, c_dp, c_u32p, d, d, i, i, pBC, pSch],
        "chimera_dev_stage_fluxes": [vp, vp, vp, d, d, i, i, pBC, pSch, vp],
        "chimera_dev_compute_fluxes": [vp, vp, d, i, i, vp],
        "chimera_ctx_create": [C.POINTER(vp), i, i, d, d, pBC, pSch, vp],
        "chimera_ctx_destroy": [vp],
        "chimera_ctx_upload": [vp, C.POINTER(vp)],
        "chimera_problem_info": [i, i, C.POINTER(d), C.POINTER(d), pBC],
        "chimera_setup_problem": [i, i, C.POINTER(vp), vp],
        "chimera_ctx_setup_problem": [vp, i],
        "chimera_ctx_download": [vp, C.POINTER(vp)],
        "chimera_ctx_steps": [vp, i, i, c_dp, i],
        "chimera_ctx_get_time": [vp, C.POINTER(d), C.POINTER(i)],
        "chimera_ctx_set_time": [vp, d, i],
        "chimera_ctx_last_dt": [vp, C.POINTER(d)],
        "chimera_ctx_cfl_dt": [vp, i, C.POINTER(d)],
        "chimera_ctx_divb": [vp, C.POINTER(d), C.POINTER(d)],
        "chimera_ctx_stage1_mask": [vp, c_u32p],
        "chimera_ctx_sync": [vp],
        "chimera_ctx_set_timing": [vp, i],
        "chimera_ctx_stage_kernel_ms": [vp, C.POINTER(d), C.POINTER(i)],
        "chimera_ctx_set_tile_shape": [vp, i],
        "chimera_ctx_get_tile_shape": [vp, C.POINTER(i), C.POINTER(i), C.POINTER(i)],
        "chimera_ctx_autotune": [vp, i, C.POINTER(i), c_dp],
        "chimera_batch_create": [C.POINTER(vp), i, C.POINTER(i), C.POINTER(i), c_dp, c_dp, pBC,
                                 pSch, vp],
        "chimera_batch_destroy": [vp],
        "chimera_batch_upload": [vp, i, C.POINTER(vp)],
        "chimera_batch_download": [vp, i, C.POINTER(vp)],
        "chimera_batch_steps": [vp, i, i, i],
        "chimera_batch_member_time": [vp, i, C.POINTER(d), C.POINTER(i), C.POINTER(d)],
        "chimera_batch_sync": [vp],
        "chimera_batch_cells": [vp, C.POINTER(C.c_longlong)],
        "chimera_batch_set_tiling": [vp, i, i],
        "chimera_batch_get_tiling": [vp, C.POINTER(i), C.POINTER(i)],
        "chimera_batch_autotune": [vp, i, C.POINTER(i), C.POINTER(i)],
        "chimera_batch_set_timing": [vp, i],
        "chimera_batch_stage_kernel_ms": [vp, C.POINTER(d), C.POINTER(i)],
        "chimera_adj_create": [C.POINTER(vp), i, i, d, d, pBC, pSch, vp],
        "chimera_adj_destroy": [vp],
        "chimera_adj_forward_step": [vp, C.POINTER(vp), d, C.POINTER(vp)],
        "chimera_adj_step_vjp": [vp, C.POINTER(vp), d, C.POINTER(vp), C.POINTER(vp)],
        "chimera_adj_segment_vjp": [vp, C.POINTER(vp), c_dp, i, C.POINTER(vp)],
        "chimera_adj_sweep": [vp, C.POINTER(vp), c_dp, i, i, C.POINTER(vp)],
        "chimera_adj_record": [vp, C.POINTER(vp), c_dp, i, i, i, C.POINTER(vp)],
        "chimera_adj_set_fused": [vp, i],
        "chimera_adj_set_tiles": [vp, i, i],
        "chimera_face_vjp": [i, i, i, C.c_double, c_dp, c_dp, c_dp],
        "chimera_badj_create": [C.POINTER(vp), i, C.POINTER(C.c_int), C.POINTER(C.c_int), c_dp, c_dp,
                                vp, vp, vp],
        "chimera_badj_destroy": [vp],
        "chimera_badj_forward": [vp, C.POINTER(vp), c_dp, i, C.POINTER(vp)],
        "chimera_badj_segment_vjp": [vp, C.POINTER(vp), c_dp, i, C.POINTER(vp)],
        "chimera_badj_sweep": [vp, C.POINTER(vp), c_dp, i, i, C.POINTER(vp)],
        "chimera_badj_record": [vp, C.POINTER(vp), c_dp, i, i, i, C.POINTER(vp)],
        "chimera_badj_set_tiles": [vp, i, i, i],
        "chimera_badj_autotune": [vp, C.POINTER(vp), c_dp, i, C.POINTER(i), C.POINTER(C.c_double)],
        "chimera_step_noalloc": [C.POINTER(vp), d, C.POINTER(vp), i, i, d, d, pBC, pSch, vp],
    }
    for name, args in sigs.items():
        f = getattr(L, name)
        f.argtypes = args
        f.restype = C.c_int
    for name in ("chimera_ctx_stream", "chimera_batch_stream", "chimera_adj_stream",
                 "chimera_badj_stream"):
        getattr(L, name).argtypes = [vp]
        getattr(L, name).restype = vp
