"""Resident single-grid driver: the reference's time loop (src/main.f90:65-303) with
the state in HBM, through the C ABI (chimera_ctx_*).

    g = Grid.from_problem(problems.orszag_tang(256))
    g.steps(100)                      # reference dt policy
    state = g.download()              # (6, ny, nx): Mass, Momx, Momy, Energy, b_x, b_y
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib
from ._lib import BC, Scheme, check, state_ptrs

DT_REFERENCE, DT_CFL, DT_PRESCRIBED = 0, 1, 2
# tile shapes of the fused stage kernel (CHIMERA_CTX_NUM_SHAPES), in shape-index order
TILE_SHAPES = ((32, 8), (16, 16), (64, 4), (32, 32), (32, 16))
NUM_TILE_SHAPES = len(TILE_SHAPES)


def problem_info(problem_type: int, n: int) -> tuple[float, float, tuple]:
    """(dx, gamma, boundary flags) of the reference's problem type 1-4 on N x N cells
    (initialize_problem, src/mhd_init.f90:259-288)."""
    L = _lib.lib()
    dx, gamma, bc = C.c_double(), C.c_double(), BC()
    check(L.chimera_problem_info(int(problem_type), int(n), C.byref(dx), C.byref(gamma),
                                 C.byref(bc)), "chimera_problem_info")
    return dx.value, gamma.value, (bc.xlo, bc.xhi, bc.ylo, bc.yhi)


class Grid:
    """One grid resident on the current GPU (a chimera_ctx_t)."""

    def __init__(self, nx: int, ny: int, dx: float, gamma: float, bc, scheme, stream=None):
        L = _lib.lib()
        self.nx, self.ny, self.dx, self.gamma = int(nx), int(ny), float(dx), float(gamma)
        self.bc, self.scheme = BC.of(bc), Scheme.of(scheme)
        h = C.c_void_p()
        check(L.chimera_ctx_create(C.byref(h), self.nx, self.ny, self.dx, self.gamma,
                                   C.byref(self.bc), C.byref(self.scheme),
                                   C.c_void_p(stream) if stream else None), "chimera_ctx_create")
        self._h = h
        self._L = L

    @classmethod
    def from_problem(cls, prob, stream=None) -> "Grid":
        g = cls(prob.nx, prob.ny, prob.dx, prob.gamma, prob.bc, prob.scheme, stream)
        g.upload(prob.state)
        return g

    @classmethod
    def from_setup(cls, problem_type: int, n: int, scheme, stream=None) -> "Grid":
        """An N x N grid holding the initial state of the reference's problem type 1-4
        (Orszag-Tang, Kelvin-Helmholtz, field loop, rotor; src/mhd_init.f90:259-450),
        generated on the device with the reference program's bits."""
        dx, gamma, bc = problem_info(problem_type, n)
        g = cls(n, n, dx, gamma, bc, scheme, stream)
        check(g._L.chimera_ctx_setup_problem(g._h, int(problem_type)), "chimera_ctx_setup_problem")
        return g

    # ---- state transfer -------------------------------------------------------
    def upload(self, state) -> None:
        """state: numpy (6, ny, nx) float64, or a CUDA float64 tensor of that shape."""
        shape = (6, self.ny, self.nx)
        if isinstance(state, np.ndarray):
            st = np.ascontiguousarray(state, dtype=np.float64)
            check(self._L.chimera_ctx_upload(self._h, state_ptrs(st, shape, "upload")),
                  "chimera_ctx_upload")
        else:
            import torch
            t = state.contiguous()
            ptrs = state_ptrs(t, shape, "upload")
            # the copy runs on the library's stream: finish torch's work on the tensor first
            torch.cuda.synchronize()
            check(self._L.chimera_ctx_upload(self._h, ptrs), "chimera_ctx_upload")

    def download(self, out=None):
        """Into out (numpy or CUDA tensor of shape (6, ny, nx), float64, contiguous; a
        new numpy array when None)."""
        if out is None:
            out = np.empty((6, self.ny, self.nx))
        ptrs = state_ptrs(out, (6, self.ny, self.nx), "download")
        if not isinstance(out, np.ndarray):
            import torch
            torch.cuda.synchronize()   # torch may still be using the destination
        check(self._L.chimera_ctx_download(self._h, ptrs), "chimera_ctx_download")
        return out

    # ---- stepping ---------------------------------------------------------------
    def steps(self, n: int, dt_mode: int = DT_REFERENCE, dt=None, graph: bool = False) -> None:
        dti = None
        if dt_mode == DT_PRESCRIBED:
            if dt is None:
                raise ValueError("DT_PRESCRIBED needs a dt sequence")
            dt = np.ascontiguousarray(dt, dtype=np.float64)
            if dt.ndim != 1 or dt.shape[0] < n:
                raise ValueError(f"DT_PRESCRIBED needs {n} dts, got shape {dt.shape}")
            dti = dt.ctypes.data_as(C.POINTER(C.c_double))
        check(self._L.chimera_ctx_steps(self._h, int(n), int(dt_mode), dti, int(bool(graph))),
              "chimera_ctx_steps")

    def sync(self) -> None:
        check(self._L.chimera_ctx_sync(self._h), "chimera_ctx_sync")

    @property
    def stream(self) -> int:
        return self._L.chimera_ctx_stream(self._h)

    def time(self) -> tuple[float, int]:
        t, oc = C.c_double(), C.c_int()
        check(self._L.chimera_ctx_get_time(self._h, C.byref(t), C.byref(oc)), "get_time")
        return t.value, oc.value

    def set_time(self, t: float = 0.0, out_count: int = 1) -> None:
        check(self._L.chimera_ctx_set_time(self._h, float(t), int(out_count)), "set_time")

    def last_dt(self) -> float:
        d = C.c_double()
        check(self._L.chimera_ctx_last_dt(self._h, C.byref(d)), "last_dt")
        return d.value

    def cfl_dt(self, dt_mode: int = DT_REFERENCE) -> float:
        d = C.c_double()
        check(self._L.chimera_ctx_cfl_dt(self._h, int(dt_mode), C.byref(d)), "cfl_dt")
        return d.value

    def divb(self) -> tuple[float, float]:
        m, s = C.c_double(), C.c_double()
        check(self._L.chimera_ctx_divb(self._h, C.byref(m), C.byref(s)), "divb")
        return m.value, s.value

    def stage1_mask(self) -> np.ndarray:
        out = np.zeros((self.ny, self.nx), dtype=np.uint32)
        check(self._L.chimera_ctx_stage1_mask(self._h, out.ctypes.data_as(_lib.c_u32p)),
              "stage1_mask")
        return out

    def set_timing(self, enable: bool) -> None:
        check(self._L.chimera_ctx_set_timing(self._h, int(bool(enable))), "set_timing")

    def stage_kernel_ms(self) -> tuple[float, int]:
        ms, n = C.c_double(), C.c_int()
        check(self._L.chimera_ctx_stage_kernel_ms(self._h, C.byref(ms), C.byref(n)),
              "stage_kernel_ms")
        return ms.value, n.value

    def set_tile_shape(self, shape: int) -> None:
        check(self._L.chimera_ctx_set_tile_shape(self._h, int(shape)), "set_tile_shape")

    def tile_shape(self) -> tuple[int, int, int]:
        s, tx, ty = C.c_int(), C.c_int(), C.c_int()
        check(self._L.chimera_ctx_get_tile_shape(self._h, C.byref(s), C.byref(tx), C.byref(ty)),
              "get_tile_shape")
        return s.value, tx.value, ty.value

    def autotune(self, reps: int = 3) -> tuple[int, list[float]]:
        best = C.c_int()
        ms = (C.c_double * NUM_TILE_SHAPES)()
        check(self._L.chimera_ctx_autotune(self._h, int(reps), C.byref(best), ms), "autotune")
        return best.value, list(ms)

    def close(self) -> None:
        if getattr(self, "_h", None):
            self._L.chimera_ctx_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()
