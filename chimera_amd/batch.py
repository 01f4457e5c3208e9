"""Batched ensemble of independent grids of uneven sizes (chimera_batch_*): every RK
stage is one persistent launch over all members, dt is per member.

    ens = Ensemble([problems.grf_member(n, 1000 + m) for m, n in enumerate(sizes)])
    ens.steps(500, dt_mode=DT_CFL, graph=True)
    state_m = ens.download(m)
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib
from ._lib import BC, Scheme, check, state_ptrs

DT_REFERENCE, DT_CFL = 0, 1
SHAPES = {0: (32, 8), 1: (16, 16), 2: (64, 4), 3: (32, 32), 4: (32, 16)}  # 3: not HLLD


class Ensemble:
    def __init__(self, probs, stream=None, upload: bool = True):
        L = _lib.lib()
        self._L = L
        m = len(probs)
        if m == 0:
            raise ValueError("empty ensemble")
        schemes = {tuple(p.scheme) for p in probs}
        if len(schemes) != 1:
            raise ValueError("all members of a batch share one scheme")
        self.members = m
        self.shapes = [(p.ny, p.nx) for p in probs]
        nx = (C.c_int * m)(*[p.nx for p in probs])
        ny = (C.c_int * m)(*[p.ny for p in probs])
        dx = (C.c_double * m)(*[p.dx for p in probs])
        gam = (C.c_double * m)(*[p.gamma for p in probs])
        bcs = (BC * m)(*[BC.of(p.bc) for p in probs])
        sch = Scheme.of(next(iter(schemes)))
        h = C.c_void_p()
        check(L.chimera_batch_create(C.byref(h), m, nx, ny, dx, gam, bcs, C.byref(sch),
                                     C.c_void_p(stream) if stream else None), "chimera_batch_create")
        self._h = h
        if upload:
            for k, p in enumerate(probs):
                self.upload(k, p.state)

    def upload(self, member: int, state) -> None:
        """state: numpy (6, ny, nx) float64 or a CUDA float64 tensor of that shape."""
        shape = (6,) + self.shapes[member]
        if isinstance(state, np.ndarray):
            st = np.ascontiguousarray(state, dtype=np.float64)
            ptrs = state_ptrs(st, shape, f"member {member}")
        else:
            import torch
            st = state.contiguous()
            ptrs = state_ptrs(st, shape, f"member {member}")
            # the copy runs on the library's stream: finish torch's work on the tensor first
            torch.cuda.synchronize()
        check(self._L.chimera_batch_upload(self._h, int(member), ptrs), "chimera_batch_upload")

    def download(self, member: int, out=None):
        """Member `member`'s state into out (numpy or CUDA float64 (6, ny, nx); a new
        numpy array when None)."""
        if out is None:
            out = np.empty((6,) + self.shapes[member])
        ptrs = state_ptrs(out, (6,) + self.shapes[member], f"member {member}")
        if not isinstance(out, np.ndarray):
            import torch
            torch.cuda.synchronize()   # torch may still be using the destination
        check(self._L.chimera_batch_download(self._h, int(member), ptrs), "chimera_batch_download")
        return out

    def steps(self, n: int, dt_mode: int = DT_REFERENCE, graph: bool = False) -> None:
        check(self._L.chimera_batch_steps(self._h, int(n), int(dt_mode), int(bool(graph))),
              "chimera_batch_steps")

    def member_time(self, member: int) -> tuple[float, int, float]:
        t, oc, dt = C.c_double(), C.c_int(), C.c_double()
        check(self._L.chimera_batch_member_time(self._h, int(member), C.byref(t), C.byref(oc),
                                                C.byref(dt)), "member_time")
        return t.value, oc.value, dt.value

    def sync(self) -> None:
        check(self._L.chimera_batch_sync(self._h), "chimera_batch_sync")

    @property
    def cells(self) -> int:
        c = C.c_longlong()
        check(self._L.chimera_batch_cells(self._h, C.byref(c)), "cells")
        return c.value

    def set_tiling(self, shape: int, blocks_per_cu: int = 0) -> None:
        check(self._L.chimera_batch_set_tiling(self._h, int(shape), int(blocks_per_cu)), "set_tiling")

    def tiling(self) -> tuple[int, int]:
        s, b = C.c_int(), C.c_int()
        check(self._L.chimera_batch_get_tiling(self._h, C.byref(s), C.byref(b)), "get_tiling")
        return s.value, b.value

    def autotune(self, reps: int = 3) -> tuple[int, int]:
        s, b = C.c_int(), C.c_int()
        check(self._L.chimera_batch_autotune(self._h, int(reps), C.byref(s), C.byref(b)), "autotune")
        return s.value, b.value

    def set_timing(self, enable: bool) -> None:
        check(self._L.chimera_batch_set_timing(self._h, int(bool(enable))), "set_timing")

    def stage_kernel_ms(self) -> tuple[float, int]:
        ms, n = C.c_double(), C.c_int()
        check(self._L.chimera_batch_stage_kernel_ms(self._h, C.byref(ms), C.byref(n)), "stage_ms")
        return ms.value, n.value

    def close(self) -> None:
        if getattr(self, "_h", None):
            self._L.chimera_batch_destroy(self._h)
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
