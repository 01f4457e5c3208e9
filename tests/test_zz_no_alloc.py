"""chimera_step_noalloc (the allocation-free step, no counterpart in the reference)
against chimera_adj_forward_step, the existing entry point with the same semantics:
bit-identical states on two calls in a row, the second on a smaller grid (so stale
scratch contents from the first call would show), in place, and on a side stream."""
from __future__ import annotations

import ctypes as C

import numpy as np
import pytest
import torch

from chimera_amd import _lib, problems
from chimera_amd.adjoint import AdjointStepper

pytestmark = pytest.mark.gpu

MAX_CELLS = 512 * 512


def _ptrs(t: torch.Tensor):
    return (C.c_void_p * 6)(*[t[q].data_ptr() for q in range(6)])


def _noalloc(L, x, out, prob, dt, stream=None):
    ny, nx = x.shape[1], x.shape[2]
    s = torch.cuda.current_stream() if stream is None else stream
    rc = L.chimera_step_noalloc(_ptrs(x), float(dt), _ptrs(out), nx, ny, float(prob.dx),
                                float(prob.gamma), C.byref(_lib.BC.of(prob.bc)),
                                C.byref(_lib.Scheme.of(prob.scheme)), C.c_void_p(s.cuda_stream))
    return rc


def _reference_step(prob, x, dt):
    adj = AdjointStepper(prob.nx, prob.ny, prob.dx, prob.gamma, prob.bc, prob.scheme)
    try:
        return adj.forward_step(x, dt)
    finally:
        adj.close()


def test_noalloc_step_matches_forward_step(gpu_lib):
    L = gpu_lib
    cases = [problems.orszag_tang(256, scheme=(1, 2, 1, 1)),          # 256 x 256, periodic
             problems.rotor(96, scheme=(2, 1, 1, 1))]                  # smaller, outflow
    for prob in cases:
        print(f"case {prob.name} {prob.nx}x{prob.ny} scheme {prob.scheme}", flush=True)
        x = torch.from_numpy(prob.state).cuda().contiguous()
        dt = 0.3 * prob.dx / 4.0
        ref = _reference_step(prob, x, dt)
        out = torch.full_like(x, float("nan"))
        assert _noalloc(L, x, out, prob, dt) == 0, L.chimera_last_error().decode()
        torch.cuda.synchronize()
        assert torch.isfinite(out).all()
        assert torch.equal(out, ref), f"{prob.name}: noalloc step differs from forward_step"
        # in place: state_out = state_in
        y = x.clone()
        assert _noalloc(L, y, y, prob, dt) == 0, L.chimera_last_error().decode()
        torch.cuda.synchronize()
        assert torch.equal(y, ref)


def test_noalloc_step_on_side_stream_and_repeated(gpu_lib):
    L = gpu_lib
    prob = problems.orszag_tang(64, scheme=(3, 2, 1, 1))
    x = torch.from_numpy(prob.state).cuda().contiguous()
    dt = 0.3 * prob.dx / 4.0
    ref = _reference_step(prob, _reference_step(prob, x, dt), dt)
    s = torch.cuda.Stream()
    a, b = torch.empty_like(x), torch.empty_like(x)
    torch.cuda.synchronize()
    with torch.cuda.stream(s):
        assert _noalloc(L, x, a, prob, dt, stream=s) == 0
        assert _noalloc(L, a, b, prob, dt, stream=s) == 0
    s.synchronize()
    assert torch.equal(b, ref)


def test_noalloc_step_rejects_oversize_grid(gpu_lib):
    L = gpu_lib
    prob = problems.orszag_tang(32)
    x = torch.zeros((6, 1, 1), dtype=torch.float64, device="cuda")
    ptrs = (C.c_void_p * 6)(*[x[q].data_ptr() for q in range(6)])
    for nx, ny in ((513, 512), (1024, 300)):
        assert nx * ny > MAX_CELLS
        rc = L.chimera_step_noalloc(ptrs, 1e-3, ptrs, nx, ny, 1.0 / nx, 5.0 / 3.0,
                                    C.byref(_lib.BC.of(prob.bc)), C.byref(_lib.Scheme.of(prob.scheme)),
                                    None)
        assert rc == -1
        assert b"exceeds" in L.chimera_last_error()
    assert L.chimera_step_noalloc(ptrs, 0.0, ptrs, 64, 64, 1 / 64, 5 / 3, C.byref(_lib.BC.of(prob.bc)),
                                  C.byref(_lib.Scheme.of(prob.scheme)), None) == -1
