"""Differentiable SSP-RK3 evolution: torch.autograd.Function over the hand-written
adjoint kernels (chimera_adj_*), with checkpointed recomputation.

The dt sequence is frozen (recorded once, e.g. by ``record_dt``, and used by the
forward pass, the adjoint and any finite-difference check alike), so the gradient
is that of the discrete map U_0 -> U_N with the branches the forward pass took.

    dts = record_dt(prob, 1000)
    U0 = field_loop_state(Az, dx)                  # differentiable in torch
    UN = evolve(U0, dts, prob, checkpoint_every=50)
    loss = ((cell_b2(UN) - target) ** 2).sum(); loss.backward()
"""
from __future__ import annotations

import ctypes as C

import numpy as np
import torch

from . import _lib
from ._lib import BC, Scheme, check, state_ptrs
from .solver import DT_CFL, DT_PRESCRIBED, Grid


def _ptrs(t: torch.Tensor, shape=None, what: str = "state"):
    """Field pointers of a contiguous float64 CUDA state (6, ny, nx), checked."""
    if not isinstance(t, torch.Tensor):
        raise TypeError(f"{what}: need a CUDA tensor, got {type(t).__name__}")
    return state_ptrs(t, shape, what)


class AdjointStepper:
    """chimera_adj_t: one-step forward and VJP on device tensors (6, ny, nx)."""

    def __init__(self, nx, ny, dx, gamma, bc, scheme):
        L = _lib.lib()
        self._L = L
        self.nx, self.ny = int(nx), int(ny)
        self.shape = (6, self.ny, self.nx)
        h = C.c_void_p()
        check(L.chimera_adj_create(C.byref(h), self.nx, self.ny, float(dx), float(gamma),
                                   C.byref(BC.of(bc)), C.byref(Scheme.of(scheme)), None),
              "chimera_adj_create")
        self._h = h

    def set_fused(self, fused: bool) -> None:
        """Fused LDS-tiled reverse kernels (default) or the stored-intermediate ones;
        both give the same bits."""
        check(self._L.chimera_adj_set_fused(self._h, int(bool(fused))), "chimera_adj_set_fused")

    def set_tiles(self, shape_a: int = -1, shape_b: int = -1) -> None:
        """Tile shapes of the fused reverse kernels: 0 = 16x16, 1 = 16x8, 2 = 8x8, 3 = 32x8."""
        check(self._L.chimera_adj_set_tiles(self._h, int(shape_a), int(shape_b)),
              "chimera_adj_set_tiles")

    def forward_step(self, x: torch.Tensor, dt: float, out: torch.Tensor | None = None):
        out = torch.empty_like(x) if out is None else out
        torch.cuda.synchronize()
        check(self._L.chimera_adj_forward_step(self._h, _ptrs(x, self.shape, "x"), float(dt),
                                               _ptrs(out, self.shape, "out")),
              "chimera_adj_forward_step")
        return out

    def vjp(self, x: torch.Tensor, dt: float, lam_out: torch.Tensor, out=None):
        out = torch.empty_like(x) if out is None else out
        torch.cuda.synchronize()
        check(self._L.chimera_adj_step_vjp(self._h, _ptrs(x, self.shape, "x"), float(dt),
                                           _ptrs(lam_out.contiguous(), self.shape, "lam_out"),
                                           _ptrs(out, self.shape, "out")), "chimera_adj_step_vjp")
        return out

    def segment_vjp(self, x0: torch.Tensor, dts, lam: torch.Tensor) -> torch.Tensor:
        """lam <- (dU^k / dU^0)^T lam in place, for the k = len(dts) steps from x0 = U^0:
        the states are recomputed with the fused forward and the VJPs run back to back
        on the device (one call per checkpoint segment)."""
        dts = np.ascontiguousarray(dts, dtype=np.float64)
        if not lam.is_contiguous():
            raise ValueError("segment_vjp: lam must be contiguous (it is updated in place)")
        torch.cuda.synchronize()
        check(self._L.chimera_adj_segment_vjp(self._h, _ptrs(x0, self.shape, "x0"),
                                              dts.ctypes.data_as(C.POINTER(C.c_double)),
                                              len(dts), _ptrs(lam, self.shape, "lam")),
              "chimera_adj_segment_vjp")
        return lam

    def sweep(self, ckpts, dts, seg_steps: int, lam: torch.Tensor) -> torch.Tensor:
        """lam <- (dU_N/dU_0)^T lam in place for a run of len(dts) steps checkpointed
        every seg_steps steps (ckpts[k] = U at step k * seg_steps): every segment
        recomputed and reversed, the recompute of segment k-1 overlapping the reverse
        of segment k."""
        dts = np.ascontiguousarray(dts, dtype=np.float64)
        nseg = -(-len(dts) // seg_steps)
        if len(ckpts) != nseg:
            raise ValueError(f"{len(ckpts)} checkpoints for {nseg} segments")
        ck = _ptrs_many(ckpts, [self.shape[1:]], "checkpoint")
        lp = _ptrs(lam, self.shape, "lam")
        # the sweep runs on the library's streams: torch's pending work on lam and the
        # checkpoints (e.g. lam cloned from the incoming gradient) must be done first
        torch.cuda.synchronize()
        check(self._L.chimera_adj_sweep(self._h, ck, dts.ctypes.data_as(C.POINTER(C.c_double)),
                                        len(dts), int(seg_steps), lp), "chimera_adj_sweep")
        return lam

    def record(self, x0: torch.Tensor, dts, seg_steps: int, nseg: int,
               out: torch.Tensor) -> torch.Tensor:
        """len(dts) (<= seg_steps) steps from x0 into out, keeping the trajectory where
        the sweep of a run of nseg segments checkpointed every seg_steps steps expects its
        last segment; that sweep (with x0 as its last checkpoint and the same dts) then
        skips the segment's recompute."""
        dts = np.ascontiguousarray(dts, dtype=np.float64)
        x0p, outp = _ptrs(x0, self.shape, "x0"), _ptrs(out, self.shape, "out")
        torch.cuda.synchronize()   # x0 may still be written by torch's stream
        check(self._L.chimera_adj_record(self._h, x0p, dts.ctypes.data_as(C.POINTER(C.c_double)),
                                         len(dts), int(seg_steps), int(nseg), outp),
              "chimera_adj_record")
        return out

    def close(self):
        if getattr(self, "_h", None):
            self._L.chimera_adj_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def record_dt(prob, nsteps: int, dt_mode: int = DT_CFL) -> np.ndarray:
    """The dt of every step of a CFL-driven run of ``prob`` (resident GPU path)."""
    dts = np.empty(nsteps)
    with Grid.from_problem(prob) as g:
        for s in range(nsteps):
            g.steps(1, dt_mode=dt_mode)
            dts[s] = g.last_dt()
    return dts


# adjoint workspaces of the autograd backward, one per (device, geometry): the
# trajectory memory of a checkpoint segment is allocated on the first backward and
# reused by the next ones
_WORKSPACES: dict = {}


def _workspace(geom) -> AdjointStepper:
    key = (torch.cuda.current_device(), geom)
    adj = _WORKSPACES.get(key)
    if adj is None:
        nx, ny, dx, gamma, bc, scheme = geom
        adj = _WORKSPACES[key] = AdjointStepper(nx, ny, dx, gamma, bc, scheme)
    return adj


def release_workspaces() -> None:
    """Free the device memory the autograd backward of `evolve` keeps between calls."""
    for adj in _WORKSPACES.values():
        adj.close()
    _WORKSPACES.clear()


class _Evolve(torch.autograd.Function):
    @staticmethod
    def forward(ctx, U0, dts, geom, checkpoint_every):
        nx, ny, dx, gamma, bc, scheme = geom
        U0 = U0.detach().contiguous()
        n = len(dts)
        K = max(1, int(checkpoint_every))
        nseg = -(-n // K)
        s_last = (nseg - 1) * K
        # with a backward to follow, the last segment runs through the adjoint workspace,
        # which keeps its trajectory for the sweep (no recompute of that segment)
        record = ctx.needs_input_grad[0]
        ckpts = [U0.clone()]
        with Grid(nx, ny, dx, gamma, bc, scheme) as g:
            g.upload(U0)
            for s0 in range(0, s_last, K):
                g.steps(K, dt_mode=DT_PRESCRIBED, dt=dts[s0:s0 + K])
                ckpts.append(g.download(torch.empty_like(U0)))
            if not record:
                g.steps(n - s_last, dt_mode=DT_PRESCRIBED, dt=dts[s_last:])
                UN = g.download(torch.empty_like(U0))
        if record:
            UN = _workspace(geom).record(ckpts[-1], dts[s_last:], K, nseg, torch.empty_like(U0))
        ctx.save_for_backward(*ckpts)
        ctx.dts, ctx.geom, ctx.K = dts, geom, K
        return UN

    @staticmethod
    def backward(ctx, gN):
        ckpts = ctx.saved_tensors
        lam = gN.contiguous().clone()
        _workspace(ctx.geom).sweep(ckpts, ctx.dts, ctx.K, lam)
        return lam, None, None, None


def evolve(U0: torch.Tensor, dts, prob_or_geom, checkpoint_every: int = 50) -> torch.Tensor:
    """U_N after len(dts) SSP-RK3 steps with the given dt sequence; differentiable
    with respect to U0 (float64 CUDA tensor (6, ny, nx))."""
    if hasattr(prob_or_geom, "dx"):
        p = prob_or_geom
        geom = (p.nx, p.ny, p.dx, p.gamma, tuple(p.bc), tuple(p.scheme))
    else:
        geom = tuple(prob_or_geom)
    dts = np.ascontiguousarray(dts, dtype=np.float64)
    return _Evolve.apply(U0, dts, geom, checkpoint_every)


# ---- batched ensembles ---------------------------------------------------------

def _ptrs_many(ts, shapes=None, what: str = "state") -> "C.Array":
    """Field pointers of several states (member m's 6 fields at 6 m .. 6 m + 5), each
    checked; shapes (optional) gives (ny, nx) per state, repeated over the list."""
    ts = list(ts)
    out = (C.c_void_p * (6 * len(ts)))()
    for k, t in enumerate(ts):
        shape = None if shapes is None else (6,) + tuple(shapes[k % len(shapes)])
        p = _ptrs(t, shape, f"{what} {k}")
        for q in range(6):
            out[6 * k + q] = p[q]
    return out


def _dt_table(dts, members: int) -> np.ndarray:
    d = np.ascontiguousarray(dts, dtype=np.float64)
    if d.ndim != 2 or d.shape[0] != members:
        raise ValueError(f"dt table of shape {d.shape}: need (members = {members}, steps)")
    return d


class BatchAdjoint:
    """chimera_badj_t: forward steps and checkpoint-segment VJPs of many grids of uneven
    sizes at once (one persistent launch per stage over all members).  geoms: one
    (nx, ny, dx, gamma, bc) per member; the scheme is shared."""

    def __init__(self, geoms, scheme):
        L = _lib.lib()
        self._L = L
        self.members = len(geoms)
        self.shapes = [(int(g[1]), int(g[0])) for g in geoms]   # (ny, nx)
        nx = (C.c_int * self.members)(*[int(g[0]) for g in geoms])
        ny = (C.c_int * self.members)(*[int(g[1]) for g in geoms])
        dx = (C.c_double * self.members)(*[float(g[2]) for g in geoms])
        gm = (C.c_double * self.members)(*[float(g[3]) for g in geoms])
        bcs = (BC * self.members)(*[BC.of(g[4]) for g in geoms])
        sch = Scheme.of(scheme)
        h = C.c_void_p()
        check(L.chimera_badj_create(C.byref(h), self.members, nx, ny, dx, gm, C.cast(bcs, C.c_void_p),
                                    C.cast(C.pointer(sch), C.c_void_p), None), "chimera_badj_create")
        self._h = h

    def _member_ptrs(self, ts, what: str) -> "C.Array":
        """Field pointers of one state per member (the library reads 6 x members)."""
        ts = list(ts)
        if len(ts) != self.members:
            raise ValueError(f"{what}: {len(ts)} states for {self.members} members")
        return _ptrs_many(ts, self.shapes, what)

    def forward(self, states, dts, out=None):
        """Every member len(dts[m]) steps with its dts, from states to out (new tensors
        unless given; may be the inputs)."""
        d = _dt_table(dts, self.members)
        out = [torch.empty_like(x) for x in states] if out is None else out
        torch.cuda.synchronize()
        check(self._L.chimera_badj_forward(self._h, self._member_ptrs(states, "states"),
                                           d.ctypes.data_as(C.POINTER(C.c_double)), d.shape[1],
                                           self._member_ptrs(out, "out")), "chimera_badj_forward")
        return out

    def segment_vjp(self, states0, dts, lams) -> None:
        """lams[m] <- (dU^k / dU^0)^T lams[m] in place for every member, k = len(dts[m])."""
        d = _dt_table(dts, self.members)
        torch.cuda.synchronize()
        check(self._L.chimera_badj_segment_vjp(self._h, self._member_ptrs(states0, "states0"),
                                               d.ctypes.data_as(C.POINTER(C.c_double)), d.shape[1],
                                               self._member_ptrs(lams, "lams")), "chimera_badj_segment_vjp")

    def sweep(self, ckpts, dts, seg_steps: int, lams) -> None:
        """lams[m] <- (dU_N / dU_0)^T lams[m] in place for a run of dts.shape[1] steps
        checkpointed every seg_steps steps (ckpts[k][m]: member m at step k * seg_steps),
        every segment recomputed and reversed, last first, in one call."""
        d = _dt_table(dts, self.members)
        nseg = -(-d.shape[1] // seg_steps)
        if len(ckpts) != nseg or any(len(c) != self.members for c in ckpts):
            raise ValueError(f"need {nseg} checkpoints of {self.members} members")
        torch.cuda.synchronize()
        check(self._L.chimera_badj_sweep(self._h, _ptrs_many([t for c in ckpts for t in c], self.shapes),
                                         d.ctypes.data_as(C.POINTER(C.c_double)), d.shape[1],
                                         int(seg_steps), self._member_ptrs(lams, "lams")), "chimera_badj_sweep")

    def record(self, states0, dts, seg_steps: int, nseg: int, out=None):
        """dts.shape[1] (<= seg_steps) steps of every member from states0 into out (new
        tensors unless given), keeping the trajectories the next sweep of a run of nseg
        segments checkpointed every seg_steps steps would recompute for its last
        segment (with states0 as that segment's checkpoint and the same dts)."""
        d = _dt_table(dts, self.members)
        out = [torch.empty_like(x) for x in states0] if out is None else out
        torch.cuda.synchronize()
        check(self._L.chimera_badj_record(self._h, self._member_ptrs(states0, "states0"),
                                          d.ctypes.data_as(C.POINTER(C.c_double)), d.shape[1],
                                          int(seg_steps), int(nseg), self._member_ptrs(out, "out")),
              "chimera_badj_record")
        return out

    def set_tiles(self, shape_a: int = -1, shape_b: int = -1, shape_f: int = -1) -> None:
        """Tile shapes of the two reverse passes and of the forward stages: 0 = 16x16,
        1 = 32x8 (-1 keeps)."""
        check(self._L.chimera_badj_set_tiles(self._h, int(shape_a), int(shape_b), int(shape_f)),
              "chimera_badj_set_tiles")

    def autotune(self, states, dts, reps: int = 2):
        """Pick the fastest forward and reverse tile shapes on one step of every member
        from states (dts: one dt per member).  Returns ((forward, A, B), ms) with ms the
        mean times of forward shapes 0, 1 and of the (A, B) pairs (0,0), (0,1), (1,0),
        (1,1)."""
        d = np.ascontiguousarray(np.asarray(dts, dtype=np.float64).reshape(self.members, -1)[:, :1])
        best = (C.c_int * 3)()
        ms = (C.c_double * 6)()
        torch.cuda.synchronize()
        check(self._L.chimera_badj_autotune(self._h, self._member_ptrs(states, "states"),
                                            d.ctypes.data_as(C.POINTER(C.c_double)), int(reps),
                                            best, ms), "chimera_badj_autotune")
        return tuple(best), list(ms)

    def close(self) -> None:
        if getattr(self, "_h", None):
            self._L.chimera_badj_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()


class _EvolveBatch(torch.autograd.Function):
    @staticmethod
    def forward(ctx, badj, dts, K, *U0s):
        states = [u.detach().contiguous().clone() for u in U0s]
        n = dts.shape[1]
        nseg = -(-n // K)
        s_last = (nseg - 1) * K
        ckpts = [[s.clone() for s in states]]
        for s0 in range(0, s_last, K):
            badj.forward(states, dts[:, s0:s0 + K], out=states)
            ckpts.append([s.clone() for s in states])
        last = np.ascontiguousarray(dts[:, s_last:])
        if any(ctx.needs_input_grad[3:]):
            # the backward's sweep finds the last segment's trajectory in place
            badj.record(ckpts[-1], last, K, nseg, out=states)
        else:
            badj.forward(states, last, out=states)
        ctx.ckpts, ctx.badj, ctx.dts, ctx.K = ckpts, badj, dts, K
        return tuple(states)

    @staticmethod
    def backward(ctx, *grads):
        ckpts, badj, dts, K = ctx.ckpts, ctx.badj, ctx.dts, ctx.K
        lams = [g.contiguous().clone() for g in grads]
        badj.sweep(ckpts, dts, K, lams)
        return (None, None, None, *lams)


def evolve_batch(U0s, dts, badj: BatchAdjoint, checkpoint_every: int = 50):
    """U_N of every member after dts.shape[1] steps (member m with the dts dts[m]),
    through the batched forward and adjoint; differentiable with respect to every U0
    (float64 CUDA tensors (6, ny, nx) of the sizes badj was made for)."""
    d = np.ascontiguousarray(dts, dtype=np.float64)
    if d.ndim != 2 or not d.shape[0] == len(U0s) == badj.members:
        raise ValueError(f"evolve_batch: {len(U0s)} states and a dt table of shape {d.shape} "
                         f"for {badj.members} members")
    return _EvolveBatch.apply(badj, d, max(1, int(checkpoint_every)), *U0s)


# ---- differentiable set-up and diagnostics (torch ops) -------------------------

def curl_2d_periodic(Az: torch.Tensor, dx: float):
    """compute_curl_2d (src/mhd_field_ops.f90:15-57) on a periodic grid, in torch."""
    inv_dx = 1.0 / dx
    bx = (Az - torch.roll(Az, 1, 0)) * inv_dx
    by = -((Az - torch.roll(Az, 1, 1)) * inv_dx)
    return bx, by


def cell_b_periodic(bx: torch.Tensor, by: torch.Tensor):
    """average_face_to_cell_B (src/mhd_field_ops.f90:120-156), periodic, in torch."""
    return 0.5 * (bx + torch.roll(bx, 1, 1)), 0.5 * (by + torch.roll(by, 1, 0))


def field_loop_state(Az: torch.Tensor, dx: float, gamma: float = 5.0 / 3.0,
                     hydro=None) -> torch.Tensor:
    """State (6, ny, nx) of the field-loop problem with face B = curl(Az),
    differentiable in Az: the torch form of problems.field_loop_from_az, operation
    for operation.  ``hydro``: optional (rho, vx, vy, p_thermal) arrays; default
    rho = p = 1, v = (2, 1)."""
    bx, by = curl_2d_periodic(Az, dx)
    Bx, By = cell_b_periodic(bx, by)
    vol = dx * dx
    if hydro is None:
        rho = torch.ones_like(Az)
        vx = torch.full_like(Az, 2.0)
        vy = torch.full_like(Az, 1.0)
        pth = torch.ones_like(Az)
    else:
        rho, vx, vy, pth = (torch.as_tensor(a, dtype=Az.dtype, device=Az.device) for a in hydro)
    half_b2 = 0.5 * (Bx * Bx + By * By)
    P = pth + half_b2
    r = torch.clamp_min(rho, 1.0e-12)
    M = r * vol
    # divide by a tensor: on the GPU torch turns division by a Python scalar into a
    # multiplication by its reciprocal, which is not the correctly rounded quotient
    gm1 = torch.full_like(P, gamma - 1.0)
    E = ((P - half_b2) / gm1 + 0.5 * r * (vx * vx + vy * vy) + half_b2) * vol
    return torch.stack([M, M * vx, M * vy, E, bx, by])


def cell_b2(U: torch.Tensor) -> torch.Tensor:
    Bx, By = cell_b_periodic(U[4], U[5])
    return Bx * Bx + By * By


def grf_member_from_noise(w: torch.Tensor, n: int, params: dict, scheme=(3, 2, 1, 1)):
    """GRF member from its white-noise fields w (2, n, n) (differentiable in w):
    returns (Problem with a detached copy of the state, state tensor)."""
    from . import problems
    st = problems.grf_state_from_noise(w, n, params)
    prob = problems.Problem(f"grf_{n}", st.detach(), 1.0 / n, params["gamma"], problems.PERIODIC,
                            tuple(scheme))
    return prob, st
