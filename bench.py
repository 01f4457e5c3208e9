#!/usr/bin/env python3
"""Benchmark: SSP-RK3 cell-updates/s of the fused HIP stage path on one MI355X.

Workload (default): MHD rotor, 2048 x 2048 on [0,1]^2, outflow on all sides,
gamma = 1.4, HLLE flux, Van Leer limiter, MUSCL + MOOD, CT, SSP-RK3, CFL 0.3 with the
pure CFL dt policy, FP64.  A step is one full SSP-RK3 step (3 fused stage kernels
plus the on-device CFL reduction) over the whole grid; the state is resident in
HBM before the timed region starts.

Printed JSON line: value = cells * steps / wall time of the timed region (the
stream is synchronised on both sides); roofline = the fused stage kernel's
algorithmic HBM bytes per launch (128 B per cell: stage 1 reads u^n and writes u^1,
stages 2 and 3 also read u^n, 384 B per cell over the 3 stages) divided by its
average launch duration, measured with HIP events on the launch stream during
the timed region; cpu_baseline = the reference Fortran (amdflang build in
oracle/_ref, OpenMP) timed on this host on a bounded sample of the same workload,
run before the GPU is touched.
"""
from __future__ import annotations

import argparse
import json
import os
import re
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

METRIC = "cell-updates/s on 1 GPU, batched GRF vs plain; adjoint/forward cost ratio"
HBM_PEAK_GBS = 8000.0
BYTES_PER_CELL_STAGE = 128.0          # (96 + 144 + 144) / 3


def parse_args():
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2000)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--n", type=int, default=2048, help="grid size (n x n)")
    ap.add_argument("--cpu-steps", type=int, default=2,
                    help="steps of the CPU baseline sample (0 disables it)")
    ap.add_argument("--no-graph", action="store_true", help="skip the hipGraph replay pass")
    ap.add_argument("--members", type=int, default=512,
                    help="batched GRF ensemble size (0 skips the batched section)")
    ap.add_argument("--grf-steps", type=int, default=500, help="steps of every GRF member")
    ap.add_argument("--plain-per-class", type=int, default=8,
                    help="members per grid size timed on the plain single-run path")
    ap.add_argument("--adj-steps", type=int, default=1000,
                    help="steps of the field-loop adjoint section (0: skip)")
    ap.add_argument("--adj-fd-nodes", type=int, default=40,
                    help="Az nodes checked against central differences in the adjoint section")
    ap.add_argument("--dgrf-members", type=int, default=64,
                    help="members of the differentiable GRF section (0: skip)")
    ap.add_argument("--dgrf-steps", type=int, default=300)
    ap.add_argument("--strip-steps", type=int, default=200,
                    help="steps of the one-rank strip of an 8-way rotor split (0: skip)")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="also run a plain device copy of 2^26 doubles (512 MiB each way) 3 "
                         "times, the reference point for rocprofv3 FETCH_SIZE / WRITE_SIZE")
    return ap.parse_args()


def cpu_baseline(n: int, steps: int) -> dict | None:
    """Reference Fortran (oracle/_ref, amdflang -O3 -fopenmp) on `steps` rotor steps."""
    if steps <= 0:
        return None
    threads = int(os.environ.get("OMP_NUM_THREADS", "0")) or min(os.cpu_count() or 1, 16)
    os.environ["OMP_NUM_THREADS"] = str(threads)
    from chimera_amd import problems
    from oracle.oracle import RefVariant, COracle, ref_available, on_big_stack
    prob = problems.rotor(n, scheme=(2, 1, 1, 1))
    st = prob.state.copy()
    if ref_available():
        ref = RefVariant(2, 1)
        t0 = time.perf_counter()
        ref.run(st, prob.dx, prob.gamma, prob.bc, steps, dt_mode=1)
        el = time.perf_counter() - t0
        kind = "reference"
        what = "reference Fortran modules (amdflang -O3 -fopenmp -ffp-contract=off) + harness"
    else:
        orc = COracle()
        t0 = time.perf_counter()
        on_big_stack(orc.run, st, prob.dx, prob.gamma, prob.bc, [2, 1, 1, 1], steps, dt_mode=1)
        el = time.perf_counter() - t0
        kind = "port"
        what = "C restatement of the reference (gcc -O2 -fopenmp)"
    return {"value": n * n * steps / el, "unit": "cell-updates/s", "cores": threads,
            "kind": kind, "sample": f"rotor {n}x{n}, HLLE/Van Leer, {steps} SSP-RK3 steps "
            f"({el:.1f} s), {what}"}


def batched_grf(args, torch) -> dict:
    """Batched GRF ensemble (members of uneven size, Rusanov / MC, per-member CFL dt,
    periodic): every member advanced together, one persistent launch per RK stage,
    steps replayed as hipGraphs -- against the plain path, which advances the same
    members one after another through the resident single-grid driver (also graph
    replayed); the plain time is extrapolated from `plain_per_class` members per size."""
    from chimera_amd import problems
    from chimera_amd.batch import Ensemble, DT_CFL, SHAPES
    from chimera_amd.solver import Grid
    sizes = problems.grf_sizes(args.members)
    probs = [problems.grf_member_device(n, 1000 + m, scheme=(1, 2, 1, 1))
             for m, n in enumerate(sizes)]
    torch.cuda.synchronize()
    cells = sum(n * n for n in sizes)
    steps = args.grf_steps
    ens = Ensemble(probs)
    shape, bpc = ens.autotune(reps=2)
    ens.steps(2, dt_mode=DT_CFL, graph=True)
    ens.sync()
    t0 = time.perf_counter()
    ens.steps(steps, dt_mode=DT_CFL, graph=True)
    ens.sync()
    el = time.perf_counter() - t0
    ens.set_timing(True)
    ens.steps(3, dt_mode=DT_CFL)
    stage_ms, launches = ens.stage_kernel_ms()
    ens.set_timing(False)
    ens.close()
    batched = cells * steps / el

    per_class = {}
    for n in sorted(set(sizes)):
        idx = [m for m, s_ in enumerate(sizes) if s_ == n][:args.plain_per_class]
        times = []
        for m in idx:
            with Grid.from_problem(probs[m]) as g:
                g.autotune(reps=2)                 # the plain path at its best tile shape
                g.steps(2, dt_mode=DT_CFL, graph=True)
                g.sync()
                t1 = time.perf_counter()
                g.steps(steps, dt_mode=DT_CFL, graph=True)
                g.sync()
                times.append(time.perf_counter() - t1)
        per_class[n] = (sizes.count(n), len(idx), sum(times) / len(times))
    plain_time = sum(cnt * t for cnt, _, t in per_class.values())
    plain = cells * steps / plain_time
    achieved = BYTES_PER_CELL_STAGE * cells / (stage_ms * 1e-3) / 1e9 if stage_ms else None
    del probs
    torch.cuda.empty_cache()
    return {"workload": f"GRF ensemble: {args.members} members, N in {{128,256,384,512}} "
                        f"(seed 20261203), Rusanov, MC, MUSCL+MOOD, CT, per-member CFL dt, "
                        f"{steps} steps each, periodic",
            "cells": cells, "steps": steps, "value": batched, "plain_value": plain,
            "ratio": batched / plain, "unit": "cell-updates/s",
            "batched_ms_per_step": el * 1e3 / steps,
            "plain_extrapolated_s": plain_time,
            "plain_members_timed": {str(n): v[1] for n, v in per_class.items()},
            "tiling": {"tile": SHAPES[shape], "blocks_per_cu": bpc},
            "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS,
                         "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS if achieved else None,
                         "avg_launch_ms": stage_ms, "launches": launches}}


def adjoint_field_loop(args, torch) -> dict:
    """Field loop 512 x 256 (periodic, HLLE, MC), dt frozen from a CFL run, loss =
    sum((|B|^2_N - |B|^2_target)^2) with the target from Az (1 + 0.01 xi), xi seeded 7;
    gradient with respect to every Az node through the checkpointed adjoint
    (checkpoint every 50 steps).  cost ratio = backward wall time / wall time of the
    same steps on the fused forward path.

    Finite-difference check: the discrete map is not differentiable everywhere on this
    state -- with MC a face value can land exactly on its MOOD bound (MC's 2*min branch),
    where the map jumps, and outside the loop B is exactly 0 and rho, p, v exactly uniform,
    where the limiter's dfL * dfR is exactly 0 (a kink) -- so central differences there do
    not converge to the branch derivative.  The check therefore runs on the same grid with
    the Van Leer limiter (faces strictly inside the MOOD bounds), Az plus 1e-4 x a smooth
    periodic field and a smooth non-uniform background: central differences
    (eps = 1e-6 ... 1e-9 max|Az|) of the forward path (bit-identical to the reference) at
    seeded nodes inside the loop plus nodes near r = 0.3."""
    import numpy as np
    from chimera_amd import problems
    from chimera_amd.adjoint import cell_b2, evolve, field_loop_state, record_dt
    from chimera_amd.solver import Grid, DT_PRESCRIBED
    nsteps, K = args.adj_steps, 50
    rng = np.random.default_rng(7)

    def setup(limiter):
        prob, Az = problems.field_loop(512, 256, scheme=(2, limiter, 1, 1), return_az=True)
        dts = record_dt(prob, nsteps)
        return prob, Az, dts

    def run_forward(g, dx, az_np, dts):
        st = field_loop_state(torch.from_numpy(az_np).cuda(), dx)
        g.upload(st)
        g.steps(nsteps, dt_mode=DT_PRESCRIBED, dt=dts, graph=True)
        out = torch.empty_like(st)
        g.download(out)
        return cell_b2(out)

    def gradient(prob, Az, dts, T):
        Azt = torch.from_numpy(Az).cuda().requires_grad_(True)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        UN = evolve(field_loop_state(Azt, prob.dx), dts, prob, checkpoint_every=K)
        loss = ((cell_b2(UN) - T) ** 2).sum()
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        loss.backward()
        torch.cuda.synchronize()
        t2 = time.perf_counter()
        return Azt.grad.cpu().numpy(), float(loss.detach()), t1 - t0, t2 - t1

    # --- cost on the HLLE / MC configuration
    prob, Az, dts = setup(2)
    xi = rng.standard_normal(Az.shape)
    Az_t = Az * (1.0 + 0.01 * xi)
    g = Grid.from_problem(prob)
    T = run_forward(g, prob.dx, Az_t, dts)
    g.upload(field_loop_state(torch.from_numpy(Az).cuda(), prob.dx))
    g.autotune(reps=3)                         # the forward at its best tile shape
    g.steps(2, dt_mode=DT_PRESCRIBED, dt=dts[:2], graph=True)
    g.sync()
    g.upload(field_loop_state(torch.from_numpy(Az).cuda(), prob.dx))
    t0 = time.perf_counter()
    g.steps(nsteps, dt_mode=DT_PRESCRIBED, dt=dts, graph=True)
    g.sync()
    t_fwd = time.perf_counter() - t0
    fwd_tile = list(g.tile_shape()[1:])
    g.close()
    gradient(prob, Az, dts[:K], T)   # warm-up: the backward's workspace for one segment
    grad, loss_v, t_evolve, t_bwd = gradient(prob, Az, dts, T)
    grad_finite = bool(np.isfinite(grad).all())

    # --- finite differences on the differentiable (Van Leer, no uniform region) variant
    fd = None
    if args.adj_fd_nodes > 0:
        ny, nx = Az.shape
        hyd = problems.field_loop_background(nx, ny, seed=20)
        Az_v = Az + 1.0e-4 * problems.smooth_periodic(nx, ny, seed=1)
        prob_v = problems.field_loop_from_az(Az_v, prob.dx, scheme=(2, 1, 1, 1), hydro=hyd)
        dts_v = record_dt(prob_v, nsteps)
        hyd_t = [torch.from_numpy(a).cuda() for a in hyd]

        def run_forward_bg(g, dx, az_np, dts):
            st = field_loop_state(torch.from_numpy(az_np).cuda(), dx, hydro=hyd_t)
            g.upload(st)
            g.steps(nsteps, dt_mode=DT_PRESCRIBED, dt=dts, graph=True)
            out = torch.empty_like(st)
            g.download(out)
            return cell_b2(out)

        g = Grid.from_problem(prob_v)
        Az_vt = Az_v * (1.0 + 0.01 * xi)
        T_v = run_forward_bg(g, prob_v.dx, Az_vt, dts_v)

        def loss_diff(ap, am):
            # L(ap) - L(am) formed per cell before the sum (no cancellation of two
            # nearly equal totals): sum (b+ - b-) (b+ + b- - 2 T)
            bp = run_forward_bg(g, prob_v.dx, ap, dts_v)
            bm = run_forward_bg(g, prob_v.dx, am, dts_v)
            return float(((bp - bm) * (bp + bm - 2.0 * T_v)).sum())

        Azt = torch.from_numpy(Az_v).cuda().requires_grad_(True)
        UN = evolve(field_loop_state(Azt, prob_v.dx, hydro=hyd_t), dts_v, prob_v, checkpoint_every=K)
        ((cell_b2(UN) - T_v) ** 2).sum().backward()
        grad_v = Azt.grad.cpu().numpy()
        x = -1.0 + (np.arange(nx) + 0.5) * prob.dx
        y = -0.5 + (np.arange(ny) + 0.5) * prob.dx
        X, Y = np.meshgrid(x, y, indexing="xy")
        r = np.sqrt(X * X + Y * Y)
        inside = np.flatnonzero(r.ravel() < 0.3)
        ring = np.argsort(np.abs(r.ravel() - 0.3))[:64]
        n_ring = min(8, args.adj_fd_nodes)
        nodes = list(rng.choice(inside, size=args.adj_fd_nodes - n_ring, replace=False)) + \
            list(rng.choice(ring, size=n_ring, replace=False))
        gmax = float(np.abs(grad_v).max())
        fd = {"variant": "Van Leer, Az + 1e-4 smooth field, smooth background (see docstring)",
              "nodes": len(nodes),
              "rel_err_definition": "|adjoint - FD| / max(|FD|, 1e-3 max|grad|)"}
        for rel_eps in (1e-6, 1e-7, 1e-8, 1e-9):
            eps = rel_eps * float(np.abs(Az_v).max())
            rels = []
            for idx in nodes:
                j, i = divmod(int(idx), nx)
                ap, am = Az_v.copy(), Az_v.copy()
                ap[j, i] += eps
                am[j, i] -= eps
                d = loss_diff(ap, am) / (2.0 * eps)
                rels.append(abs(grad_v[j, i] - d) / max(abs(d), 1e-3 * gmax))
            rels = np.array(rels)
            fd[f"eps_{rel_eps:g}_max|Az|"] = {"median_rel_err": float(np.median(rels)),
                                              "max_rel_err": float(rels.max()),
                                              "frac_le_1e-6": float(np.mean(rels <= 1e-6))}
        g.close()
    return {"workload": f"field loop 512x256 periodic, HLLE, MC, MUSCL+MOOD, CT, {nsteps} steps "
                        f"(dt frozen from a CFL run), checkpoint every {K}",
            "loss": loss_v, "grad_nodes": int(grad.size), "grad_finite": grad_finite,
            "forward_s": t_fwd, "evolve_forward_s": t_evolve, "backward_s": t_bwd,
            "cost_ratio": t_bwd / t_fwd, "forward_tile": fwd_tile, "fd_check": fd}


def shard_strip(args, torch, full_stage_ms: float) -> dict:
    """One GPU's share of the rotor n x n split into 8 strips: its n/8 owned rows plus
    the depth-3 halo rows on each side, cut from the centre of the rotor state, through
    the same resident path (outflow BCs stand in for the received halo).  Reports the
    stage kernel per owned cell against the full grid's and the halo bytes a stage
    would exchange over xGMI."""
    from chimera_amd import problems
    from chimera_amd.solver import Grid, DT_CFL
    n, halo = args.n, 3
    owned = n // 8
    rows = owned + 2 * halo
    full = problems.rotor(n, scheme=(2, 1, 1, 1))
    j0 = n // 2 - rows // 2
    prob = problems.Problem("rotor_strip", full.state[:, j0:j0 + rows, :].copy(), full.dx,
                            full.gamma, full.bc, full.scheme)
    g = Grid.from_problem(prob)
    g.steps(2, dt_mode=DT_CFL)
    g.autotune(reps=3)
    g.set_timing(True)
    g.sync()
    t0 = time.perf_counter()
    g.steps(args.strip_steps, dt_mode=DT_CFL)
    g.sync()
    elapsed = time.perf_counter() - t0
    stage_ms, launches = g.stage_kernel_ms()
    g.close()
    halo_bytes = 2 * n * halo * 6 * 8   # both sides, 6 fields, FP64
    full_scaled = full_stage_ms * (n * rows) / (n * n)
    return {"workload": f"rotor {n}x{n} split into 8 strips: one rank's {n}x{owned} owned "
                        f"cells + {halo} halo rows per side ({n}x{rows}), HLLE, Van Leer, "
                        f"{args.strip_steps} steps",
            "owned_cell_updates_per_s": n * owned * args.strip_steps / elapsed,
            "stage_ms": stage_ms, "launches": launches,
            "full_grid_stage_ms_scaled": full_scaled,
            "stage_efficiency_vs_full": full_scaled / stage_ms if stage_ms > 0 else None,
            "halo_bytes_per_stage": halo_bytes,
            "halo_us_per_stage_at_153GBps_per_side": halo_bytes / 2 / 153e9 * 1e6}


def differentiable_grf(args, torch) -> dict:
    """Differentiable GRF ensemble: members N in {384, 448, 512} (seed 20261203),
    HLLD, MC, dt frozen from a CFL run of each member; per-member loss =
    sum((rho_N - target)^2) with a seeded target (5000 + m); gradient with respect to
    the white-noise fields that generate psi and Az.  All members advance together:
    the batched forward (one persistent launch per stage over every member's tiles)
    and the batched adjoint (checkpoints every 50 steps).  Reports the batched
    forward's cell-updates/s, forward+backward cell-updates/s and backward / forward
    cost."""
    import numpy as np
    from chimera_amd import problems
    from chimera_amd.adjoint import BatchAdjoint, evolve_batch, grf_member_from_noise, record_dt
    steps = args.dgrf_steps
    scheme = (3, 2, 1, 1)
    sizes = problems.grf_sizes(args.dgrf_members, choices=(384, 448, 512))
    cells = sum(n * n for n in sizes)
    noise, params, probs, st0s, dts, targets = [], [], [], [], [], []
    for m, n in enumerate(sizes):
        gen = torch.Generator(device="cuda").manual_seed(1000 + m)
        w = torch.randn((2, n, n), generator=gen, device="cuda", dtype=torch.float64)
        par = problems.grf_params(1000 + m)
        prob, st0 = grf_member_from_noise(w, n, par, scheme=scheme)
        noise.append(w)
        params.append(par)
        probs.append(prob)
        st0s.append(st0.detach().contiguous())
        dts.append(record_dt(prob, steps))
        tg = torch.Generator(device="cuda").manual_seed(5000 + m)
        targets.append(1.0 + 0.05 * torch.randn((n, n), generator=tg, device="cuda",
                                                dtype=torch.float64))
    dts = np.stack(dts)
    badj = BatchAdjoint([(p.nx, p.ny, p.dx, p.gamma, p.bc) for p in probs], scheme)
    try:
        # tile shapes of the forward and of both reverse passes, on one step
        tiles, tiles_ms = badj.autotune(st0s, dts[:, :1], reps=2)
        badj.forward(st0s, dts[:, :1])             # warm-up (allocation, code load)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        badj.forward(st0s, dts)
        torch.cuda.synchronize()
        t_fwd = time.perf_counter() - t0
        # warm-up of the backward: one checkpoint segment allocates the trajectories
        wk = min(50, dts.shape[1])
        U0w = [u.detach().clone().requires_grad_(True) for u in st0s]
        sum(o.sum() for o in evolve_batch(U0w, np.ascontiguousarray(dts[:, :wk]), badj,
                                          checkpoint_every=50)).backward()
        ws = [w.clone().requires_grad_(True) for w in noise]
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        U0 = [grf_member_from_noise(w, n, par, scheme=scheme)[1]
              for w, n, par in zip(ws, sizes, params)]
        UN = evolve_batch(U0, dts, badj, checkpoint_every=50)
        loss = sum(((u[0] / (p.dx * p.dx) - tgt) ** 2).sum() for u, p, tgt in zip(UN, probs, targets))
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        loss.backward()
        torch.cuda.synchronize()
        t2 = time.perf_counter()
    finally:
        badj.close()
    gnorm = [float(w.grad.norm()) for w in ws]
    shape_name = ("16x16", "32x8")
    return {"workload": f"differentiable GRF ensemble: {len(sizes)} members, N in {{384,448,512}}, "
                        f"HLLD, MC, MUSCL+MOOD, CT, {steps} steps each (dt frozen), gradient wrt "
                        "the psi / Az noise fields; batched forward and adjoint (all members per "
                        "launch), checkpoints every 50 steps",
            "cells": cells, "steps": steps,
            "forward_value": cells * steps / t_fwd,
            "fwd_bwd_value": cells * steps / (t2 - t0), "unit": "cell-updates/s",
            "cost_ratio": (t2 - t1) / t_fwd,
            "tiles": {"forward": shape_name[tiles[0]], "reverse_a": shape_name[tiles[1]],
                      "reverse_b": shape_name[tiles[2]],
                      "autotune_ms": {"forward": [round(t, 3) for t in tiles_ms[:2]],
                                      "reverse_ab": [round(t, 3) for t in tiles_ms[2:]]}},
            "all_grads_finite": bool(np.all(np.isfinite(gnorm)))}


def traffic_from_profiles(n: int) -> float | None:
    """HBM bytes per stage-kernel launch from the committed rocprofv3 PMC summary."""
    p = ROOT / "profiles" / f"stage_kernel_rotor{n}_pmc.md"
    if not p.exists():
        return None
    m = re.search(r"traffic_bytes_per_launch:\s*([0-9.eE+]+)", p.read_text())
    return float(m.group(1)) if m else None


def main():
    args = parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU only; multi-GPU sharding of the "
                          "stage update is a design (see the sharding notes), not a bench "
                          "configuration", "n_gpus": args.gpus}))
        sys.exit(2)

    cpu = cpu_baseline(args.n, args.cpu_steps)

    import torch
    from chimera_amd import problems
    from chimera_amd.solver import Grid, DT_CFL, TILE_SHAPES

    assert torch.cuda.is_available(), "bench.py needs a GPU"
    torch.cuda.set_device(0)
    if args.copy_calibration:
        a = torch.ones(1 << 26, dtype=torch.float64, device="cuda")
        b = torch.empty_like(a)
        for _ in range(3):
            b.copy_(a)
        torch.cuda.synchronize()
        del a, b
    n = args.n
    cells = n * n
    prob = problems.rotor(n, scheme=(2, 1, 1, 1))
    g = Grid.from_problem(prob)
    g.steps(args.warmup, dt_mode=DT_CFL)
    g.sync()
    tile_shape, tile_ms = g.autotune(reps=3)   # times one stage per shape; state unchanged
    tile = g.tile_shape()

    # timed region: eager launches with HIP events around every fused stage kernel
    g.set_timing(True)
    torch.cuda.synchronize()
    g.sync()
    t0 = time.perf_counter()
    g.steps(args.steps, dt_mode=DT_CFL)
    g.sync()
    t1 = time.perf_counter()
    elapsed = t1 - t0
    stage_ms, launches = g.stage_kernel_ms()
    g.set_timing(False)

    graph_value = None
    if not args.no_graph:
        g.steps(2, dt_mode=DT_CFL, graph=True)   # capture
        g.sync()
        t2 = time.perf_counter()
        g.steps(args.steps, dt_mode=DT_CFL, graph=True)
        g.sync()
        graph_value = cells * args.steps / (time.perf_counter() - t2)
    st = g.download()
    finite = bool(torch.isfinite(torch.from_numpy(st)).all())
    g.close()

    strip = shard_strip(args, torch, stage_ms) if args.strip_steps > 0 and stage_ms > 0 else None
    batched = batched_grf(args, torch) if args.members > 0 else None
    adjoint = adjoint_field_loop(args, torch) if args.adj_steps > 0 else None
    dgrf = differentiable_grf(args, torch) if args.dgrf_members > 0 else None

    value = cells * args.steps / elapsed
    achieved = BYTES_PER_CELL_STAGE * cells / (stage_ms * 1e-3) / 1e9 if stage_ms > 0 else None
    out = {
        "metric": METRIC,
        "value": value,
        "unit": "cell-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": elapsed * 1e3 / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"MHD rotor {n}x{n}, outflow, gamma=1.4, HLLE, Van Leer, "
                               "MUSCL+MOOD, CT, SSP-RK3, CFL 0.3 (pure CFL dt)",
                   "nx": n, "ny": n, "path": "plain resident (fused stage kernel)",
                   "tile": [tile[1], tile[2]],
                   "tile_autotune_ms": {f"{a}x{b}": round(t, 4) for (a, b), t in
                                        zip(TILE_SHAPES, tile_ms)}},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": (achieved / HBM_PEAK_GBS) if achieved else None,
                     "traffic": traffic_from_profiles(n),
                     "kernel": "stage_kernel<HLLE,VanLeer>",
                     "avg_launch_ms": stage_ms, "launches": launches,
                     "bytes_per_launch": BYTES_PER_CELL_STAGE * cells},
        "cpu_baseline": cpu,
        "graph_value": graph_value,
        "finite": finite,
        "batched_grf": batched,
        "adjoint_field_loop": adjoint,
        "differentiable_grf": dgrf,
        "shard_strip": strip,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
