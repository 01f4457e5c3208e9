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
    from chimera_amd.solver import Grid, DT_CFL

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

    batched = batched_grf(args, torch) if args.members > 0 else None

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
                   "nx": n, "ny": n, "path": "plain resident (fused stage kernel)"},
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
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
