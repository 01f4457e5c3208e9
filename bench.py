#!/usr/bin/env python3
"""Benchmark: SSP-RK3 cell-updates/s of the HIP stage path on one MI355X.

Headline workload: the batched GRF ensemble -- 512 members with grids drawn from
{128, 256, 384, 512}^2 (seed 20261203), periodic, Rusanov flux, MC limiter, MUSCL +
MOOD, CT, per-member CFL dt, every RK stage one persistent launch over all members.  A
step is one SSP-RK3 step of every member (3 persistent stage launches plus the batched
CFL reduction and dt finalisation); the states are resident in HBM before the timed
region starts.  value = cells x steps / wall time of the timed region (the stream
synchronised on both sides); the same members advanced one at a time through the plain
single-grid path give plain_value and the batched / plain ratio.

roofline: the batched stage kernel's algorithmic HBM bytes per launch (128 B per cell:
stage 1 reads u^n and writes u^1, stages 2 and 3 also read u^n, 384 B per cell over a
step) divided by its average launch duration, measured with HIP events on the launch
stream over the timed region; traffic from the committed rocprofv3 summary.

Named sections in the same JSON line: the rotor 2048^2 (HLLE / Van Leer, outflow), the
Orszag-Tang vortex 256^2 run to t = 0.5 (checked bit for bit against the reference
Fortran's own run), the field-loop adjoint (512 x 256) and the differentiable GRF
ensemble, each with max|div B|, the one-rank strip of a sharded rotor, and the
measured HBM bandwidth of a plain device copy.

cpu_baseline: the reference Fortran (oracle/_ref: the unmodified reference modules built
with amdflang -O3 -fopenmp) timed on this host, before the GPU is touched, at every
core the process may use and at 1 thread, each on a bounded sample of its section's
workload (stated in `sample`) extrapolated to the full workload where marked.
"""
from __future__ import annotations

import argparse
import json
import os
import re
import subprocess
import sys
import tempfile
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

METRIC = "cell-updates/s on 1 GPU, batched GRF vs plain; adjoint/forward cost ratio"
HBM_PEAK_GBS = 8000.0
BYTES_PER_CELL_STAGE = 128.0          # (96 + 144 + 144) / 3
GRF_SCHEME = (1, 2, 1, 1)             # Rusanov, MC, limiting, MOOD
ROTOR_SCHEME = (2, 1, 1, 1)           # HLLE, Van Leer
OT_SCHEME = (1, 2, 1, 1)
FL_SCHEME = (2, 2, 1, 1)              # HLLE, MC


def parse_args(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=500,
                    help="timed steps of the headline batched GRF ensemble")
    ap.add_argument("--warmup", type=int, default=5, help="untimed warm-up steps (headline)")
    ap.add_argument("--members", type=int, default=512, help="batched GRF ensemble size")
    ap.add_argument("--plain-per-class", type=int, default=8,
                    help="members per grid size timed on the plain single-run path")
    ap.add_argument("--rotor-n", type=int, default=2048, help="rotor grid size (n x n)")
    ap.add_argument("--rotor-steps", type=int, default=200, help="timed rotor steps (0: skip)")
    ap.add_argument("--no-graph", action="store_true", help="skip the hipGraph replay passes")
    ap.add_argument("--ot", type=int, default=1, help="run the OT 256^2 t = 0.5 section (0: skip)")
    ap.add_argument("--adj-steps", type=int, default=1000,
                    help="steps of the field-loop adjoint section (0: skip)")
    ap.add_argument("--adj-fd-nodes", type=int, default=16,
                    help="Az nodes checked against central differences in the adjoint section")
    ap.add_argument("--dgrf-members", type=int, default=64,
                    help="members of the differentiable GRF section (0: skip)")
    ap.add_argument("--dgrf-steps", type=int, default=300)
    ap.add_argument("--strip-steps", type=int, default=200,
                    help="steps of the one-rank strip of an 8-way rotor split (0: skip)")
    ap.add_argument("--cpu", choices=("full", "quick", "off"), default="full",
                    help="CPU baseline legs: full, quick (smaller samples, no 1-thread legs) "
                         "or off")
    ap.add_argument("--cpu-leg", default=None, help=argparse.SUPPRESS)
    return ap.parse_args(argv)


# ---------------------------------------------------------------------------------
# host description and the CPU legs (run in child processes, before the GPU is touched)

def host_info() -> dict:
    model = None
    try:
        for line in Path("/proc/cpuinfo").read_text().splitlines():
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    try:
        usable = len(os.sched_getaffinity(0))
    except AttributeError:
        usable = os.cpu_count() or 1
    return {"nproc": os.cpu_count(), "usable_cpus": usable, "cpu_model": model,
            "omp_num_threads_env": os.environ.get("OMP_NUM_THREADS")}


def _leg_rotor(spec: dict) -> dict:
    from chimera_amd import problems
    from oracle.oracle import RefVariant
    n, steps = spec["n"], spec["steps"]
    prob = problems.rotor(n, scheme=ROTOR_SCHEME)
    st = prob.state.copy()
    ref = RefVariant(*ROTOR_SCHEME[:2])
    t0 = time.perf_counter()
    ref.run(st, prob.dx, prob.gamma, prob.bc, steps, dt_mode=1)
    el = time.perf_counter() - t0
    return {"cells": n * n, "steps": steps, "seconds": el, "value": n * n * steps / el}


def _leg_ot(spec: dict) -> dict:
    """OT 256^2 under the reference dt policy to t_end (full run) or for `steps`
    steps (sample); the full run saves its final state for the GPU comparison."""
    from chimera_amd import problems
    from oracle.oracle import RefVariant
    n = spec["n"]
    prob = problems.orszag_tang(n, scheme=OT_SCHEME)
    st = prob.state.copy()
    ref = RefVariant(*OT_SCHEME[:2])
    t0 = time.perf_counter()
    if spec.get("t_end"):
        r = ref.run(st, prob.dx, prob.gamma, prob.bc, 20000, dt_mode=0, t_end=spec["t_end"])
    else:
        r = ref.run(st, prob.dx, prob.gamma, prob.bc, spec["steps"], dt_mode=0)
    el = time.perf_counter() - t0
    if spec.get("save"):
        np.save(spec["save"], st)
    return {"cells": n * n, "steps": int(r["steps"]), "t": r["t"], "seconds": el,
            "value": n * n * r["steps"] / el, "max_divb": float(r["divb"][-1][0])}


def _leg_grf(spec: dict) -> dict:
    """Members of every grid size, one after another (as the reference's Monte-Carlo
    driver runs them), `steps` CFL steps each; seconds per cell-update per size."""
    from chimera_amd import problems
    from oracle.oracle import RefVariant
    ref = RefVariant(*GRF_SCHEME[:2])
    per = {}
    for n in spec["sizes"]:
        tot, cu = 0.0, 0
        for k in range(spec["per_size"]):
            prob = problems.grf_member(n, 7000 + 97 * k + n, scheme=GRF_SCHEME)
            st = prob.state.copy()
            t0 = time.perf_counter()
            ref.run(st, prob.dx, prob.gamma, prob.bc, spec["steps"], dt_mode=1)
            tot += time.perf_counter() - t0
            cu += n * n * spec["steps"]
        per[str(n)] = {"members": spec["per_size"], "steps": spec["steps"], "seconds": tot,
                       "s_per_cell_update": tot / cu}
    return {"per_size": per, "seconds": sum(v["seconds"] for v in per.values())}


def _leg_fl(spec: dict) -> dict:
    """The field-loop forward (512 x 256, HLLE / MC, prescribed CFL-like dt) on the
    oracle: the reference Fortran's grid is N x N (module N), so the rectangular
    benchmark grid runs on the C restatement (bit-pinned to the reference on square
    grids) -- the per-forward cost of a central-difference check on the CPU."""
    from chimera_amd import problems
    from oracle.oracle import COracle
    nx, ny, steps = spec["nx"], spec["ny"], spec["steps"]
    prob = problems.field_loop(nx, ny, scheme=FL_SCHEME)
    st = prob.state.copy()
    orc = COracle()
    dts = np.full(steps, 0.3 * prob.dx / 3.5)
    t0 = time.perf_counter()
    orc.run(st, prob.dx, prob.gamma, prob.bc, list(prob.scheme), steps, dt_mode=2, dt_in=dts)
    el = time.perf_counter() - t0
    return {"cells": nx * ny, "steps": steps, "seconds": el, "value": nx * ny * steps / el}


LEGS = {"rotor": _leg_rotor, "ot": _leg_ot, "grf": _leg_grf, "fl": _leg_fl}


# The reference's loops open an OpenMP parallel region per routine, several per stage:
# idle threads must spin between them (a sleeping thread's wake-up at every region
# would be timed as the reference's cost).
OMP_WAIT_POLICY = "ACTIVE"


def run_leg(name: str, spec: dict, threads: int) -> dict | None:
    """One CPU leg in a child process with OMP_NUM_THREADS = threads (the OpenMP runtime
    reads it once per process)."""
    env = dict(os.environ, OMP_NUM_THREADS=str(threads), OMP_WAIT_POLICY=OMP_WAIT_POLICY)
    cmd = [sys.executable, str(Path(__file__).resolve()), "--cpu-leg",
           json.dumps({"name": name, "spec": spec})]
    print(f"[bench] cpu leg {name} threads={threads} {spec}", file=sys.stderr, flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True, env=env, timeout=900)
    if r.returncode != 0:
        print(f"[bench] cpu leg {name} failed:\n{r.stderr[-2000:]}", file=sys.stderr, flush=True)
        return None
    out = json.loads(r.stdout.strip().splitlines()[-1])
    out["threads"] = threads
    return out


def cpu_legs(args, tmpdir: str) -> dict:
    """Every CPU baseline figure, measured before the GPU is touched."""
    from oracle.oracle import ref_available
    res = {"host": host_info(), "available": ref_available()}
    if args.cpu == "off" or not res["available"]:
        return res
    full = args.cpu == "full"
    allc = res["host"]["usable_cpus"]
    threads = [allc, 1] if full and allc > 1 else [allc]
    sizes = (128, 256, 384, 512)
    for th in threads:
        one = th == 1
        tag = f"t{th}"
        res[f"grf_{tag}"] = run_leg("grf", {"sizes": sizes, "per_size": 1 if one or not full else 8,
                                           "steps": 1 if one else 2}, th)
        if args.rotor_steps > 0:
            res[f"rotor_{tag}"] = run_leg("rotor", {"n": args.rotor_n,
                                                    "steps": 1 if one or not full else 3}, th)
        if args.ot:
            if one or not full:
                res[f"ot_{tag}"] = run_leg("ot", {"n": 256, "steps": 40}, th)
            else:
                res[f"ot_{tag}"] = run_leg("ot", {"n": 256, "t_end": 0.5,
                                                  "save": os.path.join(tmpdir, "ot256_ref.npy")}, th)
        if args.adj_steps > 0 and not one:
            res[f"fl_{tag}"] = run_leg("fl", {"nx": 512, "ny": 256, "steps": 50 if full else 10}, th)
    return res


def cpu_baseline_of(legs: dict, key: str, unit_value, sample: str, kind="reference",
                    extra=None) -> dict | None:
    """The cpu_baseline object of one section from its legs (all cores and 1 thread)."""
    host = legs.get("host", {})
    allc = host.get("usable_cpus")
    a = legs.get(f"{key}_t{allc}")
    if not a:
        return None
    b = legs.get(f"{key}_t1") if allc and allc > 1 else None
    out = {"value": unit_value(a), "unit": "cell-updates/s", "cores": allc, "kind": kind,
           "sample": sample, "seconds": a.get("seconds"),
           "one_thread": ({"value": unit_value(b), "cores": 1, "seconds": b.get("seconds")}
                          if b else None),
           "host_nproc": host.get("nproc"), "cpu_model": host.get("cpu_model"),
           "omp_wait_policy": OMP_WAIT_POLICY.lower()}
    if extra:
        out.update(extra(a, b))
    return out


# ---------------------------------------------------------------------------------
# GPU sections

def max_divb_periodic(torch, st, dx: float) -> float:
    """max |div B| of a periodic face-B state on the device (compute_divB's interior
    formula, src/mhd_field_ops.f90:60-117, with the periodic wrap)."""
    bx, by = st[4], st[5]
    d = (bx - torch.roll(bx, 1, dims=1) + by - torch.roll(by, 1, dims=0)) * (1.0 / dx)
    return float(d.abs().max())


def copy_bandwidth(torch) -> dict:
    """A plain device copy of 2^26 doubles (512 MiB read + 512 MiB written), 3 times:
    the achievable-HBM reference point (also the calibration of the PMC byte counters)."""
    a = torch.ones(1 << 26, dtype=torch.float64, device="cuda")
    b = torch.empty_like(a)
    b.copy_(a)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    e0.record()
    for _ in range(3):
        b.copy_(a)
    e1.record()
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / 3
    del a, b
    torch.cuda.empty_cache()
    gbs = 2 * (1 << 29) / (ms * 1e-3) / 1e9
    return {"bytes_per_copy": 2 * (1 << 29), "ms": ms, "GB/s": gbs,
            "frac_of_8TBps": gbs / HBM_PEAK_GBS}


def traffic_from_profile(name: str) -> float | None:
    """HBM bytes per stage-kernel launch from a committed rocprofv3 PMC summary."""
    p = ROOT / "profiles" / name
    if not p.exists():
        return None
    m = re.search(r"traffic_bytes_per_launch:\s*([0-9.eE+]+)", p.read_text())
    return float(m.group(1)) if m else None


def batched_grf(args, torch) -> dict:
    """The headline: every member advanced together (one persistent launch per RK
    stage), against the plain path (the same members one after another through the
    resident single-grid driver, graph replayed, extrapolated from plain_per_class
    members per size)."""
    from chimera_amd import problems
    from chimera_amd.batch import Ensemble, DT_CFL, SHAPES
    from chimera_amd.solver import Grid
    sizes = problems.grf_sizes(args.members)
    probs = [problems.grf_member_device(n, 1000 + m, scheme=GRF_SCHEME)
             for m, n in enumerate(sizes)]
    torch.cuda.synchronize()
    cells = sum(n * n for n in sizes)
    steps = args.steps
    ens = Ensemble(probs)
    shape, bpc = ens.autotune(reps=2)
    ens.steps(args.warmup, dt_mode=DT_CFL)
    ens.sync()
    # timed region: eager launches with HIP events around every stage launch
    ens.set_timing(True)
    ens.sync()
    t0 = time.perf_counter()
    ens.steps(steps, dt_mode=DT_CFL)
    ens.sync()
    el = time.perf_counter() - t0
    stage_ms, launches = ens.stage_kernel_ms()
    ens.set_timing(False)
    graph_value = None
    if not args.no_graph:
        ens.steps(2, dt_mode=DT_CFL, graph=True)
        ens.sync()
        t1 = time.perf_counter()
        ens.steps(steps, dt_mode=DT_CFL, graph=True)
        ens.sync()
        graph_value = cells * steps / (time.perf_counter() - t1)
    # max|div B| over the members (periodic) after all steps
    mdb = 0.0
    for m, n in enumerate(sizes):
        out = torch.empty((6, n, n), dtype=torch.float64, device="cuda")
        ens.download(m, out)
        mdb = max(mdb, max_divb_periodic(torch, out, 1.0 / n))
    ens.close()
    value = cells * steps / el

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
    return {"sizes": sizes, "cells": cells, "steps": steps, "value": value, "seconds": el,
            "graph_value": graph_value, "plain_value": plain, "ratio": value / plain,
            "ms_per_step": el * 1e3 / steps, "plain_extrapolated_s": plain_time,
            "plain_members_timed": {str(n): v[1] for n, v in per_class.items()},
            "tiling": {"tile": SHAPES[shape], "blocks_per_cu": bpc},
            "max_divb": mdb, "stage_ms": stage_ms, "launches": launches, "achieved": achieved}


def rotor_section(args, torch, legs) -> dict:
    from chimera_amd import problems
    from chimera_amd.solver import Grid, DT_CFL, TILE_SHAPES
    n = args.rotor_n
    cells = n * n
    prob = problems.rotor(n, scheme=ROTOR_SCHEME)
    g = Grid.from_problem(prob)
    g.steps(20, dt_mode=DT_CFL)
    g.sync()
    tile_shape, tile_ms = g.autotune(reps=3)   # one stage per shape; state unchanged
    tile = g.tile_shape()
    g.set_timing(True)
    g.sync()
    t0 = time.perf_counter()
    g.steps(args.rotor_steps, dt_mode=DT_CFL)
    g.sync()
    el = time.perf_counter() - t0
    stage_ms, launches = g.stage_kernel_ms()
    g.set_timing(False)
    mdb, _ = g.divb()
    graph_value = None
    if not args.no_graph:
        g.steps(2, dt_mode=DT_CFL, graph=True)
        g.sync()
        t1 = time.perf_counter()
        g.steps(args.rotor_steps, dt_mode=DT_CFL, graph=True)
        g.sync()
        graph_value = cells * args.rotor_steps / (time.perf_counter() - t1)
    st = g.download()
    g.close()
    achieved = BYTES_PER_CELL_STAGE * cells / (stage_ms * 1e-3) / 1e9 if stage_ms > 0 else None
    leg = legs.get(f"rotor_t{legs.get('host', {}).get('usable_cpus')}")
    cpu = cpu_baseline_of(
        legs, "rotor", lambda a: a["value"],
        f"rotor {n}x{n}, HLLE / Van Leer, CFL dt: {leg['steps'] if leg else 0} SSP-RK3 steps "
        "timed at all cores, 1 at 1 thread; the cost per step does not depend on the step, so "
        "the 2000-step time is extrapolated as 2000 x the measured step time",
        extra=lambda a, b: {"extrapolated_2000_steps_s": 2000 * a["seconds"] / a["steps"]})
    return {"workload": f"MHD rotor {n}x{n}, outflow, gamma=1.4, HLLE, Van Leer, MUSCL+MOOD, CT, "
                        "SSP-RK3, CFL 0.3 (pure CFL dt)",
            "value": cells * args.rotor_steps / el, "graph_value": graph_value,
            "steps": args.rotor_steps, "ms_per_step": el * 1e3 / args.rotor_steps,
            "tile": [tile[1], tile[2]],
            "tile_autotune_ms": {f"{a}x{b}": round(t, 4) for (a, b), t in zip(TILE_SHAPES, tile_ms)},
            "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                         "frac": achieved / HBM_PEAK_GBS if achieved else None,
                         "traffic": traffic_from_profile(f"stage_kernel_rotor{n}_pmc.md"),
                         "kernel": "stage_kernel<HLLE,VanLeer,32,16>",
                         "avg_launch_ms": stage_ms, "launches": launches,
                         "bytes_per_launch": BYTES_PER_CELL_STAGE * cells},
            "max_divb": mdb, "finite": bool(np.isfinite(st).all()), "cpu_baseline": cpu,
            "gpu_vs_cpu": (cells * args.rotor_steps / el) / cpu["value"] if cpu else None}


def ot_section(args, torch, legs, tmpdir) -> dict:
    """OT 256^2 under the reference dt policy to t = 0.5: the GPU takes the number of
    steps the reference Fortran took, and its final state is compared with the
    reference's bit for bit."""
    from chimera_amd import problems
    from chimera_amd.solver import Grid, DT_REFERENCE
    n = 256
    allc = legs.get("host", {}).get("usable_cpus")
    ref = legs.get(f"ot_t{allc}")
    prob = problems.orszag_tang(n, scheme=OT_SCHEME)
    full = bool(ref and ref.get("t", 0.0) >= 0.5)
    nsteps = ref["steps"] if full else None
    with Grid.from_problem(prob) as g:
        g.steps(1)                                   # warm-up on a throw-away copy
        g.upload(prob.state)
        g.set_time(0.0, 1)
        g.sync()
        t0 = time.perf_counter()
        if nsteps is None:                           # no reference run: step to t = 0.5
            nsteps = 0
            while g.time()[0] < 0.5:
                g.steps(1)
                nsteps += 1
        else:
            g.steps(nsteps, graph=True)
        g.sync()
        el = time.perf_counter() - t0
        t_end, _ = g.time()
        mdb, _ = g.divb()
        st = g.download()
    same = None
    path = os.path.join(tmpdir, "ot256_ref.npy")
    if full and os.path.exists(path):
        same = bool(np.array_equal(st, np.load(path)))
    cpu = cpu_baseline_of(
        legs, "ot", lambda a: a["value"],
        "OT 256^2, Rusanov / MC, reference dt policy: the full run to t = 0.5 at all cores; "
        "40 steps at 1 thread", extra=lambda a, b: {"steps": a["steps"], "t": a["t"],
                                                     "max_divb": a.get("max_divb")})
    return {"workload": "Orszag-Tang 256x256 periodic, gamma=5/3, Rusanov, MC, MUSCL+MOOD, CT, "
                        "SSP-RK3, CFL 0.3 with the output-time clamp, to t = 0.5",
            "steps": nsteps, "t": t_end, "seconds": el, "value": n * n * nsteps / el,
            "max_divb": mdb, "state_equals_reference_fortran": same, "cpu_baseline": cpu,
            "gpu_vs_cpu": (n * n * nsteps / el) / cpu["value"] if cpu else None}


def adjoint_field_loop(args, torch, legs) -> dict:
    """Field loop 512 x 256 (periodic, HLLE, MC), dt frozen from a CFL run, loss =
    sum((|B|^2_N - |B|^2_target)^2) with the target from Az (1 + 0.01 xi), xi seeded 7;
    gradient with respect to every Az node through the checkpointed adjoint
    (checkpoint every 50 steps).  cost ratio = backward wall time / wall time of the
    same steps on the fused forward path.

    Finite-difference sweep (a measurement of the discrete map, not the correctness gate:
    the tests check the gradient against exact autograd through the restatement at this
    grid, and against central differences of the oracle and the reference where the map
    is differentiable): on the same grid with the Van Leer limiter, Az plus 1e-4 x a
    smooth periodic field and a smooth background, central differences
    (eps = 1e-7, 1e-8 of max|Az|) of the GPU forward at seeded nodes."""
    from chimera_amd import problems
    from chimera_amd.adjoint import cell_b2, evolve, field_loop_state, record_dt
    from chimera_amd.solver import Grid, DT_PRESCRIBED
    nsteps, K = args.adj_steps, 50
    rng = np.random.default_rng(7)

    def setup(limiter):
        prob, Az = problems.field_loop(512, 256, scheme=(2, limiter, 1, 1), return_az=True)
        dts = record_dt(prob, nsteps)
        return prob, Az, dts

    def run_forward(g, dx, az_np, dts, hydro=None):
        st = field_loop_state(torch.from_numpy(az_np).cuda(), dx, hydro=hydro)
        g.upload(st)
        g.steps(nsteps, dt_mode=DT_PRESCRIBED, dt=dts, graph=True)
        out = torch.empty_like(st)
        g.download(out)
        return out

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

    prob, Az, dts = setup(2)
    xi = rng.standard_normal(Az.shape)
    Az_t = Az * (1.0 + 0.01 * xi)
    g = Grid.from_problem(prob)
    T = cell_b2(run_forward(g, prob.dx, Az_t, dts))
    g.upload(field_loop_state(torch.from_numpy(Az).cuda(), prob.dx))
    g.autotune(reps=3)                         # the forward at its best tile shape
    g.steps(2, dt_mode=DT_PRESCRIBED, dt=dts[:2], graph=True)
    g.sync()
    g.upload(field_loop_state(torch.from_numpy(Az).cuda(), prob.dx))
    t0 = time.perf_counter()
    g.steps(nsteps, dt_mode=DT_PRESCRIBED, dt=dts, graph=True)
    g.sync()
    t_fwd = time.perf_counter() - t0
    fin = torch.empty((6, prob.ny, prob.nx), dtype=torch.float64, device="cuda")
    g.download(fin)
    mdb = max_divb_periodic(torch, fin, prob.dx)
    fwd_tile = list(g.tile_shape()[1:])
    g.close()
    gradient(prob, Az, dts[:K], T)   # warm-up: the backward's workspace for one segment
    grad, loss_v, t_evolve, t_bwd = gradient(prob, Az, dts, T)
    grad_finite = bool(np.isfinite(grad).all())

    fd = None
    if args.adj_fd_nodes > 0:
        ny, nx = Az.shape
        hyd = problems.field_loop_background(nx, ny, seed=20)
        Az_v = Az + 1.0e-4 * problems.smooth_periodic(nx, ny, seed=1)
        prob_v = problems.field_loop_from_az(Az_v, prob.dx, scheme=(2, 1, 1, 1), hydro=hyd)
        dts_v = record_dt(prob_v, nsteps)
        hyd_t = [torch.from_numpy(a).cuda() for a in hyd]
        g = Grid.from_problem(prob_v)
        T_v = cell_b2(run_forward(g, prob_v.dx, Az_v * (1.0 + 0.01 * xi), dts_v, hyd_t))

        def loss_diff(ap, am):
            bp = cell_b2(run_forward(g, prob_v.dx, ap, dts_v, hyd_t))
            bm = cell_b2(run_forward(g, prob_v.dx, am, dts_v, hyd_t))
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
        n_ring = min(4, args.adj_fd_nodes)
        nodes = list(rng.choice(inside, size=args.adj_fd_nodes - n_ring, replace=False)) + \
            list(rng.choice(ring, size=n_ring, replace=False))
        gmax = float(np.abs(grad_v).max())
        fd = {"variant": "Van Leer, Az + 1e-4 smooth field, smooth background", "nodes": len(nodes),
              "rel_err_definition": "|adjoint - FD| / max(|FD|, 1e-3 max|grad|)"}
        for rel_eps in (1e-7, 1e-8):
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
    allc = legs.get("host", {}).get("usable_cpus")
    leg = legs.get(f"fl_t{allc}")
    fd_cpu = None
    if leg:
        fd_cpu = {"per_forward_s": nsteps * leg["seconds"] / leg["steps"], "cores": allc,
                  "kind": "port", "sample": f"{leg['steps']} steps of the 512x256 forward on the "
                  "oracle (the reference Fortran's grid is square), extrapolated to "
                  f"{nsteps} steps; a central difference needs 2 forwards per node",
                  "value": leg["value"], "unit": "cell-updates/s"}
    return {"workload": f"field loop 512x256 periodic, HLLE, MC, MUSCL+MOOD, CT, {nsteps} steps "
                        f"(dt frozen from a CFL run), checkpoint every {K}",
            "loss": loss_v, "grad_nodes": int(grad.size), "grad_finite": grad_finite,
            "forward_s": t_fwd, "evolve_forward_s": t_evolve, "backward_s": t_bwd,
            "cost_ratio": t_bwd / t_fwd, "forward_tile": fwd_tile, "max_divb": mdb,
            "fd_forward_gpu_s": t_fwd, "fd_forward_cpu": fd_cpu, "fd_check": fd}


def shard_strip(args, torch, full_stage_ms: float) -> dict:
    """One GPU's share of the rotor n x n split into 8 strips: its n/8 owned rows plus
    the depth-3 halo rows on each side, cut from the centre of the rotor state, through
    the same resident path (outflow BCs stand in for the received halo).  Reports the
    stage kernel per owned cell against the full grid's and the halo bytes a stage
    would exchange over xGMI."""
    from chimera_amd import problems
    from chimera_amd.solver import Grid, DT_CFL
    n, halo = args.rotor_n, 3
    owned = n // 8
    rows = owned + 2 * halo
    full = problems.rotor(n, scheme=ROTOR_SCHEME)
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
        tiles, tiles_ms = badj.autotune(st0s, dts[:, :1], reps=2)
        badj.forward(st0s, dts[:, :1])             # warm-up (allocation, code load)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        outs = badj.forward(st0s, dts)
        torch.cuda.synchronize()
        t_fwd = time.perf_counter() - t0
        mdb = max(max_divb_periodic(torch, o, p.dx) for o, p in zip(outs, probs))
        del outs
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
            "cost_ratio": (t2 - t1) / t_fwd, "max_divb": mdb,
            "tiles": {"forward": shape_name[tiles[0]], "reverse_a": shape_name[tiles[1]],
                      "reverse_b": shape_name[tiles[2]],
                      "autotune_ms": {"forward": [round(t, 3) for t in tiles_ms[:2]],
                                      "reverse_ab": [round(t, 3) for t in tiles_ms[2:]]}},
            "all_grads_finite": bool(np.all(np.isfinite(gnorm)))}


def main():
    args = parse_args()
    if args.cpu_leg:
        leg = json.loads(args.cpu_leg)
        print(json.dumps(LEGS[leg["name"]](leg["spec"])))
        return
    if args.gpus != 1:
        print(json.dumps({"error": "the benchmark runs on one GPU only: the stage update is "
                          "not sharded across GPUs (one GPU's strip of a sharded grid is the "
                          "shard_strip section)", "n_gpus": args.gpus}))
        sys.exit(2)

    tmpdir = tempfile.mkdtemp(prefix="chimera_bench_")
    legs = cpu_legs(args, tmpdir)        # before the GPU is touched

    import torch
    assert torch.cuda.is_available(), "bench.py needs a GPU"
    torch.cuda.set_device(0)
    copy = copy_bandwidth(torch)
    batched = batched_grf(args, torch)
    rotor = rotor_section(args, torch, legs) if args.rotor_steps > 0 else None
    ot = ot_section(args, torch, legs, tmpdir) if args.ot else None
    strip = (shard_strip(args, torch, rotor["roofline"]["avg_launch_ms"])
             if args.strip_steps > 0 and rotor else None)
    adjoint = adjoint_field_loop(args, torch, legs) if args.adj_steps > 0 else None
    dgrf = differentiable_grf(args, torch) if args.dgrf_members > 0 else None

    allc = legs.get("host", {}).get("usable_cpus")

    def grf_cpu_value(a):
        per = a["per_size"]
        sec = sum(batched["sizes"].count(int(n)) * int(n) ** 2 * v["s_per_cell_update"]
                  for n, v in per.items())
        return batched["cells"] / sec

    grf_leg = legs.get(f"grf_t{allc}")
    cpu = cpu_baseline_of(
        legs, "grf", grf_cpu_value,
        (f"GRF members run one after another as the reference's Monte-Carlo driver does: "
         f"{grf_leg['per_size']['128']['members']} member(s) per grid size x "
         f"{grf_leg['per_size']['128']['steps']} CFL steps each at all cores (1 x 1 at 1 thread); "
         f"seconds per cell-update per size weighted by the ensemble's cell mix "
         f"(extrapolated to {args.members} members x {args.steps} steps)") if grf_leg else "",
        extra=lambda a, b: {"extrapolated_ensemble_s":
                            batched["cells"] * args.steps / grf_cpu_value(a)})
    value = batched["value"]
    achieved = batched["achieved"]
    out = {
        "metric": METRIC,
        "value": value,
        "unit": "cell-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": batched["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"batched GRF ensemble: {args.members} members, N in "
                               "{128,256,384,512} (seed 20261203), periodic, Rusanov, MC, "
                               "MUSCL+MOOD, CT, SSP-RK3, per-member CFL dt, one persistent "
                               "launch per RK stage",
                   "members": args.members, "cells": batched["cells"],
                   "tile": batched["tiling"]["tile"],
                   "blocks_per_cu": batched["tiling"]["blocks_per_cu"]},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS if achieved else None,
                     "traffic": traffic_from_profile("batch_stage_grf512_pmc.md"),
                     "kernel": "batch_stage_kernel<Rusanov,MC,32,16>",
                     "avg_launch_ms": batched["stage_ms"], "launches": batched["launches"],
                     "bytes_per_launch": BYTES_PER_CELL_STAGE * batched["cells"],
                     "achievable_copy_GBps": copy["GB/s"]},
        "cpu_baseline": cpu,
        "gpu_vs_cpu": value / cpu["value"] if cpu else None,
        "batched_vs_plain": {"plain_value": batched["plain_value"], "ratio": batched["ratio"],
                             "graph_value": batched["graph_value"],
                             "plain_extrapolated_s": batched["plain_extrapolated_s"],
                             "plain_members_timed": batched["plain_members_timed"],
                             "max_divb": batched["max_divb"]},
        "hbm_copy": copy,
        "host": legs.get("host"),
        "rotor": rotor,
        "ot256": ot,
        "adjoint_field_loop": adjoint,
        "differentiable_grf": dgrf,
        "shard_strip": strip,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
