"""The reference's command line on the resident GPU path.

    python -m chimera_amd.cli <problem_type> <N> <seed> [output_path] [h5_filename] [M_s] [beta]

Arguments, messages and output follow src/mhd_init.f90:167-256 (get_user_options)
and src/main.f90:24-330: problem 1 Orszag-Tang, 2 Kelvin-Helmholtz, 3 field loop,
4 rotor, 5 random (GRF) problem; the scheme of the reference's mhd_config (Rusanov,
MC, limiting, MOOD, CFL 0.3); SSP-RK3 steps with the reference dt policy (output-time
clamp) until t = tEnd = 1; a snapshot of rho, thermal P, cell B and v every
tOut = 0.01 and at the end; the progress line with avg(|div B|) every 1% (not for
problem 5); one HDF5 file in the reference's layout at output_path // h5_filename.

Problems 1-4 are set up on the GPU by the library (problems.setup_device), with the
bits of the reference program's initial state.  Problem 5 is the reference's GRF set-up
(problems.grf_problem) generated on the GPU with uniforms drawn from numpy's PCG64 seeded with `seed`: the reference draws them
with the Fortran compiler's random_number, which no other build reproduces.  Its
realized M_s and beta are written as the reference computes them; problems 1-4 write
0 for both (the reference never sets them).  Like the reference's `stop`, an invalid
command line prints the message and exits with status 0.
"""
from __future__ import annotations

import sys

import numpy as np

from . import problems
from .solver import DT_REFERENCE, Grid

T_END = 1.0                 # src/mhd_config.f90:16
T_OUT = 0.01                # src/mhd_config.f90:52
OUT_INTERVAL = 0.01 * T_END  # src/mhd_config.f90:53
SCHEME = (1, 2, 1, 1)       # src/mhd_config.f90:26-37: Rusanov, MC, limiting, MOOD
USAGE = [" ERROR: Usage: ./chimera <problem_type> <N> <seed> [output_path] [h5_filename]",
         "   problem_type : 1=Orszag-Tang 2=KH 3=Field-Loop 4=Rotor 5=MC",
         "   N            : grid size (integer >= 4)",
         "   seed         : random seed (integer)",
         "   output_path  : optional (default: './outputs/')",
         "   h5_filename  : optional (default: 'primitive_snaps.h5')"]


class UsageError(Exception):
    pass


def parse_args(argv: list[str]) -> dict:
    """get_user_options (src/mhd_init.f90:167-256)."""
    if len(argv) < 3:
        raise UsageError("\n".join(USAGE))
    try:
        p = int(argv[0])
    except ValueError:
        p = -1
    if not 1 <= p <= 5:
        raise UsageError(f" ERROR: problem_type must be 1-5, got: '{argv[0]}'")
    try:
        n = int(argv[1])
    except ValueError:
        n = -1
    if n < 4:
        raise UsageError(f" ERROR: N must be an integer >= 4, got: '{argv[1]}'")
    try:
        seed = int(argv[2])
    except ValueError:
        raise UsageError(f" ERROR: Invalid seed: '{argv[2]}'") from None

    def opt_float(k):
        try:
            return float(argv[k]) if len(argv) > k else -1.0
        except ValueError:
            return -1.0

    return {"problem_type": p, "N": n, "seed": seed,
            "out_path": argv[3].strip() if len(argv) > 3 else "./outputs/",
            "h5_filename": argv[4] if len(argv) > 4 else "primitive_snaps.h5",
            "target_M_s": opt_float(5), "target_beta": opt_float(6)}


def setup(problem_type: int, n: int, seed: int, target_M_s: float = -1.0,
          target_beta: float = -1.0, device="cuda"):
    """(Problem, M_s, beta) for the reference's problem types (src/mhd_init.f90:259-291);
    every state is generated on the GPU."""
    if problem_type in problems.SETUP_NAMES:
        return problems.setup_device(problem_type, n, scheme=SCHEME), 0.0, 0.0
    prob, info = problems.grf_problem(n, seed, target_M_s, target_beta, scheme=SCHEME,
                                      device=device)
    return prob, info["M_s_actual"], info["beta_actual"]


def _snapshot(g: Grid, prob) -> dict:
    """Cell B and primitives of the current state (src/main.f90:278-290) through the
    drop-in kernels; P is stored as thermal pressure."""
    from . import dropin
    st = g.download()
    Bx, By = dropin.average_face_to_cell_B(st[4], st[5], prob.bc)
    rho, vx, vy, P = dropin.get_primitive(st[:4], Bx, By, prob.gamma, prob.dx * prob.dx)
    return {"rho": rho, "P": P - 0.5 * (Bx * Bx + By * By), "Bx": Bx, "By": By, "Vx": vx, "Vy": vy}


def run(opts: dict, out=print) -> str:
    from . import h5io
    prob, m_s, beta = setup(opts["problem_type"], opts["N"], opts["seed"], opts["target_M_s"],
                            opts["target_beta"])
    snaps, times = [], []
    t, count, stamp = 0.0, 1, 1
    n = opts["N"]
    with Grid.from_problem(prob) as g:
        g.autotune(reps=2)   # fastest tile shape for this grid; every shape gives the same bits
        while t < T_END:
            # one step with dt clamped to the next output time on the device
            g.steps(1, dt_mode=DT_REFERENCE, graph=True)
            t, _ = g.time()
            if t >= count * T_OUT or t >= T_END:
                snaps.append(_snapshot(g, prob))
                times.append(t)
                count += 1
            if t >= stamp * OUT_INTERVAL and t < T_END + 1.0e-8 and opts["problem_type"] != 5:
                _, divb_sum = g.divb()
                pct = int(np.rint(100.0 * (stamp * OUT_INTERVAL) / T_END))
                out(f"{pct:3d}%:  t ={t:7.3f}  , avg(|divB|) ={divb_sum / (n * n):12.4E}")
                stamp += 1
    path = opts["out_path"].strip() + opts["h5_filename"].strip()
    fields = {k: np.stack([s[k] for s in snaps]) for k in h5io.FIELDS}
    h5io.write_snapshots(path, fields, np.array(times), prob.gamma, m_s, beta)
    return path


def main(argv: list[str] | None = None) -> int:
    argv = sys.argv[1:] if argv is None else argv
    try:
        opts = parse_args(argv)
    except UsageError as e:
        print(e)
        return 0
    run(opts)
    return 0


if __name__ == "__main__":
    sys.exit(main())
